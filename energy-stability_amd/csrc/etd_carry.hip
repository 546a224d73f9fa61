// etd_carry.hip -- the AC spectrum carried from one step to the next (gfx950).
//
// After a step, u = V (x) V (x) V u-hat exactly as the backward transform
// computed it from u-hat.  The next step transforms the ghost-filled field
// u_f = u + d, where the ghost fill (transfer, trans.c:1939-1996) changed
// only the shell of 2g cells, so
//   S^-1 u_f = u-hat + S^-1 d,  S^-1 = Vi (x) Vi (x) Vi,
// and S^-1 d costs a few small products instead of three full mode products
// over every AC field.  d splits into the cells each sweep of the fill wrote:
//   z sweep: planes k in G (all j, i),  y sweep: k interior, rows j in G,
//   x sweep: k, j interior, columns i in G    (G = {0 .. g-1, n-g .. n-1}),
// and for u-hat[k^][j^][i^] (x fastest, the layout every forward transform
// ends in)
//   S^-1 d = sum_{s} Vi[k^][G_s] P_s[j^][i^] + Vi[j^][G_s] R_s[k^][i^]
//                                           + Vi[i^][G_s] X_s[k^][j^]
//   P_s = Vi d_z,s Vi^T,  R_s = Vi_I d_y,s Vi^T,  X_s = Vi_I d_x,s Vi_I^T
// with Vi_I the columns of Vi at the interior cells.  Recorded as n x n
// matrices with zeros outside the interior rows (y) or rows and columns (x),
// every one of them is Vi D Vi^T, so all three families are two plain
// strided-batched GEMMs (rocBLAS) over the nac 2g 3 matrices, and
// k_carry_expand adds the three sums
// to the carried spectrum in one pass, which in a step also applies the AC
// phi update (carry_correct_phi).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <rocblas/rocblas.h>

#include "etd_internal.h"

// phi1 and 1 - phi1 H' of one AC point from H and e = exp(-dt H), with
// phi_at's formula and |H| > 1e-8 switch
__device__ __forceinline__ phi_pt phi_sep(double H, double e, double dt) {
  phi_pt r;
  const bool big = fabs(H) > 1.0e-8;
  r.phi1 = big ? (1.0 - e) / H : dt;
  r.phi2 = 0.0;
  r.lin = big ? 1.0 - r.phi1 * H : 1.0;
  return r;
}

// exp(-dt H) of an AC point factors over the axes: -dt H = dt LE epn2
// (kkz Dz + kky Dy + kkx Dx) - dt LE KE, so with ax[a n + i] = exp(dt LE epn2
// kk_a D_a[i]) (a = x, y, z; and at dt2 from 3 n) and e0 = exp(-dt LE KE),
// exp(-dt H) = e0 ax_z[k^] ax_y[j^] ax_x[i^].  Every factor is at most 1
// (D <= 0; LE, epn2, kk >= 0), so the product overflows only where e0 =
// exp(-dt H) of the zero mode does, and then the update is not finite either
// way.
__global__ void k_phi_axes(const double *DDX, const double *DDY, const double *DDZ, int n, phi_args a,
                           int two, double *ax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double c = a.LE * a.epn2, c2 = c * a.dt2;
  ax[i] = exp(c * a.dt * a.kkx * DDX[i]);
  ax[n + i] = exp(c * a.dt * a.kky * DDY[i]);
  ax[2 * n + i] = exp(c * a.dt * a.kkz * DDZ[i]);
  if (two) {
    ax[3 * n + i] = exp(c2 * a.kkx * DDX[i]);
    ax[4 * n + i] = exp(c2 * a.kky * DDY[i]);
    ax[5 * n + i] = exp(c2 * a.kkz * DDZ[i]);
  }
}

// u-hat = alpha s1 + beta s2 (s2 may be null) + S^-1 d over nac fields, and
// out = u-hat (PH = 0), or the first-order phi update of u-hat and N-hat at
// step a.dt (PH = 1; and into out2 at step a.dt2, PH = 2) as k_phiscale
// forms it from u-hat stored in T, with exp(-dt H) from the axis factors
// ax (k_phi_axes) instead of one exp per point.  A block owns KB values of k^
// and JB values of j^ of one field, a thread one i^ (and i^ + 256 ... where n
// exceeds the block).  The thread's R values of its KB rows k^ stay in
// registers across the JB rows j^, its P values of one j^ across the KB rows
// k^; X, Vi[k^][G] and Vi[j^][G] are uniform.  All loads of JU rows j^ (KB
// values of s1, s2 and N-hat each, their P values) are issued before any is
// used.  The rows a block writes are JB consecutive rows of KB planes.  out
// may be s1 or s2: every element is read by the thread that writes it,
// before it writes it, and by no other thread (so no __restrict__ on them).
template <typename T, int KB, int JB, int JU, int PH>
__global__ __launch_bounds__(256) void k_carry_expand(const T *s1, T alpha, const T *s2, T beta, T *out,
                                                      const T *__restrict__ P, const T *__restrict__ R,
                                                      const T *__restrict__ X, const T *__restrict__ G4, int n,
                                                      int with_corr, const T *nhat, T *out2,
                                                      const double *__restrict__ DDX,
                                                      const double *__restrict__ DDY,
                                                      const double *__restrict__ DDZ, phi_args a,
                                                      const double *__restrict__ ax, double2 e0) {
  constexpr int G = 4;  // 2 g ghost layers, g = 2 (carry_possible)
  const int j0 = blockIdx.x * JB, k0 = blockIdx.y * KB, v = blockIdx.z;
  const int nk = n - k0 < KB ? n - k0 : KB, nj = n - j0 < JB ? n - j0 : JB;
  const unsigned n2 = (unsigned)n * n;
  const size_t fo = (size_t)v * n2 * n;
  const T *Pv = P + (size_t)v * G * n2, *Rv = R + (size_t)v * G * n2, *Xv = X + (size_t)v * G * n2;
  const T *s1v = s1 + fo, *s2v_ = s2 ? s2 + fo : nullptr, *nv_ = PH ? nhat + fo : nullptr;
  T *ov = out + fo, *ov2 = PH == 2 ? out2 + fo : nullptr;
  T gk[KB][G];  // Vi[k^][G] of the block's rows (uniform)
  double ez[KB], ez2[KB];  // the z factors of exp(-dt H) of the block's rows
  const double E[2] = {e0.x, e0.y};
#pragma unroll
  for (int kb = 0; kb < KB; kb++) {
    const int kh = k0 + (kb < nk ? kb : 0);
#pragma unroll
    for (int s = 0; s < G; s++) gk[kb][s] = with_corr ? G4[kh * G + s] : (T)0;
    ez[kb] = PH ? ax[2 * n + kh] : 0.0;
    ez2[kb] = PH == 2 ? ax[5 * n + kh] : 0.0;
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) {  // n may exceed the block
    T rv[KB][G], gi[G];
#pragma unroll
    for (int kb = 0; kb < KB; kb++) {
      const unsigned kh = (unsigned)(k0 + (kb < nk ? kb : 0));
#pragma unroll
      for (int s = 0; s < G; s++) rv[kb][s] = with_corr ? Rv[s * n2 + kh * n + i] : (T)0;
    }
#pragma unroll
    for (int s = 0; s < G; s++) gi[s] = with_corr ? G4[i * G + s] : (T)0;
    const double dxi = PH ? a.kkx * DDX[i] : 0.0, exi = PH ? ax[i] : 0.0, exi2 = PH == 2 ? ax[3 * n + i] : 0.0;
    for (int jb = 0; jb < nj; jb += JU) {  // JU rows j^ per pass, all their loads in flight together
      T sv[JU][KB], s2v[JU][KB], nv[JU][KB], pj[JU][G];
      double ey[JU], ey2[JU];
#pragma unroll
      for (int u = 0; u < JU; u++) {
        const int jh = j0 + (jb + u < nj ? jb + u : jb);
        ey[u] = PH ? ax[n + jh] : 0.0;
        ey2[u] = PH == 2 ? ax[4 * n + jh] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < JU; u++) {
        const unsigned l0 = (unsigned)(j0 + (jb + u < nj ? jb + u : jb)) * n + i;  // past the edge: valid, not stored
#pragma unroll
        for (int kb = 0; kb < KB; kb++) {
          const unsigned l = l0 + (unsigned)(k0 + (kb < nk ? kb : 0)) * n2;
          sv[u][kb] = s1v[l];
          s2v[u][kb] = s2v_ ? s2v_[l] : (T)0;
          nv[u][kb] = PH ? nv_[l] : (T)0;
        }
#pragma unroll
        for (int s = 0; s < G; s++) pj[u][s] = with_corr ? Pv[s * n2 + l0] : (T)0;
      }
#pragma unroll
      for (int u = 0; u < JU; u++) {
        if (jb + u >= nj) break;
        const int jh = j0 + jb + u;
        const unsigned l0 = (unsigned)jh * n + i;
#pragma unroll
        for (int kb = 0; kb < KB; kb++) {
          if (kb >= nk) break;
          const unsigned kh = (unsigned)(k0 + kb);
          T corr = (T)0;
          if (with_corr) {
#pragma unroll
            for (int s = 0; s < G; s++) {
              corr += gk[kb][s] * pj[u][s];
              corr += G4[jh * G + s] * rv[kb][s];
              corr += gi[s] * Xv[s * n2 + kh * n + jh];
            }
          }
          T val = alpha * sv[u][kb] + corr;
          if (s2v_) val += beta * s2v[u][kb];
          if constexpr (PH == 0) {
            ov[l0 + kh * n2] = val;
          } else {
            // H of this point as k_phiscale forms it (AC kind); exp(-dt H)
            // from the per-axis factors
            const double tmp = a.kkz * DDZ[kh] + a.kky * DDY[jh] + dxi;
            const double H = -a.LE * (tmp * a.epn2 - a.KE);
            const phi_pt ph = phi_sep(H, E[0] * ez[kb] * ey[u] * exi, a.dt);
            ov[l0 + kh * n2] = (T)(ph.lin * (double)val + ph.phi1 * (double)nv[u][kb]);
            if constexpr (PH == 2) {
              const phi_pt p2 = phi_sep(H, E[1] * ez2[kb] * ey2[u] * exi2, a.dt2);
              ov2[l0 + kh * n2] = (T)(p2.lin * (double)val + p2.phi1 * (double)nv[u][kb]);
            }
          }
        }
      }
    }
  }
}

// buffer layout (elements of T): dz, dy, dx (the recorded changes, nac 2g
// n x n matrices each, in that order), t1 (3 nac 2g n^2), P, R, X (nac 2g
// n^2 each, in that order), Vi and Vi^T (row-major n^2 each), G4 (n x 2g)
template <typename T>
static size_t carry_elems(const etd_ctx *c, size_t *off) {
  const size_t n = c->n, g = c->g, G = 2 * g, nac = c->nac, n2 = n * n;
  size_t o = 0;
  off[0] = o, o += nac * G * n2;      // dz
  off[1] = o, o += nac * G * n2;      // dy
  off[2] = o, o += nac * G * n2;      // dx
  off[3] = o, o += 3 * nac * G * n2;  // t1
  off[4] = o, o += nac * G * n2;      // P
  off[5] = o, o += nac * G * n2;      // R
  off[6] = o, o += nac * G * n2;      // X
  off[7] = o, o += n2;                // Vi
  off[8] = o, o += n2;                // Vi^T
  off[9] = o, o += n * G;             // G4
  return o;
}

// one box, two ghost layers, and coefficients whose exp(-dt H) factors stay
// at most 1 (k_phi_axes)
bool carry_possible(const etd_ctx *c) {
  const etd_params &p = c->p;
  return c->P == 1 && c->nac > 0 && c->g == 2 && !c->no_carry && p.LE >= 0 && p.epn2 >= 0 && p.kkx >= 0 &&
         p.kky >= 0 && p.kkz >= 0;
}

// all or nothing: on any failure nothing stays allocated (see carry_ready)
template <typename T>
int carry_prepare(etd_ctx *c) {
  if (c->cr_buf) return ETD_OK;
  const int n = c->n, g = c->g, G = 2 * g;
  size_t off[10];
  const size_t tot = carry_elems<T>(c, off);
  const size_t nh = 2 * (size_t)n * n + (size_t)n * G;
  T *h = (T *)malloc(sizeof(T) * nh);
  if (!h) return ETD_ERR_NOMEM;
  // hVi is column-major: Vi(r, c) = hVi[r + n c]
  for (int r = 0; r < n; r++)
    for (int q = 0; q < n; q++) {
      h[(size_t)r * n + q] = (T)c->hVi[r + (size_t)n * q];                  // Vi[r][q]
      h[(size_t)n * n + (size_t)q * n + r] = (T)c->hVi[r + (size_t)n * q];  // Vi^T[q][r]
    }
  for (int a = 0; a < n; a++)
    for (int s = 0; s < G; s++) {
      const int gi = s < g ? s : n - 2 * g + s;
      h[2 * (size_t)n * n + (size_t)a * G + s] = (T)c->hVi[a + (size_t)n * gi];
    }
  void *buf = nullptr;
  double *phi = nullptr;
  rocblas_handle hb = nullptr;
  int rc = ETD_OK;
  if (hipMalloc(&buf, sizeof(T) * tot) != hipSuccess || hipMalloc(&phi, sizeof(double) * 6 * (size_t)n) != hipSuccess)
    rc = ETD_ERR_NOMEM;
  else if (hipMemcpy((T *)buf + off[7], h, sizeof(T) * nh, hipMemcpyHostToDevice) != hipSuccess ||
           // the zeros of dy and dx outside what the fill records
           hipMemset((T *)buf + off[1], 0, sizeof(T) * (off[3] - off[1])) != hipSuccess ||
           rocblas_create_handle(&hb) != rocblas_status_success ||
           rocblas_set_stream(hb, c->stream) != rocblas_status_success)
    rc = ETD_ERR_DEVICE;
  free(h);
  if (rc) {
    if (hb) rocblas_destroy_handle(hb);
    (void)hipFree(buf);
    (void)hipFree(phi);
    return rc;
  }
  c->cr_blas = hb;
  c->cr_phi = phi;
  c->cr_buf = buf;
  return ETD_OK;
}

// whether this context carries from here on: the buffers exist or were just
// made.  If they cannot be made (device memory short), the context stops
// carrying for good and steps transform u, as with ETD_CARRY=0; the failed
// allocation's error is cleared so that it does not fail the step.
bool carry_ready(etd_ctx *c) {
  if (!carry_possible(c)) return false;
  if ((c->mixed ? carry_prepare<float>(c) : carry_prepare<double>(c)) == ETD_OK) return true;
  (void)hipGetLastError();
  c->no_carry = 1;
  c->carry_ok = 0;
  return false;
}

template <typename T>
void carry_deltas(etd_ctx *c, T **dz, T **dy, T **dx) {
  size_t off[10];
  carry_elems<T>(c, off);
  T *b = (T *)c->cr_buf;
  *dz = b + off[0];
  *dy = b + off[1];
  *dx = b + off[2];
}

// row-major C (M x N, row stride ldc) = A (M x K, lda) B (K x N, ldb) for
// `batch` matrices (strides sa, sb, sc; 0 shares an operand), through the
// column-major BLAS as C^T = B^T A^T; a failure is kept in c->blas_status
// for the step to report
static void gemm_rm(etd_ctx *c, const double *A, long lda, long sa, const double *B, long ldb, long sb, double *C,
                    long ldc, long sc, int M, int N, int K, int batch) {
  const double one = 1.0, zero = 0.0;
  const rocblas_status st =
      rocblas_dgemm_strided_batched((rocblas_handle)c->cr_blas, rocblas_operation_none, rocblas_operation_none, N, M,
                                    K, &one, B, ldb, sb, A, lda, sa, &zero, C, ldc, sc, batch);
  if (st != rocblas_status_success && !c->blas_status) c->blas_status = (int)st;
}
static void gemm_rm(etd_ctx *c, const float *A, long lda, long sa, const float *B, long ldb, long sb, float *C,
                    long ldc, long sc, int M, int N, int K, int batch) {
  const float one = 1.0f, zero = 0.0f;
  const rocblas_status st =
      rocblas_sgemm_strided_batched((rocblas_handle)c->cr_blas, rocblas_operation_none, rocblas_operation_none, N, M,
                                    K, &one, B, ldb, sb, A, lda, sa, &zero, C, ldc, sc, batch);
  if (st != rocblas_status_success && !c->blas_status) c->blas_status = (int)st;
}

int carry_status(etd_ctx *c) {
  const int st = c->blas_status;
  c->blas_status = 0;
  if (!st) return ETD_OK;
  snprintf(c->err, sizeof(c->err), "carry: rocBLAS strided-batched GEMM failed (status %d)", st);
  return ETD_ERR_DEVICE;
}

// P, R, X from the recorded changes: T = D Vi^T, then Vi T, over the
// 3 nac 2g matrices D of dz, dy, dx (nothing to do on physical sides, which
// have no ghost fill)
template <typename T>
static void carry_gemms(etd_ctx *c) {
  if (!c->p.periodic) return;
  const int n = c->n, nb = 3 * c->nac * 2 * c->g;
  const long n2 = (long)n * n;
  size_t off[10];
  carry_elems<T>(c, off);
  T *b = (T *)c->cr_buf;
  gemm_rm(c, b + off[0], n, n2, b + off[8], n, 0, b + off[3], n, n2, n, n, n, nb);
  gemm_rm(c, b + off[7], n, 0, b + off[3], n, n2, b + off[4], n, n2, n, n, n, nb);
}

template <typename T, int PH>
static void carry_expand(etd_ctx *c, const T *s1, double alpha, const T *s2, double beta, T *out, const T *nhat,
                         T *out2, const phi_args &a) {
  const int n = c->n, nac = c->nac;
  size_t off[10];
  carry_elems<T>(c, off);
  const T *b = (const T *)c->cr_buf;
  double2 e0 = make_double2(0.0, 0.0);
  if (PH) {
    k_phi_axes<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(c->DDX, c->DDY, c->DDZ, n, a, PH == 2,
                                                                   c->cr_phi);
    e0 = make_double2(exp(-a.dt * a.LE * a.KE), exp(-a.dt2 * a.LE * a.KE));
  }
  constexpr int KB = 4, JB = 8, JU = 2;
  const dim3 grid((unsigned)((n + JB - 1) / JB), (unsigned)((n + KB - 1) / KB), (unsigned)nac);
  k_carry_expand<T, KB, JB, JU, PH><<<grid, 256, 0, c->stream>>>(s1, (T)alpha, s2, (T)beta, out, b + off[4],
                                                                 b + off[5], b + off[6], b + off[9], n,
                                                                 c->p.periodic, nhat, out2, c->DDX, c->DDY,
                                                                 c->DDZ, a, c->cr_phi, e0);
}

template <typename T>
void carry_correct(etd_ctx *c, const T *s1, double alpha, const T *s2, double beta, T *out) {
  etd_prof_begin(c, PC_CARRY);
  carry_gemms<T>(c);
  carry_expand<T, 0>(c, s1, alpha, s2, beta, out, nullptr, nullptr, phi_args_of(c, 0.0, 0.0));
  etd_prof_end(c, PC_CARRY);
}

template <typename T>
void carry_correct_phi(etd_ctx *c, const T *s1, double alpha, const T *s2, double beta, const T *nhat, T *out,
                       double dt, T *out2, double dt2) {
  etd_prof_begin(c, PC_CARRY);
  carry_gemms<T>(c);
  if (out2)
    carry_expand<T, 2>(c, s1, alpha, s2, beta, out, nhat, out2, phi_args_of(c, dt, dt2));
  else
    carry_expand<T, 1>(c, s1, alpha, s2, beta, out, nhat, nullptr, phi_args_of(c, dt, 0.0));
  etd_prof_end(c, PC_CARRY);
}

template int carry_prepare<double>(etd_ctx *);
template int carry_prepare<float>(etd_ctx *);
template void carry_deltas<double>(etd_ctx *, double **, double **, double **);
template void carry_deltas<float>(etd_ctx *, float **, float **, float **);
template void carry_correct<double>(etd_ctx *, const double *, double, const double *, double, double *);
template void carry_correct<float>(etd_ctx *, const float *, double, const float *, double, float *);
template void carry_correct_phi<double>(etd_ctx *, const double *, double, const double *, double, const double *,
                                        double *, double, double *, double);
template void carry_correct_phi<float>(etd_ctx *, const float *, double, const float *, double, const float *,
                                       float *, double, float *, double);
