// etd_adapt.hip -- adaptive ETD with step doubling, the discrete-energy
// acceptance test, and the plain / learned dt controllers.  No counterpart
// in the reference, which steps with a fixed dt (ScLETD.c:84-191).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "etd_internal.h"

#define RBLOCKS 2048
#define RCOLS 9
#define ERR_COL 8  // k_maxdiff_extrap's column, beside the acceptance statistics' eight

// ------------------------------------------------------------------ reductions
// Interior cells visited one (z, y) row per wave, lanes along x: one 32-bit
// division per row instead of a 64-bit one per cell.  Fixed mapping, so the
// partial sums are deterministic.
#define FOR_INTERIOR_CELLS(N, x, y, z)                                                          \
  for (int r_ = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), z = r_ / (N), y = r_ - z * (N); \
       r_ < (N) * (N); r_ += gridDim.x * (blockDim.x >> 6), z = r_ / (N), y = r_ - z * (N))        \
    for (int x = threadIdx.x & 63; x < (N); x += 64)

// Run features over the interior cells of the AC fields: max |u - u0|,
// max |eta|, count of cells with max_p |eta_p| > 0.1, and the sum over p of
// |grad eta_p| (central differences on the periodic images).  Fields go in
// batches of FB whose loads are all issued before any is used (a batch past
// the last field re-reads the last field and ignores it), so a thread waits
// for memory once per batch, not once per field.
#define FB 4
struct cell_nb {
  size_t l, xp, xm, yp, ym, zp, zm;
};
__device__ __forceinline__ cell_nb neighbours(int n, int g, int N, int x, int y, int z) {
  const int xp = x + 1 == N ? 0 : x + 1, xm = x == 0 ? N - 1 : x - 1;
  const int yp = y + 1 == N ? 0 : y + 1, ym = y == 0 ? N - 1 : y - 1;
  const int zp = z + 1 == N ? 0 : z + 1, zm = z == 0 ? N - 1 : z - 1;
  auto at = [&](int zz, int yy, int xx) { return (((size_t)(zz + g)) * n + (yy + g)) * n + (xx + g); };
  cell_nb c;
  c.l = at(z, y, x);
  c.xp = at(z, y, xp);
  c.xm = at(z, y, xm);
  c.yp = at(z, yp, x);
  c.ym = at(z, ym, x);
  c.zp = at(zp, y, x);
  c.zm = at(zm, y, x);
  return c;
}

template <typename T>
__global__ __launch_bounds__(256) void k_features(const T *__restrict__ u, const T *__restrict__ u0, int n, int g,
                                                  int nac, double ihx, double ihy, double ihz,
                                                  double *__restrict__ part) {
  const int N = n - 2 * g;
  const size_t n3 = (size_t)n * n * n;
  double dmax = 0.0, emax = 0.0, cnt = 0.0, gsum = 0.0;
  FOR_INTERIOR_CELLS(N, x, y, z) {
    const cell_nb c = neighbours(n, g, N, x, y, z);
    double cell = 0.0;
    for (int p0 = 0; p0 < nac; p0 += FB) {
      T v[FB], v0[FB], vxp[FB], vxm[FB], vyp[FB], vym[FB], vzp[FB], vzm[FB];
#pragma unroll
      for (int e = 0; e < FB; e++) {
        const int p = p0 + e < nac ? p0 + e : nac - 1;
        const T *f = u + n3 * p;
        v[e] = f[c.l];
        v0[e] = u0[n3 * p + c.l];
        vxp[e] = f[c.xp];
        vxm[e] = f[c.xm];
        vyp[e] = f[c.yp];
        vym[e] = f[c.ym];
        vzp[e] = f[c.zp];
        vzm[e] = f[c.zm];
      }
#pragma unroll
      for (int e = 0; e < FB; e++) {
        if (p0 + e >= nac) break;
        const double vv = v[e];
        dmax = fmax(dmax, fabs(vv - (double)v0[e]));
        cell = fmax(cell, fabs(vv));
        const double gx = ((double)vxp[e] - (double)vxm[e]) * 0.5 * ihx;
        const double gy = ((double)vyp[e] - (double)vym[e]) * 0.5 * ihy;
        const double gz = ((double)vzp[e] - (double)vzm[e]) * 0.5 * ihz;
        gsum += (double)sqrtf((float)(gx * gx + gy * gy + gz * gz));  // a feature: FP32 suffices
      }
    }
    emax = fmax(emax, cell);
    cnt += cell > 0.1 ? 1.0 : 0.0;
  }
  for (int o = 32; o > 0; o >>= 1) {
    dmax = fmax(dmax, __shfl_down(dmax, o, 64));
    emax = fmax(emax, __shfl_down(emax, o, 64));
    cnt += __shfl_down(cnt, o, 64);
    gsum += __shfl_down(gsum, o, 64);
  }
  __shared__ double sh[4][4];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[0][w] = dmax;
    sh[1][w] = emax;
    sh[2][w] = cnt;
    sh[3][w] = gsum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double *o = part + blockIdx.x * RCOLS;
    o[0] = fmax(fmax(sh[0][0], sh[0][1]), fmax(sh[0][2], sh[0][3]));
    o[1] = fmax(fmax(sh[1][0], sh[1][1]), fmax(sh[1][2], sh[1][3]));
    o[2] = ((sh[2][0] + sh[2][1]) + sh[2][2]) + sh[2][3];
    o[3] = ((sh[3][0] + sh[3][1]) + sh[3][2]) + sh[3][3];
  }
}

// block sums of two per-thread values into part[block * RCOLS + 0 / 1]
__device__ inline void block_sum2(double a, double b, double *__restrict__ part) {
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    b += __shfl_down(b, o, 64);
  }
  __shared__ double sh[2][4];
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = a;
    sh[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[blockIdx.x * RCOLS] = ((sh[0][0] + sh[0][1]) + sh[0][2]) + sh[0][3];
    part[blockIdx.x * RCOLS + 1] = ((sh[1][0] + sh[1][1]) + sh[1][2]) + sh[1][3];
  }
}

// sum over AC fields and interior cells of eta_p felas_p (column 0) and of
// |eta_p felas_p| (column 1), felas from the ring slots of the latest elastic
// evaluation (its base linear index Lbase)
template <typename T>
__global__ __launch_bounds__(256) void k_elastic_energy(const T *__restrict__ u, const T *__restrict__ E, int n, int g,
                                                        int nac, int ring, int slot0,
                                                        double *__restrict__ part) {
  const int N = n - 2 * g;
  const size_t N3 = (size_t)N * N * N, n3 = (size_t)n * n * n;
  double acc = 0.0, aabs = 0.0;
  FOR_INTERIOR_CELLS(N, x, y, z) {
    const size_t l = (((size_t)(z + g)) * n + (y + g)) * n + (x + g);
    const size_t m = ((size_t)z * N + y) * N + x;
    for (int p0 = 0; p0 < nac; p0 += FB) {
      T a[FB], b[FB];
#pragma unroll
      for (int e = 0; e < FB; e++) {
        const int p = p0 + e < nac ? p0 + e : nac - 1;
        int slot = slot0 + p;  // p < nac < ring
        if (slot >= ring) slot -= ring;
        a[e] = u[n3 * p + l];
        b[e] = E[(size_t)slot * N3 + m];
      }
#pragma unroll
      for (int e = 0; e < FB; e++) {
        if (p0 + e >= nac) break;
        const double t = (double)a[e] * (double)b[e];
        acc += t;
        aabs += fabs(t);
      }
    }
  }
  block_sum2(acc, aabs, part);
}

// The scheme's own discrete free energy over the interior (periodic single
// subdomain: the truncated operators plus the explicit outside coupling are
// the periodic 7-point Laplacian), per cell h^3 times
//   AC: sum_p epn2/2 (kkx (D+x eta_p)^2 + kky (D+y)^2 + kkz (D+z)^2)
//       + epn2/2 grad_c eta_p . Lambda_p grad_c eta_p (anisotropic, central)
//       + a2/2 S2 - a3/3 S3 + a4/4 S2^2,   S2 = sum eta^2, S3 = sum eta^3
//   CH: epn2/2 |D+ c|^2 + (c^2 - 1)^2 / 4
// The semi-discrete dynamics is the gradient flow of this energy (plus the
// elastic energy), unlike the 8-corner check_soln_new diagnostic.  The sum of
// the absolute values of its terms is the scale of its rounding.
struct scheme_args {
  int n, g, nac, nch, aniso;
  double ihx, ihy, ihz;  // 1 / h: products, not FP64 divisions, in the cell loop
  double epn2, kkx, kky, kkz, a2, a3, a4;
};

// The acceptance statistics of a candidate in one pass over its fields
// (scheme energy, elastic energy, run features; columns of part):
//   0, 1  the scheme energy above and the sum of its terms' absolute values
//   2, 3  sum_p eta_p felas_p and sum |eta_p felas_p| (with_el)
//   4, 5  max |u - u0| and max |eta| over the AC fields
//   6, 7  count of cells with max_p |eta_p| > 0.1, sum_p |grad eta_p|
// (the features as k_features computes them).
template <typename T>
__global__ __launch_bounds__(256) void k_accept_stats(const T *__restrict__ u, const T *__restrict__ u0,
                                                      const T *__restrict__ E, const double *__restrict__ lam,
                                                      scheme_args a, int with_el, int ring, int slot0,
                                                      double *__restrict__ part) {
  const int n = a.n, g = a.g, N = n - 2 * g;
  const size_t N3 = (size_t)N * N * N, n3 = (size_t)n * n * n;
  const int V = a.nac + a.nch;
  double acc = 0.0, aabs = 0.0, el = 0.0, elabs = 0.0, dmax = 0.0, emax = 0.0, cnt = 0.0, gsum = 0.0;
  FOR_INTERIOR_CELLS(N, x, y, z) {
    const cell_nb c = neighbours(n, g, N, x, y, z);
    const size_t m = ((size_t)z * N + y) * N + x;
    double s2 = 0.0, s3 = 0.0, e = 0.0, ea = 0.0, cell = 0.0;
    for (int p0 = 0; p0 < V; p0 += FB) {
      T v[FB], v0[FB], vxp[FB], vyp[FB], vzp[FB], vxm[FB], vym[FB], vzm[FB], fe[FB];
#pragma unroll
      for (int q = 0; q < FB; q++) {
        const int p = p0 + q < V ? p0 + q : V - 1;
        const T *f = u + n3 * p;
        v[q] = f[c.l];
        vxp[q] = f[c.xp];
        vyp[q] = f[c.yp];
        vzp[q] = f[c.zp];
        vxm[q] = f[c.xm];
        vym[q] = f[c.ym];
        vzm[q] = f[c.zm];
        const int pa = p < a.nac ? p : 0;  // AC-only reads stay inside the AC fields
        v0[q] = u0[n3 * pa + c.l];
        if (with_el) {
          int slot = slot0 + pa;  // pa < nac < ring
          if (slot >= ring) slot -= ring;
          fe[q] = E[(size_t)slot * N3 + m];
        }
      }
#pragma unroll
      for (int q = 0; q < FB; q++) {
        const int p = p0 + q;
        if (p >= V) break;
        const double vv = v[q];
        const double dx = ((double)vxp[q] - vv) * a.ihx, dy = ((double)vyp[q] - vv) * a.ihy,
                     dz = ((double)vzp[q] - vv) * a.ihz;
        if (p < a.nac) {
          e += 0.5 * a.epn2 * (a.kkx * dx * dx + a.kky * dy * dy + a.kkz * dz * dz);
          const double gxc = ((double)vxp[q] - (double)vxm[q]) * 0.5 * a.ihx;
          const double gyc = ((double)vyp[q] - (double)vym[q]) * 0.5 * a.ihy;
          const double gzc = ((double)vzp[q] - (double)vzm[q]) * 0.5 * a.ihz;
          if (a.aniso) {
            // the reference's F2 convention: gy, gz as (u[-1] - u[+1]) / 2h
            const double gx = gxc, gy = -gyc, gz = -gzc;
            const double *L = lam + 9 * p;
            const double lx = L[0] * gx + L[1] * gy + L[2] * gz, ly = L[3] * gx + L[4] * gy + L[5] * gz,
                         lz = L[6] * gx + L[7] * gy + L[8] * gz;
            const double t = 0.5 * a.epn2 * (gx * lx + gy * ly + gz * lz);
            e += t;
            ea += fabs(t) - t;
          }
          s2 += vv * vv;
          s3 += vv * vv * vv;
          dmax = fmax(dmax, fabs(vv - (double)v0[q]));
          cell = fmax(cell, fabs(vv));
          gsum += (double)sqrtf((float)(gxc * gxc + gyc * gyc + gzc * gzc));
          if (with_el) {
            const double t = vv * (double)fe[q];
            el += t;
            elabs += fabs(t);
          }
        } else {
          e += 0.5 * a.epn2 * (dx * dx + dy * dy + dz * dz) + 0.25 * (vv * vv - 1.0) * (vv * vv - 1.0);
        }
      }
    }
    ea += e + fabs(0.5 * a.a2 * s2) + fabs(a.a3 / 3.0 * s3) + 0.25 * fabs(a.a4) * s2 * s2;
    e += 0.5 * a.a2 * s2 - a.a3 / 3.0 * s3 + 0.25 * a.a4 * s2 * s2;
    acc += e;
    aabs += ea;
    if (a.nac > 0) {
      emax = fmax(emax, cell);
      cnt += cell > 0.1 ? 1.0 : 0.0;
    }
  }
  double vals[8] = {acc, aabs, el, elabs, dmax, emax, cnt, gsum};
#pragma unroll
  for (int k = 0; k < 8; k++)
    for (int o = 32; o > 0; o >>= 1) {
      const double w = __shfl_down(vals[k], o, 64);
      vals[k] = (k == 4 || k == 5) ? fmax(vals[k], w) : vals[k] + w;
    }
  __shared__ double sh[8][4];
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < 8; k++) sh[k][threadIdx.x >> 6] = vals[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 8; k++)
      part[blockIdx.x * RCOLS + k] = (k == 4 || k == 5) ? fmax(fmax(sh[k][0], sh[k][1]), fmax(sh[k][2], sh[k][3]))
                                                        : ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
}

// max that keeps a NaN once it has seen one (fmax drops NaN operands, which
// would hide a NaN cell from the error estimate)
__host__ __device__ __forceinline__ double nanmax(double a, double b) { return (b > a || b != b) ? b : a; }

// max |u - b| per block into part[block * RCOLS + ERR_COL] (NaN if any cell is NaN) and,
// with extrapolate, u = 2 u - b (local extrapolation of the step-doubling
// pair) in the same pass
template <typename T>
__global__ __launch_bounds__(256) void k_maxdiff_extrap(T *__restrict__ u, const T *__restrict__ b, size_t cnt,
                                                        int extrapolate, double *__restrict__ part) {
  double m = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x) {
    const double x = u[i], y = b[i];
    m = nanmax(m, fabs(x - y));
    if (extrapolate) u[i] = (T)(2.0 * x - y);
  }
  for (int o = 32; o > 0; o >>= 1) m = nanmax(m, __shfl_down(m, o, 64));
  __shared__ double sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x * RCOLS + ERR_COL] = nanmax(nanmax(sh[0], sh[1]), nanmax(sh[2], sh[3]));
}



// ------------------------------------------------------------------ host helpers
#define ACHK(call)                                                                              \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) {                                                                     \
      snprintf(c->err, sizeof(c->err), "%s: %s", #call, hipGetErrorString(e_));                 \
      return ETD_ERR_DEVICE;                                                                    \
    }                                                                                           \
  } while (0)

static int fetch_part(etd_ctx *c, int cols) {
  ACHK(hipMemcpyAsync(c->ad_hpart, c->ad_part, sizeof(double) * RBLOCKS * RCOLS, hipMemcpyDeviceToHost, c->stream));
  ACHK(hipStreamSynchronize(c->stream));
  (void)cols;
  c->ad_fetches++;
  return ETD_OK;
}

struct run_feat {
  double dmax, emax, ifrac, gmean;
};

template <typename T>
static int features_of(etd_ctx *c, const T *u, const T *u0, run_feat *f) {
  if (c->nac == 0) {
    memset(f, 0, sizeof(*f));
    return ETD_OK;
  }
  k_features<T><<<RBLOCKS, 256, 0, c->stream>>>(u, u0, c->n, c->g, c->nac, 1.0 / c->p.hx, 1.0 / c->p.hy,
                                                1.0 / c->p.hz, c->ad_part);
  int rc = fetch_part(c, 4);
  if (rc) return rc;
  double dm = 0.0, em = 0.0, cn = 0.0, gs = 0.0;
  for (int i = 0; i < RBLOCKS; i++) {
    const double *o = c->ad_hpart + i * RCOLS;
    dm = fmax(dm, o[0]);
    em = fmax(em, o[1]);
    cn += o[2];
    gs += o[3];
  }
  f->dmax = dm;
  f->emax = em;
  f->ifrac = cn / (double)c->N3;
  f->gmean = gs / ((double)c->N3 * c->nac);
  return ETD_OK;
}

// The acceptance energy of the current fields, FP64.  kind 0: the
// check_soln_new energy of the AC and CH fields (the reference's diagnostic);
// kind 1: the scheme's discrete free energy (k_accept_stats).  With
// with_elastic, plus the elastic energy ES/2 h^3 sum_p eta_p felas_p from a
// fresh elastic evaluation of these fields, which the next step reuses
// (c->elastic_fresh) if they are kept.  *Eabs is the sum of the absolute
// values of the energy's terms (kind 1) or |E| (kind 0).
template <typename T>
static int energy_of(etd_ctx *c, int kind, int with_elastic, double *E, double *Eabs, const T *u0, run_feat *feat,
                     int *have_feat) {
  double e = 0.0, ea = 0.0;
  int rc = ETD_OK;
  *have_feat = 0;
  const bool el_on = with_elastic && c->p.elastic == 1 && c->nac > 0;
  const double w = 0.5 * c->p.elastic_scale * c->p.hx * c->p.hy * c->p.hz;
  const etd_fields<T> F = fields<T>(c);
  if (kind == 0) {
    double *st = (double *)malloc(sizeof(double) * (4 * c->V + 1));
    if (!st) return ETD_ERR_NOMEM;
    rc = launch_diag<T>(c, st);
    if (!rc) {
      e = c->nac > 0 ? st[3] : 0.0;
      for (int q = 0; q < c->nch; q++) e += st[4 * (c->nac + q) + 3];
      ea = fabs(e);
    }
    free(st);
    if (rc) return rc;
    if (el_on) {
      // the evaluation packs interior cells only: the fields' ghost cells are
      // left as the step wrote them (plain stepping refreshes them next step)
      launch_elastic<T>(c);
      const long long Lbase = (c->elastic_evals - 1) * (long long)c->nac;
      k_elastic_energy<T><<<RBLOCKS, 256, 0, c->stream>>>(F.u, F.E, c->n, c->g, c->nac, c->ring,
                                                          (int)(Lbase % c->ring), c->ad_part);
      if ((rc = fetch_part(c, 2))) return rc;
      double s = 0.0;
      for (int i = 0; i < RBLOCKS; i++) s += c->ad_hpart[i * RCOLS];
      e += w * s;
      ea = fabs(e);
      c->elastic_fresh = 1;  // the ring's latest evaluation is of these fields
    }
  } else {
    scheme_args a;
    a.n = c->n;
    a.g = c->g;
    a.nac = c->nac;
    a.nch = c->nch;
    a.aniso = c->p.anisotropic;
    a.ihx = 1.0 / c->p.hx;
    a.ihy = 1.0 / c->p.hy;
    a.ihz = 1.0 / c->p.hz;
    a.epn2 = c->p.epn2;
    a.kkx = c->p.kkx;
    a.kky = c->p.kky;
    a.kkz = c->p.kkz;
    a.a2 = c->p.a2;
    a.a3 = c->p.a3;
    a.a4 = c->p.a4;
    int slot0 = 0;
    if (el_on) {
      launch_elastic<T>(c);
      slot0 = (int)(((c->elastic_evals - 1) * (long long)c->nac) % c->ring);
    }
    k_accept_stats<T><<<RBLOCKS, 256, 0, c->stream>>>(F.u, u0, F.E, c->lam, a, el_on ? 1 : 0, c->ring, slot0,
                                                      c->ad_part);
    if ((rc = fetch_part(c, 8))) return rc;
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < RBLOCKS; i++) {
      const double *o = c->ad_hpart + i * RCOLS;
      for (int k = 0; k < 8; k++) s[k] = (k == 4 || k == 5) ? fmax(s[k], o[k]) : s[k] + o[k];
    }
    const double h3 = c->p.hx * c->p.hy * c->p.hz;
    e = s[0] * h3;
    ea = s[1] * h3;
    if (el_on) {
      e += w * s[2];
      ea += fabs(w) * s[3];
      c->elastic_fresh = 1;  // the ring's latest evaluation is of these fields
    }
    if (feat && c->nac > 0) {
      feat->dmax = s[4];
      feat->emax = s[5];
      feat->ifrac = s[6] / (double)c->N3;
      feat->gmean = s[7] / ((double)c->N3 * c->nac);
      *have_feat = 1;
    }
  }
  *E = e;
  *Eabs = ea;
  return ETD_OK;
}

static double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }
static double lg(double x, double floor_) { return log10(fabs(x) + floor_); }

extern "C" void etd_adapt_defaults(etd_adapt_opts *o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->t_end = 1.0;
  o->dt0 = 1e-3;
  o->dt_min = 1e-6;
  o->dt_max = 1.0;
  o->tol = 1e-4;
  o->safety = 0.9;
  o->learned_safety = 0.97;
  o->learned_online = 1;
  o->grow = 2.0;
  o->shrink = 0.2;
  o->controller = ETD_CTRL_PLAIN;
  o->extrapolate = 1;
  o->energy_test = 1;
  o->energy_elastic = 1;
  o->energy_kind = 1;
  o->max_attempts = 1000000;
}

extern "C" int etd_set_controller(etd_ctx *c, etd_mlp *m) {
  if (!c) return ETD_ERR_ARG;
  c->mlp = m;
  return ETD_OK;
}

template <typename T>
static int run_adaptive(etd_ctx *c, const etd_adapt_opts *o, etd_adapt_stats *st, double *log, int log_cap,
                        int *log_rows) {
  const int V = c->V;
  const size_t n3 = c->n3, cnt = n3 * V;
  const etd_fields<T> F = fields<T>(c);
  if (!c->ad_S0) {
    ACHK(hipMalloc(&c->ad_S0, sizeof(T) * cnt));
    ACHK(hipMalloc(&c->ad_SB, sizeof(T) * cnt));
    ACHK(hipMalloc(&c->ad_spec, sizeof(T) * 2 * cnt));  // the phi-scaled pair (dt, dt / 2)
    ACHK(hipMalloc(&c->ad_part, sizeof(double) * RBLOCKS * RCOLS));
    ACHK(hipHostMalloc(&c->ad_hpart, sizeof(double) * RBLOCKS * RCOLS, hipHostMallocDefault));
  }
  T *S0 = (T *)c->ad_S0, *SB = (T *)c->ad_SB, *SP = (T *)c->ad_spec, *SP2 = SP + cnt;
  // The attempts carry the AC spectrum: the second half step starts from the
  // first's spectrum (ETD1: phi-scaled, SP2; ETD2: its corrector's result),
  // and after an accepted attempt u is the backward transform of spec (u_h2)
  // or, extrapolated, of 2 spec - SP, SP the full step's spectrum (not
  // written again before the next attempt's forward stage).  A rejection
  // moves u back to S0, whose spectrum is not kept: the next attempt
  // transforms it.
  const bool carry_run = carry_ready(c);
  c->carry_ok = 0;
  carry_base(c, nullptr, 1.0, nullptr, 0.0);
  const double dt_fixed = c->p.dt;
  // energy-increase slack: the rounding level of the storage type times the
  // energy's scale, for the scheme energy the larger absolute term sum of
  // the two states (the same for one run or for the run cut into pieces),
  // for check_soln_new the largest |E| of the run
  const double slack = sizeof(T) == 4 ? 1e-7 : 1e-12;
  const int with_el = o->energy_elastic;
  const int learned = o->controller == ETD_CTRL_LEARNED && c->mlp;
  memset(st, 0, sizeof(*st));
  if (log_rows) *log_rows = 0;

  // state of the run
  double E, Eabs;
  const int ekind = o->energy_kind;
  run_feat rf;
  int hf = 0;
  int rc = energy_of<T>(c, ekind, with_el, &E, &Eabs, F.u, &rf, &hf);
  double Escale = fabs(E);
  if (rc) return rc;
  if (!isfinite(E)) {
    snprintf(c->err, sizeof(c->err), "adaptive: the energy of the initial state is not finite (%g)", E);
    return ETD_ERR_NONFINITE;
  }
  if (!hf && (rc = features_of<T>(c, F.u, F.u, &rf))) return rc;
  double hist[4] = {o->dt0, o->dt0, o->dt0, o->dt0};
  double err_last = 0.1 * o->tol, dEdt = 0.0, plain_next = o->dt0;
  int rej_since = 0, fallback = 0;
  // online correction of the network (decades): the largest recent
  // over-prediction y_pred - y_obs, decaying by 0.8 per learned step, is
  // taken off the next prediction (learned_online)
  double lbias = 0.0;
  double dt = o->dt0;
  int attempts = 0;
  rc = ETD_OK;
  while (c->t < o->t_end * (1.0 - 1e-12) && attempts < o->max_attempts) {
    // features seen by the proposal
    double feat[ETD_NFEAT];
    for (int k = 0; k < 4; k++) feat[k] = log10(hist[k]);
    feat[4] = log10(fmax(err_last, 1e-30) / o->tol);
    feat[5] = lg(dEdt, 1e-12);
    feat[6] = lg(E, 1e-30);
    feat[7] = lg(rf.dmax, 1e-12);
    feat[8] = rf.ifrac;
    feat[9] = rf.gmean;
    feat[10] = log10(c->t + 1e-6);
    feat[11] = log10(1.0 + st->accepted);
    feat[12] = rej_since;
    feat[13] = lg(rf.emax, 1e-30);
    feat[14] = log10(o->tol);
    feat[15] = log10(plain_next);
    int used_learned = 0;
    double y_pred = 0.0;
    const double plain_at_proposal = plain_next;
    if (rej_since == 0 && st->accepted + attempts > 0) {
      dt = plain_next;
      if (learned && !fallback) {
        // the network predicts log10 of the error model's largest step over
        // the plain proposal; at most a decade either way
        double y;
        if (etd_mlp_predict(c->mlp, feat, &y) == ETD_OK && isfinite(y)) {
          // no more than a decade above the plain proposal; never below the
          // smaller of it and the last step, so a network that errs low
          // cannot shrink the step where the plain rule would not
          // y predicts log10(dt_best / dt_plain), dt_best the error model's
          // largest step; its accuracy allows a smaller margin than the
          // plain rule's
          const double yc = o->learned_online ? y - lbias : y;
          dt = clampd(o->learned_safety * plain_next * pow(10.0, yc), fmin(plain_next, hist[0]), 10.0 * plain_next);
          st->learned_proposals++;
          used_learned = 1;
          y_pred = y;
        } else {
          st->learned_fallbacks++;
        }
      }
    }
    dt = clampd(dt, o->dt_min, o->dt_max);
    const double remain = o->t_end - c->t;
    if (dt > remain) dt = remain;
    attempts++;
    // attempt: full step and two half steps from u_n
    ACHK(hipMemcpyAsync(S0, F.u, sizeof(T) * cnt, hipMemcpyDeviceToDevice, c->stream));
    if (!c->p.etd2) {
      if constexpr (sizeof(T) == 4) ch_mean_mark(c, F.u);
      stage_forward<T>(c);  // N(u_n) and the forward transforms, shared
      // the full step and the first half step from one phi pass
      phi_stage<T>(c, SP, dt, SP2, 0.5 * dt);
      modeprod3<T>(c, 0, SP, SB, V, 0);
      modeprod3<T>(c, 0, SP2, F.u, V, 0);
      c->p.dt = 0.5 * dt;
      if constexpr (sizeof(T) == 4) {
        ch_mean_restore_centered(c, SB);
        ch_mean_restore_centered(c, F.u);
      }
      if (carry_run) {
        c->carry_ok = 1;
        carry_base(c, SP2, 1.0, nullptr, 0.0);
      }
      if ((rc = step_once_t<T>(c))) break;
    } else {
      // ETD2: the full step and the first half step share the predictor's
      // forward stage and one phi pass; each corrector runs on its own
      // predictor (and, carrying, leaves its step's spectrum: the full
      // step's in SP, the half step's in spec)
      if constexpr (sizeof(T) == 4) ch_mean_mark(c, F.u);
      stage_forward<T>(c);
      phi_stage<T>(c, SP, dt, SP2, 0.5 * dt);
      c->p.dt = dt;
      modeprod3<T>(c, 0, SP, F.u, V, 0);
      etd2_stage<T>(c, SP, carry_run ? SP : (T *)nullptr);
      ACHK(hipMemcpyAsync(SB, F.u, sizeof(T) * cnt, hipMemcpyDeviceToDevice, c->stream));
      c->p.dt = 0.5 * dt;
      modeprod3<T>(c, 0, SP2, F.u, V, 0);
      etd2_stage<T>(c, SP2, carry_run ? F.spec : (T *)nullptr);
      if (carry_run) {
        c->carry_ok = 1;
        carry_base(c, nullptr, 1.0, nullptr, 0.0);
      }
      c->carry_next = carry_run;
      rc = step_once_t<T>(c);
      c->carry_next = 0;
      if (rc) break;
    }
    st->step_evals += 3;
    ACHK(hipGetLastError());
    if ((rc = carry_status(c))) break;
    double err = 0.0;
    // the error's partial maxima ride along with the acceptance statistics'
    // fetch (one host synchronisation per attempt instead of two)
    k_maxdiff_extrap<T><<<RBLOCKS, 256, 0, c->stream>>>(F.u, SB, cnt, o->extrapolate, c->ad_part);
    double Enew, Eabs_new;
    run_feat nf;
    int nhf = 0;
    const long long f0 = c->ad_fetches;
    if ((rc = energy_of<T>(c, ekind, with_el, &Enew, &Eabs_new, S0, &nf, &nhf))) break;
    if (c->ad_fetches == f0 && (rc = fetch_part(c, RCOLS))) break;  // that energy fetched nothing
    for (int i = 0; i < RBLOCKS; i++) err = nanmax(err, c->ad_hpart[i * RCOLS + ERR_COL]);
    const double scale = ekind == 0 ? fmax(Escale, fabs(E)) : fmax(Eabs, Eabs_new);
    // a non-finite candidate (error estimate or energy) is never accepted:
    // it is rejected with the largest shrink, and at dt_min the run stops
    // with an error, the fields restored to the last accepted state
    const int finite = isfinite(err) && isfinite(Enew);
    const double fac = !finite ? o->shrink : (err > 0.0 ? o->safety * sqrt(o->tol / err) : o->grow);
    const int err_ok = err <= o->tol && finite;
    const int energy_ok = finite && (!o->energy_test || Enew <= E + slack * scale);
    const int at_min = dt <= o->dt_min * (1.0 + 1e-12);
    if (at_min && !finite) {
      ACHK(hipMemcpyAsync(F.u, S0, sizeof(T) * cnt, hipMemcpyDeviceToDevice, c->stream));
      c->elastic_fresh = 0;
      snprintf(c->err, sizeof(c->err),
               "adaptive: the candidate at dt_min = %g is not finite (err %g, energy %g) at t = %g", dt, err, Enew,
               c->t);
      rc = ETD_ERR_NONFINITE;
      break;
    }
    const int forced = at_min && finite;
    if (used_learned && finite && err > 0.0) {
      // what the network should have predicted for this attempt: the error
      // model's largest step at the observed err, over the plain proposal
      const double y_obs = log10(dt * sqrt(o->tol / err)) - log10(plain_at_proposal);
      lbias = fmax(0.8 * lbias, y_pred - y_obs);
    }
    if ((err_ok && energy_ok) || forced) {
      if (!nhf && (rc = features_of<T>(c, F.u, S0, &nf))) break;
      c->t += dt;
      st->accepted++;
      if (log && log_rows && *log_rows < log_cap) {
        double *row = log + (size_t)(*log_rows) * ETD_LOG_COLS;
        row[0] = c->t;
        row[1] = dt;
        row[2] = err;
        row[3] = Enew;
        memcpy(row + 4, feat, sizeof(feat));
        (*log_rows)++;
      }
      for (int k = 3; k > 0; k--) hist[k] = hist[k - 1];
      hist[0] = dt;
      dEdt = (Enew - E) / dt;
      E = Enew;
      Eabs = Eabs_new;
      Escale = fmax(Escale, fabs(E));
      err_last = err;
      rf = nf;
      plain_next = dt * clampd(fac, o->shrink, o->grow);
      rej_since = 0;
      fallback = 0;
      if (carry_run) {
        c->carry_ok = 1;
        if (o->extrapolate)
          carry_base(c, nullptr, 2.0, SP, -1.0);
        else
          carry_base(c, nullptr, 1.0, nullptr, 0.0);
      }
    } else {
      ACHK(hipMemcpyAsync(F.u, S0, sizeof(T) * cnt, hipMemcpyDeviceToDevice, c->stream));
      c->elastic_fresh = 0;  // the latest evaluation is of the rejected candidate
      c->carry_ok = 0;
      if (!err_ok) {
        st->rejected_err++;
        dt = dt * clampd(fac, o->shrink, 1.0);
      } else {
        st->rejected_energy++;
        dt = 0.5 * dt;
      }
      rej_since++;
      if (learned) {
        fallback = 1;  // retry with the plain controller
        st->learned_fallbacks++;
      }
    }
  }
  c->p.dt = dt_fixed;
  c->carry_ok = 0;  // the base may be 2 spec - SP: fixed steps start afresh
  carry_base(c, nullptr, 1.0, nullptr, 0.0);
  st->t = c->t;
  st->dt_last = hist[0];
  st->energy = E;
  if (rc) return rc;
  ACHK(hipStreamSynchronize(c->stream));
  return ETD_OK;
}

extern "C" int etd_run_adaptive(etd_ctx *c, const etd_adapt_opts *o, etd_adapt_stats *st, double *log, int log_cap,
                                int *log_rows) {
  if (!c || !o || !st) return ETD_ERR_ARG;
  if (!(o->tol > 0.0) || !(o->dt0 > 0.0) || !(o->dt_min > 0.0) || !(o->dt_max >= o->dt_min) || !(o->safety > 0.0) || !(o->learned_safety > 0.0) ||
      !(o->grow >= 1.0) || !(o->shrink > 0.0 && o->shrink <= 1.0))
    return ETD_ERR_ARG;
  if (o->controller == ETD_CTRL_LEARNED && !c->mlp) {
    snprintf(c->err, sizeof(c->err), "learned controller requested without etd_set_controller");
    return ETD_ERR_ARG;
  }
  if (c->P > 1) {
    snprintf(c->err, sizeof(c->err), "%s: one box only (px,py,pz = 1,1,1)", __func__);
    return ETD_ERR_UNSUPPORTED;
  }
  if (!c->p.periodic) {
    snprintf(c->err, sizeof(c->err), "adaptive stepping needs a periodic box (its energy is the periodic one)");
    return ETD_ERR_UNSUPPORTED;
  }
  if (c->p.free_energy != 0 || c->mask || c->p.variable_mobility || c->p.elastic == 2) {
    // the acceptance test needs the free energy the step descends; it is the Landau / double-well one
    snprintf(c->err, sizeof(c->err),
             "adaptive stepping accepts on the Landau / double-well energy: not with the CALPHAD driving force, "
             "a grain mask, variable mobility or elasticity on the CH fields");
    return ETD_ERR_UNSUPPORTED;
  }
  return c->mixed ? run_adaptive<float>(c, o, st, log, log_cap, log_rows)
                  : run_adaptive<double>(c, o, st, log, log_cap, log_rows);
}
