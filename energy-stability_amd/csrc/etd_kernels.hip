// etd_kernels.hip -- hand-written CDNA4 (gfx950) kernels of the ScLETD step.
//
// Every kernel takes n^3 fields in the reference's xyz layout (x fastest)
// unless stated; spectral data is in the rotated x'y'z' layout produced by
// three rotating mode products.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <stdint.h>
#include <stdio.h>
#include <mutex>
#include <type_traits>
#include <utility>

#include "etd_internal.h"

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int wrap_e(int e, int n, int g) {
  const int N = n - 2 * g;
  return e < g ? e + N : (e >= n - g ? e - N : e);
}
__device__ __forceinline__ size_t IDX(int n, int k, int j, int i) { return ((size_t)k * n + j) * n + i; }

// Workgroups go to the 8 XCDs in turn, each XCD with its own L2.  For a
// stencil whose neighbours lie in nearby workgroups, renumber a grid of a
// multiple of 8 workgroups so that each XCD runs one contiguous range:
// workgroup b works on block (b % 8) (G / 8) + b / 8, and a neighbouring
// block's rows are then in the same L2.
__device__ __forceinline__ size_t xcd_block() {
  const unsigned G = gridDim.x, b = blockIdx.x;
  return (G & 7u) ? b : (size_t)(b & 7u) * (G >> 3) + (b >> 3);
}
__host__ __forceinline__ unsigned round8(size_t blocks) { return (unsigned)((blocks + 7) & ~(size_t)7); }

// wave-wide (64 lanes) reductions; lane 0 holds the result
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_down(v, o, 64));
  return v;
}

// l = (k n + j) n + i, for l < 2^32: each quotient from a double-precision
// reciprocal (within one of the true quotient: the relative error is about
// 2^-52) and one correction step, instead of 64-bit integer division.
__device__ __forceinline__ void idx3(size_t l, int n, int &i, int &j, int &k) {
  const double inv = 1.0 / n;
  const unsigned un = (unsigned)n, lu = (unsigned)l;
  unsigned r = (unsigned)((double)lu * inv);
  int rem = (int)(lu - r * un);
  if (rem < 0) {
    r--;
    rem += n;
  } else if (rem >= n) {
    r++;
    rem -= n;
  }
  unsigned kk = (unsigned)((double)r * inv);
  int jj = (int)(r - kk * un);
  if (jj < 0) {
    kk--;
    jj += n;
  } else if (jj >= n) {
    kk++;
    jj -= n;
  }
  i = rem;
  j = jj;
  k = (int)kk;
}

// ------------------------------------------------------------------ halo
// Periodic self-exchange of transfer() (trans.c:1939-1996): every ghost value
// comes from the pre-transfer state, z faces override y faces override x
// faces.  In place: z phase, then y phase (reads x-ghost cells of interior
// planes, not yet rewritten), then x phase (reads interior cells only).
// In a decomposed run (px x py x pz boxes) the source of the low (high)
// ghosts is the neighbouring box before (after) this one along the sweep's
// axis, lo / hi elements away: the z and y exchanges carry whole planes and
// rows of the neighbour, its own ghost cells included, and every sweep reads
// only cells the sweeps before it do not write, so running each sweep over
// all boxes before the next reads pre-transfer data, as the reference's
// send-everything-then-unpack order does.
// dlt (or null): the change new - old of every cell a sweep writes, per
// field as 2g contiguous n x n matrices: z [plane][j][i], y [row][k][i], x
// [column][k][j] (y and x write their interior rows / columns only; the
// rest of their matrices stays zero).
template <typename T>
__global__ void k_ghost_z(T *base, size_t fs, int n, int g, long long lo, long long hi, T *dlt) {
  T *u = base + fs * blockIdx.y;
  const size_t n2 = (size_t)n * n;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)2 * g * n2) return;
  int pl = (int)(t / n2);
  size_t r = t % n2;
  int k = pl < g ? pl : n - 2 * g + pl;
  const T v = u[(pl < g ? lo : hi) + (long long)((size_t)wrap_e(k, n, g) * n2 + r)];
  if (dlt) dlt[(size_t)2 * g * n2 * blockIdx.y + t] = (T)((double)v - (double)u[(size_t)k * n2 + r]);
  u[(size_t)k * n2 + r] = v;
}
template <typename T>
__global__ void k_ghost_y(T *base, size_t fs, int n, int g, long long lo, long long hi, T *dlt) {
  T *u = base + fs * blockIdx.y;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int rows = 2 * g;
  if (t >= (size_t)(n - 2 * g) * rows * n) return;
  int i = (int)(t % n);
  int rr = (int)((t / n) % rows);
  int k = g + (int)(t / ((size_t)n * rows));
  int j = rr < g ? rr : n - 2 * g + rr;
  const T v = u[(rr < g ? lo : hi) + (long long)IDX(n, k, wrap_e(j, n, g), i)];
  if (dlt)
    dlt[(size_t)rows * n * n * blockIdx.y + ((size_t)rr * n + k) * n + i] = (T)((double)v - (double)u[IDX(n, k, j, i)]);
  u[IDX(n, k, j, i)] = v;
}
template <typename T>
__global__ void k_ghost_x(T *base, size_t fs, int n, int g, long long lo, long long hi, T *dlt) {
  T *u = base + fs * blockIdx.y;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int cols = 2 * g, m = n - 2 * g;
  if (t >= (size_t)m * m * cols) return;
  int cc = (int)(t % cols);
  int j = g + (int)((t / cols) % m);
  int k = g + (int)(t / ((size_t)cols * m));
  int i = cc < g ? cc : n - 2 * g + cc;
  const T v = u[(cc < g ? lo : hi) + (long long)IDX(n, k, j, wrap_e(i, n, g))];
  if (dlt)
    dlt[(size_t)cols * n * n * blockIdx.y + ((size_t)cc * n + k) * n + j] = (T)((double)v - (double)u[IDX(n, k, j, i)]);
  u[IDX(n, k, j, i)] = v;
}

// element offset from box b's fields to those of the box (sx, sy, sz) away
static long long host_box_off(const etd_ctx *c, int b, int sx, int sy, int sz) {
  const int bx = b % c->px, by = (b / c->px) % c->py, bz = b / (c->px * c->py);
  const int x = (bx + sx + c->px) % c->px, y = (by + sy + c->py) % c->py, z = (bz + sz + c->pz) % c->pz;
  return (long long)((z * c->py + y) * c->px + x - b) * (long long)c->bs;
}

// base: box 0's first field; the boxes' fields follow at stride bs.  dz, dy,
// dx (one box only): record each sweep's changes (k_ghost_z/y/x)
template <typename T>
void launch_ghost_fill(etd_ctx *c, T *base, int nfields, size_t fstride, T *dz, T *dy, T *dx) {
  if (!c->p.periodic) return;  // physical sides: no neighbour, unpack writes nothing (trans.c:154-251)
  const int n = c->n, g = c->g;
  const size_t n2 = (size_t)n * n;
  size_t tz = 2 * g * n2, ty = (size_t)(n - 2 * g) * 2 * g * n, tx = (size_t)(n - 2 * g) * (n - 2 * g) * 2 * g;
  etd_prof_begin(c, PC_HALO);
  for (int b = 0; b < c->P; b++)
    k_ghost_z<T><<<dim3((unsigned)((tz + 255) / 256), nfields), 256, 0, c->stream>>>(
        base + b * c->bs, fstride, n, g, host_box_off(c, b, 0, 0, -1), host_box_off(c, b, 0, 0, 1), dz);
  for (int b = 0; b < c->P; b++)
    k_ghost_y<T><<<dim3((unsigned)((ty + 255) / 256), nfields), 256, 0, c->stream>>>(
        base + b * c->bs, fstride, n, g, host_box_off(c, b, 0, -1, 0), host_box_off(c, b, 0, 1, 0), dy);
  for (int b = 0; b < c->P; b++)
    k_ghost_x<T><<<dim3((unsigned)((tx + 255) / 256), nfields), 256, 0, c->stream>>>(
        base + b * c->bs, fstride, n, g, host_box_off(c, b, -1, 0, 0), host_box_off(c, b, 1, 0, 0), dx);
  etd_prof_end(c, PC_HALO);
}

// ------------------------------------------------------------------ mode product
// out = A x_1 in for each of nfields fields: the contraction over the fastest
// axis a of in[a + n*b + n^2*c], C(m, col) = sum_a A(m,a) in[a + n*col],
// col = b + n*c, written rotated (out[col + n^2*m], the reference's
// cblas_dgemm + xyz_yzx, calc.c:631-684) or in GEMM layout (out[m + n*col]).
//
// FP64 MFMA 16x16x4: a wave owns 16*NB columns and all MT = ceil(n/16) row
// tiles.  K runs in chunks of 16; inside a chunk lane group q = lane/16
// supplies k = 16 kc + 4 q + s at MFMA step s, so each lane's four B values
// are contiguous in memory (one 32-byte read).  A chunks, pre-arranged as
// [s][t][lane] rows of 64 doubles (conflict-free LDS reads), are staged
// through LDS (double-buffered) and shared by the workgroup's 4 waves; B is
// read straight from HBM one chunk ahead.
template <int MT, int NB, bool VEC>
__global__ __launch_bounds__(256) void k_modeprod_f64(const double *__restrict__ A4, const double *__restrict__ in,
                                                      size_t in_fs, double *__restrict__ out, size_t out_fs, int n,
                                                      int rotate, int accumulate) {
  // A chunk: MT*64 double4 = MT*128 double2, padded to APT*256 (zeros) so
  // every thread moves exactly APT double2 per chunk
  constexpr int A2 = MT * 128;
  constexpr int APT = (A2 + 255) / 256;
  __shared__ double2 sA2[2][APT * 256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane >> 4, r16 = lane & 15;
  const double *src = in + in_fs * blockIdx.y;
  double *dst = out + out_fs * blockIdx.y;
  const long n2 = (long)n * n;
  const int nkc = (n + 15) / 16;
  const long colbase = (long)blockIdx.x * (64 * NB) + wave * (16 * NB);

  v4d acc[MT][NB];
#pragma unroll
  for (int t = 0; t < MT; t++)
#pragma unroll
    for (int u = 0; u < NB; u++) acc[t][u] = (v4d){0.0, 0.0, 0.0, 0.0};

  // B fragment of column tile u for K chunk kc (zero outside the array)
  auto loadB = [&](int kc, int u) -> v4d {
    const long col = colbase + 16 * u + r16;
    const int k0 = 16 * kc + 4 * q;
    v4d v = (v4d){0.0, 0.0, 0.0, 0.0};
    if (col < n2) {
      const double *p = src + k0 + (long)n * col;
      if (VEC && k0 + 3 < n) {
        const double2 lo = *reinterpret_cast<const double2 *>(p);
        const double2 hi = *reinterpret_cast<const double2 *>(p + 2);
        v = (v4d){lo.x, lo.y, hi.x, hi.y};
      } else {
        if (k0 + 0 < n) v.x = p[0];
        if (k0 + 1 < n) v.y = p[1];
        if (k0 + 2 < n) v.z = p[2];
        if (k0 + 3 < n) v.w = p[3];
      }
    }
    return v;
  };
  // register staging of the next A chunk (fold expressions keep every index a
  // compile-time constant so the array lives in VGPRs)
  // staging of the next A chunk in a vector register tuple (a plain array
  // here was kept in scratch memory, which serialised every prefetch)
  typedef double stage_t __attribute__((ext_vector_type(2 * APT)));
  stage_t aprf;
  constexpr auto AIDX = std::make_index_sequence<APT>{};
  auto fetch_a = [&](int kc, auto seq) {
    const double2 *g2 = reinterpret_cast<const double2 *>(A4) + (size_t)kc * (APT * 256);
    [&]<size_t... I>(std::index_sequence<I...>) {
      ((aprf[2 * I] = g2[threadIdx.x + 256 * I].x, aprf[2 * I + 1] = g2[threadIdx.x + 256 * I].y), ...);
    }(seq);
  };
  auto store_a = [&](int b, auto seq) {
    [&]<size_t... I>(std::index_sequence<I...>) {
      ((sA2[b][threadIdx.x + 256 * I] = make_double2(aprf[2 * I], aprf[2 * I + 1])), ...);
    }(seq);
  };

  v4d bcur[NB], bnext[NB];
  fetch_a(0, AIDX);
#pragma unroll
  for (int u = 0; u < NB; u++) bcur[u] = loadB(0, u);
  store_a(0, AIDX);
  __syncthreads();
  for (int kc = 0; kc < nkc; kc++) {
    const int buf = kc & 1;
    const bool more = kc + 1 < nkc;
    if (more) {
      fetch_a(kc + 1, AIDX);
#pragma unroll
      for (int u = 0; u < NB; u++) bnext[u] = loadB(kc + 1, u);
    }
    // step s outermost: the 4 * MT * NB MFMAs of a chunk cycle through MT * NB
    // independent accumulators, so no MFMA waits on its predecessor
    const double *sa = reinterpret_cast<const double *>(&sA2[buf][0]);
#pragma unroll
    for (int st = 0; st < 4; st++) {
#pragma unroll
      for (int t = 0; t < MT; t++) {
        const double a = sa[(st * MT + t) * 64 + lane];
#pragma unroll
        for (int u = 0; u < NB; u++) acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bcur[u][st], acc[t][u], 0, 0, 0);
      }
    }
    if (more) {
      store_a(buf ^ 1, AIDX);
#pragma unroll
      for (int u = 0; u < NB; u++) bcur[u] = bnext[u];
    }
    __syncthreads();
  }
  // D(16x16) of v_mfma_f64_16x16x4f64: register rr of lane (q, r16) holds
  // row q + 4 rr, column r16
#pragma unroll
  for (int t = 0; t < MT; t++)
#pragma unroll
    for (int u = 0; u < NB; u++) {
      const long col = colbase + 16 * u + r16;
      if (col >= n2) continue;
#pragma unroll
      for (int rr = 0; rr < 4; rr++) {
        const int m = 16 * t + q + 4 * rr;
        if (m >= n) continue;
        const double val = acc[t][u][rr];
        const size_t o = rotate ? (size_t)col + (size_t)n2 * m : (size_t)m + (size_t)n * col;
        if (accumulate) dst[o] += val;
        else dst[o] = val;
      }
    }
}

// Same contraction on v_mfma_f64_4x4x4f64, which this part issues at about
// 1.5x the FP64 rate of the 16x16x4 form.  The instruction holds four
// independent 4x4x4 products; lane = x + 4 blk + 16 y addresses A[x][y],
// B[y][x] and D[y][x] of block blk.  Blocks are stacked along rows, so one
// instruction covers rows 16t + lane%16 (the same A staging as above) and the
// 4 columns colbase + 4u + lane%4; a wave owns 4*NB4 columns.
template <int MT, int NB4, bool VEC>
__global__ __launch_bounds__(256) void k_modeprod_f64m4(const double *__restrict__ A4, const double *__restrict__ in,
                                                        size_t in_fs, double *__restrict__ out, size_t out_fs, int n,
                                                        int rotate, int accumulate) {
  constexpr int A2 = MT * 128;
  constexpr int APT = (A2 + 255) / 256;
  __shared__ double2 sA2[2][APT * 256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane >> 4, j4 = lane & 3;
  const double *src = in + in_fs * blockIdx.y;
  double *dst = out + out_fs * blockIdx.y;
  const long n2 = (long)n * n;
  const int nkc = (n + 15) / 16;
  // 4 waves x 4*NB4 columns per workgroup
  const long wcol = (long)blockIdx.x * (16 * NB4) + wave * (4 * NB4);
  double acc[MT][NB4];
#pragma unroll
  for (int t = 0; t < MT; t++)
#pragma unroll
    for (int u = 0; u < NB4; u++) acc[t][u] = 0.0;

  auto loadB = [&](int kc, int u) -> v4d {
    const long col = wcol + 4 * u + j4;
    const int k0 = 16 * kc + 4 * q;
    v4d v = (v4d){0.0, 0.0, 0.0, 0.0};
    if (col < n2) {
      const double *p = src + k0 + (long)n * col;
      if (VEC && k0 + 3 < n) {
        const double2 lo = *reinterpret_cast<const double2 *>(p);
        const double2 hi = *reinterpret_cast<const double2 *>(p + 2);
        v = (v4d){lo.x, lo.y, hi.x, hi.y};
      } else {
        if (k0 + 0 < n) v.x = p[0];
        if (k0 + 1 < n) v.y = p[1];
        if (k0 + 2 < n) v.z = p[2];
        if (k0 + 3 < n) v.w = p[3];
      }
    }
    return v;
  };
  // staging of the next A chunk in a vector register tuple (a plain array
  // here was kept in scratch memory, which serialised every prefetch)
  typedef double stage_t __attribute__((ext_vector_type(2 * APT)));
  stage_t aprf;
  constexpr auto AIDX = std::make_index_sequence<APT>{};
  auto fetch_a = [&](int kc, auto seq) {
    const double2 *g2 = reinterpret_cast<const double2 *>(A4) + (size_t)kc * (APT * 256);
    [&]<size_t... I>(std::index_sequence<I...>) {
      ((aprf[2 * I] = g2[threadIdx.x + 256 * I].x, aprf[2 * I + 1] = g2[threadIdx.x + 256 * I].y), ...);
    }(seq);
  };
  auto store_a = [&](int b, auto seq) {
    [&]<size_t... I>(std::index_sequence<I...>) {
      ((sA2[b][threadIdx.x + 256 * I] = make_double2(aprf[2 * I], aprf[2 * I + 1])), ...);
    }(seq);
  };

  v4d bcur[NB4], bnext[NB4];
  fetch_a(0, AIDX);
#pragma unroll
  for (int u = 0; u < NB4; u++) bcur[u] = loadB(0, u);
  store_a(0, AIDX);
  __syncthreads();
  for (int kc = 0; kc < nkc; kc++) {
    const int buf = kc & 1;
    // the last iteration re-reads chunk 0 (in bounds, unused) so the staging
    // registers are written and read unconditionally
    const int kn = kc + 1 < nkc ? kc + 1 : 0;
    fetch_a(kn, AIDX);
#pragma unroll
    for (int u = 0; u < NB4; u++) bnext[u] = loadB(kn, u);
    const double *sa = reinterpret_cast<const double *>(&sA2[buf][0]);
    // A operands of step st+1 are read from LDS while step st's MFMAs run
    double acur[MT], anext[MT];
#pragma unroll
    for (int t = 0; t < MT; t++) acur[t] = sa[t * 64 + lane];
#pragma unroll
    for (int st = 0; st < 4; st++) {
      if (st < 3) {
#pragma unroll
        for (int t = 0; t < MT; t++) anext[t] = sa[((st + 1) * MT + t) * 64 + lane];
      }
#pragma unroll
      for (int t = 0; t < MT; t++)
#pragma unroll
        for (int u = 0; u < NB4; u++)
          acc[t][u] = __builtin_amdgcn_mfma_f64_4x4x4f64(acur[t], bcur[u][st], acc[t][u], 0, 0, 0);
      if (st < 3) {
#pragma unroll
        for (int t = 0; t < MT; t++) acur[t] = anext[t];
      }
    }
    store_a(buf ^ 1, AIDX);
#pragma unroll
    for (int u = 0; u < NB4; u++) bcur[u] = bnext[u];
    __syncthreads();
  }
  // D of block blk = (lane/4)%4: lane holds row 16t + 4 blk + lane/16,
  // column wcol + 4u + lane%4
  const int mrow = 4 * ((lane >> 2) & 3) + q;
#pragma unroll
  for (int t = 0; t < MT; t++)
#pragma unroll
    for (int u = 0; u < NB4; u++) {
      const long col = wcol + 4 * u + j4;
      const int m = 16 * t + mrow;
      if (col >= n2 || m >= n) continue;
      const size_t o = rotate ? (size_t)col + (size_t)n2 * m : (size_t)m + (size_t)n * col;
      if (accumulate) dst[o] += acc[t][u];
      else dst[o] = acc[t][u];
    }
}

// ---- 3-stage direct-to-LDS pipeline (default) ----
// A and B chunks move global -> LDS with buffer_load ... lds (no VGPR
// staging), two chunks ahead of the MFMAs; the buffer descriptor's range
// check returns 0 for out-of-range offsets, which also supplies the K tail.
typedef int i32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ i32x4 make_rsrc(const void *base, uint32_t bytes) {
  const uint64_t a = (uint64_t)base;
  i32x4 r;
  r.x = (int)(uint32_t)a;
  r.y = (int)(uint32_t)(a >> 32);  // stride 0
  r.z = (int)bytes;                // num_records (bytes)
  r.w = 0x00020000;                // gfx9 data format word
  return r;
}

// One 16-byte global -> LDS copy per lane: lane L's bytes land at
// lds_wave_base + 16 L.  Written as asm so the compiler does not treat the
// LDS write as aliasing every later LDS read (which would force a wait for
// all outstanding copies before each read); completion is tracked by the
// explicit vmcnt waits in the pipeline.  M0 needs one wait state before the
// copy reads it.  The instruction has no destination VGPR; its address
// operands are read at issue.
// M0 is reserved to the compiler; the compiled kernels of this file use it
// only inside these asm blocks (checked in the ISA), so the clobber holds.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void dma16(i32x4 rs, uint32_t lds_wave_base, int voff) {
  const uint32_t b = __builtin_amdgcn_readfirstlane(lds_wave_base);
  asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tbuffer_load_dwordx4 %1, %2, 0 offen lds"
               :
               : "s"(b), "v"(voff), "s"(rs)
               : "memory", "m0");
}
#pragma clang diagnostic pop

// Wait until at most VM copies of this wave are in flight, then barrier.
template <int VM>
__device__ __forceinline__ void wait_copies_barrier() {
  static_assert(VM >= 0 && VM < 64, "vmcnt");
  asm volatile("s_waitcnt vmcnt(%0)\n\ts_barrier" : : "n"(VM) : "memory");
}

template <int MT, int NB4, int NSTAGE>
struct mp_dl_cfg {
  static constexpr int APT = (MT * 128 + 255) / 256;   // 16-byte A pieces per thread per chunk
  static constexpr int ASZ = APT * 256 * 16;           // A chunk bytes
  static constexpr int WCOLS = 4 * NB4;                // columns per wave
  static constexpr int BSZ = 4 * WCOLS * 128;          // B chunk bytes (16 k x 8 B per column)
  static constexpr int BPW = WCOLS / 8;                // B pieces (1 KB) per wave per chunk
  static constexpr int STAGE = ASZ + BSZ;
  static constexpr int NST = NSTAGE;                   // LDS stages (chunks in flight + 1)
  static_assert(NST == 2 || NST == 3, "stages");
  static constexpr int LDS = NST * STAGE;
  static constexpr int LPC = APT + BPW;                // loads per thread per chunk
};

template <int MT, int NB4, int NSTAGE, int OCC>
__global__ __launch_bounds__(256, OCC) void k_modeprod_dl(const double *__restrict__ A4, const double *__restrict__ in,
                                                          size_t in_fs, double *__restrict__ out, size_t out_fs, int n,
                                                          int rotate, int accumulate) {
  using CF = mp_dl_cfg<MT, NB4, NSTAGE>;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane >> 4, j4 = lane & 3;
  const double *src = in + in_fs * blockIdx.y;
  double *dst = out + out_fs * blockIdx.y;
  const long n2 = (long)n * n;
  const int nkc = (n + 15) / 16;
  const long wgcol = (long)blockIdx.x * (4 * CF::WCOLS);
  const long wcol = wgcol + wave * CF::WCOLS;
  const i32x4 rsA = make_rsrc(A4, (uint32_t)((size_t)nkc * CF::APT * 256 * 16));
  const i32x4 rsB = make_rsrc(src, (uint32_t)((size_t)n * n2 * 8));
  const uint32_t lds0 = (uint32_t)(uintptr_t)smem;

  auto issue = [&](int kc, int stg) {
    const uint32_t sbase = lds0 + stg * CF::STAGE;
#pragma unroll
    for (int e = 0; e < CF::APT; e++)
      dma16(rsA, sbase + (wave * 64 + 256 * e) * 16, (int)(((size_t)kc * CF::APT * 256 + threadIdx.x + 256 * e) * 16));
#pragma unroll
    for (int i = 0; i < CF::BPW; i++) {
      const long col = wcol + 8 * i + (lane >> 3);
      const int k = 16 * kc + 2 * (lane & 7);
      const bool ok = col < n2 && k < n;
      const int voff = ok ? (int)((k + (long)n * col) * 8) : (int)0x80000000;
      dma16(rsB, sbase + CF::ASZ + (wave * CF::WCOLS + 8 * i) * 128, voff);
    }
  };

  double acc[MT][NB4];
#pragma unroll
  for (int t = 0; t < MT; t++)
#pragma unroll
    for (int u = 0; u < NB4; u++) acc[t][u] = 0.0;

  issue(0, 0);
  if (CF::NST == 3 && nkc > 1) issue(1, 1);
  for (int kc = 0; kc < nkc; kc++) {
    // chunk kc is in LDS for every wave, and every wave is past chunk kc-1,
    // whose stage the next copy overwrites
    if (CF::NST == 3 && kc + 1 < nkc) wait_copies_barrier<(CF::NST - 2) * CF::LPC>();
    else wait_copies_barrier<0>();
    if (kc + CF::NST - 1 < nkc) issue(kc + CF::NST - 1, (kc + CF::NST - 1) % CF::NST);
    const char *stg = smem + (kc % CF::NST) * CF::STAGE;
    const double *sa = reinterpret_cast<const double *>(stg);
    const double *sb = reinterpret_cast<const double *>(stg + CF::ASZ);
    v4d bv[NB4];
#pragma unroll
    for (int u = 0; u < NB4; u++) {
      const double *bp = sb + (wave * CF::WCOLS + 4 * u + j4) * 16 + 4 * q;
      const double2 lo = *reinterpret_cast<const double2 *>(bp);
      const double2 hi = *reinterpret_cast<const double2 *>(bp + 2);
      bv[u] = (v4d){lo.x, lo.y, hi.x, hi.y};
    }
    double acur[MT], anext[MT];
#pragma unroll
    for (int t = 0; t < MT; t++) acur[t] = sa[t * 64 + lane];
#pragma unroll
    for (int st = 0; st < 4; st++) {
      if (st < 3) {
#pragma unroll
        for (int t = 0; t < MT; t++) anext[t] = sa[((st + 1) * MT + t) * 64 + lane];
      }
#pragma unroll
      for (int t = 0; t < MT; t++)
#pragma unroll
        for (int u = 0; u < NB4; u++)
          acc[t][u] = __builtin_amdgcn_mfma_f64_4x4x4f64(acur[t], bv[u][st], acc[t][u], 0, 0, 0);
      if (st < 3) {
#pragma unroll
        for (int t = 0; t < MT; t++) acur[t] = anext[t];
      }
    }
  }
  const int mrow = 4 * ((lane >> 2) & 3) + q;
#pragma unroll
  for (int t = 0; t < MT; t++)
#pragma unroll
    for (int u = 0; u < NB4; u++) {
      const long col = wcol + 4 * u + j4;
      const int m = 16 * t + mrow;
      if (col >= n2 || m >= n) continue;
      const size_t o = rotate ? (size_t)col + (size_t)n2 * m : (size_t)m + (size_t)n * col;
      if (accumulate) dst[o] += acc[t][u];
      else dst[o] = acc[t][u];
    }
}

// host: A(m,k) column-major -> A4[kc*CH + (s*MT + t)*64 + lane]
//       = A(16t + lane%16, 16kc + 4(lane/16) + s), zero outside n x n;
// CH = per-chunk stride (MT*256 doubles rounded up to 512)
size_t arrange_A_chunk(int MT) { return (size_t)((MT * 128 + 255) / 256) * 256 * 2; }
void arrange_A(int n, const double *Acm, double *A4, int MT, int nkc) {
  const size_t CH = arrange_A_chunk(MT);
  memset(A4, 0, sizeof(double) * CH * nkc);
  for (int kc = 0; kc < nkc; kc++)
    for (int t = 0; t < MT; t++)
      for (int lane = 0; lane < 64; lane++)
        for (int s = 0; s < 4; s++) {
          int m = 16 * t + (lane & 15), k = 16 * kc + 4 * (lane >> 4) + s;
          A4[(size_t)kc * CH + ((size_t)s * MT + t) * 64 + lane] = (m < n && k < n) ? Acm[m + (size_t)n * k] : 0.0;
        }
}

template <int MT>
static void modeprod_mt(etd_ctx *c, const double *A4, const double *in, size_t in_fs, double *out, size_t out_fs,
                        int nfields, int rotate, int accumulate) {
  const int n = c->n;
  const long n2 = (long)n * n;
  auto dl = [&](auto nb4, auto nst, auto occ) {
    constexpr int NB4 = decltype(nb4)::value, NST = decltype(nst)::value, OCC = decltype(occ)::value;
    using CF = mp_dl_cfg<MT, NB4, NST>;
    static_assert(CF::LDS * OCC <= 160 * 1024, "LDS budget");
    static bool attr_set = false;
    if (!attr_set) {
      (void)hipFuncSetAttribute((const void *)k_modeprod_dl<MT, NB4, NST, OCC>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, CF::LDS);
      attr_set = true;
    }
    dim3 grid((unsigned)((n2 + 4 * CF::WCOLS - 1) / (4 * CF::WCOLS)), nfields);
    k_modeprod_dl<MT, NB4, NST, OCC><<<grid, 256, CF::LDS, c->stream>>>(A4, in, in_fs, out, out_fs, n, rotate,
                                                                        accumulate);
  };
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I1 = std::integral_constant<int, 1>;
  using I4 = std::integral_constant<int, 4>;
  using I8 = std::integral_constant<int, 8>;
  if constexpr (MT <= 13) {
    if (n % 2 == 0 && c->mp_variant <= 2) {
      if (c->mp_variant == 0)
        dl(I8{}, I3{}, I1{});  // 32 columns per wave, 3 stages, one workgroup per CU
      else if (c->mp_variant == 1)
        dl(I4{}, I2{}, I2{});  // 16 columns per wave, 2 stages, two workgroups per CU
      else
        dl(I4{}, I3{}, I1{});
      return;
    }
  }
  if (c->mp_variant == 16) {
    constexpr int NB = 2;
    dim3 grid((unsigned)((n2 + 64 * NB - 1) / (64 * NB)), nfields);
    if (n % 2 == 0)
      k_modeprod_f64<MT, NB, true><<<grid, 256, 0, c->stream>>>(A4, in, in_fs, out, out_fs, n, rotate, accumulate);
    else
      k_modeprod_f64<MT, NB, false><<<grid, 256, 0, c->stream>>>(A4, in, in_fs, out, out_fs, n, rotate, accumulate);
  } else {
    // 4 column groups per wave keep VGPR + AGPR at ~224 (two waves per SIMD,
    // so one workgroup's prologue / epilogue overlaps another's MFMAs);
    // 8 groups halve the LDS reads per MFMA at one wave per SIMD
    auto go = [&](auto nbtag) {
      constexpr int NB4 = decltype(nbtag)::value;
      dim3 grid((unsigned)((n2 + 16 * NB4 - 1) / (16 * NB4)), nfields);
      if (n % 2 == 0)
        k_modeprod_f64m4<MT, NB4, true><<<grid, 256, 0, c->stream>>>(A4, in, in_fs, out, out_fs, n, rotate, accumulate);
      else
        k_modeprod_f64m4<MT, NB4, false><<<grid, 256, 0, c->stream>>>(A4, in, in_fs, out, out_fs, n, rotate, accumulate);
    };
    if (c->mp_nb == 8 && MT <= 13) go(std::integral_constant<int, 8>{});
    else go(std::integral_constant<int, 4>{});
  }
}

void launch_modeprod(etd_ctx *c, const double *A4, const double *in, size_t in_fs, double *out, size_t out_fs,
                     int nfields, int rotate, int accumulate) {
  const int MT = (c->n + 15) / 16;
  etd_prof_begin(c, PC_MODEPROD);
  switch (MT) {
#define MTC(X) case X: modeprod_mt<X>(c, A4, in, in_fs, out, out_fs, nfields, rotate, accumulate); break;
    MTC(1) MTC(2) MTC(3) MTC(4) MTC(5) MTC(6) MTC(7) MTC(8) MTC(9) MTC(10) MTC(11) MTC(12) MTC(13) MTC(14)
    MTC(15) MTC(16) MTC(17)
#undef MTC
    default: break;
  }
  etd_prof_end(c, PC_MODEPROD);
}

// ------------------------------------------------------------------ phi scaling
// ac_updateU_new / ch_updateU_new (calc.c:489-535) and prepare_U2_new
// (553-567) on the rotated spectral layout l = i' + n j' + n^2 k'.  phi1 and
// phi2 (vars.c:302-359) are recomputed in registers from D instead of read
// from n^3 tables.
// One thread per spectral point, looping over the fields: every AC field
// shares H (and so phi), as do the CH fields, so exp and the divisions run
// once per point and kind instead of once per field.
// With outp2 (first order only) the same spectra also go to outp2 scaled for
// step a.dt2: one read of u-hat and N-hat for two step sizes, each output
// computed exactly as a launch of its own would.
template <typename T>
__global__ void k_phiscale(const T *uhat, const T *nhat, T *outp, size_t fs, int first_var, int nvars,
                           const double *DDX, const double *DDY, const double *DDZ, phi_args a, int order2,
                           T *outp2) {
  const int n = a.n;
  const size_t n3 = (size_t)n * n * n;
  const size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n3) return;
  int i, j, k;
  idx3(l, n, i, j, k);
  const int last = first_var + nvars;
  phi_pt pac, pch, pac2, pch2;
  if (first_var < a.nac) {
    const double tmp = a.kkz * DDZ[k] + a.kky * DDY[j] + a.kkx * DDX[i];
    pac = phi_at(-a.LE * (tmp * a.epn2 - a.KE), a.dt);
    if (outp2) pac2 = phi_at(-a.LE * (tmp * a.epn2 - a.KE), a.dt2);
  }
  if (last > a.nac) {
    const double tmp = DDZ[k] + DDY[j] + DDX[i];
    pch = phi_at(a.LCI * (tmp * tmp * a.epn2 - a.KCI * tmp), a.dt);
    if (outp2) pch2 = phi_at(a.LCI * (tmp * tmp * a.epn2 - a.KCI * tmp), a.dt2);
  }
  // fields in batches of FB whose loads are issued before any is used
  constexpr int FB = 4;
  for (int v0 = first_var; v0 < last; v0 += FB) {
    T uv[FB], nv[FB];
#pragma unroll
    for (int e = 0; e < FB; e++) {
      const int v = v0 + e < last ? v0 + e : last - 1;
      uv[e] = uhat[fs * (v - first_var) + l];
      if (!order2) nv[e] = nhat[fs * (v - first_var) + l];
    }
#pragma unroll
    for (int e = 0; e < FB; e++) {
      const int v = v0 + e;
      if (v >= last) break;
      const phi_pt &ph = v < a.nac ? pac : pch;
      T *o = outp + fs * (v - first_var);
      if (order2)
        o[l] = (T)(ph.phi2 * (double)uv[e]);
      else
        o[l] = (T)(ph.lin * (double)uv[e] + ph.phi1 * (double)nv[e]);
      if (outp2) {
        const phi_pt &p2 = v < a.nac ? pac2 : pch2;
        outp2[fs * (v - first_var) + l] = (T)(p2.lin * (double)uv[e] + p2.phi1 * (double)nv[e]);
      }
    }
  }
}

// out = (1 - phi1 H) uhat + phi1 nhat (or phi2 uhat), phi at the context's
// current dt; out may be uhat
template <typename T>
void launch_phiscale(etd_ctx *c, const T *uhat, const T *nhat, T *out, int first_var, int nvars, int order2) {
  const phi_args a = phi_args_of(c, c->p.dt, 0.0);
  etd_prof_begin(c, PC_PHI);
  k_phiscale<T><<<(unsigned)((c->n3 + 255) / 256), 256, 0, c->stream>>>(uhat, nhat, out, c->n3, first_var, nvars,
                                                                      c->DDX, c->DDY, c->DDZ, a, order2, nullptr);
  etd_prof_end(c, PC_PHI);
}

// first order over fields first_var .. first_var + nvars: out = phi-update
// at step dt, out2 at step dt2, from one read of uhat and nhat (the
// step-doubling pair)
template <typename T>
void launch_phiscale_pair(etd_ctx *c, const T *uhat, const T *nhat, T *out, double dt, T *out2, double dt2,
                          int first_var, int nvars) {
  const phi_args a = phi_args_of(c, dt, dt2);
  etd_prof_begin(c, PC_PHI);
  k_phiscale<T><<<(unsigned)((c->n3 + 255) / 256), 256, 0, c->stream>>>(uhat, nhat, out, c->n3, first_var, nvars,
                                                                      c->DDX, c->DDY, c->DDZ, a, 0, out2);
  etd_prof_end(c, PC_PHI);
}

// ------------------------------------------------------------------ nonlinear
struct nl_args {
  int n, g, N, nac, nch, aniso, elastic, ring;
  long long Lbase;  // linear index of AC field 0 of the latest elastic evaluation
  int slot0;        // Lbase mod ring
  double LE, KE, a2, a3, a4, epn2, hx, hy, hz, LCI, KCI, es;
  double ihx2, ihy2, ihz2;  // 1 / h^2, formed once on the host
  int phys;                 // physical boundaries on every side (periodic = 0, one box)
  int fe;                   // free energy: 0 Landau polynomial, 1 CALPHAD (alpha)
  double gbcc, ghcp;        // CALPHAD Gibbs energies (fe == 1)
  const unsigned char *mask;  // grain mask or NULL
  // decomposed run: px x py x pz boxes (P of them), this box at (bx, by, bz),
  // bs elements between the same field of consecutive boxes, and this box's
  // first interior cell in the global grid of edge N (the elastic grid)
  int P, px, py, pz, bx, by, bz;
  long long bs;
  int bxo, byo, bzo;
};

// -1 / 0 / +1: the extended coordinate e lies before, in or after the interior
__device__ __forceinline__ int box_shift(int e, int n, int g) { return e < g ? -1 : (e >= n - g ? 1 : 0); }
// element offset from this box's array to the same array of the box (sx, sy,
// sz) away, periodically; 0 for one box
__device__ __forceinline__ long long box_off(const nl_args &a, int sx, int sy, int sz) {
  if (a.P == 1) return 0;
  int x = a.bx + sx, y = a.by + sy, z = a.bz + sz;
  x += x < 0 ? a.px : (x >= a.px ? -a.px : 0);
  y += y < 0 ? a.py : (y >= a.py ? -a.py : 0);
  z += z < 0 ? a.pz : (z >= a.pz ? -a.pz : 0);
  return (long long)((z * a.py + y) * a.px + x - ((a.bz * a.py + a.by) * a.px + a.bx)) * a.bs;
}
// The periodic interior image of extended position (z, y, x), -2 <= z, y, x
// < n + 2, which the exchanged and relayed slabs hold (trans.c:7-925): in a
// decomposed run it lies in the box the position points into (the diagonal
// box for an edge or corner position).
template <typename T>
__device__ __forceinline__ double Pimg(const T *u, const nl_args &a, int z, int y, int x) {
  const int n = a.n, g = a.g;
  return u[box_off(a, box_shift(x, n, g), box_shift(y, n, g), box_shift(z, n, g)) +
           (long long)IDX(n, wrap_e(z, n, g), wrap_e(y, n, g), wrap_e(x, n, g))];
}
// the global interior coordinate of extended coordinate e of a box whose
// interior starts at o, periodic in a grid of edge Ng
__device__ __forceinline__ int gwrap(int e, int o, int g, int Ng) {
  const int X = o + e - g;
  return X < 0 ? X + Ng : (X >= Ng ? X - Ng : X);
}

// Elastic driving force of AC field p at extended cell (k,j,i).  The
// reference writes IFFT results into the interior of ONE buffer f and
// refreshes its ghosts with the same self-exchange before copying it to
// felas_p (elastic.c:664-687, 714-723), so ghost cells whose source is itself
// a ghost of f carry values of earlier variants: with d = number of ghost
// coordinates, the value is the interior of variant p - max(0, d-1) at the
// fully wrapped position (variants before 0 come from the previous call).
template <typename T>
__device__ __forceinline__ double felas_at(const T *E, const nl_args &a, int p, int k, int j, int i) {
  if (a.phys && (k < a.g || k >= a.n - a.g || j < a.g || j >= a.n - a.g || i < a.g || i >= a.n - a.g))
    return 0.0;  // physical sides: the force buffer is never exchanged, 0 outside the interior
  const int n = a.n, g = a.g, N = a.N;
  const int gk = (k < g || k >= n - g), gj = (j < g || j >= n - g), gi = (i < g || i >= n - g);
  int lag = gk + gj + gi - 1;
  lag = lag < 0 ? 0 : lag;
  // (Lbase + p - lag) mod ring, with p < nac <= ring - 2 and lag <= 2
  int slot = a.slot0 + p - lag;
  if (slot >= a.ring) slot -= a.ring;
  if (slot < 0) slot += a.ring;
  const int wk = gwrap(k, a.bzo, g, N), wj = gwrap(j, a.byo, g, N), wi = gwrap(i, a.bxo, g, N);
  return E[(size_t)slot * N * N * N + ((size_t)wk * N + wj) * N + wi];
}

// The field-independent part of felas_at: the ring lag of cell (k,j,i) and
// its offset within a ring slot; felas of AC field p is then
// E[felas_slot(a, p, lag) N^3 + off].
struct felas_cell {
  int lag;
  size_t off;
};
__device__ __forceinline__ felas_cell felas_locate(const nl_args &a, int k, int j, int i) {
  const int n = a.n, g = a.g, N = a.N;
  const int gk = (k < g || k >= n - g), gj = (j < g || j >= n - g), gi = (i < g || i >= n - g);
  int lag = gk + gj + gi - 1;
  felas_cell c;
  c.lag = lag < 0 ? 0 : lag;
  const int wk = gwrap(k, a.bzo, g, N), wj = gwrap(j, a.byo, g, N), wi = gwrap(i, a.bxo, g, N);
  c.off = ((size_t)wk * N + wj) * N + wi;
  return c;
}
__device__ __forceinline__ int felas_slot(const nl_args &a, int p, int lag) {
  int slot = a.slot0 + p - lag;
  if (slot >= a.ring) slot -= a.ring;
  if (slot < 0) slot += a.ring;
  return slot;
}

// ac_calc_FU (calc.c:8-68) -> ac_calc_F1 (anisotropic.c:9-161): all AC fields
// of one point per thread (sum u^2 couples them).  Fields go in batches of FB
// whose loads are issued before any is used (past the last field a batch
// re-reads it and ignores it); FB >= nac puts every field's loads of a pass in
// flight at once, which the FP32 fields need to keep HBM busy (a wave's load
// moves half the bytes of an FP64 one).  Sums in field order for any FB.
#define CAL_WMEGA 100.0    // vars.c:144
#define CAL_GNORMAL 50000.0  // ScLETD.h:27
#define CAL_ALPHA 0.02       // ScLETD.h:29
// mask_here: the grain-mask term is applied here (no gradient pass follows)
template <typename T, int FB>
__global__ __launch_bounds__(256) void k_nonlin_ac(const T *__restrict__ u, T *__restrict__ out,
                                                   const T *__restrict__ sub, const T *__restrict__ E, nl_args a,
                                                   int with_felas, int mask_here) {
  const int n = a.n, g = a.g;
  const unsigned n3 = (unsigned)n * n * n;  // n <= 272: every offset below fits 32 bits
  const size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n3) return;
  int i, j, k;
  idx3(l, n, i, j, k);
  const unsigned lo = (unsigned)l;
  // explicit outside coupling (the reference's calc_mu values): a cell lies on
  // at most one box face per axis (fx, fy, fz), and its image offset is the
  // same for every field; in a decomposed run the image lies in the
  // neighbouring, possibly diagonal, box
  const int wk = wrap_e(k, n, g), wj = wrap_e(j, n, g), wi = wrap_e(i, n, g);
  const int sk = box_shift(k, n, g), sj = box_shift(j, n, g), si = box_shift(i, n, g);
  const bool fx = !a.phys && (i == 0 || i == n - 1), fy = !a.phys && (j == 0 || j == n - 1),
             fz = !a.phys && (k == 0 || k == n - 1);  // no outside values on physical sides
  const long long ox = !fx ? 0 : i == 0 ? (long long)IDX(n, wk, wj, wrap_e(-1, n, g)) + box_off(a, -1, sj, sk)
                                       : (long long)IDX(n, wk, wj, wrap_e(n, n, g)) + box_off(a, 1, sj, sk);
  const long long oy = !fy ? 0 : j == 0 ? (long long)IDX(n, wk, wrap_e(-1, n, g), wi) + box_off(a, si, -1, sk)
                                       : (long long)IDX(n, wk, wrap_e(n, n, g), wi) + box_off(a, si, 1, sk);
  const long long oz = !fz ? 0 : k == 0 ? (long long)IDX(n, wrap_e(-1, n, g), wj, wi) + box_off(a, si, sj, -1)
                                       : (long long)IDX(n, wrap_e(n, n, g), wj, wi) + box_off(a, si, sj, 1);
  // physical sides: felas is 0 outside the interior
  const bool fel_zero = a.phys && (k < g || k >= n - g || j < g || j >= n - g || i < g || i >= n - g);
  const felas_cell fc = felas_locate(a, k, j, i);
  const unsigned N3 = (unsigned)a.N * a.N * a.N, foff = (unsigned)fc.off;
  const double ce = a.LE * a.epn2;
  // batch p0 of the fields: u, felas and the ETD2 difference, all loads issued
  // before any is used
  T upv[FB], fel[FB], sbv[FB];
  auto load_batch = [&](int p0) {
#pragma unroll
    for (int e = 0; e < FB; e++) {
      const unsigned p = (unsigned)(p0 + e < a.nac ? p0 + e : a.nac - 1);
      upv[e] = u[n3 * p + lo];
      if (with_felas) fel[e] = fel_zero ? (T)0 : E[(unsigned)felas_slot(a, (int)p, fc.lag) * N3 + foff];
      if (sub) sbv[e] = sub[n3 * p + lo];
    }
  };
  // one batch holds every field (nac <= FB): its loads serve both passes
  const bool one = a.nac <= FB;
  if (one) load_batch(0);
  double s = 0.0;
  for (int m0 = 0; m0 < a.nac; m0 += FB) {
    T um[FB];
#pragma unroll
    for (int e = 0; e < FB; e++) um[e] = one ? upv[e] : u[n3 * (unsigned)(m0 + e < a.nac ? m0 + e : a.nac - 1) + lo];
#pragma unroll
    for (int e = 0; e < FB; e++)
      if (m0 + e < a.nac) s = s + (double)um[e] * (double)um[e];
  }
  for (int p0 = 0; p0 < a.nac; p0 += FB) {
    if (!one) load_batch(p0);
#pragma unroll
    for (int e = 0; e < FB; e++) {
      const int p = p0 + e;
      if (p >= a.nac) break;
      const T *up_f = u + n3 * (unsigned)p;
      const double up = upv[e];
      double f = a.LE * a.KE * up;
      if (a.fe == 1) {
        // CALPHAD driving force (anisotropic_hip.c:2075-2140); sum_{m != p} u_m^2 = s - u_p^2
        const double dp = 30.0 * up * up * (1 - up) * (1 - up);
        const double dq = 4.0 * CAL_WMEGA * up * (1 - up) * (2 - up);
        const double dfeta = -(a.gbcc * dp) + (a.ghcp * dp) + dq;
        f -= a.LE * (dfeta / CAL_GNORMAL + 4.0 * CAL_ALPHA * (up * (s - up * up)));
      } else {
        const double fp = a.a2 * up - a.a3 * up * up + a.a4 * s * up;
        f -= a.LE * fp;
      }
      if (fx) f += ce * (double)up_f[ox] / a.hx / a.hx;
      if (fy) f += ce * (double)up_f[oy] / a.hy / a.hy;
      if (fz) f += ce * (double)up_f[oz] / a.hz / a.hz;
      if (with_felas) f -= a.LE * a.es * (double)fel[e];
      if (mask_here && a.mask && a.mask[lo] == 0) f -= 0.5 * up;
      if (sub) f -= sbv[e];
      out[n3 * (unsigned)p + lo] = (T)f;
    }
  }
}

// ac_calc_F2 (anisotropic.c:164-627), pass 1: G = Lambda (gx, gy, gz) with
// gy, gz taken as (u[-1] - u[+1]) / 2h; out-of-box neighbours read 0.
template <typename T>
__global__ void k_f2_grad(const T *__restrict__ u, T *__restrict__ G, const double *__restrict__ lam, nl_args a) {
  const int n = a.n;
  const size_t n3 = (size_t)n * n * n;
  const size_t l = xcd_block() * blockDim.x + threadIdx.x;  // neighbours +-n, +-n^2 on this XCD
  if (l >= n3) return;
  int i, j, k;
  idx3(l, n, i, j, k);
  // low neighbour of cell 0: 0 beside a neighbour, the mirror cell 1 on a physical side; high: 0
  const double a_left = i > 0 ? u[l - 1] : (a.phys ? (double)u[l + 1] : 0.0);
  const double a_right = i < n - 1 ? u[l + 1] : 0.0;
  const double a_top = j > 0 ? u[l - n] : (a.phys ? (double)u[l + n] : 0.0);
  const double a_bottom = j < n - 1 ? u[l + n] : 0.0;
  const double a_front = k > 0 ? u[l - (size_t)n * n] : (a.phys ? (double)u[l + (size_t)n * n] : 0.0);
  const double a_back = k < n - 1 ? u[l + (size_t)n * n] : 0.0;
  const double gx = (a_right - a_left) / (2.0 * a.hx);
  const double gy = (a_top - a_bottom) / (2.0 * a.hy);
  const double gz = (a_front - a_back) / (2.0 * a.hz);
  G[l] = (T)(lam[0] * gx + lam[1] * gy + lam[2] * gz);
  G[n3 + l] = (T)(lam[3] * gx + lam[4] * gy + lam[5] * gz);
  G[2 * n3 + l] = (T)(lam[6] * gx + lam[7] * gy + lam[8] * gz);
}

// pass 2: out -= epn2 LE div(G) (then the elastic term, keeping the
// reference's order f1 - epn2 LE F2 - LE ES felas)
template <typename T>
__global__ void k_f2_div(const T *__restrict__ G, T *__restrict__ out, const T *__restrict__ E, nl_args a, int p,
                         int with_felas, const T *__restrict__ u) {
  const int n = a.n;
  const size_t n3 = (size_t)n * n * n;
  const size_t l = xcd_block() * blockDim.x + threadIdx.x;  // neighbours +-n, +-n^2 on this XCD
  if (l >= n3) return;
  int i, j, k;
  idx3(l, n, i, j, k);
  const T *gx = G, *gy = G + n3, *gz = G + 2 * n3;
  const double a_left = i > 0 ? gx[l - 1] : (a.phys ? (double)gx[l + 1] : 0.0);
  const double a_right = i < n - 1 ? gx[l + 1] : 0.0;
  const double a_top = j > 0 ? gy[l - n] : (a.phys ? (double)gy[l + n] : 0.0);
  const double a_bottom = j < n - 1 ? gy[l + n] : 0.0;
  const double a_front = k > 0 ? gz[l - (size_t)n * n] : (a.phys ? (double)gz[l + (size_t)n * n] : 0.0);
  const double a_back = k < n - 1 ? gz[l + (size_t)n * n] : 0.0;
  double tmp = 0.0;
  tmp += (a_right - a_left) / 2.0 / a.hx;
  tmp += (a_top - a_bottom) / 2.0 / a.hy;
  tmp += (a_front - a_back) / 2.0 / a.hz;
  double f = out[l];
  if (a.fe == 1 && with_felas) f -= a.LE * a.es * felas_at(E, a, p, k, j, i);  // the alpha driver's order
  f -= a.epn2 * a.LE * tmp;
  if (a.fe != 1 && with_felas) f -= a.LE * a.es * felas_at(E, a, p, k, j, i);
  if (a.mask && a.mask[l] == 0) f -= 0.5 * (double)u[l];
  out[l] = (T)f;
}

// The beta->alpha driver's gradient term (anisotropic_calc_dev_2,
// anisotropic_hip.c:528-647): f2 = sum Lambda_ij d_i d_j u with three-point
// second differences and four-corner cross differences on every extended
// cell, neighbours outside the box read from the periodic image (the
// enlarged slabs); then, in that driver's order (ac_calc_F2_dev, 2190-2211
// for fe == 1; omega's order otherwise), the elastic term, the grain mask and
// the ETD2 difference.
template <typename T>
__global__ __launch_bounds__(256) void k_f2_full(const T *__restrict__ u, T *__restrict__ out,
                                                 const T *__restrict__ sub, const T *__restrict__ E,
                                                 const double *__restrict__ lam, nl_args a, int p, int with_felas) {
  const int n = a.n;
  const size_t n3 = (size_t)n * n * n;
  const size_t l = xcd_block() * blockDim.x + threadIdx.x;
  if (l >= n3) return;
  int i, j, k;
  idx3(l, n, i, j, k);
  const bool inner = i > 0 && i < n - 1 && j > 0 && j < n - 1 && k > 0 && k < n - 1;
  // cells inside the box read the field (its edge and corner ghosts are the ghost fill's, not pure
  // images); cells outside it the periodic image, as the enlarged slabs hold
  auto U = [&](int dz, int dy, int dx) -> double {
    if (inner) return u[l + ((long)dz * n + dy) * n + dx];
    int z = k + dz, y = j + dy, x = i + dx;
    if (a.phys) {  // physical sides mirror: n_left = i + 1 at i == 0, n_right = i - 1 at i == n - 1
      z = z < 0 ? -z : (z >= n ? 2 * (n - 1) - z : z);
      y = y < 0 ? -y : (y >= n ? 2 * (n - 1) - y : y);
      x = x < 0 ? -x : (x >= n ? 2 * (n - 1) - x : x);
    }
    if (z >= 0 && z < n && y >= 0 && y < n && x >= 0 && x < n) return u[IDX(n, z, y, x)];
    return Pimg(u, a, z, y, x);
  };
  const double m = U(0, 0, 0);
  const double fxx = (U(0, 0, 1) - 2 * m + U(0, 0, -1)) * a.ihx2;
  const double fyy = (U(0, 1, 0) - 2 * m + U(0, -1, 0)) * a.ihy2;
  const double fzz = (U(1, 0, 0) - 2 * m + U(-1, 0, 0)) * a.ihz2;
  const double fxy = ((U(0, 1, 1) - U(0, 1, -1)) - (U(0, -1, 1) - U(0, -1, -1))) / (4.0 * a.hx * a.hy);
  const double fxz = ((U(1, 0, 1) - U(1, 0, -1)) - (U(-1, 0, 1) - U(-1, 0, -1))) / (4.0 * a.hx * a.hz);
  const double fyz = ((U(1, 1, 0) - U(1, -1, 0)) - (U(-1, 1, 0) - U(-1, -1, 0))) / (4.0 * a.hy * a.hz);
  double t = 0.0;
  t += lam[0] * fxx;
  t += lam[1] * fxy;
  t += lam[2] * fxz;
  t += lam[3] * fxy;
  t += lam[4] * fyy;
  t += lam[5] * fyz;
  t += lam[6] * fxz;
  t += lam[7] * fyz;
  t += lam[8] * fzz;
  double f = out[l];
  if (a.fe == 1 && with_felas) f -= a.LE * a.es * felas_at(E, a, p, k, j, i);
  f -= a.epn2 * a.LE * t;
  if (a.fe != 1 && with_felas) f -= a.LE * a.es * felas_at(E, a, p, k, j, i);
  if (a.mask && a.mask[l] == 0) f -= 0.5 * (double)u[l];
  if (sub) f -= (double)sub[l];
  out[l] = (T)f;
}

__device__ __forceinline__ double ch_f(double cc) { return cc * (cc * cc - 1.0); }

// ch_calc_FU_ConstantMobility (calc.c:70-473) in one pass per CH field:
//   t   = -LCI KCI c + LCI CH_f(c), minus LCI epn2 c_out / h^2 on the faces
//   out = Lap_T t (out-of-box neighbours 0) - mu_out / h^2 on the faces
// The reference stores t and then takes its Laplacian; here each thread forms
// t at its 7 stencil points from c in FP64 (a few dozen FP64 operations a
// point, cheap next to a pass of t through HBM), so t is never rounded to
// the storage type.  1 / h^2 is applied as a product (one rounding away from
// the reference's two divisions).
struct ch_h {
  double x, y, z;  // 1 / h^2
};
// E, q: the elastic force ring and the CH field's index in it (ELASTIC = 2: t += LCI felas, calc.c:90-101)
template <typename T>
__device__ __forceinline__ double ch_t_at(const T *c, const T *E, int q, const nl_args &a, const ch_h &ih, int k, int j,
                                          int i) {
  const int n = a.n;
  const double cc = c[IDX(n, k, j, i)];
  double f = -a.LCI * a.KCI * cc;
  f += a.LCI * ch_f(cc);
  if (a.elastic == 2) f += a.LCI * felas_at(E, a, q, k, j, i);
  const double ce = a.LCI * a.epn2;
  if (a.phys) return f;  // no outside values on physical sides
  if (i == 0) f -= ce * Pimg(c, a, k, j, -1) * ih.x;
  if (i == n - 1) f -= ce * Pimg(c, a, k, j, n) * ih.x;
  if (j == 0) f -= ce * Pimg(c, a, k, -1, i) * ih.y;
  if (j == n - 1) f -= ce * Pimg(c, a, k, n, i) * ih.y;
  if (k == 0) f -= ce * Pimg(c, a, -1, j, i) * ih.z;
  if (k == n - 1) f -= ce * Pimg(c, a, n, j, i) * ih.z;
  return f;
}

// chemical potential LCI (epn2 Lap c - CH_f(c)) at an outside cell, from the
// periodic image (ch_mu, trans.c:1433-1937)
template <typename T>
__device__ double ch_mu_out(const T *c, const T *E, int q, const nl_args &a, const ch_h &ih, int z, int y, int x) {
  const int g = a.g;
  const double v0 = Pimg(c, a, z, y, x);
  double mu = (Pimg(c, a, z, y, x - 1) + (-2.0) * v0 + Pimg(c, a, z, y, x + 1)) * ih.x;
  mu += (Pimg(c, a, z, y - 1, x) + (-2.0) * v0 + Pimg(c, a, z, y + 1, x)) * ih.y;
  mu += (Pimg(c, a, z - 1, y, x) + (-2.0) * v0 + Pimg(c, a, z + 1, y, x)) * ih.z;
  mu = a.LCI * a.epn2 * mu;
  mu -= a.LCI * ch_f(v0);
  if (a.elastic == 2) {
    // ELASTIC = 2 (SWITCH_CH_FIELDMU, ScLETD.h:153-156): the force of the latest evaluation at the interior image
    const int N = a.N;
    const size_t off = ((size_t)gwrap(z, a.bzo, g, N) * N + gwrap(y, a.byo, g, N)) * N + gwrap(x, a.bxo, g, N);
    mu -= a.LCI * (double)E[(size_t)felas_slot(a, q, 0) * N * N * N + off];
  }
  return mu;
}

// out = N_c - sub (ETD2 difference, when sub != NULL) - nc (the mixed path's
// N(c_mean), when nc != NULL: the centred nonlinear term).
//
// Points whose whole stencil is at least one cell from the box faces (the
// inner (n - 4)^3) need t only pointwise; the others (the shell, two cells
// thick) need the face terms, the outside values and mu_out.  The grid keeps
// them apart so every wave runs one of the two paths: blocks [0, nbi8) take
// the inner points (x fastest, blocks spread over the XCDs in contiguous
// ranges so the +-n^2 neighbours share an L2), the rest the shell points in
// the order of shell_point.
__device__ __forceinline__ void shell_point(long s, int n, int &i, int &j, int &k) {
  const long n2 = (long)n * n, ring = 8L * n - 16;  // points of a middle plane's shell
  if (s < 4 * n2) {                                   // planes 0, 1, n - 2, n - 1
    const int p = (int)(s / n2);
    k = p < 2 ? p : n - 4 + p;
    const int r = (int)(s - p * n2);
    j = r / n;
    i = r - j * n;
    return;
  }
  s -= 4 * n2;
  const int p = (int)(s / ring);
  k = 2 + p;
  int r = (int)(s - p * ring);
  if (r < 4 * n) {  // rows 0, 1, n - 2, n - 1
    const int q = r / n;
    j = q < 2 ? q : n - 4 + q;
    i = r - q * n;
    return;
  }
  r -= 4 * n;
  j = 2 + r / 4;
  const int e = r & 3;
  i = e < 2 ? e : n - 4 + e;
}
__host__ __device__ __forceinline__ long shell_count(int n) {
  return 4L * n * n + (long)(n - 4) * (8L * n - 16);
}

template <typename T>
__global__ __launch_bounds__(256) void k_ch_nl(const T *__restrict__ c, T *__restrict__ out, const T *__restrict__ sub,
                                               const float *__restrict__ nc, const T *__restrict__ E, int q, nl_args a,
                                               unsigned nbi8) {
  const int n = a.n;
  const ch_h ih = {a.ihx2, a.ihy2, a.ihz2};
  const unsigned b = blockIdx.x;
  if (b < nbi8) {
    const int NI = n - 4;
    const size_t bb = (size_t)(b & 7u) * (nbi8 >> 3) + (b >> 3);
    const size_t li = bb * blockDim.x + threadIdx.x;
    if (li >= (size_t)NI * NI * NI) return;
    int i, j, k;
    idx3(li, NI, i, j, k);
    const size_t dz = (size_t)n * n, l = IDX(n, k + 2, j + 2, i + 2);
    double tc, r;
    if (a.elastic == 2) {  // t carries LCI felas (calc.c:90-101)
      auto te = [&](int dk, int dj, int di) { return ch_t_at(c, E, q, a, ih, k + 2 + dk, j + 2 + dj, i + 2 + di); };
      tc = te(0, 0, 0);
      r = (te(0, 0, -1) + te(0, 0, 1) - 2.0 * tc) * ih.x;
      r += (te(0, -1, 0) + te(0, 1, 0) - 2.0 * tc) * ih.y;
      r += (te(-1, 0, 0) + te(1, 0, 0) - 2.0 * tc) * ih.z;
    } else {
      auto tp = [&](double cc) { return -a.LCI * a.KCI * cc + a.LCI * ch_f(cc); };
      tc = tp(c[l]);
      r = (tp(c[l - 1]) + tp(c[l + 1]) - 2.0 * tc) * ih.x;
      r += (tp(c[l - n]) + tp(c[l + n]) - 2.0 * tc) * ih.y;
      r += (tp(c[l - dz]) + tp(c[l + dz]) - 2.0 * tc) * ih.z;
    }
    if (sub) r -= sub[l];
    if (nc) r -= nc[l];
    out[l] = (T)r;
    return;
  }
  const long s = (long)(b - nbi8) * blockDim.x + threadIdx.x;
  if (s >= shell_count(n)) return;
  int i, j, k;
  shell_point(s, n, i, j, k);
  const size_t l = IDX(n, k, j, i);
  const double tc = ch_t_at(c, E, q, a, ih, k, j, i);
  const double f_right = i < n - 1 ? ch_t_at(c, E, q, a, ih, k, j, i + 1) : 0.0;
  const double f_bottom = j < n - 1 ? ch_t_at(c, E, q, a, ih, k, j + 1, i) : 0.0;
  const double f_back = k < n - 1 ? ch_t_at(c, E, q, a, ih, k + 1, j, i) : 0.0;
  // low neighbour of cell 0: 0 beside a neighbour, the mirror (cell 1) on a physical side
  const double f_left = i > 0 ? ch_t_at(c, E, q, a, ih, k, j, i - 1) : (a.phys ? f_right : 0.0);
  const double f_top = j > 0 ? ch_t_at(c, E, q, a, ih, k, j - 1, i) : (a.phys ? f_bottom : 0.0);
  const double f_front = k > 0 ? ch_t_at(c, E, q, a, ih, k - 1, j, i) : (a.phys ? f_back : 0.0);
  double r = (f_left + f_right - 2.0 * tc) * ih.x;
  r += (f_top + f_bottom - 2.0 * tc) * ih.y;
  r += (f_front + f_back - 2.0 * tc) * ih.z;
  if (a.phys) {
    if (sub) r -= sub[l];
    if (nc) r -= nc[l];
    out[l] = (T)r;
    return;
  }
  if (i == 0) r -= ch_mu_out(c, E, q, a, ih, k, j, -1) * ih.x;
  if (i == n - 1) r -= ch_mu_out(c, E, q, a, ih, k, j, n) * ih.x;
  if (j == 0) r -= ch_mu_out(c, E, q, a, ih, k, -1, i) * ih.y;
  if (j == n - 1) r -= ch_mu_out(c, E, q, a, ih, k, n, i) * ih.y;
  if (k == 0) r -= ch_mu_out(c, E, q, a, ih, -1, j, i) * ih.z;
  if (k == n - 1) r -= ch_mu_out(c, E, q, a, ih, n, j, i) * ih.z;
  if (sub) r -= sub[l];
  if (nc) r -= nc[l];
  out[l] = (T)r;
}

template <typename T>
__global__ void k_sub(T *out, const T *sub, size_t cnt) {
  const size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l < cnt) out[l] = out[l] - sub[l];
}
template <typename T>
void launch_sub(etd_ctx *c, T *out, const T *sub, size_t cnt) {
  k_sub<T><<<(unsigned)((cnt + 255) / 256), 256, 0, c->stream>>>(out, sub, cnt);
}

static nl_args make_nl(etd_ctx *c) {
  nl_args a;
  a.n = c->n;
  a.g = c->g;
  a.N = c->Ne;  // the elastic grid: the felas ring holds its interior
  a.nac = c->nac;
  a.nch = c->nch;
  a.aniso = c->p.anisotropic;
  a.elastic = c->p.elastic;
  a.ring = c->ring > 0 ? c->ring : 1;
  a.Lbase = (c->elastic_evals - 1) * (long long)c->nel;
  a.slot0 = (int)(((a.Lbase % a.ring) + a.ring) % a.ring);
  a.LE = c->p.LE;
  a.KE = c->p.KE;
  a.a2 = c->p.a2;
  a.a3 = c->p.a3;
  a.a4 = c->p.a4;
  a.epn2 = c->p.epn2;
  a.hx = c->p.hx;
  a.hy = c->p.hy;
  a.hz = c->p.hz;
  a.LCI = c->p.LCI;
  a.KCI = c->p.KCI;
  a.es = c->p.elastic_scale;
  a.ihx2 = 1.0 / (a.hx * a.hx);
  a.ihy2 = 1.0 / (a.hy * a.hy);
  a.ihz2 = 1.0 / (a.hz * a.hz);
  a.phys = !c->p.periodic;
  a.fe = c->p.free_energy;
  a.gbcc = c->gbcc;
  a.ghcp = c->ghcp;
  a.mask = c->mask;
  a.P = c->P;
  a.px = c->px;
  a.py = c->py;
  a.pz = c->pz;
  a.bx = c->box % c->px;
  a.by = (c->box / c->px) % c->py;
  a.bz = c->box / (c->px * c->py);
  a.bs = (long long)c->bs;
  a.bxo = a.bx * c->N;
  a.byo = a.by * c->N;
  a.bzo = a.bz * c->N;
  return a;
}

// ------------------------------------------------------------------ variable-mobility CH
// ch_calc_FU_VariableMobility (mobility.c:838-858) with update_M_C
// (860-922), one periodic box, CH fields q = 0, 1, 2 (Cu, Mn, Ni):
//   M_q = x_q (1 - x_q) (D_qA (1 - eta_0) + D_qG eta_0), C_q = max_box M_q
//   f0  = -LCI KCI c + LCI CH_f(c)
//   ft1 = -f0 + KCI c,            + epn2 c_out / h^2 on the faces
//   ft2 = epn2 Lap_T c - f0,      + epn2 c_out / h^2 on the faces
//   N   = LCI (-C Lap_T ft1 - C mu_out / h^2 (faces) - div((M - C) grad ft2))
// Lap_T and the divergence read 0 outside the box.  Three passes: the box
// maxima of M (order-free, so exact), ft1 and ft2 into scratch, and N; M at
// the stencil points is recomputed from eta_0 instead of being stored.
#define VM_BLOCKS 256
struct vm_args {
  double X[3], DA[3], DG[3];
};
__device__ __forceinline__ double vm_M(const vm_args &m, int q, double eta) {
  return m.X[q] * (m.DA[q] * (1.0 - eta) + m.DG[q] * eta);
}

template <typename T>
__global__ __launch_bounds__(256) void k_vm_max(const T *__restrict__ eta, size_t n3, vm_args m,
                                                double *__restrict__ part) {
  double mx[3] = {-1.0e30, -1.0e30, -1.0e30};
  for (size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x; l < n3; l += (size_t)gridDim.x * blockDim.x) {
    const double e = eta[l];
#pragma unroll
    for (int q = 0; q < 3; q++) mx[q] = fmax(mx[q], vm_M(m, q, e));
  }
  __shared__ double sh[3][4];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const double v = wave_max(mx[q]);
    if ((threadIdx.x & 63) == 0) sh[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int q = threadIdx.x;
    part[q * VM_BLOCKS + blockIdx.x] = fmax(fmax(sh[q][0], sh[q][1]), fmax(sh[q][2], sh[q][3]));
  }
}
// C_q = the maximum of the block maxima, into part[3 VM_BLOCKS + q]
__global__ __launch_bounds__(256) void k_vm_cmax(double *__restrict__ part) {
  double v = -1.0e30;
  for (int b = threadIdx.x; b < VM_BLOCKS; b += blockDim.x) v = fmax(v, part[blockIdx.x * VM_BLOCKS + b]);
  v = wave_max(v);
  __shared__ double sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) part[3 * VM_BLOCKS + blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

template <typename T>
__device__ __forceinline__ double lap_zero_at(const T *a, int n, size_t l, int k, int j, int i, const nl_args &q) {
  const size_t dz = (size_t)n * n;
  const double ac = a[l];
  const double ar = i < n - 1 ? (double)a[l + 1] : 0.0, ab = j < n - 1 ? (double)a[l + n] : 0.0;
  const double ak = k < n - 1 ? (double)a[l + dz] : 0.0;
  const double al = i > 0 ? (double)a[l - 1] : (q.phys ? ar : 0.0);
  const double at = j > 0 ? (double)a[l - n] : (q.phys ? ab : 0.0);
  const double af = k > 0 ? (double)a[l - dz] : (q.phys ? ak : 0.0);
  double t = 0.0;
  t += (al + ar - 2.0 * ac) * q.ihx2;
  t += (at + ab - 2.0 * ac) * q.ihy2;
  t += (af + ak - 2.0 * ac) * q.ihz2;
  return t;
}

template <typename T>
__global__ __launch_bounds__(256) void k_vm_ft(const T *__restrict__ c, T *__restrict__ ft1, T *__restrict__ ft2,
                                               nl_args a) {
  const int n = a.n;
  const size_t n3 = (size_t)n * n * n;
  const size_t l = xcd_block() * blockDim.x + threadIdx.x;
  if (l >= n3) return;
  int i, j, k;
  idx3(l, n, i, j, k);
  const double cc = c[l];
  const double f0 = -a.LCI * a.KCI * cc + a.LCI * ch_f(cc);
  double t1 = -f0 + a.KCI * cc;
  double t2 = a.epn2 * lap_zero_at(c, n, l, k, j, i, a) - f0;
  double fc = 0.0;  // epn2 c_out / h^2 on the faces, left, right, top, bottom, front, back
  if (a.phys) {
    ft1[l] = (T)t1;
    ft2[l] = (T)t2;
    return;
  }
  if (i == 0) fc += a.epn2 * Pimg(c, a, k, j, -1) * a.ihx2;
  if (i == n - 1) fc += a.epn2 * Pimg(c, a, k, j, n) * a.ihx2;
  if (j == 0) fc += a.epn2 * Pimg(c, a, k, -1, i) * a.ihy2;
  if (j == n - 1) fc += a.epn2 * Pimg(c, a, k, n, i) * a.ihy2;
  if (k == 0) fc += a.epn2 * Pimg(c, a, -1, j, i) * a.ihz2;
  if (k == n - 1) fc += a.epn2 * Pimg(c, a, n, j, i) * a.ihz2;
  ft1[l] = (T)(t1 + fc);
  ft2[l] = (T)(t2 + fc);
}

template <typename T>
__global__ __launch_bounds__(256) void k_vm_out(const T *__restrict__ c, const T *__restrict__ eta,
                                                const T *__restrict__ ft1, const T *__restrict__ ft2,
                                                T *__restrict__ out, const T *__restrict__ sub,
                                                const double *__restrict__ cmax, vm_args m, int q, nl_args a,
                                                const T *__restrict__ E) {
  const int n = a.n;
  const size_t n3 = (size_t)n * n * n;
  const size_t l = xcd_block() * blockDim.x + threadIdx.x;
  if (l >= n3) return;
  int i, j, k;
  idx3(l, n, i, j, k);
  const double C = cmax[q];
  const ch_h ih = {a.ihx2, a.ihy2, a.ihz2};
  double f1 = -C * lap_zero_at(ft1, n, l, k, j, i, a);
  if (!a.phys) {
    if (i == 0) f1 -= C * ch_mu_out(c, E, q, a, ih, k, j, -1) * ih.x;
    if (i == n - 1) f1 -= C * ch_mu_out(c, E, q, a, ih, k, j, n) * ih.x;
    if (j == 0) f1 -= C * ch_mu_out(c, E, q, a, ih, k, -1, i) * ih.y;
    if (j == n - 1) f1 -= C * ch_mu_out(c, E, q, a, ih, k, n, i) * ih.y;
    if (k == 0) f1 -= C * ch_mu_out(c, E, q, a, ih, -1, j, i) * ih.z;
    if (k == n - 1) f1 -= C * ch_mu_out(c, E, q, a, ih, n, j, i) * ih.z;
  }
  // div((M - C) grad ft2), mobility.c:435-722: (m_nb + m_mid)(a_nb - a_mid) / 2h^2, 0 outside the box
  const size_t dz = (size_t)n * n;
  const double am = ft2[l], mm = vm_M(m, q, (double)eta[l]) - C;
  auto term = [&](bool in, size_t o, double ih2) {
    const double an = in ? (double)ft2[o] : 0.0, mn = in ? vm_M(m, q, (double)eta[o]) - C : 0.0;
    return (mn + mm) * (an - am) * 0.5 * ih2;
  };
  // the low neighbour of cell 0 on a physical side is the mirror cell 1
  double d = 0.0;
  d += term(i > 0 || a.phys, i > 0 ? l - 1 : l + 1, a.ihx2);
  d += term(i < n - 1, l + 1, a.ihx2);
  d += term(j > 0 || a.phys, j > 0 ? l - n : l + n, a.ihy2);
  d += term(j < n - 1, l + n, a.ihy2);
  d += term(k > 0 || a.phys, k > 0 ? l - dz : l + dz, a.ihz2);
  d += term(k < n - 1, l + dz, a.ihz2);
  double r = a.LCI * (f1 + -d);
  if (sub) r -= (double)sub[l];
  out[l] = (T)r;
}

template <typename T>
static void launch_vm(etd_ctx *c, const T *u, T *out, const T *sub, T *scratch, const nl_args &a) {
  vm_args m;
  for (int q = 0; q < 3; q++) m.X[q] = c->vm_X[q], m.DA[q] = c->vm_DA[q], m.DG[q] = c->vm_DG[q];
  const size_t n3 = c->n3;
  const unsigned nb = round8((n3 + 255) / 256);
  k_vm_max<T><<<VM_BLOCKS, 256, 0, c->stream>>>(u, n3, m, c->vm_part);  // eta_0 = AC field 0
  k_vm_cmax<<<3, 256, 0, c->stream>>>(c->vm_part);
  for (int q = 0; q < 3; q++) {
    const T *cf = u + n3 * (c->nac + q);
    k_vm_ft<T><<<nb, 256, 0, c->stream>>>(cf, scratch, scratch + n3, a);
    k_vm_out<T><<<nb, 256, 0, c->stream>>>(cf, u, scratch, scratch + n3, out + n3 * (c->nac + q),
                                           sub ? sub + n3 * (c->nac + q) : nullptr, c->vm_part + 3 * VM_BLOCKS, m,
                                           q, a, fields<T>(c).E);
  }
}

template <typename T>
static void launch_ch_nl(etd_ctx *c, const T *cf, T *out, const T *sub, const float *nc, int q, const nl_args &a) {
  const long NI = c->n - 4;
  const unsigned nbi8 = round8((size_t)((NI * NI * NI + 255) / 256));
  const unsigned nbs = (unsigned)((shell_count(c->n) + 255) / 256);
  k_ch_nl<T><<<nbi8 + nbs, 256, 0, c->stream>>>(cf, out, sub, nc, fields<T>(c).E, q, a, nbi8);
}

// N for every variable into out (V x n3); when sub != NULL the result is
// N - sub (the ETD2 difference prepare_U1_new, calc.c:537-551).  ch_centre
// (mixed path): the CH outputs are N - N(c_mean), the centred term.
template <typename T>
void launch_nonlinear(etd_ctx *c, T *out, const T *sub, int ch_centre) {
  nl_args a = make_nl(c);
  const etd_fields<T> F = fields<T>(c);
  const size_t n3 = c->n3;
  const unsigned nb = (unsigned)((n3 + 255) / 256);
  const int felas = c->p.elastic == 1 && c->nac > 0;
  etd_prof_begin(c, PC_NONLIN);
  if (c->nac > 0) {
    const int aniso = c->p.anisotropic;
    // with F2 the felas and ETD2 subtraction move to the divergence pass
    const T *sb = aniso ? nullptr : sub;
    const int fl = felas && !aniso;
    // FP32: all fields' loads at once; FP64 loads move twice the bytes, and
    // batches of 4 keep 8 waves per SIMD
    int fb = c->nl_batch;
    if (!fb) fb = sizeof(T) == 8 ? 4 : (c->nac <= 4 ? 4 : c->nac <= 8 ? 8 : c->nac <= 12 ? 12 : 16);
    const int mh = aniso ? 0 : 1;  // the grain mask goes with the last pass over the field
    if (fb == 4)
      k_nonlin_ac<T, 4><<<nb, 256, 0, c->stream>>>(F.u, out, sb, F.E, a, fl, mh);
    else if (fb == 8)
      k_nonlin_ac<T, 8><<<nb, 256, 0, c->stream>>>(F.u, out, sb, F.E, a, fl, mh);
    else if (fb == 12)
      k_nonlin_ac<T, 12><<<nb, 256, 0, c->stream>>>(F.u, out, sb, F.E, a, fl, mh);
    else
      k_nonlin_ac<T, 16><<<nb, 256, 0, c->stream>>>(F.u, out, sb, F.E, a, fl, mh);
    if (aniso == 1)
      for (int p = 0; p < c->nac; p++) {
        k_f2_grad<T><<<round8(nb), 256, 0, c->stream>>>(F.u + n3 * p, F.gtmp, c->lam + 9 * p, a);
        k_f2_div<T><<<round8(nb), 256, 0, c->stream>>>(F.gtmp, out + n3 * p, F.E, a, p, felas, F.u + n3 * p);
        if (sub) launch_sub(c, out + n3 * p, sub + n3 * p, n3);  // after the F2 / felas terms
      }
    if (aniso == 2)
      for (int p = 0; p < c->nac; p++)
        k_f2_full<T><<<round8(nb), 256, 0, c->stream>>>(F.u + n3 * p, out + n3 * p, sub ? sub + n3 * p : nullptr,
                                                        F.E, c->lam + 9 * p, a, p, felas);
  }
  if (c->p.variable_mobility)
    launch_vm<T>(c, F.u, out, sub, F.gtmp, a);
  else
    for (int q = 0; q < c->nch; q++) {
      const int v = c->nac + q;
      const float *nc = ch_centre && sizeof(T) == 4 && c->ch_const ? c->ch_const + n3 * q : nullptr;
      launch_ch_nl<T>(c, F.u + n3 * v, out + n3 * v, sub ? sub + n3 * v : nullptr, nc, q, a);
    }
  etd_prof_end(c, PC_NONLIN);
}


// ------------------------------------------------------------------ elastic
// Own FFT: each workgroup transforms LINES lines of length N held in LDS with
// mixed-radix (2,3,4,5) Stockham passes:
//   pass (radix R, p = product of earlier radices), i in [0, N/R):
//   k = i mod p, a_t = x[i + t N/R] w^(t k N/(pR)), y = DFT_R(a),
//   x'[(i-k) R + k + t p] = y_t
// Forward uses w = exp(-2 pi i/N); the inverse conjugates (scaling 1/N^3 is
// applied once when results leave the transform).
#define FFT_LINES 8
struct fft_plan {
  int N, npass;
  int radix[16];
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by -i (sgn = +1, forward) or +i (sgn = -1, inverse)
__device__ __forceinline__ double2 cmi(double2 a, int sgn) {
  return sgn > 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

template <int R, int SGN>
__device__ __forceinline__ void dft_r(double2 *a) {
  if constexpr (R == 2) {
    double2 y0 = cadd(a[0], a[1]), y1 = csub(a[0], a[1]);
    a[0] = y0;
    a[1] = y1;
  } else if constexpr (R == 4) {
    double2 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
    double2 s13 = cadd(a[1], a[3]), d13 = cmi(csub(a[1], a[3]), SGN);
    a[0] = cadd(s02, s13);
    a[1] = cadd(d02, d13);
    a[2] = csub(s02, s13);
    a[3] = csub(d02, d13);
  } else if constexpr (R == 3) {
    const double h = 0.86602540378443864676;  // sin(2 pi / 3)
    double2 s = cadd(a[1], a[2]), d = csub(a[1], a[2]);
    double2 y0 = cadd(a[0], s);
    double2 t = make_double2(a[0].x - 0.5 * s.x, a[0].y - 0.5 * s.y);
    double2 id = cmi(make_double2(h * d.x, h * d.y), SGN);  // -i h d (forward)
    a[0] = y0;
    a[1] = cadd(t, id);
    a[2] = csub(t, id);
  } else {  // 5
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
    const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
    double2 b1 = cadd(a[1], a[4]), d1 = csub(a[1], a[4]);
    double2 b2 = cadd(a[2], a[3]), d2 = csub(a[2], a[3]);
    double2 y0 = cadd(a[0], cadd(b1, b2));
    double2 t1 = make_double2(a[0].x + c1 * b1.x + c2 * b2.x, a[0].y + c1 * b1.y + c2 * b2.y);
    double2 t2 = make_double2(a[0].x + c2 * b1.x + c1 * b2.x, a[0].y + c2 * b1.y + c1 * b2.y);
    // forward: y1 = t1 - i (s1 d1 + s2 d2), y4 = t1 + i(...); y2 = t2 - i (s2 d1 - s1 d2), y3 = t2 + i(...)
    double2 e1 = cmi(make_double2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y), SGN);
    double2 e2 = cmi(make_double2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y), SGN);
    a[0] = y0;
    a[1] = cadd(t1, e1);
    a[4] = csub(t1, e1);
    a[2] = cadd(t2, e2);
    a[3] = csub(t2, e2);
  }
}

// one Stockham pass of radix R over the FFT_LINES lines held in LDS
template <int R, int SGN>
__device__ __forceinline__ void fft_pass(const double2 *__restrict__ src, double2 *__restrict__ dst, int N, int p,
                                         const double2 *__restrict__ tw) {
  const unsigned M = (unsigned)(N / R);
  const unsigned stw = (unsigned)(N / (p * R));
  const unsigned up = (unsigned)p;
  for (unsigned b = threadIdx.x; b < FFT_LINES * M; b += blockDim.x) {
    const unsigned ln = b / M, i = b - ln * M;
    const unsigned k = i % up;
    double2 a[R];
    const double2 *s = src + ln * (unsigned)N;
#pragma unroll
    for (int t = 0; t < R; t++) {
      double2 x = s[i + t * M];
      if (t > 0 && k > 0) {
        double2 w = tw[t * k * stw];  // t k stw < N
        if (SGN < 0) w.y = -w.y;
        x = cmul(x, w);
      }
      a[t] = x;
    }
    dft_r<R, SGN>(a);
    double2 *o = dst + ln * (unsigned)N;
    const unsigned base = (i - k) * R + k;
#pragma unroll
    for (int t = 0; t < R; t++) o[base + t * up] = a[t];
  }
}

// T2 = double2 or float2 in HBM; the lines are transformed in FP64 in LDS
template <int SGN, typename T2>
__global__ __launch_bounds__(256) void k_fft_axis(T2 *data, size_t fs, int axis, fft_plan pl,
                                                  const double2 *__restrict__ tw) {
  extern __shared__ double2 sbuf[];
  const int N = pl.N;
  double2 *buf0 = sbuf, *buf1 = sbuf + FFT_LINES * N;
  T2 *d = data + fs * blockIdx.y;
  const unsigned uN = (unsigned)N, N2 = uN * uN;
  const unsigned line0 = blockIdx.x * FFT_LINES;  // first line of the group
  const unsigned tot = FFT_LINES * uN;
  // line L = x + N * o: axis 0 runs along x of row L; axis 1 along y at
  // (x, z = o); axis 2 along z at (x, y = o)
  for (unsigned e = threadIdx.x; e < tot; e += blockDim.x) {
    unsigned ln, t;
    size_t addr;
    if (axis == 0) {
      t = e % uN;
      ln = e / uN;
      addr = (size_t)(line0 + ln) * uN + t;
    } else {
      ln = e % FFT_LINES;
      t = e / FFT_LINES;
      const unsigned L = line0 + ln, x = L % uN, o = L / uN;
      addr = axis == 1 ? x + (size_t)uN * t + (size_t)N2 * o : x + (size_t)uN * o + (size_t)N2 * t;
    }
    double2 v = make_double2(0.0, 0.0);
    if (line0 + ln < N2) {
      const T2 x = d[addr];
      v = make_double2(x.x, x.y);
    }
    buf0[ln * uN + t] = v;
  }
  __syncthreads();
  double2 *src = buf0, *dst = buf1;
  int p = 1;
  for (int ps = 0; ps < pl.npass; ps++) {
    const int R = pl.radix[ps];
    switch (R) {
      case 4: fft_pass<4, SGN>(src, dst, N, p, tw); break;
      case 2: fft_pass<2, SGN>(src, dst, N, p, tw); break;
      case 3: fft_pass<3, SGN>(src, dst, N, p, tw); break;
      default: fft_pass<5, SGN>(src, dst, N, p, tw); break;
    }
    __syncthreads();
    double2 *tmp = src;
    src = dst;
    dst = tmp;
    p *= R;
  }
  for (unsigned e = threadIdx.x; e < tot; e += blockDim.x) {
    unsigned ln, t;
    size_t addr;
    if (axis == 0) {
      t = e % uN;
      ln = e / uN;
      addr = (size_t)(line0 + ln) * uN + t;
    } else {
      ln = e % FFT_LINES;
      t = e / FFT_LINES;
      const unsigned L = line0 + ln, x = L % uN, o = L / uN;
      addr = axis == 1 ? x + (size_t)uN * t + (size_t)N2 * o : x + (size_t)uN * o + (size_t)N2 * t;
    }
    if (line0 + ln < N2) {
      const double2 v = src[ln * uN + t];
      T2 x;
      x.x = v.x;
      x.y = v.y;
      d[addr] = x;
    }
  }
}

// ---- two-stage line FFT for N = 16 R2 (R2 in 2..16 with factors 2, 3, 4, 5)
// x[n], n = n1 + 16 n2; Y[n1][k2] = sum_n2 x[n1 + 16 n2] W_R2^(n2 k2),
// times W_N^(n1 k2); X[k2 + R2 k1] = sum_n1 Y[n1][k2] W_16^(n1 k1).  Both
// stages are DFTs in registers (composite sizes by one more split); the only
// LDS trip is the transpose between them, stored [line][k2][n1] with a
// one-element pad per k2 row and per line so that neither stage conflicts.
// A workgroup transforms 16 lines with 16 threads each; lanes run along the
// line for the x axis and across lines (consecutive x) for y and z, so every
// global access covers whole 128-byte segments.
template <int R, int SGN>
__device__ __forceinline__ void dft_any(double2 *a);

template <int A, int B, int SGN>
__device__ __forceinline__ void dft_split(double2 *x) {  // size A B, in place, natural order
  double2 y[A][B];
#pragma unroll
  for (int n1 = 0; n1 < A; n1++) {
    double2 t[B];
#pragma unroll
    for (int n2 = 0; n2 < B; n2++) t[n2] = x[n1 + A * n2];
    dft_any<B, SGN>(t);
#pragma unroll
    for (int k2 = 0; k2 < B; k2++) {
      if (n1 * k2 == 0) {
        y[n1][k2] = t[k2];
      } else {
        const double ang = 2.0 * 3.14159265358979323846 * (double)(n1 * k2) / (double)(A * B);
        y[n1][k2] = cmul(t[k2], make_double2(cos(ang), -SGN * sin(ang)));
      }
    }
  }
#pragma unroll
  for (int k2 = 0; k2 < B; k2++) {
    double2 s[A];
#pragma unroll
    for (int n1 = 0; n1 < A; n1++) s[n1] = y[n1][k2];
    dft_any<A, SGN>(s);
#pragma unroll
    for (int k1 = 0; k1 < A; k1++) x[k2 + B * k1] = s[k1];
  }
}

template <int R, int SGN>
__device__ __forceinline__ void dft_any(double2 *a) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2 || R == 3 || R == 4 || R == 5) {
    dft_r<R, SGN>(a);
  } else if constexpr (R == 16) {
    dft_split<4, 4, SGN>(a);
  } else if constexpr (R % 4 == 0) {
    dft_split<4, R / 4, SGN>(a);
  } else if constexpr (R % 3 == 0) {
    dft_split<3, R / 3, SGN>(a);
  } else if constexpr (R % 2 == 0) {
    dft_split<2, R / 2, SGN>(a);
  } else if constexpr (R % 5 == 0) {
    dft_split<5, R / 5, SGN>(a);
  }
}

template <int R2>
struct fft16_cfg {
  static constexpr int N = 16 * R2;
  static constexpr int ROW = 16 + 1;          // one k2 row of Y (16 n1 values + pad)
  static constexpr int LINE = R2 * ROW + 1;   // one line (+ pad)
  static constexpr int LDS = 16 * LINE * 16;  // bytes (double2)
};

// IO = 0: in place on data.  IO = 1 (forward x pass of an elastic
// evaluation): the input is read from the fields, eta_2r + i eta_2r+1 on the
// interior of the extended layout (the pack).  IO = 2 (inverse z pass): the
// output goes to the felas ring, real and imaginary parts to the slots of
// fields 2r and 2r+1, times 1/N^3 (the unpack); each value is rounded as the
// separate pack / unpack kernels round it.
struct fft_io {
  const void *u;
  void *E;
  int n, g, nac, ring, slot0;
};

template <int SGN, int R2, typename T2, int IO = 0>
__global__ __launch_bounds__(256) void k_fft16(T2 *data, size_t fs, int axis, const double2 *__restrict__ tw,
                                               fft_io io = fft_io{}) {
  using T = decltype(T2{}.x);
  using CF = fft16_cfg<R2>;
  constexpr int N = CF::N;
  // the transpose between the stages holds FP32 pairs for FP32 data (one
  // rounding of the intermediate, within the FP32 transform's accuracy; half
  // the LDS, so twice the workgroups per CU) and FP64 pairs for FP64 data
  using LT = typename std::conditional<sizeof(T2) == 8, float2, double2>::type;
  __shared__ LT buf[16 * CF::LINE];
  T2 *d = data + fs * blockIdx.y;
  const int tid = threadIdx.x;
  const int ll = axis == 0 ? tid >> 4 : tid & 15;  // line within the group
  const int j = axis == 0 ? tid & 15 : tid >> 4;   // thread within the line
  // The FP32 unpack writes each real field's row segment of 16 x as 64 bytes,
  // half a 128-byte line whose other half the next group of 16 lines writes.
  // Workgroups go to the 8 XCDs in turn and each XCD has its own L2, so there
  // each XCD takes a contiguous range of line groups (gridDim.x = N^2 / 16 is
  // a multiple of 8): both halves of a line meet in one L2.
  unsigned bx = blockIdx.x;
  if constexpr (IO == 2 && sizeof(T2) == 8) bx = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  const unsigned L = bx * 16 + ll;                  // N^2 % 16 == 0: every line exists
  const unsigned x = L % N, o = L / N;
  const size_t base = axis == 0 ? (size_t)L * N : (axis == 1 ? x + (size_t)N * N * o : x + (size_t)N * o);
  const size_t stride = axis == 0 ? 1 : (axis == 1 ? (size_t)N : (size_t)N * N);
  LT *bl = buf + ll * CF::LINE;
  {  // stage 1: thread n1 = j, DFT over n2 of x[n1 + 16 n2], twiddle, to LDS [k2][n1]
    const int n1 = j;
    double2 a[R2];
    if constexpr (IO == 1) {
      // axis 0, pack: the group's 16 x rows of fields 2r and 2r+1 go to LDS
      // with consecutive threads on consecutive x (coalesced; the rows of the
      // extended layout start two ghost cells in), padded so that the strided
      // reads below take no extra bank passes
      constexpr int SP = N + R2;  // fits: the buffer holds 16 (17 R2 + 1) pairs
      static_assert(16 * SP * sizeof(T2) <= sizeof(buf), "staging fits the transpose buffer");
      T2 *stg = reinterpret_cast<T2 *>(buf);
      const T *u = (const T *)io.u;
      const size_t n3 = (size_t)io.n * io.n * io.n;
      const int pa = 2 * (int)blockIdx.y, pb = pa + 1;
      // 16 N / 256 = R2 elements per thread, all loads issued before any
      // store (the loop is unrolled)
      T va[R2], vb[R2];
#pragma unroll
      for (int k = 0; k < R2; k++) {
        const int e = tid + 256 * k, l2 = e / N, xx = e - l2 * N;
        const unsigned Lg = bx * 16 + l2;
        const size_t src = IDX(io.n, (int)(Lg / N) + io.g, (int)(Lg % N) + io.g, xx + io.g);
        va[k] = u[n3 * pa + src];
        vb[k] = pb < io.nac ? u[n3 * pb + src] : (T)0;
      }
#pragma unroll
      for (int k = 0; k < R2; k++) {
        const int e = tid + 256 * k, l2 = e / N, xx = e - l2 * N;
        T2 v;
        v.x = va[k];
        v.y = vb[k];
        stg[l2 * SP + xx] = v;
      }
      __syncthreads();
#pragma unroll
      for (int n2 = 0; n2 < R2; n2++) {
        const T2 v = stg[ll * SP + n1 + 16 * n2];
        a[n2] = make_double2(v.x, v.y);
      }
      __syncthreads();  // the transpose below reuses the buffer
    } else {
#pragma unroll
      for (int n2 = 0; n2 < R2; n2++) {
        const T2 v = d[base + (size_t)(n1 + 16 * n2) * stride];
        a[n2] = make_double2(v.x, v.y);
      }
    }
    dft_any<R2, SGN>(a);
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) {
      double2 w = tw[(n1 * k2) % N];  // exp(-2 pi i n1 k2 / N)
      if (SGN < 0) w.y = -w.y;
      const double2 y = (n1 * k2 == 0) ? a[k2] : cmul(a[k2], w);
      LT v;
      v.x = y.x;
      v.y = y.y;
      bl[k2 * CF::ROW + n1] = v;
    }
  }
  __syncthreads();
  if (j < R2) {  // stage 2: thread k2 = j, DFT over n1, out X[k2 + R2 k1]
    const int k2 = j;
    double2 b[16];
#pragma unroll
    for (int n1 = 0; n1 < 16; n1++) {
      const LT v = bl[k2 * CF::ROW + n1];
      b[n1] = make_double2(v.x, v.y);
    }
    dft_any<16, SGN>(b);
#pragma unroll
    for (int k1 = 0; k1 < 16; k1++) {
      T2 v;
      v.x = b[k1].x;
      v.y = b[k1].y;
      const size_t m = base + (size_t)(k2 + R2 * k1) * stride;
      if constexpr (IO == 2) {
        T *E = (T *)io.E;
        const size_t N3 = (size_t)N * N * N;
        const double inv = 1.0 / ((double)N * N * N);
        const int r = (int)blockIdx.y;
        int sa = io.slot0 + 2 * r, sb = sa + 1;  // (slot0 + p) mod ring, p < nac < ring
        if (sa >= io.ring) sa -= io.ring;
        if (sb >= io.ring) sb -= io.ring;
        E[(size_t)sa * N3 + m] = (T)((double)v.x * inv);
        if (2 * r + 1 < io.nac) E[(size_t)sb * N3 + m] = (T)((double)v.y * inv);
      } else {
        d[m] = v;
      }
    }
  }
}

// N = 16 R2 with R2 one of the instantiated register DFT lengths below
bool fft16_supported(int N) {
  if (N % 16) return false;
  switch (N / 16) {
    case 2: case 3: case 4: case 5: case 6: case 8: case 9: case 10: case 12: case 15: case 16:
      return true;
    default:
      return false;
  }
}

template <int SGN, typename T2, int IO = 0>
static bool launch_fft16(etd_ctx *c, T2 *data, int npair, int axis, fft_io io = fft_io{}) {
  const int N = c->Ne;
  if (!fft16_supported(N)) return false;
  dim3 grid((unsigned)((size_t)N * N / 16), npair);
  const double2 *tw = (const double2 *)c->tw;
  switch (N / 16) {
#define F16(R)                                                                                   \
  case R:                                                                                        \
    k_fft16<SGN, R, T2, IO><<<grid, 256, 0, c->stream>>>(data, c->Ne3, axis, tw, io);            \
    return true;
    F16(2) F16(3) F16(4) F16(5) F16(6) F16(8) F16(9) F16(10) F16(12) F16(15) F16(16)
#undef F16
    default: return false;
  }
}

template <typename T>
void fft3d(etd_ctx *c, T *data, int npair, int sgn) {
  using T2 = typename cplx<T>::t;
  fft_plan pl;
  pl.N = c->Ne;
  pl.npass = c->fft_npass;
  for (int i = 0; i < 16; i++) pl.radix[i] = c->fft_radix[i];
  const size_t N2 = (size_t)c->Ne * c->Ne;
  etd_prof_begin(c, PC_FFT);
  dim3 grid((unsigned)((N2 + FFT_LINES - 1) / FFT_LINES), npair);
  size_t lds = 2 * FFT_LINES * (size_t)c->Ne * sizeof(double2);
  for (int ax = 0; ax < 3; ax++) {
    if (c->fft_variant == 1 && (sgn > 0 ? launch_fft16<1, T2>(c, (T2 *)data, npair, ax)
                                        : launch_fft16<-1, T2>(c, (T2 *)data, npair, ax)))
      continue;
    if (sgn > 0)
      k_fft_axis<1, T2><<<grid, 256, lds, c->stream>>>((T2 *)data, c->Ne3, ax, pl, (const double2 *)c->tw);
    else
      k_fft_axis<-1, T2><<<grid, 256, lds, c->stream>>>((T2 *)data, c->Ne3, ax, pl, (const double2 *)c->tw);
  }
  etd_prof_end(c, PC_FFT);
}

// pack interior eta_2r + i eta_2r+1
template <typename T>
__global__ void k_el_pack(const T *__restrict__ u, typename cplx<T>::t *__restrict__ Z, int n, int g, int N, int nac) {
  const size_t N3 = (size_t)N * N * N;
  const size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= N3) return;
  const int r = blockIdx.y;
  int x, y, z;
  idx3(m, N, x, y, z);
  const size_t n3 = (size_t)n * n * n;
  const size_t src = IDX(n, z + g, y + g, x + g);
  const int pa = 2 * r, pb = 2 * r + 1;
  typename cplx<T>::t zz;
  zz.x = u[n3 * pa + src];
  zz.y = pb < nac ? u[n3 * pb + src] : (T)0;
  Z[N3 * r + m] = zz;
}

// the same over a decomposed run's boxes (px per axis, fields of box b at
// b bs): global cell (x, y, z) of the (px N)^3 grid is interior cell
// (x % N, y % N, z % N) of box (x / N, y / N, z / N)
template <typename T>
__global__ void k_el_pack_boxes(const T *__restrict__ u, typename cplx<T>::t *__restrict__ Z, int n, int g, int N,
                                int nac, int px, long long bs) {
  const int Ng = px * N;
  const size_t N3 = (size_t)Ng * Ng * Ng;
  const size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= N3) return;
  const int r = blockIdx.y;
  int x, y, z;
  idx3(m, Ng, x, y, z);
  const size_t n3 = (size_t)n * n * n;
  const int b = ((z / N) * px + y / N) * px + x / N;
  const long long src = b * bs + (long long)IDX(n, z % N + g, y % N + g, x % N + g);
  const int pa = 2 * r, pb = 2 * r + 1;
  typename cplx<T>::t zz;
  zz.x = u[n3 * pa + src];
  zz.y = pb < nac ? u[n3 * pb + src] : (T)0;
  Z[N3 * r + m] = zz;
}

// k-space contraction acc_p = sum_q B_pq(n) F_q with B computed on the fly
// (elastic_calculate_BN, elastic.c:460-605: n from signed wavenumbers
// g < N/2 ? g : g - N, k = 0 uses n = (0,1,0); Omega = (C n n)^-1 by the
// adjugate as inverse_3x3).  The reference keeps Re IFFT(B F)
// (fft3d.c:133-136).  That signed-wavenumber rule is not odd under k -> -k
// on the Nyquist planes (and next to them for odd N), so B F is not
// Hermitian there; Re IFFT(X) = IFFT((X(k) + conj X(-k)) / 2), so the
// contraction uses the Hermitian part B~ = (B(k) + B(-k)) / 2 wherever the
// mirror's signed wavenumber is not the negative, and the two packed
// variants of a pair then come back exactly as the real and imaginary parts.
#define EL_MAXV 16
__device__ __forceinline__ int signed_wave(int x, int N) { return x < N / 2 ? x : x - N; }

__device__ void b_geometry(int sx, int sy, int sz, const etd_dev_consts *__restrict__ K, double *nv, double *om) {
  if (sx == 0 && sy == 0 && sz == 0) {
    nv[0] = 0.0;
    nv[1] = 1.0;
    nv[2] = 0.0;
  } else {
    const double n1 = sx, n2 = sy, n3 = sz;
    // one division per point (a sign flip of n flips nv exactly, so the
    // geometry at -k stays the exact mirror of the one at k)
    const double ir = 1.0 / sqrt(n1 * n1 + n2 * n2 + n3 * n3);
    nv[0] = n1 * ir;
    nv[1] = n2 * ir;
    nv[2] = n3 * ir;
  }
  double oi[9];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) {
      double s = 0.0;
      for (int j = 0; j < 3; ++j)
        for (int l = 0; l < 3; ++l) s += K->C4[((i * 3 + j) * 3 + k) * 3 + l] * nv[j] * nv[l];
      oi[i * 3 + k] = s;
    }
  double nr = 0.0;
  nr += oi[0] * (oi[4] * oi[8] - oi[5] * oi[7]);
  nr -= oi[1] * (oi[3] * oi[8] - oi[5] * oi[6]);
  nr += oi[2] * (oi[3] * oi[7] - oi[4] * oi[6]);
  const double inr = 1.0 / nr;
  om[0] = (oi[4] * oi[8] - oi[5] * oi[7]) * inr;
  om[1] = -(oi[3] * oi[8] - oi[6] * oi[5]) * inr;
  om[2] = (oi[3] * oi[7] - oi[6] * oi[4]) * inr;
  om[3] = -(oi[1] * oi[8] - oi[7] * oi[2]) * inr;
  om[4] = (oi[0] * oi[8] - oi[2] * oi[6]) * inr;
  om[5] = -(oi[0] * oi[7] - oi[6] * oi[1]) * inr;
  om[6] = (oi[1] * oi[5] - oi[4] * oi[2]) * inr;
  om[7] = -(oi[0] * oi[5] - oi[3] * oi[2]) * inr;
  om[8] = (oi[0] * oi[4] - oi[1] * oi[3]) * inr;
}

// w = (sigma_p^T n) Omega for variant p
__device__ __forceinline__ void b_wrow(const double *sp, const double *nv, const double *om, double *w) {
  double av[3];
  for (int j = 0; j < 3; j++) av[j] = nv[0] * sp[0 * 3 + j] + nv[1] * sp[1 * 3 + j] + nv[2] * sp[2 * 3 + j];
  for (int k = 0; k < 3; k++) w[k] = av[0] * om[0 * 3 + k] + av[1] * om[1 * 3 + k] + av[2] * om[2 * 3 + k];
}

// One k-point of the contraction for NAC variants held in registers.
// b_q = sigma_q n, c_p = Omega b_p, B_pq = t0_pq - c_p . b_q (equal to the
// reference's n_i s_p,ij Omega_jk s_q,kl n_l for symmetric Omega and
// sigma); with a symmetric Voigt matrix B is symmetric and each unordered
// pair is formed once.  MIRROR: B is replaced by (B(k) + B(-k)) / 2.
// WRITE_NEG (with MIRROR false): also write the point -k, whose values are
// the conjugates: F_q(-k) = conj F_q(k) for real fields and B(-k) = B(k) (the
// geometry at -n is an exact sign flip of the one at n, squared away), so
// acc_p(-k) = conj acc_p(k) bit for bit and Y_r(-k) = (a_x + b_y, b_x - a_y)
// with a = acc_2r, b = acc_2r+1.
// Arithmetic in the storage precision: FP64 for FP64 fields, FP32 for the
// mixed path's FP32 spectra (half the registers: 4 waves per SIMD instead of
// 2 at NAC = 12); the geometry Omega(n) is always formed in FP64.
template <int NAC, bool MIRROR, typename T2, bool WRITE_NEG = false>
__device__ __forceinline__ void contract_point(const T2 *__restrict__ Z, T2 *__restrict__ Y, int N,
                                               int npair, size_t m, const etd_dev_consts *__restrict__ K) {
  const size_t N3 = (size_t)N * N * N;
  int x, y, z;
  idx3(m, N, x, y, z);
  const int xm = (N - x) % N, ym = (N - y) % N, zm = (N - z) % N;
  const size_t mm = ((size_t)zm * N + ym) * N + xm;
  using CT = typename std::conditional<sizeof(T2) == 8, float, double>::type;
  struct C2 {
    CT x, y;
  };
  C2 F[NAC];
#pragma unroll
  for (int r = 0; r < (NAC + 1) / 2; r++) {
    const T2 za = Z[N3 * r + m], zb = Z[N3 * r + mm];
    const C2 a = {(CT)za.x, (CT)za.y}, b = {(CT)zb.x, (CT)zb.y};
    // F_2r = (Z(m) + conj Z(-m)) / 2, F_2r+1 = (Z(m) - conj Z(-m)) / 2i
    F[2 * r] = C2{(CT)0.5 * (a.x + b.x), (CT)0.5 * (a.y - b.y)};
    if (2 * r + 1 < NAC) F[2 * r + 1] = C2{(CT)0.5 * (a.y + b.y), (CT)-0.5 * (a.x - b.x)};
  }
  CT nv[3], om[9];
  {
    double nvd[3], omd[9];
    b_geometry(signed_wave(x, N), signed_wave(y, N), signed_wave(z, N), K, nvd, omd);
#pragma unroll
    for (int k = 0; k < 3; k++) nv[k] = (CT)nvd[k];
#pragma unroll
    for (int k = 0; k < 9; k++) om[k] = (CT)omd[k];
  }
  // b_q = sigma_q n for every variant; c_p = b_p Omega only for the row p
  // being summed (3 live values instead of 3 NAC: fewer VGPRs, more waves)
  CT bv[NAC][3];
#pragma unroll
  for (int q = 0; q < NAC; q++) {
    const double *sg = K->sig + 9 * q;
#pragma unroll
    for (int k = 0; k < 3; k++)
      bv[q][k] = (CT)sg[k * 3 + 0] * nv[0] + (CT)sg[k * 3 + 1] * nv[1] + (CT)sg[k * 3 + 2] * nv[2];
  }
  CT bm[MIRROR ? NAC : 1][3], cm[MIRROR ? NAC : 1][3];
  if constexpr (MIRROR) {
    CT nv2[3], om2[9];
    {
      double nvd[3], omd[9];
      b_geometry(signed_wave(xm, N), signed_wave(ym, N), signed_wave(zm, N), K, nvd, omd);
#pragma unroll
      for (int k = 0; k < 3; k++) nv2[k] = (CT)nvd[k];
#pragma unroll
      for (int k = 0; k < 9; k++) om2[k] = (CT)omd[k];
    }
#pragma unroll
    for (int q = 0; q < NAC; q++) {
      const double *sg = K->sig + 9 * q;
#pragma unroll
      for (int k = 0; k < 3; k++)
        bm[q][k] = (CT)sg[k * 3 + 0] * nv2[0] + (CT)sg[k * 3 + 1] * nv2[1] + (CT)sg[k * 3 + 2] * nv2[2];
#pragma unroll
      for (int k = 0; k < 3; k++)
        cm[q][k] = bm[q][0] * om2[0 * 3 + k] + bm[q][1] * om2[1 * 3 + k] + bm[q][2] * om2[2 * 3 + k];
    }
  }
  C2 acc[NAC];
#pragma unroll
  for (int p = 0; p < NAC; p++) acc[p] = C2{(CT)0, (CT)0};
#pragma unroll
  for (int p = 0; p < NAC; p++) {
    CT cvp[3];
#pragma unroll
    for (int k = 0; k < 3; k++) cvp[k] = bv[p][0] * om[0 * 3 + k] + bv[p][1] * om[1 * 3 + k] + bv[p][2] * om[2 * 3 + k];
#pragma unroll
    for (int q = p; q < NAC; q++) {
      const CT t0 = (CT)K->t0[p * ETD_MAX_VAR + q];
      CT B = t0 - (cvp[0] * bv[q][0] + cvp[1] * bv[q][1] + cvp[2] * bv[q][2]);
      if constexpr (MIRROR)
        B = (CT)0.5 * (B + (t0 - (cm[p][0] * bm[q][0] + cm[p][1] * bm[q][1] + cm[p][2] * bm[q][2])));
      acc[p].x += F[q].x * B;
      acc[p].y += F[q].y * B;
      if (q != p) {
        acc[q].x += F[p].x * B;
        acc[q].y += F[p].y * B;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < (NAC + 1) / 2; r++) {
    T2 y;
    if (2 * r + 1 < NAC) {  // Y_r = acc_2r + i acc_2r+1
      y.x = acc[2 * r].x - acc[2 * r + 1].y;
      y.y = acc[2 * r].y + acc[2 * r + 1].x;
    } else {
      y.x = acc[2 * r].x;
      y.y = acc[2 * r].y;
    }
    Y[N3 * r + m] = y;
    if constexpr (WRITE_NEG) {
      if (mm != m) {
        T2 w;
        if (2 * r + 1 < NAC) {
          w.x = acc[2 * r].x + acc[2 * r + 1].y;
          w.y = acc[2 * r + 1].x - acc[2 * r].y;
        } else {
          w.x = acc[2 * r].x;
          w.y = -acc[2 * r].y;
        }
        Y[N3 * r + mm] = w;
      }
    }
  }
}

// One thread per pair {k, -k}: planes z = 0 .. N/2 hold a member of every
// pair; the point with the smaller index computes both (within the planes
// z = 0 and z = N/2, which map to themselves, half the threads return).
template <int NAC, typename T2>
__global__ __launch_bounds__(256) void k_el_contract_t(const T2 *__restrict__ Z, T2 *__restrict__ Y, int N,
                                                       int npair, const etd_dev_consts *__restrict__ K) {
  const size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= (size_t)N * N * (N / 2 + 1)) return;
  int x, y, z;
  idx3(m, N, x, y, z);
  const int xm = (N - x) % N, ym = (N - y) % N, zm = (N - z) % N;
  if (((size_t)zm * N + ym) * N + xm < m) return;
  contract_point<NAC, false, T2, true>(Z, Y, N, npair, m, K);
}

// points whose mirror's signed wavenumber is not the negative: the index
// planes c = N/2 (and c = N/2 + 1 for odd N) in each direction; one thread
// per (plane, position), duplicates across the planes write equal values
template <int NAC, typename T2>
__global__ __launch_bounds__(256) void k_el_contract_mirror(const T2 *__restrict__ Z, T2 *__restrict__ Y,
                                                            int N, int npair, const etd_dev_consts *__restrict__ K) {
  const int nplane = (N & 1) ? 2 : 1;
  const size_t N2 = (size_t)N * N;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * nplane * N2) return;
  const int axis = (int)(t / (nplane * N2));
  const int which = (int)((t / N2) % nplane);
  const size_t r = t % N2;
  const int c = N / 2 + which;
  const int a = (int)(r % N), b = (int)(r / N);
  int x, y, z;
  if (axis == 0) { x = c; y = a; z = b; }
  else if (axis == 1) { x = a; y = c; z = b; }
  else { x = a; y = b; z = c; }
  contract_point<NAC, true, T2>(Z, Y, N, npair, ((size_t)z * N + y) * N + x, K);
}

template <typename T>
__global__ void k_el_unpack(const typename cplx<T>::t *__restrict__ Y, T *__restrict__ E, int N, int nac, int ring,
                            int slot0) {
  const size_t N3 = (size_t)N * N * N;
  const size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= N3) return;
  const int r = blockIdx.y;
  const double inv = 1.0 / ((double)N * N * N);
  const typename cplx<T>::t v = Y[N3 * r + m];
  // slots (slot0 + p) mod ring, p < nac < ring
  int sa = slot0 + 2 * r, sb = sa + 1;
  if (sa >= ring) sa -= ring;
  if (sb >= ring) sb -= ring;
  E[(size_t)sa * N3 + m] = (T)((double)v.x * inv);
  if (2 * r + 1 < nac) E[(size_t)sb * N3 + m] = (T)((double)v.y * inv);
}

template <typename T>
void launch_elastic(etd_ctx *c) {
  using T2 = typename cplx<T>::t;
  const etd_fields<T> F = fields<T>(c);
  const int N = c->Ne;  // the global interior of a decomposed run (fftMPI's distributed transform)
  const size_t N3 = c->Ne3;
  const unsigned nb = (unsigned)((N3 + 255) / 256);
  etd_prof_begin(c, PC_ELASTIC);
  const long long Lbase = c->elastic_evals * (long long)c->nel;
  const T *uel = F.u + c->n3 * c->el_v0;  // the fields the force acts on (AC or CH)
  const int slot0 = (int)(Lbase % c->ring);
  // two-stage FFT: the pack is the forward x pass's input and the unpack the
  // inverse z pass's output
  // (the fused passes exist only where the two-stage kernel does: N = 16 R2
  // with an instantiated R2; elsewhere pack, radix passes and unpack)
  const bool fused = c->fft_variant == 1 && fft16_supported(N) && c->P == 1;
  fft_io io{uel, F.E, c->n, c->g, c->nel, c->ring, slot0};
  if (fused) {
    etd_prof_begin(c, PC_FFT);
    launch_fft16<1, T2, 1>(c, (T2 *)F.Z, c->npair, 0, io);
    launch_fft16<1, T2>(c, (T2 *)F.Z, c->npair, 1);
    launch_fft16<1, T2>(c, (T2 *)F.Z, c->npair, 2);
    etd_prof_end(c, PC_FFT);
  } else {
    if (c->P == 1)
      k_el_pack<T><<<dim3(nb, c->npair), 256, 0, c->stream>>>(uel, (T2 *)F.Z, c->n, c->g, N, c->nel);
    else
      k_el_pack_boxes<T><<<dim3(nb, c->npair), 256, 0, c->stream>>>(uel, (T2 *)F.Z, c->n, c->g, c->N, c->nel, c->px,
                                                                     (long long)c->bs);
    fft3d<T>(c, F.Z, c->npair, +1);
  }
  {
    const int nplane = (N & 1) ? 2 : 1;
    const unsigned nbm = (unsigned)((3 * nplane * (size_t)N * N + 255) / 256);
    const unsigned nbh = (unsigned)(((size_t)N * N * (N / 2 + 1) + 255) / 256);
    const T2 *Zc = (const T2 *)F.Z;
    T2 *Yc = (T2 *)F.Y;
    switch (c->nel) {
#define NACC(X)                                                                                    case X:                                                                                            k_el_contract_t<X, T2><<<nbh, 256, 0, c->stream>>>(Zc, Yc, N, c->npair, c->dconst);                  k_el_contract_mirror<X, T2><<<nbm, 256, 0, c->stream>>>(Zc, Yc, N, c->npair, c->dconst);            break;
      NACC(1) NACC(2) NACC(3) NACC(4) NACC(5) NACC(6) NACC(7) NACC(8) NACC(9) NACC(10) NACC(11) NACC(12) NACC(13)
      NACC(14) NACC(15) NACC(16)
#undef NACC
      default: break;
    }
  }
  if (fused) {
    etd_prof_begin(c, PC_FFT);
    launch_fft16<-1, T2>(c, (T2 *)F.Y, c->npair, 0);
    launch_fft16<-1, T2>(c, (T2 *)F.Y, c->npair, 1);
    launch_fft16<-1, T2, 2>(c, (T2 *)F.Y, c->npair, 2, io);
    etd_prof_end(c, PC_FFT);
  } else {
    fft3d<T>(c, F.Y, c->npair, -1);
    k_el_unpack<T><<<dim3(nb, c->npair), 256, 0, c->stream>>>((const T2 *)F.Y, F.E, N, c->nel, c->ring, slot0);
  }
  c->elastic_evals++;
  etd_prof_end(c, PC_ELASTIC);
}

template <typename T>
__global__ void k_felas_mat(T *felas, const T *E, nl_args a) {
  const int n = a.n;
  const size_t n3 = (size_t)n * n * n;
  const size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n3) return;
  const int p = blockIdx.y;
  int i, j, k;
  idx3(l, n, i, j, k);
  felas[n3 * p + l] = (T)felas_at(E, a, p, k, j, i);
}

template <typename T>
void launch_felas_materialize(etd_ctx *c, T *felas) {
  nl_args a = make_nl(c);
  k_felas_mat<T><<<dim3((unsigned)((c->n3 + 255) / 256), c->nel), 256, 0, c->stream>>>(felas, fields<T>(c).E, a);
}

// ------------------------------------------------------------------ energy
// check_soln_new (field.c:407-678) in one kernel that reads every field
// once: the statistics (max, min, sum of u + 1 over [g, n - g)^3) and the
// energy sums (cells [g, n - g - 1)^3) as FP64 partial sums per block,
// added up in a fixed order by diag_reduce (host) or k_diag_finish (device).
struct diag_args {
  int n, g, nac, nch;
  int lo;  // first cell of the sums: g, or 0 when the low side is physical (ix2 = 0, vars.c:58-157)
  double hx, hy, hz, epn2, a2, a3, a4;
  double ihx, ihy, ihz;  // 1 / h
};


// One block per (row j, chunk of planes), threads along i, 63 cells per wave
// (lane 63 only supplies the i + 1 corners of lane 62).  The work of a block
// is a list of tasks (field, sub-range of E = 8 planes): the AC
// fields sub-range by sub-range (all AC fields of a sub-range in turn), then
// each CH field's sub-ranges.  A task loads rows j and j + 1 of the E + 1
// planes it needs into registers; the loads of task t + 1 are issued before
// task t is computed (two register buffers, used in turn).  Every row of
// every field is read from HBM once per chunk, plus one shared plane per
// sub-range; the i + 1 corners come by lane shifts.
//
// Per plane p the thread forms, from its four corners a = (j, i),
// a' = (j, i + 1), b = (j + 1, i), b' = (j + 1, i + 1):
//   Dx = (a' - a) + (b' - b),  Dy = (b - a) + (b' - a'),  P = (a + a') + (b + b')
// and a cell between planes p and p + 1 has the corner sums S = P_p + P_{p+1},
// SX = Dx_p + Dx_{p+1}, SY = Dy_p + Dy_{p+1}, SZ = P_{p+1} - P_p: mean S / 8,
// gradient (SX / 4hx, SY / 4hy, SZ / 4hz).  These are the reference's
// eight-corner sums (field.c:466-486) regrouped; the gradient energy
// epn2/2 grad.Lambda grad is the quadratic form of (SX, SY, SZ) with the
// symmetric matrix Q = epn2/2 D (Lambda + Lambda^T)/2 D, D = diag(1/4h), and
// the Landau sums take S^2 = 64 mean^2 and S^3 = 512 mean^3 (exact scalings).
// About 26 FP64 operations per AC cell and field instead of about 40: the
// kernel is bound by FP64 issue, not bytes.  The regrouping changes each
// cell's terms by a few roundings.  The per-cell arithmetic runs on every
// lane; the lanes that hold no cell or point (lane 63, lanes past the row)
// are masked once per task, not per cell.
//
// Energy: kind 0 the AC gradient energy plus the Landau terms of the cell
// means (S2, S3 of each plane wait in registers while the AC fields of the
// sub-range pass), kind 1 + q the CH field's gradient + double-well energy,
// kind 1 + nch the couple term (gradient energy of the AC fields q < nch and
// of the CH fields, from the same loads).  Statistics: a point is counted by
// the block whose row j is its row (the last row, n - g - 1, by the last row
// block as its row j + 1), the chunk holding its plane (plane n - g - 1 by
// the last chunk) and one lane; sum(u + 1) is kept as sum(u) plus the count.
// Each task's statistics are wave-reduced into a per-wave slot in LDS, the
// slots combined at the end.
//
// Partials: b = row ECHUNK + chunk of nb = m ECHUNK; statistics at
// part[(v nb + b) 3 + {max, min, sum}], energies at part[3 V nb + kind nb + b].
// The 1-D grid is numbered so that each XCD (workgroups go to the 8 XCDs in
// turn) runs a contiguous range of rows: row j + 1, read by blocks j and
// j + 1, stays in one L2.
#define ECHUNK 8
template <typename T>
__global__ __launch_bounds__(320) void k_diag(const T *__restrict__ u, const double *__restrict__ lam,
                                              double *__restrict__ part, diag_args a) {
  constexpr int E = 8;  // planes per sub-range: 152-192 VGPRs, 2-3 waves per SIMD (12 and 6 measured slower)
  extern __shared__ double wst[];             // [wave][field][max, min, sum]
  const int n = a.n, g = a.g, lo = a.lo, m = n - g - 1 - lo, V = a.nac + a.nch, nac = a.nac, nch = a.nch;
  const int per = (int)gridDim.x / 8;  // gridDim.x = m ECHUNK, a multiple of 8
  const int w = ((int)blockIdx.x % 8) * per + (int)blockIdx.x / 8;
  const int row = w % m, chunk = w / m;
  const int nb = m * ECHUNK, bi = row * ECHUNK + chunk;
  const int j = lo + row;
  const int kb = lo + (int)(((long)m * chunk) / ECHUNK), ke = lo + (int)(((long)m * (chunk + 1)) / ECHUNK);
  const int nsub = (ke - kb + E - 1) / E;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (int)blockDim.x >> 6;
  const int i = lo + 63 * wv + lane;
  const bool cell = lane < 63 && i < lo + m;
  const bool st_lane = i < n - g && (lane < 63 || wv == nw - 1);  // each i of [lo, n - g) once
  const bool st_row1 = row == m - 1;                              // also count row j + 1 = n - g - 1
  const bool last_chunk = chunk == ECHUNK - 1;
  const size_t n2 = (size_t)n * n, n3 = n2 * n;
  const size_t r0 = (size_t)j * n + (i < n ? i : 0);
  const double he = 0.5 * a.epn2, dx = 0.25 * a.ihx, dy = 0.25 * a.ihy, dz = 0.25 * a.ihz;
  const double wx = he * dx * dx, wy = he * dy * dy, wz = he * dz * dz;  // |grad|^2 weights of SX^2, SY^2, SZ^2
  for (int t = threadIdx.x; t < nw * V; t += blockDim.x) {
    wst[3 * t + 0] = -1.0e30;
    wst[3 * t + 1] = 1.0e30;
    wst[3 * t + 2] = 0.0;
  }
  __syncthreads();
  __shared__ double sh[5];  // up to 5 waves: m <= 315 (contexts have n <= 272)
  auto block_sum_out = [&](double v, int kind) {
    v = wave_sum(cell ? v : 0.0);
    __syncthreads();
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
      for (int q = 0; q < nw; q++) s += sh[q];
      part[(size_t)3 * V * nb + (size_t)kind * nb + bi] = s;
    }
  };
  // task (v, s): field v, sub-range s (planes kb + E s ..); AC fields first, all of a sub-range in turn, then
  // the CH fields one by one
  struct tk {
    int v, s;
  };
  auto next = [&](tk x) {
    if (x.v < nac) {
      if (++x.v == nac) {
        x.v = 0;
        if (++x.s == nsub) x.v = nac, x.s = 0;
      }
    } else if (++x.s == nsub) {
      x.s = 0;
      ++x.v;
    }
    return x;
  };
  const int nt = V * nsub;
  // every lane loads (lanes past the box read column 0: their values feed no counted cell or point)
  auto load = [&](tk x, T (&x0)[E + 1], T (&x1)[E + 1]) {
    const int k0 = kb + E * x.s, last = ke - k0;
    const T *f0 = u + n3 * x.v + (size_t)k0 * n2 + r0, *f1 = f0 + n;
#pragma unroll
    for (int p = 0; p <= E; p++) {
      const size_t o = (size_t)(p < last ? p : last) * n2;
      x0[p] = f0[o];
      x1[p] = f1[o];
    }
  };
  const bool couple = nac > 0 && nch > 0;
  double e = 0.0, ec = 0.0, sq = 0.0;
  double s2[E], s3[E];
#pragma unroll
  for (int p = 0; p < E; p++) s2[p] = s3[p] = 0.0;
  auto process = [&](tk x, const T (&x0)[E + 1], const T (&x1)[E + 1]) {
    const int v = x.v, k0 = kb + E * x.s;
    const int np = ke - k0 < E ? ke - k0 : E;
    const bool ac = v < nac, cp = couple && (v >= nac || v < nch);
    // the planes whose points this block counts: [k0, k0 + np), and plane ke in the last chunk's last sub-range
    const int nst = np + ((last_chunk && k0 + np == ke) ? 1 : 0);
    double Qxx = 0, Qyy = 0, Qzz = 0, Qxy = 0, Qxz = 0, Qyz = 0;
    if (ac) {
      const double *L = lam + 9 * v;
      Qxx = he * dx * dx * L[0];
      Qyy = he * dy * dy * L[4];
      Qzz = he * dz * dz * L[8];
      Qxy = he * dx * dy * (L[1] + L[3]);  // off-diagonal pairs: the form counts each once
      Qxz = he * dx * dz * (L[2] + L[6]);
      Qyz = he * dy * dz * (L[5] + L[7]);
    }
    T mx = (T)-1.0e30, mn = (T)1.0e30;
    double sm = 0.0;
    double Dx0 = 0, Dy0 = 0, P0 = 0;
#pragma unroll
    for (int p = 0; p <= E; p++) {
      const T a1 = __shfl_down(x0[p], 1, 64), b1 = __shfl_down(x1[p], 1, 64);  // shifted in the field's type
      if (p < nst) {
        mx = fmax(mx, x0[p]);
        mn = fmin(mn, x0[p]);
        sm += (double)x0[p];
        if (st_row1) {
          mx = fmax(mx, x1[p]);
          mn = fmin(mn, x1[p]);
          sm += (double)x1[p];
        }
      }
      const double A = x0[p], B = x1[p], A1 = a1, B1 = b1;
      const double Dx1 = (A1 - A) + (B1 - B), Dy1 = (B - A) + (B1 - A1), P1 = (A + A1) + (B + B1);
      if (p > 0 && p - 1 < np) {  // the cell between planes p - 1 and p (uniform condition)
        const int c = p - 1;
        const double S = P0 + P1, SX = Dx0 + Dx1, SY = Dy0 + Dy1, SZ = P1 - P0;
        if (ac) {
          const double qa = Qxx * SX + Qxy * SY + Qxz * SZ, qb = Qyy * SY + Qyz * SZ, qc = Qzz * SZ;
          e = fma(SX, qa, e);
          e = fma(SY, qb, e);
          e = fma(SZ, qc, e);
          const double S2 = S * S;
          s2[c] += S2;
          s3[c] = fma(S2, S, s3[c]);
        }
        if (cp || !ac) {
          const double g2 = fma(wx * SX, SX, fma(wy * SY, SY, wz * SZ * SZ));
          if (cp) ec += g2;
          if (!ac) {
            const double w2 = fma(S * S, 1.0 / 64.0, -1.0);
            sq += g2 + w2 * w2 * 0.25;
          }
        }
      }
      Dx0 = Dx1, Dy0 = Dy1, P0 = P1;
    }
    // this task's statistics: lanes that count no point give the identity
    const double cnt = (double)(nst * (st_row1 ? 2 : 1));
    double dmx = st_lane ? (double)mx : -1.0e30, dmn = st_lane ? (double)mn : 1.0e30;
    sm = st_lane ? sm + cnt : 0.0;
    dmx = wave_max(dmx);
    dmn = wave_min(dmn);
    sm = wave_sum(sm);
    if (lane == 0) {
      double *o = wst + 3 * (wv * V + v);
      o[0] = fmax(o[0], dmx);
      o[1] = fmin(o[1], dmn);
      o[2] += sm;
    }
    if (ac && v == nac - 1) {  // the sub-range's Landau terms, after its last AC field
#pragma unroll
      for (int p = 0; p < E; p++) {
        if (p < np)
          e += (a.a2 / 128.0) * s2[p] - (a.a3 / 1536.0) * s3[p] + (a.a4 / 16384.0) * s2[p] * s2[p];
        s2[p] = s3[p] = 0.0;
      }
      if (x.s == nsub - 1) block_sum_out(e, 0);
    }
    if (!ac && x.s == nsub - 1) {  // the CH field's last sub-range
      block_sum_out(sq, 1 + v - nac);
      sq = 0.0;
    }
  };
  T xa0[E + 1], xa1[E + 1], xb0[E + 1], xb1[E + 1];
  tk ta = {nac > 0 ? 0 : nac, 0};
  if (nt > 0) load(ta, xa0, xa1);
  for (int t = 0; t < nt; t += 2) {
    const tk tb = next(ta);
    if (t + 1 < nt) load(tb, xb0, xb1);
    process(ta, xa0, xa1);
    if (t + 1 < nt) {
      ta = next(tb);
      if (t + 2 < nt) load(ta, xa0, xa1);
      process(tb, xb0, xb1);
    }
  }
  if (nsub == 0) {  // an empty chunk (fewer than ECHUNK planes in all): zero partials
    for (int q = 0; q <= nch; q++) block_sum_out(0.0, q);
  } else if (nac == 0) {
    block_sum_out(0.0, 0);  // no AC fields: kind 0 is read by nobody, kept defined
  }
  if (couple) block_sum_out(ec, 1 + nch);
  __syncthreads();
  for (int v = threadIdx.x; v < V; v += blockDim.x) {
    double mx = -1.0e30, mn = 1.0e30, sm = 0.0;
    for (int q = 0; q < nw; q++) {
      const double *o = wst + 3 * (q * V + v);
      mx = fmax(mx, o[0]);
      mn = fmin(mn, o[1]);
      sm += o[2];
    }
    double *o = part + ((size_t)v * nb + bi) * 3;
    o[0] = mx;
    o[1] = mn;
    o[2] = sm;
  }
}

static int diag_kinds(const etd_ctx *c) { return 1 + c->nch + ((c->nac > 0 && c->nch > 0) ? 1 : 0); }
static int diag_lo(const etd_ctx *c) { return c->p.periodic ? c->g : 0; }
static int diag_blocks(const etd_ctx *c) { return (c->n - c->g - 1 - diag_lo(c)) * ECHUNK; }
// doubles of partial sums one diagnostics call writes
static size_t diag_partials(const etd_ctx *c) { return ((size_t)3 * c->V + diag_kinds(c)) * diag_blocks(c); }
size_t etd_diag_partials(const etd_ctx *c) { return diag_partials(c); }

static diag_args make_diag_args(const etd_ctx *c) {
  diag_args a;
  a.n = c->n;
  a.g = c->g;
  a.nac = c->nac;
  a.nch = c->nch;
  a.lo = diag_lo(c);
  a.hx = c->p.hx;
  a.hy = c->p.hy;
  a.hz = c->p.hz;
  a.epn2 = c->p.epn2;
  a.a2 = c->p.a2;
  a.a3 = c->p.a3;
  a.a4 = c->p.a4;
  a.ihx = 1.0 / c->p.hx;
  a.ihy = 1.0 / c->p.hy;
  a.ihz = 1.0 / c->p.hz;
  return a;
}

template <typename T>
static void launch_diag_kernel(etd_ctx *c, double *part) {
  static_assert(ECHUNK % 8 == 0, "the XCD mapping needs a multiple of 8 work items");
  const int me = c->n - c->g - 1 - diag_lo(c), nw = (me + 62) / 63;
  k_diag<T><<<(unsigned)(me * ECHUNK), 64 * nw, sizeof(double) * 3 * nw * c->V, c->stream>>>(fields<T>(c).u, c->lam,
                                                                                           part, make_diag_args(c));
}

// Reduction of the partials, one result per thread index t (0 .. V + nch + 1),
// each a sequential sum in block order; shared by the host (diag_reduce) and
// k_diag_finish so the two give the same bits.  t < V: max, min, vol of field
// t; t == V: the AC energy (reported for every AC field); V < t <= V + nch:
// CH field t - V - 1; t == V + nch + 1: the couple energy, which continues
// from the last CH field's energy as check_soln_new's accumulator does.
__host__ __device__ static inline void diag_reduce_one(const double *part, int nb, int V, int nac, int nch, double h3,
                                                       int t, double *stats) {
#pragma clang fp contract(off)
  const double *he = part + (size_t)3 * V * nb;
  if (t < V) {
    double mx = -1.0e30, mn = 1.0e30, vol = 0.0;
    for (int b = 0; b < nb; b++) {
      const double *o = part + ((size_t)t * nb + b) * 3;
      if (o[0] > mx) mx = o[0];
      if (o[1] < mn) mn = o[1];
      vol += o[2];
    }
    stats[4 * t + 0] = mx;
    stats[4 * t + 1] = mn;
    stats[4 * t + 2] = vol * h3;  // the energy column [4 t + 3] belongs to the energy indices below
  } else if (t == V) {
    if (nac > 0) {
      double e = 0.0;
      for (int b = 0; b < nb; b++) e += he[b];
      e = e * h3;
      for (int v = 0; v < nac; v++) stats[4 * v + 3] = e;
    }
  } else if (t <= V + nch) {
    const int q = t - V - 1;
    double e = 0.0;
    for (int b = 0; b < nb; b++) e += he[(size_t)(1 + q) * nb + b];
    stats[4 * (nac + q) + 3] = e * h3;
  } else if (t == V + nch + 1) {
    double couple = 0.0;
    if (nac > 0 && nch > 0) {
      double last = 0.0;
      for (int b = 0; b < nb; b++) last += he[(size_t)nch * nb + b];
      couple = last * h3;
      for (int b = 0; b < nb; b++) couple += he[(size_t)(1 + nch) * nb + b];
      couple = couple * h3;
    }
    stats[4 * V] = couple;
  }
}

template <typename T>
int launch_diag(etd_ctx *c, double *stats) {
  etd_prof_begin(c, PC_ENERGY);
  launch_diag_kernel<T>(c, c->part);
  etd_prof_end(c, PC_ENERGY);
  if (hipGetLastError() != hipSuccess) return ETD_ERR_DEVICE;
  if (hipMemcpyAsync(c->hpart, c->part, diag_partials(c) * sizeof(double), hipMemcpyDeviceToHost, c->stream) !=
      hipSuccess)
    return ETD_ERR_DEVICE;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return ETD_ERR_DEVICE;
  const double h3 = c->p.hx * c->p.hy * c->p.hz;
  for (int v = 0; v < c->V; v++) stats[4 * v + 3] = 0.0;
  for (int t = 0; t <= c->V + c->nch + 1; t++)
    diag_reduce_one(c->hpart, diag_blocks(c), c->V, c->nac, c->nch, h3, t, stats);
  return ETD_OK;
}

// ------------------------------------------------------------------ allocation-free diagnostics
// check_soln_new into a caller's device buffer with no allocation, no host
// synchronisation and no memset: the per-block partial sums go to a
// file-scope device array (one copy per GPU, kept for the process's
// lifetime) and k_diag_finish adds them up in the order launch_diag uses on
// the host, so the statistics are bit for bit those of etd_diag.
#define DIAG_SCRATCH_DOUBLES 65536  // 512 KiB
#define DIAG_SCRATCH_DEVICES 64
static __device__ __align__(256) double scratch[DIAG_SCRATCH_DOUBLES];
static double *scratch_dev_ptr[DIAG_SCRATCH_DEVICES];
static std::mutex scratch_mu;

// the device address of `scratch` on device dev (looked up once per device,
// with that device current; the host-side name is not a device address)
static int scratch_address(int dev, double **out) {
  if (dev < 0 || dev >= DIAG_SCRATCH_DEVICES) return ETD_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev_ptr[dev]) {
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return ETD_ERR_DEVICE;
    if (cur != dev && hipSetDevice(dev) != hipSuccess) return ETD_ERR_DEVICE;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !ptr) return ETD_ERR_DEVICE;
    scratch_dev_ptr[dev] = (double *)ptr;
  }
  *out = scratch_dev_ptr[dev];
  return ETD_OK;
}

// One thread per result index of diag_reduce_one (the host's reduction, so
// the statistics are bit for bit those of etd_diag).
__global__ __launch_bounds__(256) void k_diag_finish(const double *__restrict__ part, int nb, int V, int nac, int nch,
                                                     double h3, double *__restrict__ stats) {
  if ((int)threadIdx.x <= V + nch + 1) diag_reduce_one(part, nb, V, nac, nch, h3, (int)threadIdx.x, stats);
}

extern "C" int etd_diag_noalloc(etd_ctx *c, double *stats_dev) {
  if (!c || !stats_dev) return ETD_ERR_ARG;
  if (c->P > 1) {
    snprintf(c->err, sizeof(c->err), "%s: one box only (px,py,pz = 1,1,1)", __func__);
    return ETD_ERR_UNSUPPORTED;
  }
  const size_t need = diag_partials(c);
  if (need > DIAG_SCRATCH_DOUBLES || c->V + c->nch + 2 > 256) {
    snprintf(c->err, sizeof(c->err),
             "etd_diag_noalloc: this context needs %zu doubles of partial sums ((3 V + 1 + nch + couple) x cells "
             "per axis x 8); the static buffer holds %d",
             need, DIAG_SCRATCH_DOUBLES);
    return ETD_ERR_UNSUPPORTED;
  }
  double *sc = nullptr;
  int rc = scratch_address(c->device, &sc);
  if (rc) {
    snprintf(c->err, sizeof(c->err), "etd_diag_noalloc: no address for the static scratch on device %d", c->device);
    return rc;
  }
  if (c->mixed)
    launch_diag_kernel<float>(c, sc);
  else
    launch_diag_kernel<double>(c, sc);
  k_diag_finish<<<1, 256, 0, c->stream>>>(sc, diag_blocks(c), c->V, c->nac, c->nch, c->p.hx * c->p.hy * c->p.hz,
                                          stats_dev);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    snprintf(c->err, sizeof(c->err), "etd_diag_noalloc: %s", hipGetErrorString(e));
    return ETD_ERR_DEVICE;
  }
  return ETD_OK;
}

// ------------------------------------------------------------------ counter calibration
// A plain 16-byte-per-lane copy: reads and writes exactly 16 * n16 bytes.
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ in, uint4 *__restrict__ out, size_t n16) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) out[i] = in[i];
}

void launch_copy16(etd_ctx *c, const void *in, void *out, size_t n16) {
  const size_t blocks = (n16 + 255) / 256;
  k_copy16<<<(unsigned)(blocks < 8192 ? blocks : 8192), 256, 0, c->stream>>>((const uint4 *)in, (uint4 *)out, n16);
}

// ------------------------------------------------------------------ instantiations (FP64 and FP32 storage)
// ------------------------------------------------------------------ CH mean
// Interior sum of CH field blockIdx.y per block (one (z, y) row per wave,
// lanes along x), into part[blockIdx.y * CH_BLOCKS + blockIdx.x].
__device__ __forceinline__ double block_sum256(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  __shared__ double sh[4];
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  return ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// block sum over CH_WAVES waves of 64 (the row kernels below)
#define CH_WAVES 16
__device__ __forceinline__ double block_sum_w(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  __shared__ double sh[CH_WAVES];
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < CH_WAVES; w++) s += sh[w];
  return s;
}

// Interior rows (z, y) one per wave, CH_WAVES waves per block; each lane
// loads its (up to 5) cells of a row before adding them, so a wave has the
// whole row in flight (N <= 268 < 5 x 64).
__global__ __launch_bounds__(64 * CH_WAVES) void k_ch_sum(const float *__restrict__ u, size_t fs, int n, int g,
                                                          double *__restrict__ part) {
  const float *f = u + fs * blockIdx.y;
  const int N = n - 2 * g, lane = threadIdx.x & 63;
  double acc = 0.0;
  for (int r = blockIdx.x * CH_WAVES + (threadIdx.x >> 6); r < N * N; r += gridDim.x * CH_WAVES) {
    const int z = r / N, y = r - z * N;
    const float *row = f + IDX(n, z + g, y + g, g);
    float v[5];
#pragma unroll
    for (int e = 0; e < 5; e++) v[e] = lane + 64 * e < N ? row[lane + 64 * e] : 0.0f;
#pragma unroll
    for (int e = 0; e < 5; e++) acc += (double)v[e];
  }
  const double t = block_sum_w(acc);
  if (threadIdx.x == 0) part[blockIdx.y * CH_BLOCKS + blockIdx.x] = t;
}

// the partials of field q = blockIdx.x reduced in a fixed order into target[q]
__global__ __launch_bounds__(256) void k_ch_target(const double *__restrict__ part, double *__restrict__ target) {
  const double s = block_sum256(threadIdx.x < CH_BLOCKS ? part[blockIdx.x * CH_BLOCKS + threadIdx.x] : 0.0);
  if (threadIdx.x == 0) target[blockIdx.x] = s;
}

// u_q += (target_q - sum_q) / N^3 on the interior of CH field q = blockIdx.y;
// every block reduces the CH_BLOCKS partials of sum_q in the same order
__global__ __launch_bounds__(64 * CH_WAVES) void k_ch_shift(float *__restrict__ u, size_t fs, int n, int g,
                                                            const double *__restrict__ target,
                                                            const double *__restrict__ part) {
  float *f = u + fs * blockIdx.y;
  const int N = n - 2 * g, lane = threadIdx.x & 63;
  const int q = blockIdx.y;
  const double a = block_sum_w(threadIdx.x < CH_BLOCKS ? part[q * CH_BLOCKS + threadIdx.x] : 0.0);
  const double shift = (target[q] - a) / ((double)N * N * N);
  for (int r = blockIdx.x * CH_WAVES + (threadIdx.x >> 6); r < N * N; r += gridDim.x * CH_WAVES) {
    const int z = r / N, y = r - z * N;
    float *row = f + IDX(n, z + g, y + g, g);
    float v[5];
#pragma unroll
    for (int e = 0; e < 5; e++) v[e] = lane + 64 * e < N ? row[lane + 64 * e] : 0.0f;
#pragma unroll
    for (int e = 0; e < 5; e++)
      if (lane + 64 * e < N) row[lane + 64 * e] = (float)((double)v[e] + shift);
  }
}

// The undo of the centring and the mean restoration of CH field q =
// blockIdx.y in one pass over the extended box: ghost cells get the mean m
// added, interior cells (target_q - a_q) / N^3 with a_q the interior sum of the
// centred field (the partials of k_ch_sum, reduced in the same order by every
// block).
__global__ __launch_bounds__(64 * CH_WAVES) void k_ch_shift_center(float *__restrict__ u, size_t fs, int n, int g,
                                                                   const double *__restrict__ target,
                                                                   const double *__restrict__ part, double invN3) {
  float *f = u + fs * blockIdx.y;
  const int N = n - 2 * g, lane = threadIdx.x & 63;
  const int q = blockIdx.y;
  const double a = block_sum_w(threadIdx.x < CH_BLOCKS ? part[q * CH_BLOCKS + threadIdx.x] : 0.0);
  const double shift = (target[q] - a) / ((double)N * N * N);
  const float m = (float)(target[q] * invN3);
  for (int r = blockIdx.x * CH_WAVES + (threadIdx.x >> 6); r < n * n; r += gridDim.x * CH_WAVES) {
    const int z = r / n, y = r - z * n;
    const bool row_in = z >= g && z < n - g && y >= g && y < n - g;
    float *row = f + IDX(n, z, y, 0);
    float v[5];
#pragma unroll
    for (int e = 0; e < 5; e++) v[e] = lane + 64 * e < n ? row[lane + 64 * e] : 0.0f;
#pragma unroll
    for (int e = 0; e < 5; e++) {
      const int x = lane + 64 * e;
      if (x < n) row[x] = (row_in && x >= g && x < n - g) ? (float)((double)v[e] + shift) : v[e] + m;
    }
  }
}

// f = target[q] / N^3 on the whole extended box
__global__ void k_ch_fill(float *__restrict__ f, size_t n3, const double *__restrict__ target, int q, double invN3) {
  const size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l < n3) f[l] = (float)(target[q] * invN3);
}

// u += sign c_mean (whole box), N += sign N(c_mean), CH field blockIdx.y
__global__ void k_ch_center(float *__restrict__ u, float *__restrict__ N, size_t n3, const double *__restrict__ target,
                            const float *__restrict__ nconst, double invN3, int sign) {
  const size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n3) return;
  const int q = blockIdx.y;
  if (u) {
    float *f = u + n3 * q;
    const float m = (float)(target[q] * invN3);
    f[l] = sign > 0 ? f[l] + m : f[l] - m;
  }
  if (N) {
    float *f = N + n3 * q;
    const float k = nconst[n3 * q + l];
    f[l] = sign > 0 ? f[l] + k : f[l] - k;
  }
}

void ch_mean_mark(etd_ctx *c, const float *u) {
  if (!c->ch_sum || !c->ch_stale) return;
  double *part = c->ch_sum + c->nch;
  k_ch_sum<<<dim3(CH_BLOCKS, c->nch), 64 * CH_WAVES, 0, c->stream>>>(u + c->n3 * c->nac, c->n3, c->n, c->g, part);
  k_ch_target<<<c->nch, 256, 0, c->stream>>>(part, c->ch_sum);
  // N(c_mean) of each CH field, through the same kernels as the step, on a
  // uniform field in the mode-product scratch
  const double invN3 = 1.0 / ((double)c->N * c->N * c->N);
  nl_args a = make_nl(c);
  const unsigned nb = (unsigned)((c->n3 + 255) / 256);
  float *uni = c->f32.w1;
  for (int q = 0; q < c->nch; q++) {
    k_ch_fill<<<nb, 256, 0, c->stream>>>(uni, c->n3, c->ch_sum, q, invN3);
    launch_ch_nl<float>(c, uni, c->ch_const + c->n3 * q, nullptr, nullptr, q, a);
  }
  c->ch_stale = 0;
}

void ch_center(etd_ctx *c, float *u, float *N, int sign) {
  if (!c->ch_sum) return;
  const double invN3 = 1.0 / ((double)c->N * c->N * c->N);
  const size_t off = c->n3 * c->nac;
  k_ch_center<<<dim3((unsigned)((c->n3 + 255) / 256), c->nch), 256, 0, c->stream>>>(
      u ? u + off : nullptr, N ? N + off : nullptr, c->n3, c->ch_sum, c->ch_const, invN3, sign);
}

void ch_mean_restore_centered(etd_ctx *c, float *u) {
  if (!c->ch_sum) return;
  double *part = c->ch_sum + c->nch;
  const double invN3 = 1.0 / ((double)c->N * c->N * c->N);
  k_ch_sum<<<dim3(CH_BLOCKS, c->nch), 64 * CH_WAVES, 0, c->stream>>>(u + c->n3 * c->nac, c->n3, c->n, c->g, part);
  k_ch_shift_center<<<dim3(CH_BLOCKS, c->nch), 64 * CH_WAVES, 0, c->stream>>>(u + c->n3 * c->nac, c->n3, c->n,
                                                                              c->g, c->ch_sum, part, invN3);
}

void ch_mean_restore(etd_ctx *c, float *u) {
  if (!c->ch_sum) return;
  double *part = c->ch_sum + c->nch;
  k_ch_sum<<<dim3(CH_BLOCKS, c->nch), 64 * CH_WAVES, 0, c->stream>>>(u + c->n3 * c->nac, c->n3, c->n, c->g, part);
  k_ch_shift<<<dim3(CH_BLOCKS, c->nch), 64 * CH_WAVES, 0, c->stream>>>(u + c->n3 * c->nac, c->n3, c->n, c->g,
                                                                       c->ch_sum, part);
}

#define ETD_INST(T)                                                                   \
  template void launch_ghost_fill<T>(etd_ctx *, T *, int, size_t, T *, T *, T *);    \
  template void launch_phiscale<T>(etd_ctx *, const T *, const T *, T *, int, int, int); \
  template void launch_phiscale_pair<T>(etd_ctx *, const T *, const T *, T *, double, T *, double, int, int); \
  template void launch_nonlinear<T>(etd_ctx *, T *, const T *, int);                 \
  template void launch_elastic<T>(etd_ctx *);                                        \
  template void fft3d<T>(etd_ctx *, T *, int, int);                                  \
  template void launch_felas_materialize<T>(etd_ctx *, T *);                         \
  template int launch_diag<T>(etd_ctx *, double *);
ETD_INST(double)
ETD_INST(float)
#undef ETD_INST
