// etd_kernels_mixed.hip -- mode products of the mixed-precision step on
// gfx950: FP32 fields, V / V^-1 split into BF16 terms, FP32 MFMA
// accumulation (v_mfma_f32_16x16x32_bf16).
//
// A value x is split exactly into BF16 terms x = x0 + x1 + x2 (x0 = bf16(x),
// x1 = bf16(x - x0), x2 = x - x0 - x1; each residual is exact in FP32 and the
// last one has at most 8 significant bits, so three terms hold every FP32
// value).  With A = a0 + a1 + a2 (split from FP64) the product keeps the six
// terms a_i x_j with i + j <= 2 ("split-3"); the dropped ones are 2^-24
// relative, below FP32 rounding.  Two terms (i + j <= 1: three products,
// "split-2") keep 2^-16.
//
// Same contraction as the FP64 kernel: C(m, col) = sum_a A(m,a) in[a + n col],
// written rotated (out[col + n^2 m]) or in GEMM layout (out[m + n col]).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <math.h>

#include "etd_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4m __attribute__((ext_vector_type(4)));
typedef unsigned u32x4v __attribute__((ext_vector_type(4)));

// ---- global -> LDS copies (buffer_load ... lds), as in the FP64 kernel
__device__ __forceinline__ i32x4m mk_rsrc(const void *base, uint32_t bytes) {
  const uint64_t a = (uint64_t)base;
  i32x4m r;
  r.x = (int)(uint32_t)a;
  r.y = (int)(uint32_t)(a >> 32);
  r.z = (int)bytes;
  r.w = 0x00020000;
  return r;
}

// lane L's 16 bytes land at lds_wave_base + 16 L; offsets outside the
// descriptor's range read as zero.  asm so the compiler does not make every
// later LDS read wait for all copies in flight; the explicit vmcnt wait below
// orders them.  One wait state between the M0 write and the copy.  M0 is
// reserved to the compiler; the compiled kernels of this file use it only
// inside these asm blocks (checked in the ISA), so the clobber holds.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void copy16_to_lds(i32x4m rs, uint32_t lds_wave_base, int voff) {
  const uint32_t b = __builtin_amdgcn_readfirstlane(lds_wave_base);
  asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tbuffer_load_dwordx4 %1, %2, 0 offen lds"
               :
               : "s"(b), "v"(voff), "s"(rs)
               : "memory", "m0");
}
#pragma clang diagnostic pop

__device__ __forceinline__ void wait_all_copies_barrier() { asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" : : : "memory"); }

// ---- tiling
// A workgroup of WAVES waves owns 16 UT WAVES columns (UT column tiles of 16
// per wave) and all MT row tiles of 16; K runs in chunks of 32 (one MFMA).
// Per chunk the LDS stage holds
//   A: terms x MT x 64 lanes x 16 B, pre-arranged in operand order, and
//   X: 32 k x COLS columns FP32 as [q = k/8][col][8 k] (a lane's 8
//      consecutive k of one column are 32 contiguous bytes, and the 16 lanes
//      of an operand read 16 consecutive 32-byte blocks).
// Each A fragment read from LDS feeds the MFMAs of UT column tiles.
template <int MT, int NT, int WAVES, int UT = 2>
struct mps_cfg {
  static constexpr int THREADS = 64 * WAVES;
  static constexpr int COLS = 16 * UT * WAVES;
  static constexpr int APIECES = NT * MT * 64;                // 16-byte pieces of one A chunk
  static constexpr int APT = (APIECES + THREADS - 1) / THREADS;  // per thread
  static constexpr int ASZ = APT * THREADS * 16;               // A chunk bytes (zero padded)
  static constexpr int XSZ = 32 * COLS * 4;                    // X chunk bytes
  static constexpr int XPT = XSZ / 16 / THREADS;               // = 2 UT
  static constexpr int STAGE = ASZ + XSZ;
  static constexpr int LDS = 2 * STAGE;
  static_assert(XPT == 2 * UT && UT % 2 == 0, "X pieces");
};

__device__ __forceinline__ void split_bf16(float x, __bf16 &t0, __bf16 &t1, __bf16 &t2) {
  t0 = (__bf16)x;
  const float r1 = x - (float)t0;
  t1 = (__bf16)r1;
  t2 = (__bf16)(r1 - (float)t1);
}

// Epilogue of a workgroup tile: lane (q, r16) holds rows 16 t + 4 q + i of
// columns col0 + 16 u.  Raw buffer stores with 32-bit byte offsets (n^3 * 4 <
// 2^31), advanced by one add per row; a lane whose column or row is outside
// the matrix gets an offset past the buffer's end, which the hardware drops
// (rows m >= n would alias the next column in the GEMM layout).  No branches
// per element, no 64-bit address arithmetic.
template <int MT, int UT, bool ACCUMULATE>
__device__ __forceinline__ void store_tile_(const f32x4 (&acc)[MT][UT], float *dst, int n, long col0, int q,
                                            int rotate) {
  const long n2 = (long)n * n;
  const __amdgpu_buffer_rsrc_t rD =
      __builtin_amdgcn_make_buffer_rsrc((void *)dst, 0, (int)((size_t)n * n2 * 4), 0x00020000);
  const uint32_t s_row = rotate ? (uint32_t)n2 * 4u : 4u, s_col = rotate ? 4u : (uint32_t)n * 4u;
  const uint32_t OOB = 0x80000000u;
  // MT = ceil(n / 16): only the last row tile holds rows >= n
  const int last_rows = n - 16 * (MT - 1) - 4 * q;  // valid i in that tile: i < last_rows
#pragma unroll
  for (int u = 0; u < UT; u++) {
    const long col = col0 + 16 * u;
    const bool cok = col < n2;
    uint32_t o = (uint32_t)col * s_col + (uint32_t)(4 * q) * s_row;
#pragma unroll
    for (int t = 0; t < MT; t++) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const bool rok = t < MT - 1 || i < last_rows;
        const int off = (int)((cok && rok) ? o : OOB);
        float v = acc[t][u][i];
        if constexpr (ACCUMULATE) v = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rD, off, 0, 0)) + v;
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rD, off, 0, 0);
        o += s_row;
      }
      o += 12u * s_row;
    }
  }
}

// Epilogue for a wave that owns row tiles t0 .. t0 + MTH - 1 of MT (tiles at
// MT and beyond are skipped; only tile MT - 1 holds rows >= n)
template <int MTH, int UT, bool ACCUMULATE>
__device__ __forceinline__ void store_rows_(const f32x4 (&acc)[MTH][UT], float *dst, int n, long col0, int q,
                                           int rotate, int t0, int MT) {
  const long n2 = (long)n * n;
  const __amdgpu_buffer_rsrc_t rD =
      __builtin_amdgcn_make_buffer_rsrc((void *)dst, 0, (int)((size_t)n * n2 * 4), 0x00020000);
  const uint32_t s_row = rotate ? (uint32_t)n2 * 4u : 4u, s_col = rotate ? 4u : (uint32_t)n * 4u;
  const uint32_t OOB = 0x80000000u;
  const int last_rows = n - 16 * (MT - 1) - 4 * q;
#pragma unroll
  for (int u = 0; u < UT; u++) {
    const long col = col0 + 16 * u;
    const bool cok = col < n2;
    uint32_t o = (uint32_t)col * s_col + (uint32_t)(16 * t0 + 4 * q) * s_row;
#pragma unroll
    for (int tt = 0; tt < MTH; tt++) {
      const int t = t0 + tt;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const bool rok = t < MT - 1 || (t == MT - 1 && i < last_rows);
        const int off = (int)((cok && rok) ? o : OOB);
        float v = acc[tt][u][i];
        if constexpr (ACCUMULATE) v = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rD, off, 0, 0)) + v;
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rD, off, 0, 0);
        o += s_row;
      }
      o += 12u * s_row;
    }
  }
}

// Epilogue with out[m + n col]: a lane's 4 rows 16 t + 4 q + [0, 4) of one
// column are 16 contiguous bytes, one 128-bit store (or load-add-store).  n is
// a multiple of 4, so the 4 rows are inside or outside the matrix together.
template <int MT, int UT>
__device__ __forceinline__ void store_tile_m4(const f32x4 (&acc)[MT][UT], float *dst, int n, long col0, int q,
                                              int accumulate) {
  const long n2 = (long)n * n;
  const __amdgpu_buffer_rsrc_t rD =
      __builtin_amdgcn_make_buffer_rsrc((void *)dst, 0, (int)((size_t)n * n2 * 4), 0x00020000);
  const uint32_t OOB = 0x80000000u;
#pragma unroll
  for (int u = 0; u < UT; u++) {
    const long col = col0 + 16 * u;
    const bool cok = col < n2;
    const uint32_t o0 = (uint32_t)col * (uint32_t)n * 4u + (uint32_t)(16 * q);
#pragma unroll
    for (int t = 0; t < MT; t++) {
      const bool rok = 16 * t + 4 * q < n;
      const int off = (int)((cok && rok) ? o0 + 64u * (uint32_t)t : OOB);
      f32x4 v = acc[t][u];
      if (accumulate) {
        const f32x4 o = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rD, off, 0, 0));
        v = v + o;
      }
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4v, v), rD, off, 0, 0);
    }
  }
}

template <int MT, int UT>
__device__ __forceinline__ void store_tile(const f32x4 (&acc)[MT][UT], float *dst, int n, long col0, int q,
                                           int rotate, int accumulate) {
  if (accumulate)
    store_tile_<MT, UT, true>(acc, dst, n, col0, q, rotate);
  else
    store_tile_<MT, UT, false>(acc, dst, n, col0, q, rotate);
}

template <int MT, int NT, int WAVES>
__global__ __launch_bounds__(64 * WAVES, WAVES / 4) void k_modeprod_split(const uint16_t *__restrict__ A3, const float *__restrict__ in,
                                                                  size_t in_fs, float *__restrict__ out, size_t out_fs,
                                                                  int n, int rotate, int accumulate, int a_chunk) {
  using CF = mps_cfg<MT, NT, WAVES>;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float *src = in + in_fs * blockIdx.y;
  float *dst = out + out_fs * blockIdx.y;
  const long n2 = (long)n * n;
  const int nkc = (n + 31) / 32;
  const long wgcol = (long)blockIdx.x * CF::COLS;
  // a_chunk: byte stride of the arranged operand's chunks (>= CF::ASZ)
  const i32x4m rsA = mk_rsrc(A3, (uint32_t)((size_t)nkc * a_chunk));
  const i32x4m rsX = mk_rsrc(src, (uint32_t)((size_t)n * n2 * 4));
  const uint32_t lds0 = (uint32_t)(uintptr_t)smem;

  auto issue = [&](int kc, int stg) {
    const uint32_t sbase = lds0 + stg * CF::STAGE;
#pragma unroll
    for (int e = 0; e < CF::APT; e++)
      copy16_to_lds(rsA, sbase + (e * CF::THREADS + wave * 64) * 16,
                    (int)((size_t)kc * a_chunk + ((size_t)e * CF::THREADS + threadIdx.x) * 16));
    // X piece e of this lane: q = e, column 32 wave + lane/2, k half lane%2
#pragma unroll
    for (int e = 0; e < CF::XPT; e++) {
      const long col = wgcol + 32 * wave + (lane >> 1);
      const int k = 32 * kc + 8 * e + 4 * (lane & 1);
      const bool ok = col < n2 && k < n;
      const int voff = ok ? (int)((k + (long)n * col) * 4) : (int)0x80000000;
      copy16_to_lds(rsX, sbase + CF::ASZ + (e * CF::THREADS + wave * 64) * 16, voff);
    }
  };

  f32x4 acc[MT][2];
#pragma unroll
  for (int t = 0; t < MT; t++)
#pragma unroll
    for (int u = 0; u < 2; u++) acc[t][u] = (f32x4){0.f, 0.f, 0.f, 0.f};

  const int q = lane >> 4, r16 = lane & 15;
  issue(0, 0);
  for (int kc = 0; kc < nkc; kc++) {
    // chunk kc is in LDS for every wave, and every wave is past chunk kc-1,
    // whose stage the next copies overwrite
    wait_all_copies_barrier();
    if (kc + 1 < nkc) issue(kc + 1, (kc + 1) & 1);
    const char *stg = smem + (kc & 1) * CF::STAGE;
    const bf16x8 *sa = reinterpret_cast<const bf16x8 *>(stg);
    const float *sx = reinterpret_cast<const float *>(stg + CF::ASZ);
    bf16x8 b[2][NT];
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const float *bp = sx + ((size_t)q * CF::COLS + 32 * wave + 16 * u + r16) * 8;
      const float4 lo = *reinterpret_cast<const float4 *>(bp);
      const float4 hi = *reinterpret_cast<const float4 *>(bp + 4);
      const float x[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
      for (int e = 0; e < 8; e++) {
        __bf16 t0, t1, t2;
        split_bf16(x[e], t0, t1, t2);
        b[u][0][e] = t0;
        if constexpr (NT > 1) b[u][1][e] = t1;
        if constexpr (NT > 2) b[u][2][e] = t2;
      }
    }
#pragma unroll
    for (int t = 0; t < MT; t++) {
      bf16x8 a[NT];
#pragma unroll
      for (int tm = 0; tm < NT; tm++) a[tm] = sa[(tm * MT + t) * 64 + lane];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        // smallest terms first
        if constexpr (NT > 2) {
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[u][0], acc[t][u], 0, 0, 0);
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[u][1], acc[t][u], 0, 0, 0);
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[u][2], acc[t][u], 0, 0, 0);
        }
        if constexpr (NT > 1) {
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[u][0], acc[t][u], 0, 0, 0);
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[u][1], acc[t][u], 0, 0, 0);
        }
        acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[u][0], acc[t][u], 0, 0, 0);
      }
    }
  }
  // D: lane holds rows 4 q + i of column r16
  store_tile<MT, 2>(acc, dst, n, wgcol + 32 * wave + r16, q, rotate, accumulate);
}

// Persistent variant: gridDim.x workgroups (one per CU) loop over the tiles
// (field, 32 WAVES columns); the double-buffered chunk pipeline runs across
// tile boundaries, so the next tile's first chunk loads during the current
// tile's last chunk and no workgroup start-up is exposed.
template <int MT, int NT, int WAVES, int UT = 2, int XL = 1>
__global__ __launch_bounds__(64 * WAVES, WAVES / 4) void k_modeprod_split_p(
    const uint16_t *__restrict__ A3, const float *__restrict__ in, size_t in_fs, float *__restrict__ out,
    size_t out_fs, int n, int rotate, int accumulate, int a_chunk, int nfields) {
  using CF = mps_cfg<MT, NT, WAVES, UT>;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long n2 = (long)n * n;
  const int nkc = (n + 31) / 32;
  const int ncb = (int)((n2 + CF::COLS - 1) / CF::COLS);
  const int ntiles = ncb * nfields;
  int tile = blockIdx.x;
  if (tile >= ntiles) return;  // uniform
  const i32x4m rsA = mk_rsrc(A3, (uint32_t)((size_t)nkc * a_chunk));
  const uint32_t lds0 = (uint32_t)(uintptr_t)smem;

  auto issue = [&](int tl, int kc, int stg) {
    const int f = tl / ncb;
    const long wgc = (long)(tl % ncb) * CF::COLS;
    i32x4m rsX = mk_rsrc(in + in_fs * f, (uint32_t)((size_t)n * n2 * 4));
    // the tile is the same for the whole workgroup: keep the descriptor scalar
    rsX.x = __builtin_amdgcn_readfirstlane(rsX.x);
    rsX.y = __builtin_amdgcn_readfirstlane(rsX.y);
    rsX.z = __builtin_amdgcn_readfirstlane(rsX.z);
    const uint32_t sbase = lds0 + stg * CF::STAGE;
#pragma unroll
    for (int e = 0; e < CF::APT; e++)
      copy16_to_lds(rsA, sbase + (e * CF::THREADS + wave * 64) * 16,
                    (int)((size_t)kc * a_chunk + ((size_t)e * CF::THREADS + threadIdx.x) * 16));
    // XL = 0: X piece e is k group q = e % 4 of 32-column block cb; lane L
    // brings column 32 cb + L / 2, k half L % 2, to [q][col][8 k] (32 bytes
    // from each of 32 columns per instruction).  XL = 1: 16-byte piece p of
    // the chunk holds k quad (p % 8) ^ (c % 8) of column c = p / 8, so each
    // instruction brings 8 whole 128-byte column runs.
#pragma unroll
    for (int e = 0; e < CF::XPT; e++) {
      long col;
      int k;
      uint32_t ldsoff;
      if constexpr (XL) {
        const int c = (e * CF::THREADS + wave * 64 + lane) >> 3;
        col = wgc + c;
        k = 32 * kc + 4 * ((lane & 7) ^ (c & 7));
        ldsoff = (uint32_t)(e * CF::THREADS + wave * 64) * 16;
      } else {
        const int qe = e & 3, cb = wave * (UT / 2) + (e >> 2);
        col = wgc + 32 * cb + (lane >> 1);
        k = 32 * kc + 8 * qe + 4 * (lane & 1);
        ldsoff = (uint32_t)(qe * CF::COLS + 32 * cb) * 32;
      }
      const bool ok = col < n2 && k < n;
      const int voff = ok ? (int)((k + (long)n * col) * 4) : (int)0x80000000;
      copy16_to_lds(rsX, sbase + CF::ASZ + ldsoff, voff);
    }
  };

  const int q = lane >> 4, r16 = lane & 15;
  issue(tile, 0, 0);
  int g = 0;  // chunks consumed so far: stage g & 1
#pragma unroll 1
  for (; tile < ntiles; tile += gridDim.x) {
    const int f = tile / ncb;
    const long wgcol = (long)(tile % ncb) * CF::COLS;
    float *dst = out + out_fs * f;
    f32x4 acc[MT][UT];
#pragma unroll
    for (int t = 0; t < MT; t++)
#pragma unroll
      for (int u = 0; u < UT; u++) acc[t][u] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (int kc = 0; kc < nkc; kc++, g++) {
      // this chunk is in LDS for every wave, and every wave is past the
      // previous one, whose stage the next copies overwrite
      wait_all_copies_barrier();
      if (kc + 1 < nkc)
        issue(tile, kc + 1, (g + 1) & 1);
      else if (tile + (int)gridDim.x < ntiles)
        issue(tile + gridDim.x, 0, (g + 1) & 1);
      const char *stg = smem + (g & 1) * CF::STAGE;
      const bf16x8 *sa = reinterpret_cast<const bf16x8 *>(stg);
      const float *sx = reinterpret_cast<const float *>(stg + CF::ASZ);
      bf16x8 b[UT][NT];
      // column tile u's split; tiles after the first are split between the
      // MFMAs of row tile 0, so that their VALU work overlaps MFMA work
      auto split_tile = [&](int u) {
        float4 lo, hi;
        if constexpr (XL) {
          const int c = 16 * UT * wave + 16 * u + r16;
          const float4 *cp = reinterpret_cast<const float4 *>(sx) + 8 * c;
          lo = cp[(2 * q) ^ (c & 7)];
          hi = cp[(2 * q + 1) ^ (c & 7)];
        } else {
          const float *bp = sx + ((size_t)q * CF::COLS + 16 * UT * wave + 16 * u + r16) * 8;
          lo = *reinterpret_cast<const float4 *>(bp);
          hi = *reinterpret_cast<const float4 *>(bp + 4);
        }
        const float x[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
        for (int e = 0; e < 8; e++) {
          __bf16 t0, t1, t2;
          split_bf16(x[e], t0, t1, t2);
          b[u][0][e] = t0;
          if constexpr (NT > 1) b[u][1][e] = t1;
          if constexpr (NT > 2) b[u][2][e] = t2;
        }
      };
      for (int u = 0; u < UT; u++) split_tile(u);
      // row tiles in groups of TP: TP UT independent accumulation chains in
      // flight, so consecutive MFMAs on one accumulator are TP UT issues apart
#pragma unroll
      for (int t = 0; t < MT; t += 2) {
        constexpr int TP = 2;
        bf16x8 a[TP][NT];
#pragma unroll
        for (int d = 0; d < TP; d++)
          if (t + d < MT)
#pragma unroll
            for (int tm = 0; tm < NT; tm++) a[d][tm] = sa[(tm * MT + t + d) * 64 + lane];
        // smallest terms first, as in the other kernels
        constexpr int NP = NT == 3 ? 6 : (NT == 2 ? 3 : 1);
        constexpr int ia[6] = {2, 1, 0, 1, 0, 0}, ib[6] = {0, 1, 2, 0, 1, 0};
        constexpr int off = NT == 3 ? 0 : (NT == 2 ? 3 : 5);
#pragma unroll
        for (int pi = 0; pi < NP; pi++)
#pragma unroll
          for (int d = 0; d < TP; d++)
#pragma unroll
            for (int u = 0; u < UT; u++)
              if (t + d < MT)
                acc[t + d][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[d][ia[off + pi]], b[u][ib[off + pi]],
                                                                        acc[t + d][u], 0, 0, 0);
      }
    }
    store_tile<MT, UT>(acc, dst, n, wgcol + 16 * UT * wave + r16, q, rotate, accumulate);
  }
}

// Two-workgroups-per-CU variant (persistent): 4 waves x 32 columns, ONE
// V / V^-1 stage (40 KB) and two field stages (16 KB each), 72 KB of LDS, so
// two workgroups share a CU.  Each SIMD then holds waves of two workgroups
// that do not meet at the same barriers: while one waits for its next
// operand chunk (issued only after every wave of the workgroup is past the
// previous one) or its epilogue stores, the other issues MFMAs.
template <int N_>
__device__ __forceinline__ void wait_vm_barrier() {
  if constexpr (N_ == 0)
    asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" : : : "memory");
  else
    asm volatile("s_waitcnt vmcnt(%0)\n\ts_barrier" : : "n"(N_) : "memory");
}
__device__ __forceinline__ void plain_barrier() { asm volatile("s_barrier" : : : "memory"); }

template <int MT, int NT>
__global__ __launch_bounds__(256, 2) void k_modeprod_split_2wg(const uint16_t *__restrict__ A3,
                                                               const float *__restrict__ in, size_t in_fs,
                                                               float *__restrict__ out, size_t out_fs, int n,
                                                               int rotate, int accumulate, int a_chunk, int nfields) {
  using CF = mps_cfg<MT, NT, 4, 2>;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long n2 = (long)n * n;
  const int nkc = (n + 31) / 32;
  const int ncb = (int)((n2 + CF::COLS - 1) / CF::COLS);
  const int ntiles = ncb * nfields;
  if ((int)blockIdx.x >= ntiles) return;  // uniform
  const int my_tiles = (ntiles - (int)blockIdx.x + (int)gridDim.x - 1) / (int)gridDim.x;
  const int total = my_tiles * nkc;
  const i32x4m rsA = mk_rsrc(A3, (uint32_t)((size_t)nkc * a_chunk));
  const uint32_t lds0 = (uint32_t)(uintptr_t)smem;
  auto tile_of = [&](int g) { return (int)blockIdx.x + (g / nkc) * (int)gridDim.x; };

  auto issue_a = [&](int kc) {
#pragma unroll
    for (int e = 0; e < CF::APT; e++)
      copy16_to_lds(rsA, lds0 + (e * CF::THREADS + wave * 64) * 16,
                    (int)((size_t)kc * a_chunk + ((size_t)e * CF::THREADS + threadIdx.x) * 16));
  };
  auto issue_x = [&](int g, int stg) {
    const int tl = tile_of(g), kc = g % nkc;
    const int f = tl / ncb;
    const long wgc = (long)(tl % ncb) * CF::COLS;
    i32x4m rsX = mk_rsrc(in + in_fs * f, (uint32_t)((size_t)n * n2 * 4));
    rsX.x = __builtin_amdgcn_readfirstlane(rsX.x);
    rsX.y = __builtin_amdgcn_readfirstlane(rsX.y);
    rsX.z = __builtin_amdgcn_readfirstlane(rsX.z);
    const uint32_t sbase = lds0 + CF::ASZ + stg * CF::XSZ;
#pragma unroll
    for (int e = 0; e < CF::XPT; e++) {
      const long col = wgc + 32 * wave + (lane >> 1);
      const int k = 32 * kc + 8 * e + 4 * (lane & 1);
      const bool ok = col < n2 && k < n;
      const int voff = ok ? (int)((k + (long)n * col) * 4) : (int)0x80000000;
      copy16_to_lds(rsX, sbase + (e * CF::COLS + 32 * wave) * 32, voff);
    }
  };

  const int q = lane >> 4, r16 = lane & 15;
  issue_a(0);
  issue_x(0, 0);
  const bool two = total > 1;
  if (two) issue_x(1, 1);
  f32x4 acc[MT][2];
#pragma unroll
  for (int t = 0; t < MT; t++)
#pragma unroll
    for (int u = 0; u < 2; u++) acc[t][u] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
  for (int g = 0; g < total; g++) {
    // A(g) and X(g) are in LDS for every wave; only X(g + 1), issued after
    // them, may still be in flight
    if (g + 1 < total)
      wait_vm_barrier<CF::XPT>();
    else
      wait_vm_barrier<0>();
    const bf16x8 *sa = reinterpret_cast<const bf16x8 *>(smem);
    const float *sx = reinterpret_cast<const float *>(smem + CF::ASZ + (g & 1) * CF::XSZ);
    bf16x8 b[2][NT];
    auto split_tile = [&](int u) {
      const float *bp = sx + ((size_t)q * CF::COLS + 32 * wave + 16 * u + r16) * 8;
      const float4 lo = *reinterpret_cast<const float4 *>(bp);
      const float4 hi = *reinterpret_cast<const float4 *>(bp + 4);
      const float x[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
      for (int e = 0; e < 8; e++) {
        __bf16 t0, t1, t2;
        split_bf16(x[e], t0, t1, t2);
        b[u][0][e] = t0;
        if constexpr (NT > 1) b[u][1][e] = t1;
        if constexpr (NT > 2) b[u][2][e] = t2;
      }
    };
    split_tile(0);
#pragma unroll
    for (int t = 0; t < MT; t++) {
      bf16x8 a[NT];
#pragma unroll
      for (int tm = 0; tm < NT; tm++) a[tm] = sa[(tm * MT + t) * 64 + lane];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        if (t == 0 && u > 0) split_tile(u);
        if constexpr (NT > 2) {
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[u][0], acc[t][u], 0, 0, 0);
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[u][1], acc[t][u], 0, 0, 0);
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[u][2], acc[t][u], 0, 0, 0);
        }
        if constexpr (NT > 1) {
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[u][0], acc[t][u], 0, 0, 0);
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[u][1], acc[t][u], 0, 0, 0);
        }
        acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[u][0], acc[t][u], 0, 0, 0);
      }
    }
    if (g % nkc == nkc - 1) {
      const int tl = tile_of(g);
      store_tile<MT, 2>(acc, out + out_fs * (tl / ncb), n, (long)(tl % ncb) * CF::COLS + 32 * wave + r16, q, rotate,
                        accumulate);
#pragma unroll
      for (int t = 0; t < MT; t++)
#pragma unroll
        for (int u = 0; u < 2; u++) acc[t][u] = (f32x4){0.f, 0.f, 0.f, 0.f};
    }
    // every wave is past chunk g: its A stage and X stage may be refilled
    plain_barrier();
    if (g + 1 < total) issue_a((g + 1) % nkc);
    if (g + 2 < total) issue_x(g + 2, g & 1);
  }
}

// Wave-grid variant (persistent): 8 waves as 2 row halves x 4 column groups
// of 64 columns.  A wave reads only its half of the split V / V^-1 chunk
// from LDS and feeds each fragment to four column tiles, which cuts the LDS
// reads per chunk by a third; the two waves of a column group split the same
// field values (twice the split VALU).  Same LDS stages as the default.
template <int MT, int NT>
__global__ __launch_bounds__(512, 2) void k_modeprod_split_rc(const uint16_t *__restrict__ A3,
                                                              const float *__restrict__ in, size_t in_fs,
                                                              float *__restrict__ out, size_t out_fs, int n,
                                                              int rotate, int accumulate, int a_chunk, int nfields) {
  using CF = mps_cfg<MT, NT, 8, 2>;  // COLS = 256, same staging as the default kernel
  constexpr int MTH = (MT + 1) / 2;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rh = wave & 1, cg = wave >> 1;
  const int t0 = rh * MTH;
  const long n2 = (long)n * n;
  const int nkc = (n + 31) / 32;
  const int ncb = (int)((n2 + CF::COLS - 1) / CF::COLS);
  const int ntiles = ncb * nfields;
  int tile = blockIdx.x;
  if (tile >= ntiles) return;  // uniform
  const i32x4m rsA = mk_rsrc(A3, (uint32_t)((size_t)nkc * a_chunk));
  const uint32_t lds0 = (uint32_t)(uintptr_t)smem;

  auto issue = [&](int tl, int kc, int stg) {
    const int f = tl / ncb;
    const long wgc = (long)(tl % ncb) * CF::COLS;
    i32x4m rsX = mk_rsrc(in + in_fs * f, (uint32_t)((size_t)n * n2 * 4));
    rsX.x = __builtin_amdgcn_readfirstlane(rsX.x);
    rsX.y = __builtin_amdgcn_readfirstlane(rsX.y);
    rsX.z = __builtin_amdgcn_readfirstlane(rsX.z);
    const uint32_t sbase = lds0 + stg * CF::STAGE;
#pragma unroll
    for (int e = 0; e < CF::APT; e++)
      copy16_to_lds(rsA, sbase + (e * CF::THREADS + wave * 64) * 16,
                    (int)((size_t)kc * a_chunk + ((size_t)e * CF::THREADS + threadIdx.x) * 16));
#pragma unroll
    for (int e = 0; e < CF::XPT; e++) {
      const long col = wgc + 32 * wave + (lane >> 1);
      const int k = 32 * kc + 8 * e + 4 * (lane & 1);
      const bool ok = col < n2 && k < n;
      const int voff = ok ? (int)((k + (long)n * col) * 4) : (int)0x80000000;
      copy16_to_lds(rsX, sbase + CF::ASZ + (e * CF::THREADS + wave * 64) * 16, voff);
    }
  };

  const int q = lane >> 4, r16 = lane & 15;
  issue(tile, 0, 0);
  int g = 0;
#pragma unroll 1
  for (; tile < ntiles; tile += gridDim.x) {
    const int f = tile / ncb;
    const long wgcol = (long)(tile % ncb) * CF::COLS;
    f32x4 acc[MTH][4];
#pragma unroll
    for (int t = 0; t < MTH; t++)
#pragma unroll
      for (int u = 0; u < 4; u++) acc[t][u] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (int kc = 0; kc < nkc; kc++, g++) {
      wait_all_copies_barrier();
      if (kc + 1 < nkc)
        issue(tile, kc + 1, (g + 1) & 1);
      else if (tile + (int)gridDim.x < ntiles)
        issue(tile + gridDim.x, 0, (g + 1) & 1);
      const char *stg = smem + (g & 1) * CF::STAGE;
      const bf16x8 *sa = reinterpret_cast<const bf16x8 *>(stg);
      const float *sx = reinterpret_cast<const float *>(stg + CF::ASZ);
      bf16x8 b[4][NT];
      auto split_tile = [&](int u) {
        const float *bp = sx + ((size_t)q * CF::COLS + 64 * cg + 16 * u + r16) * 8;
        const float4 lo = *reinterpret_cast<const float4 *>(bp);
        const float4 hi = *reinterpret_cast<const float4 *>(bp + 4);
        const float x[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
        for (int e = 0; e < 8; e++) {
          __bf16 t0_, t1_, t2_;
          split_bf16(x[e], t0_, t1_, t2_);
          b[u][0][e] = t0_;
          if constexpr (NT > 1) b[u][1][e] = t1_;
          if constexpr (NT > 2) b[u][2][e] = t2_;
        }
      };
      split_tile(0);
#pragma unroll
      for (int tt = 0; tt < MTH; tt++) {
        const int t = t0 + tt;
        if (t >= MT) break;  // wave-uniform
        bf16x8 a[NT];
#pragma unroll
        for (int tm = 0; tm < NT; tm++) a[tm] = sa[(tm * MT + t) * 64 + lane];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          if (tt == 0 && u > 0) split_tile(u);
          if constexpr (NT > 2) {
            acc[tt][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[u][0], acc[tt][u], 0, 0, 0);
            acc[tt][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[u][1], acc[tt][u], 0, 0, 0);
            acc[tt][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[u][2], acc[tt][u], 0, 0, 0);
          }
          if constexpr (NT > 1) {
            acc[tt][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[u][0], acc[tt][u], 0, 0, 0);
            acc[tt][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[u][1], acc[tt][u], 0, 0, 0);
          }
          acc[tt][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[u][0], acc[tt][u], 0, 0, 0);
        }
      }
    }
    if (accumulate)
      store_rows_<MTH, 4, true>(acc, out + out_fs * f, n, wgcol + 64 * cg + r16, q, rotate, t0, MT);
    else
      store_rows_<MTH, 4, false>(acc, out + out_fs * f, n, wgcol + 64 * cg + r16, q, rotate, t0, MT);
  }
}

// FP16 split-2 (persistent, the default kernel's tiling and staging): A is
// stored as two FP16 terms of A / s, s a power of two (a_inv = 1 / s); each
// lane's column chunk of 32 field values (its 8 and those of the three lanes
// with the same column) is scaled by 2^(14 - e), e the exponent of its largest
// magnitude, and split into x0 = fp16(xs), x1 = fp16(xs - x0).  Three
// v_mfma_f32_16x16x32_f16 (a1 x0 + a0 x1 + a0 x0) per tile and chunk go into a
// chunk-local FP32 tile, added to the accumulator times 2^(e - 14) / s (exact
// powers of two).  Values below 2^-14 of their chunk's scale lose relative
// precision, which is below FP32 rounding of the chunk's sum.
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

template <int MT>
__global__ __launch_bounds__(512, 2) void k_modeprod_f16_p(const uint16_t *__restrict__ A2, const float *__restrict__ in,
                                                           size_t in_fs, float *__restrict__ out, size_t out_fs, int n,
                                                           int rotate, int accumulate, int a_chunk, int nfields,
                                                           float a_inv) {
  constexpr int NT = 2;
  using CF = mps_cfg<MT, NT, 8, 2>;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long n2 = (long)n * n;
  const int nkc = (n + 31) / 32;
  const int ncb = (int)((n2 + CF::COLS - 1) / CF::COLS);
  const int ntiles = ncb * nfields;
  int tile = blockIdx.x;
  if (tile >= ntiles) return;  // uniform
  const i32x4m rsA = mk_rsrc(A2, (uint32_t)((size_t)nkc * a_chunk));
  const uint32_t lds0 = (uint32_t)(uintptr_t)smem;

  auto issue = [&](int tl, int kc, int stg) {
    const int f = tl / ncb;
    const long wgc = (long)(tl % ncb) * CF::COLS;
    i32x4m rsX = mk_rsrc(in + in_fs * f, (uint32_t)((size_t)n * n2 * 4));
    rsX.x = __builtin_amdgcn_readfirstlane(rsX.x);
    rsX.y = __builtin_amdgcn_readfirstlane(rsX.y);
    rsX.z = __builtin_amdgcn_readfirstlane(rsX.z);
    const uint32_t sbase = lds0 + stg * CF::STAGE;
#pragma unroll
    for (int e = 0; e < CF::APT; e++)
      copy16_to_lds(rsA, sbase + (e * CF::THREADS + wave * 64) * 16,
                    (int)((size_t)kc * a_chunk + ((size_t)e * CF::THREADS + threadIdx.x) * 16));
#pragma unroll
    for (int e = 0; e < CF::XPT; e++) {
      const long col = wgc + 32 * wave + (lane >> 1);
      const int k = 32 * kc + 8 * e + 4 * (lane & 1);
      const bool ok = col < n2 && k < n;
      const int voff = ok ? (int)((k + (long)n * col) * 4) : (int)0x80000000;
      copy16_to_lds(rsX, sbase + CF::ASZ + (e * CF::THREADS + wave * 64) * 16, voff);
    }
  };

  const int q = lane >> 4, r16 = lane & 15;
  issue(tile, 0, 0);
  int g = 0;
#pragma unroll 1
  for (; tile < ntiles; tile += gridDim.x) {
    const int f = tile / ncb;
    const long wgcol = (long)(tile % ncb) * CF::COLS;
    f32x4 acc[MT][2];
#pragma unroll
    for (int t = 0; t < MT; t++)
#pragma unroll
      for (int u = 0; u < 2; u++) acc[t][u] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (int kc = 0; kc < nkc; kc++, g++) {
      wait_all_copies_barrier();
      if (kc + 1 < nkc)
        issue(tile, kc + 1, (g + 1) & 1);
      else if (tile + (int)gridDim.x < ntiles)
        issue(tile + gridDim.x, 0, (g + 1) & 1);
      const char *stg = smem + (g & 1) * CF::STAGE;
      const f16x8 *sa = reinterpret_cast<const f16x8 *>(stg);
      const float *sx = reinterpret_cast<const float *>(stg + CF::ASZ);
      f16x8 b[2][2];
      float back[2];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const float *bp = sx + ((size_t)q * CF::COLS + 32 * wave + 16 * u + r16) * 8;
        const float4 lo = *reinterpret_cast<const float4 *>(bp);
        const float4 hi = *reinterpret_cast<const float4 *>(bp + 4);
        const float x[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
        float m = 0.f;
#pragma unroll
        for (int e = 0; e < 8; e++) m = fmaxf(m, fabsf(x[e]));
        // the column's 32 values of this chunk sit in lanes r16 + 16 q
        m = fmaxf(m, __shfl_xor(m, 16, 64));
        m = fmaxf(m, __shfl_xor(m, 32, 64));
        const int ex = m > 0.f ? __builtin_amdgcn_frexp_expf(m) : 0;
        const float sc = __builtin_ldexpf(1.0f, 14 - ex);
        back[u] = __builtin_ldexpf(a_inv, ex - 14);
#pragma unroll
        for (int e = 0; e < 8; e++) {
          const float xs = x[e] * sc;
          const _Float16 h0 = (_Float16)xs;
          b[u][0][e] = h0;
          b[u][1][e] = (_Float16)(xs - (float)h0);
        }
      }
#pragma unroll
      for (int t = 0; t < MT; t += 2) {
        f16x8 a[2][2];
#pragma unroll
        for (int d = 0; d < 2; d++)
          if (t + d < MT)
#pragma unroll
            for (int tm = 0; tm < 2; tm++) a[d][tm] = sa[(tm * MT + t + d) * 64 + lane];
        f32x4 dch[2][2];
#pragma unroll
        for (int d = 0; d < 2; d++)
#pragma unroll
          for (int u = 0; u < 2; u++)
            if (t + d < MT) dch[d][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[d][1], b[u][0], (f32x4){0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
        for (int d = 0; d < 2; d++)
#pragma unroll
          for (int u = 0; u < 2; u++)
            if (t + d < MT) dch[d][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[d][0], b[u][1], dch[d][u], 0, 0, 0);
#pragma unroll
        for (int d = 0; d < 2; d++)
#pragma unroll
          for (int u = 0; u < 2; u++)
            if (t + d < MT) dch[d][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[d][0], b[u][0], dch[d][u], 0, 0, 0);
#pragma unroll
        for (int d = 0; d < 2; d++)
#pragma unroll
          for (int u = 0; u < 2; u++)
            if (t + d < MT)
#pragma unroll
              for (int i = 0; i < 4; i++) acc[t + d][u][i] = fmaf(dch[d][u][i], back[u], acc[t + d][u][i]);
      }
    }
    store_tile<MT, 2>(acc, out + out_fs * f, n, wgcol + 32 * wave + r16, q, rotate, accumulate);
  }
}

// FP16 split-2 with field chunks prefetched two chunks ahead: the V / V^-1
// operand (L2-resident) keeps a two-slot LDS ring, the field operand (from
// HBM) a three-slot ring, so while chunk g is multiplied the copies of field
// chunks g + 1 and g + 2 are in flight (2 x 32 KB per CU instead of one).
// Arithmetic, operand layout and results are those of k_modeprod_f16_p.
// Chunk g of a workgroup is K chunk g % nkc of its tile
// blockIdx.x + (g / nkc) gridDim.x.  Per chunk, after the previous chunk's
// reads: wait until the copies of chunk g are complete (vmcnt counts this
// wave's copies, which complete in issue order; only field chunk g + 1's XPT
// copies may remain), barrier, then issue V chunk g + 1 and field chunk g + 2
// into the slots that chunk g - 1 used.  After a tile's stores, and where no
// field chunk g + 1 was issued, the wait drains everything.
// XL = 0: a field chunk sits in LDS as [k/8][col][8 k] and each copy
// instruction takes 32 bytes from each of 32 columns.  XL = 1: it sits as
// [col][k/4 ^ (col & 7)] in 16-byte pieces, so each copy instruction takes
// 8 whole 128-byte lines (8 columns x 32 k) and the XOR keeps the 16 lanes of
// an operand read on distinct banks as in XL = 0.
// XL = 2 contracts the SLOWEST axis instead: X(k, col) = in[col + n^2 k]
// (col = i + n j), written rotated as out[m + n col] (or in place of k,
// out[col + n^2 m], without rotation).  Three rotated products still run
// x, y, z -> x', y', z' (z first) and end in the same spectral layout.  The
// 32 rows of a field chunk are 32 runs of 256 consecutive floats (1 KB), so
// every copy instruction reads whole 128-byte lines and no line is shared
// between K chunks: the contract-fastest kernels' 784-byte columns start
// mid-line for 7 of 8 columns, and the half line that the next chunk needs
// is fetched again when it has left the cache.  A row sits in LDS as 64
// 16-byte pieces, piece L holding the row's piece L ^ 4 (r / 8) for row r of
// the chunk: a lane reads the 8 rows 8 q .. 8 q + 7 of its column, and the
// XOR puts the four q groups of an operand read on four distinct groups of
// 16 banks.
template <int MT>
struct f16r_cfg {
  static constexpr int THREADS = 512, WAVES = 8, COLS = 256;
  static constexpr int APIECES = 2 * MT * 64;             // 16-byte pieces of one V chunk (2 terms)
  static constexpr int AFULL = APIECES / THREADS;         // passes every thread copies
  static constexpr int AREM = APIECES - AFULL * THREADS;  // pieces of the last pass
  static constexpr int ASZ = APIECES * 16;
  static constexpr int XSZ = 32 * COLS * 4;
  static constexpr int XPT = XSZ / 16 / THREADS;
  static constexpr int LDS = 2 * ASZ + 3 * XSZ;
  static constexpr bool FITS = LDS <= 160 * 1024;  // MT <= 16 (n <= 256)
  static_assert(AREM % 64 == 0, "the last pass of V pieces is whole waves");
  static_assert(XPT == 4, "the partial wait below counts 4 copies");
};

__device__ __forceinline__ void wait_copies4_barrier() { asm volatile("s_waitcnt vmcnt(4)\n\ts_barrier" : : : "memory"); }

template <int MT, int XL>
__global__ __launch_bounds__(512, 2) void k_modeprod_f16_r3(const uint16_t *__restrict__ A2, const float *__restrict__ in,
                                                            size_t in_fs, float *__restrict__ out, size_t out_fs, int n,
                                                            int rotate, int accumulate, int a_chunk, int nfields,
                                                            float a_inv) {
  using CF = f16r_cfg<MT>;
  static_assert(CF::FITS, "LDS budget");
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long n2 = (long)n * n;
  const int nkc = (n + 31) / 32;
  const int ncb = (int)((n2 + CF::COLS - 1) / CF::COLS);
  const int ntiles = ncb * nfields;
  if ((int)blockIdx.x >= ntiles) return;  // uniform
  const int my_tiles = (ntiles - (int)blockIdx.x + (int)gridDim.x - 1) / (int)gridDim.x;
  const int total = my_tiles * nkc;
  const i32x4m rsA = mk_rsrc(A2, (uint32_t)((size_t)nkc * a_chunk));
  const uint32_t lds0 = (uint32_t)(uintptr_t)smem;

  auto issue_a = [&](int gg) {
    const int kc = gg % nkc;
    const uint32_t sbase = lds0 + (gg & 1) * CF::ASZ;
#pragma unroll
    for (int e = 0; e < CF::AFULL; e++)
      copy16_to_lds(rsA, sbase + (e * CF::THREADS + wave * 64) * 16,
                    (int)((size_t)kc * a_chunk + ((size_t)e * CF::THREADS + threadIdx.x) * 16));
    if (CF::AREM > 0 && wave < CF::AREM / 64)  // wave-uniform
      copy16_to_lds(rsA, sbase + (CF::AFULL * CF::THREADS + wave * 64) * 16,
                    (int)((size_t)kc * a_chunk + ((size_t)CF::AFULL * CF::THREADS + threadIdx.x) * 16));
  };
  auto issue_x = [&](int gg) {
    const int tl = (int)blockIdx.x + (gg / nkc) * (int)gridDim.x;
    const int kc = gg % nkc;
    const int f = tl / ncb;
    const long wgc = (long)(tl % ncb) * CF::COLS;
    i32x4m rsX = mk_rsrc(in + in_fs * f, (uint32_t)((size_t)n * n2 * 4));
    rsX.x = __builtin_amdgcn_readfirstlane(rsX.x);
    rsX.y = __builtin_amdgcn_readfirstlane(rsX.y);
    rsX.z = __builtin_amdgcn_readfirstlane(rsX.z);
    const uint32_t sbase = lds0 + 2 * CF::ASZ + (gg % 3) * CF::XSZ;
    if constexpr (XL == 2) {
      // row r = 8 e + wave of the chunk: one 1 KB run, lane L -> LDS piece L
      // holding the run's piece L ^ 4 e (n^2 is a multiple of 4: a piece is
      // inside or outside the matrix as a whole)
#pragma unroll
      for (int e = 0; e < CF::XPT; e++) {
        const int r = 8 * e + wave;
        const int k = 32 * kc + r;
        const long col = wgc + 4 * (lane ^ (4 * e));
        const bool ok = col < n2 && k < n;
        const int voff = ok ? (int)((col + n2 * k) * 4) : (int)0x80000000;
        copy16_to_lds(rsX, sbase + r * (CF::COLS * 4), voff);
      }
      return;
    }
#pragma unroll
    for (int e = 0; e < CF::XPT; e++) {
      long col;
      int k;
      if constexpr (XL) {
        const int c = 64 * e + 8 * wave + (lane >> 3);  // piece index e 512 + 64 wave + lane = 8 c + slot
        col = wgc + c;
        k = 32 * kc + 4 * ((lane & 7) ^ (c & 7));
      } else {
        col = wgc + 32 * wave + (lane >> 1);
        k = 32 * kc + 8 * e + 4 * (lane & 1);
      }
      const bool ok = col < n2 && k < n;
      const int voff = ok ? (int)((k + (long)n * col) * 4) : (int)0x80000000;
      copy16_to_lds(rsX, sbase + (e * CF::THREADS + wave * 64) * 16, voff);
    }
  };

  const int q = lane >> 4, r16 = lane & 15;
  issue_x(0);
  issue_a(0);
  if (total > 1) issue_x(1);
  f32x4 acc[MT][2];
  // XL = 2: one scale per column and tile instead of one per chunk: the
  // column's values are scaled by 2^(12 - cexp[u]), cexp the exponent of the
  // largest magnitude seen so far in the column (-1000 while all are zero),
  // so the three MFMAs of every chunk accumulate straight into acc.  A chunk
  // whose largest value exceeds 2^(cexp + 3) (its scaled values would leave
  // FP16's range) raises cexp and rescales the column's acc first, in a
  // branch taken only when a lane of the wave needs it.  The tile's acc is
  // multiplied by 2^(cexp - 12) / s once, in the epilogue.
  int cexp[2] = {-1000, -1000};
  int xs = 0;  // g % 3
#pragma unroll 1
  for (int g = 0; g < total; g++, xs = xs == 2 ? 0 : xs + 1) {
    const int kc = g % nkc;
    const int tile = (int)blockIdx.x + (g / nkc) * (int)gridDim.x;
    if (kc == 0) {
#pragma unroll
      for (int t = 0; t < MT; t++)
#pragma unroll
        for (int u = 0; u < 2; u++) acc[t][u] = (f32x4){0.f, 0.f, 0.f, 0.f};
      cexp[0] = cexp[1] = -1000;
    }
    if (g + 1 < total && (kc != 0 || g == 0))  // uniform
      wait_copies4_barrier();
    else
      wait_all_copies_barrier();
    if (g + 1 < total) issue_a(g + 1);
    if (g + 2 < total) issue_x(g + 2);
    const f16x8 *sa = reinterpret_cast<const f16x8 *>(smem + (g & 1) * CF::ASZ);
    const float *sx = reinterpret_cast<const float *>(smem + 2 * CF::ASZ + xs * CF::XSZ);
    f16x8 b[2][2];
    float back[2];
#pragma unroll
    for (int u = 0; u < 2; u++) {
      float x[8];
      if constexpr (XL == 2) {
        const int c = 32 * wave + 16 * u + r16;
        const float *rp = sx + 8 * q * CF::COLS + 4 * ((c >> 2) ^ (4 * q)) + (c & 3);
#pragma unroll
        for (int e = 0; e < 8; e++) x[e] = rp[e * CF::COLS];
      } else {
        float4 lo, hi;
        if constexpr (XL) {
          const int c = 32 * wave + 16 * u + r16;
          const float4 *cp = reinterpret_cast<const float4 *>(sx) + 8 * c;
          lo = cp[(2 * q) ^ (c & 7)];
          hi = cp[(2 * q + 1) ^ (c & 7)];
        } else {
          const float *bp = sx + ((size_t)q * CF::COLS + 32 * wave + 16 * u + r16) * 8;
          lo = *reinterpret_cast<const float4 *>(bp);
          hi = *reinterpret_cast<const float4 *>(bp + 4);
        }
        x[0] = lo.x, x[1] = lo.y, x[2] = lo.z, x[3] = lo.w, x[4] = hi.x, x[5] = hi.y, x[6] = hi.z, x[7] = hi.w;
      }
      float m = 0.f;
#pragma unroll
      for (int e = 0; e < 8; e++) m = fmaxf(m, fabsf(x[e]));
      m = fmaxf(m, __shfl_xor(m, 16, 64));
      m = fmaxf(m, __shfl_xor(m, 32, 64));
      float sc;
      if constexpr (XL == 2) {
        const int ex = m > 0.f ? __builtin_amdgcn_frexp_expf(m) : -1000;
        const bool grow = ex > cexp[u] + 3;
        if (__any(grow)) {  // wave-uniform; lanes that keep their scale multiply by 1
          const float f = grow ? __builtin_ldexpf(1.0f, cexp[u] - ex) : 1.0f;
#pragma unroll
          for (int t = 0; t < MT; t++) acc[t][u] = acc[t][u] * f;
        }
        if (grow) cexp[u] = ex;
        sc = __builtin_ldexpf(1.0f, 12 - cexp[u]);
      } else {
        const int ex = m > 0.f ? __builtin_amdgcn_frexp_expf(m) : 0;
        sc = __builtin_ldexpf(1.0f, 14 - ex);
        back[u] = __builtin_ldexpf(a_inv, ex - 14);
      }
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const float xsv = x[e] * sc;
        const _Float16 h0 = (_Float16)xsv;
        b[u][0][e] = h0;
        b[u][1][e] = (_Float16)(xsv - (float)h0);
      }
    }
    if constexpr (XL == 2) {
#pragma unroll
      for (int t = 0; t < MT; t++) {
        f16x8 a0 = sa[t * 64 + lane], a1 = sa[(MT + t) * 64 + lane];
#pragma unroll
        for (int u = 0; u < 2; u++) {
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, b[u][0], acc[t][u], 0, 0, 0);
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, b[u][1], acc[t][u], 0, 0, 0);
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, b[u][0], acc[t][u], 0, 0, 0);
        }
      }
    } else {
#pragma unroll
      for (int t = 0; t < MT; t += 2) {
        f16x8 a[2][2];
#pragma unroll
        for (int d = 0; d < 2; d++)
          if (t + d < MT)
#pragma unroll
            for (int tm = 0; tm < 2; tm++) a[d][tm] = sa[(tm * MT + t + d) * 64 + lane];
        f32x4 dch[2][2];
#pragma unroll
        for (int d = 0; d < 2; d++)
#pragma unroll
          for (int u = 0; u < 2; u++)
            if (t + d < MT) dch[d][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[d][1], b[u][0], (f32x4){0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
        for (int d = 0; d < 2; d++)
#pragma unroll
          for (int u = 0; u < 2; u++)
            if (t + d < MT) dch[d][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[d][0], b[u][1], dch[d][u], 0, 0, 0);
#pragma unroll
        for (int d = 0; d < 2; d++)
#pragma unroll
          for (int u = 0; u < 2; u++)
            if (t + d < MT) dch[d][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[d][0], b[u][0], dch[d][u], 0, 0, 0);
#pragma unroll
        for (int d = 0; d < 2; d++)
#pragma unroll
          for (int u = 0; u < 2; u++)
            if (t + d < MT)
#pragma unroll
              for (int i = 0; i < 4; i++) acc[t + d][u][i] = fmaf(dch[d][u][i], back[u], acc[t + d][u][i]);
      }
    }
    if (kc == nkc - 1) {
      const int f = tile / ncb;
      const long wgcol = (long)(tile % ncb) * CF::COLS;
      if constexpr (XL == 2) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
          const float bk = __builtin_ldexpf(a_inv, cexp[u] - 12);
#pragma unroll
          for (int t = 0; t < MT; t++) acc[t][u] = acc[t][u] * bk;
        }
        if (rotate)
          store_tile_m4<MT, 2>(acc, out + out_fs * f, n, wgcol + 32 * wave + r16, q, accumulate);
        else  // k replaced by m in place: the strides of the other kernels' rotated store
          store_tile<MT, 2>(acc, out + out_fs * f, n, wgcol + 32 * wave + r16, q, 1, accumulate);
      } else {
        store_tile<MT, 2>(acc, out + out_fs * f, n, wgcol + 32 * wave + r16, q, rotate, accumulate);
      }
    }
  }
}

// Register-prefetch variant (persistent): the field operand goes straight
// from HBM into registers with compiler-visible buffer loads issued two K
// chunks ahead, so HBM latency is covered by two chunks of MFMA work; only
// the split V / V^-1 operand (L2-resident, shared by the waves) passes
// through LDS, staged through registers one chunk ahead in a two-stage ring
// (2 x 40 KB for MT = 13).  A workgroup's chunk sequence runs over its tiles
// without a break: chunk g is K chunk g % nkc of the workgroup's tile
// blockIdx.x + (g / nkc) gridDim.x.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

template <int MT, int NT>
struct msr_cfg {
  static constexpr int WAVES = 8;
  static constexpr int THREADS = 64 * WAVES;
  static constexpr int COLS = 32 * WAVES;
  static constexpr int APIECES = NT * MT * 64;
  static constexpr int APT = (APIECES + THREADS - 1) / THREADS;
  static constexpr int ASZ = APT * THREADS * 16;
  static constexpr int LDS = 2 * ASZ;
};

__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" : : : "memory"); }

template <int MT, int NT>
__global__ __launch_bounds__(512, 1) void k_modeprod_sr(const uint16_t *__restrict__ A3, const float *__restrict__ in,
                                                        size_t in_fs, float *__restrict__ out, size_t out_fs, int n,
                                                        int rotate, int accumulate, int a_chunk, int nfields) {
  using CF = msr_cfg<MT, NT>;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane >> 4, r16 = lane & 15;
  const long n2 = (long)n * n;
  const int nkc = (n + 31) / 32;
  const int ncb = (int)((n2 + CF::COLS - 1) / CF::COLS);
  const int ntiles = ncb * nfields;
  if ((int)blockIdx.x >= ntiles) return;  // uniform
  const int my_tiles = (ntiles - (int)blockIdx.x + (int)gridDim.x - 1) / (int)gridDim.x;
  const int total = my_tiles * nkc;  // chunks of this workgroup
  const __amdgpu_buffer_rsrc_t rA =
      __builtin_amdgcn_make_buffer_rsrc((void *)A3, 0, (int)((size_t)nkc * a_chunk), 0x00020000);

  // A chunk kc -> registers (this thread's pieces)
  auto load_a = [&](int kc, u32x4 (&ra)[CF::APT]) {
#pragma unroll
    for (int e = 0; e < CF::APT; e++)
      ra[e] = __builtin_amdgcn_raw_buffer_load_b128(rA, (int)((size_t)kc * a_chunk + (size_t)(e * CF::THREADS + threadIdx.x) * 16), 0, 0);
  };
  auto store_a = [&](int stg, const u32x4 (&ra)[CF::APT]) {
    u32x4 *d = reinterpret_cast<u32x4 *>(smem + stg * CF::ASZ);
#pragma unroll
    for (int e = 0; e < CF::APT; e++) d[e * CF::THREADS + threadIdx.x] = ra[e];
  };
  // field chunk g -> registers: per column tile u, k = 32 kc + 8 q + [0, 8)
  auto load_x = [&](int g, u32x4 (&rx)[4]) {
    if (g >= total) return;
    const int tl = (int)blockIdx.x + (g / nkc) * (int)gridDim.x, kc = g % nkc;
    const int f = tl / ncb;
    const long wgc = (long)(tl % ncb) * CF::COLS;
    const __amdgpu_buffer_rsrc_t rX =
        __builtin_amdgcn_make_buffer_rsrc((void *)(in + in_fs * f), 0, (int)((size_t)n * n2 * 4), 0x00020000);
    const int k = 32 * kc + 8 * q;
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const long col = wgc + 32 * wave + 16 * u + r16;
      const bool ok = col < n2 && k < n;
      const int off = ok ? (int)((k + (long)n * col) * 4) : (int)0x80000000;
      rx[2 * u] = __builtin_amdgcn_raw_buffer_load_b128(rX, off, 0, 0);
      rx[2 * u + 1] = __builtin_amdgcn_raw_buffer_load_b128(rX, ok && k + 4 < n ? off + 16 : (int)0x80000000, 0, 0);
    }
  };

  f32x4 acc[MT][2];
#pragma unroll
  for (int t = 0; t < MT; t++)
#pragma unroll
    for (int u = 0; u < 2; u++) acc[t][u] = (f32x4){0.f, 0.f, 0.f, 0.f};

  u32x4 ra[CF::APT];
  u32x4 x0[4], x1[4], x2[4];
  load_a(0, ra);
  load_x(0, x0);
  load_x(1, x1);
  store_a(0, ra);
  if (total > 1) load_a(1 % nkc, ra);

  // one chunk: consumes field slot xs (chunk g), refills slot xl with g + 2
  auto chunk = [&](int g, u32x4 (&xs)[4], u32x4 (&xl)[4]) {
    const int kc = g % nkc;
    // stage g & 1 is complete, and every wave is past chunk g - 1, whose
    // stage the next store overwrites
    lds_barrier();
    if (g + 1 < total) {
      store_a((g + 1) & 1, ra);
      if (g + 2 < total) load_a((g + 2) % nkc, ra);
    }
    load_x(g + 2, xl);
    const bf16x8 *sa = reinterpret_cast<const bf16x8 *>(smem + (g & 1) * CF::ASZ);
    bf16x8 b[2][NT];
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const u32x4 lo = xs[2 * u], hi = xs[2 * u + 1];
      const float x[8] = {__uint_as_float(lo.x), __uint_as_float(lo.y), __uint_as_float(lo.z), __uint_as_float(lo.w),
                          __uint_as_float(hi.x), __uint_as_float(hi.y), __uint_as_float(hi.z), __uint_as_float(hi.w)};
#pragma unroll
      for (int e = 0; e < 8; e++) {
        __bf16 t0, t1, t2;
        split_bf16(x[e], t0, t1, t2);
        b[u][0][e] = t0;
        if constexpr (NT > 1) b[u][1][e] = t1;
        if constexpr (NT > 2) b[u][2][e] = t2;
      }
    }
#pragma unroll
    for (int t = 0; t < MT; t++) {
      bf16x8 a[NT];
#pragma unroll
      for (int tm = 0; tm < NT; tm++) a[tm] = sa[(tm * MT + t) * 64 + lane];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        if constexpr (NT > 2) {
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[u][0], acc[t][u], 0, 0, 0);
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[u][1], acc[t][u], 0, 0, 0);
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[u][2], acc[t][u], 0, 0, 0);
        }
        if constexpr (NT > 1) {
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[u][0], acc[t][u], 0, 0, 0);
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[u][1], acc[t][u], 0, 0, 0);
        }
        acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[u][0], acc[t][u], 0, 0, 0);
      }
    }
    if (kc == nkc - 1) {
      const int tl = (int)blockIdx.x + (g / nkc) * (int)gridDim.x;
      const int f = tl / ncb;
      const long wgcol = (long)(tl % ncb) * CF::COLS;
      store_tile<MT, 2>(acc, out + out_fs * f, n, wgcol + 32 * wave + r16, q, rotate, accumulate);
#pragma unroll
      for (int t = 0; t < MT; t++)
#pragma unroll
        for (int u = 0; u < 2; u++) acc[t][u] = (f32x4){0.f, 0.f, 0.f, 0.f};
    }
  };

#pragma unroll 1
  for (int g = 0; g < total; g += 3) {
    chunk(g, x0, x2);
    if (g + 1 < total) chunk(g + 1, x1, x0);
    if (g + 2 < total) chunk(g + 2, x2, x1);
  }
}

// ---- host side
// the arranged operand's chunk stride serves both workgroup shapes: the
// larger of the two copy granularities
template <int MT, int NT>
static size_t chunk_bytes() {
  const size_t a = mps_cfg<MT, NT, 8>::ASZ, b = mps_cfg<MT, NT, 4>::ASZ;
  return a > b ? a : b;
}

static size_t chunk_bytes_rt(int MT, int NT) {
  switch (MT) {
#define C1(X) \
  case X: return NT == 3 ? chunk_bytes<X, 3>() : chunk_bytes<X, 2>();
    C1(1) C1(2) C1(3) C1(4) C1(5) C1(6) C1(7) C1(8) C1(9) C1(10) C1(11) C1(12) C1(13) C1(14) C1(15) C1(16) C1(17)
#undef C1
    default: return 0;
  }
}

size_t split_A_bytes(int n, int terms) {
  const int MT = (n + 15) / 16, nkc = (n + 31) / 32;
  return chunk_bytes_rt(MT, terms) * nkc;
}

static uint16_t bf16_rn(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static double bf16_val(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return (double)f;
}

// A (column-major FP64) -> per chunk kc: [term][row tile t][lane][8 bf16],
// element e of lane l = term of A(16 t + l % 16, 32 kc + 8 (l / 16) + e),
// zero outside n x n; chunks padded to the kernel's copy granularity.
void arrange_split_A(int n, const double *Acm, uint16_t *A3, int terms) {
  const int MT = (n + 15) / 16, nkc = (n + 31) / 32;
  const size_t cb = chunk_bytes_rt(MT, terms);
  memset(A3, 0, cb * nkc);
  for (int kc = 0; kc < nkc; kc++) {
    uint16_t *ch = A3 + (size_t)kc * cb / 2;
    for (int t = 0; t < MT; t++)
      for (int l = 0; l < 64; l++)
        for (int e = 0; e < 8; e++) {
          const int m = 16 * t + (l & 15), k = 32 * kc + 8 * (l >> 4) + e;
          const double a = (m < n && k < n) ? Acm[m + (size_t)n * k] : 0.0;
          double r = a;
          for (int tm = 0; tm < terms; tm++) {
            const uint16_t h = bf16_rn((float)r);
            ch[(((size_t)tm * MT + t) * 64 + l) * 8 + e] = h;
            r -= bf16_val(h);
          }
        }
  }
}

static uint16_t f16_bits(double v) {
  const _Float16 h = (_Float16)(float)v;
  uint16_t b;
  memcpy(&b, &h, 2);
  return b;
}

float arrange_split_A_f16(int n, const double *Acm, uint16_t *A2) {
  const int MT = (n + 15) / 16, nkc = (n + 31) / 32;
  const size_t cb = chunk_bytes_rt(MT, 2);
  double amax = 0.0;
  for (size_t i = 0; i < (size_t)n * n; i++) amax = fmax(amax, fabs(Acm[i]));
  int ex = 0;
  if (amax > 0.0) (void)frexp(amax, &ex);       // amax in [2^(ex-1), 2^ex)
  const double s = ldexp(1.0, 14 - ex);         // max |A s| < 2^14
  memset(A2, 0, cb * nkc);
  for (int kc = 0; kc < nkc; kc++) {
    uint16_t *ch = A2 + (size_t)kc * cb / 2;
    for (int t = 0; t < MT; t++)
      for (int l = 0; l < 64; l++)
        for (int e = 0; e < 8; e++) {
          const int m = 16 * t + (l & 15), k = 32 * kc + 8 * (l >> 4) + e;
          const double a = (m < n && k < n) ? Acm[m + (size_t)n * k] * s : 0.0;
          const _Float16 h0 = (_Float16)(float)a;
          ch[(((size_t)0 * MT + t) * 64 + l) * 8 + e] = f16_bits((double)h0);
          ch[(((size_t)1 * MT + t) * 64 + l) * 8 + e] = f16_bits(a - (double)h0);
        }
  }
  return (float)(1.0 / s);
}

template <int MT>
static void f16_launch(etd_ctx *c, const uint16_t *A2, const float *in, size_t in_fs, float *out, size_t out_fs,
                       int nfields, int rotate, int accumulate, float a_inv) {
  using CF = mps_cfg<MT, 2, 8, 2>;
  static_assert(CF::LDS <= 160 * 1024, "LDS budget");
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void *)k_modeprod_f16_p<MT>, hipFuncAttributeMaxDynamicSharedMemorySize, CF::LDS);
    attr = true;
  }
  const long ncb = ((long)c->n * c->n + CF::COLS - 1) / CF::COLS;
  const long ntiles = ncb * nfields;
  const long nwg = ntiles < c->num_cu ? ntiles : c->num_cu;
  if constexpr (f16r_cfg<MT>::FITS) {
    if (c->split_kernel != 7) {  // 7: the two-slot kernel below
      using CR = f16r_cfg<MT>;
      static_assert(CR::COLS == CF::COLS && CR::THREADS == CF::THREADS, "same tiling");
      static bool attr_r = false;
      if (!attr_r) {
        (void)hipFuncSetAttribute((const void *)k_modeprod_f16_r3<MT, 0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  CR::LDS);
        (void)hipFuncSetAttribute((const void *)k_modeprod_f16_r3<MT, 1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  CR::LDS);
        (void)hipFuncSetAttribute((const void *)k_modeprod_f16_r3<MT, 2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  CR::LDS);
        attr_r = true;
      }
      if (rotate >= 2)  // contraction over the slowest axis: 2 rotated, 3 not
        k_modeprod_f16_r3<MT, 2><<<(unsigned)nwg, CR::THREADS, CR::LDS, c->stream>>>(
            A2, in, in_fs, out, out_fs, c->n, rotate == 2 ? 1 : 0, accumulate, (int)chunk_bytes<MT, 2>(), nfields,
            a_inv);
      else if (c->split_kernel != 5)
        k_modeprod_f16_r3<MT, 1><<<(unsigned)nwg, CR::THREADS, CR::LDS, c->stream>>>(
            A2, in, in_fs, out, out_fs, c->n, rotate, accumulate, (int)chunk_bytes<MT, 2>(), nfields, a_inv);
      else
        k_modeprod_f16_r3<MT, 0><<<(unsigned)nwg, CR::THREADS, CR::LDS, c->stream>>>(
            A2, in, in_fs, out, out_fs, c->n, rotate, accumulate, (int)chunk_bytes<MT, 2>(), nfields, a_inv);
      return;
    }
  }
  k_modeprod_f16_p<MT><<<(unsigned)nwg, CF::THREADS, CF::LDS, c->stream>>>(
      A2, in, in_fs, out, out_fs, c->n, rotate, accumulate, (int)chunk_bytes<MT, 2>(), nfields, a_inv);
}

template <int MT, int NT>
static void mps_launch(etd_ctx *c, const uint16_t *A3, const float *in, size_t in_fs, float *out, size_t out_fs,
                       int nfields, int rotate, int accumulate) {
  const int n = c->n;
  const long n2 = (long)n * n;
  auto go = [&](auto wtag) {
    constexpr int W = decltype(wtag)::value;
    using CF = mps_cfg<MT, NT, W>;
    static_assert(CF::LDS <= 160 * 1024, "LDS budget");
    static bool attr_set = false;
    if (!attr_set) {
      (void)hipFuncSetAttribute((const void *)k_modeprod_split<MT, NT, W>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                CF::LDS);
      (void)hipFuncSetAttribute((const void *)k_modeprod_split_p<MT, NT, W>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, CF::LDS);
      (void)hipFuncSetAttribute((const void *)k_modeprod_split_p<MT, NT, W, 2, 0>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, CF::LDS);
      attr_set = true;
    }
    const long ncb = (n2 + CF::COLS - 1) / CF::COLS;
    if constexpr (W == 8) {
      if (c->split_kernel == 1) {
        using CR = msr_cfg<MT, NT>;
        static_assert(CR::COLS == CF::COLS && CR::ASZ == CF::ASZ, "same tiling");
        static bool attr_r = false;
        if (!attr_r) {
          (void)hipFuncSetAttribute((const void *)k_modeprod_sr<MT, NT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    CR::LDS);
          attr_r = true;
        }
        const long ntiles = ncb * nfields;
        const long nwg = ntiles < c->num_cu ? ntiles : c->num_cu;
        k_modeprod_sr<MT, NT><<<(unsigned)nwg, CR::THREADS, CR::LDS, c->stream>>>(
            A3, in, in_fs, out, out_fs, n, rotate, accumulate, (int)chunk_bytes<MT, NT>(), nfields);
        return;
      }
    }
    if (c->split_persistent) {
      const long ntiles = ncb * nfields;
      const long nwg = ntiles < c->num_cu ? ntiles : c->num_cu;  // one workgroup per CU (LDS-bound)
      if (c->split_kernel == 9)  // 32-byte field copies, [k/8][col][8 k] layout
        k_modeprod_split_p<MT, NT, W, 2, 0><<<(unsigned)nwg, CF::THREADS, CF::LDS, c->stream>>>(
            A3, in, in_fs, out, out_fs, n, rotate, accumulate, (int)chunk_bytes<MT, NT>(), nfields);
      else
        k_modeprod_split_p<MT, NT, W><<<(unsigned)nwg, CF::THREADS, CF::LDS, c->stream>>>(
            A3, in, in_fs, out, out_fs, n, rotate, accumulate, (int)chunk_bytes<MT, NT>(), nfields);
    } else {
      dim3 grid((unsigned)ncb, nfields);
      k_modeprod_split<MT, NT, W><<<grid, CF::THREADS, CF::LDS, c->stream>>>(A3, in, in_fs, out, out_fs, n, rotate,
                                                                            accumulate, (int)chunk_bytes<MT, NT>());
    }
  };
  if constexpr (mps_cfg<MT, NT, 8, 2>::LDS <= 160 * 1024 && MT >= 2) {
    if (c->split_kernel == 4) {
      // 2 row halves x 4 column groups of waves
      using CF = mps_cfg<MT, NT, 8, 2>;
      static bool attr5 = false;
      if (!attr5) {
        (void)hipFuncSetAttribute((const void *)k_modeprod_split_rc<MT, NT>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, CF::LDS);
        attr5 = true;
      }
      const long ncb = ((long)n * n + CF::COLS - 1) / CF::COLS;
      const long ntiles = ncb * nfields;
      const long nwg = ntiles < c->num_cu ? ntiles : c->num_cu;
      k_modeprod_split_rc<MT, NT><<<(unsigned)nwg, CF::THREADS, CF::LDS, c->stream>>>(
          A3, in, in_fs, out, out_fs, n, rotate, accumulate, (int)chunk_bytes<MT, NT>(), nfields);
      return;
    }
  }
  if constexpr (mps_cfg<MT, NT, 4, 2>::ASZ + 2 * mps_cfg<MT, NT, 4, 2>::XSZ <= 80 * 1024) {
    if (c->split_kernel == 3) {
      // two 4-wave workgroups per CU, one operand stage each
      using CF = mps_cfg<MT, NT, 4, 2>;
      constexpr int lds = CF::ASZ + 2 * CF::XSZ;
      static bool attr3 = false;
      if (!attr3) {
        (void)hipFuncSetAttribute((const void *)k_modeprod_split_2wg<MT, NT>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        attr3 = true;
      }
      const long ncb = ((long)n * n + CF::COLS - 1) / CF::COLS;
      const long ntiles = ncb * nfields;
      const long nwg = ntiles < 2L * c->num_cu ? ntiles : 2L * c->num_cu;
      k_modeprod_split_2wg<MT, NT><<<(unsigned)nwg, CF::THREADS, lds, c->stream>>>(
          A3, in, in_fs, out, out_fs, n, rotate, accumulate, (int)chunk_bytes<MT, NT>(), nfields);
      return;
    }
  }
  if constexpr (mps_cfg<MT, NT, 4, 4>::LDS <= 160 * 1024) {
    if (c->split_kernel == 2) {
      // 4 waves x 64 columns: each A fragment feeds four column tiles
      using CF = mps_cfg<MT, NT, 4, 4>;
      static bool attr4 = false;
      if (!attr4) {
        (void)hipFuncSetAttribute((const void *)k_modeprod_split_p<MT, NT, 4, 4>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, CF::LDS);
        attr4 = true;
      }
      const long ncb = ((long)n * n + CF::COLS - 1) / CF::COLS;
      const long ntiles = ncb * nfields;
      const long nwg = ntiles < c->num_cu ? ntiles : c->num_cu;
      k_modeprod_split_p<MT, NT, 4, 4><<<(unsigned)nwg, CF::THREADS, CF::LDS, c->stream>>>(
          A3, in, in_fs, out, out_fs, n, rotate, accumulate, (int)chunk_bytes<MT, NT>(), nfields);
      return;
    }
  }
  if constexpr (MT <= 14) {
    if (c->split_waves == 4)
      go(std::integral_constant<int, 4>{});
    else
      go(std::integral_constant<int, 8>{});
  } else {
    go(std::integral_constant<int, 4>{});
  }
}

bool f16_zs_supported(int n) { return f16r_cfg<16>::FITS && (n + 15) / 16 <= 16; }

void launch_modeprod_split(etd_ctx *c, const uint16_t *A3, const float *in, size_t in_fs, float *out, size_t out_fs,
                           int nfields, int rotate, int accumulate, float a_inv) {
  const int MT = (c->n + 15) / 16;
  etd_prof_begin(c, PC_MODEPROD);
  if (c->split_f16) {
    switch (MT) {
#define MF(X)                                                                             \
  case X:                                                                                 \
    f16_launch<X>(c, A3, in, in_fs, out, out_fs, nfields, rotate, accumulate, a_inv);     \
    break;
      MF(1) MF(2) MF(3) MF(4) MF(5) MF(6) MF(7) MF(8) MF(9) MF(10) MF(11) MF(12) MF(13) MF(14) MF(15) MF(16) MF(17)
#undef MF
      default: break;
    }
    etd_prof_end(c, PC_MODEPROD);
    return;
  }
  switch (MT) {
#define MC(X)                                                                                      \
  case X:                                                                                          \
    if (c->split_terms == 2)                                                                       \
      mps_launch<X, 2>(c, A3, in, in_fs, out, out_fs, nfields, rotate, accumulate);                \
    else                                                                                           \
      mps_launch<X, 3>(c, A3, in, in_fs, out, out_fs, nfields, rotate, accumulate);                \
    break;
    MC(1) MC(2) MC(3) MC(4) MC(5) MC(6) MC(7) MC(8) MC(9) MC(10) MC(11) MC(12) MC(13) MC(14) MC(15) MC(16) MC(17)
#undef MC
    default: break;
  }
  etd_prof_end(c, PC_MODEPROD);
}
