// etd_alpha_io.hip -- the beta -> alpha driver's output and restart formats
// (source_beta_to_alpha/src): the binary checkpoint, the uint16 field dumps
// and the detector feed (2x downsample, binarise at 0.1, 8 cells per byte).
// The conversions run on the GPU; the host writes the files in the
// reference's byte layout.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "etd_internal.h"

#define AIO_CHK(c, call)                                                                     \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      snprintf((c)->err, sizeof((c)->err), "%s: %s", #call, hipGetErrorString(e_));          \
      return ETD_ERR_DEVICE;                                                                 \
    }                                                                                        \
  } while (0)

// Dtype2B_E (anisotropic_hip.c:3920-3931): (Stype)((E + 0.1) * 30000), the
// double -> unsigned short cast as the GPU executes it: truncation toward
// zero saturating to [0, 2^32 - 1], then the low 16 bits.
template <typename T>
__global__ __launch_bounds__(256) void k_field2b(const T *__restrict__ u, uint16_t *__restrict__ out, size_t cnt) {
  for (size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x; l < cnt; l += (size_t)gridDim.x * blockDim.x) {
    const double x = ((double)u[l] + 0.1) * 30000.0;
    const uint32_t w = x != x ? 0u : (x <= 0.0 ? 0u : (x >= 4294967295.0 ? 0xffffffffu : (uint32_t)x));
    out[l] = (uint16_t)(w & 0xffffu);
  }
}

// skip_copyin_new_2 (anisotropic_hip.c:3897-3918) + the binarisation of
// anisotropic_hip.c:4264-4278 + float2bin (4400-4427): byte t of the packed
// output holds cells 8 t .. 8 t + 7 of the (N/2)^3 downsampled volume (x
// fastest), bit e set when that cell's 2x2x2 mean of the interior exceeds
// 0.1; binary (optional) gets 1.0f / 0.0f per downsampled cell.  The mean is
// the reference's left-to-right FP64 sum of the 8 cells times 1/8.
template <typename T>
__global__ __launch_bounds__(256) void k_detector(const T *__restrict__ u, int n, int g, uint8_t *__restrict__ bits,
                                                  float *__restrict__ binary, long nbytes) {
  const int d = (n - 2 * g) / 2;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbytes) return;
  uint32_t byte = 0;
  for (int e = 0; e < 8; e++) {
    const long f = 8 * t + e;
    const int i = (int)(f % d), j = (int)((f / d) % d), k = (int)(f / ((long)d * d));
    const size_t n2 = (size_t)n * n;
    const size_t b = (size_t)(2 * k + g) * n2 + (size_t)(2 * j + g) * n + (size_t)(2 * i + g);
    const double s = (double)u[b] + (double)u[b + 1] + (double)u[b + n] + (double)u[b + n + 1] + (double)u[b + n2] +
                     (double)u[b + n2 + 1] + (double)u[b + n2 + n] + (double)u[b + n2 + n + 1];
    const double m = (1.0 / 8.0) * s;
    const bool on = m > 0.1;
    if (on) byte |= 1u << e;
    if (binary) binary[f] = on ? 1.0f : 0.0f;
  }
  bits[t] = (uint8_t)byte;
}

static int put_file(etd_ctx *c, const char *fn, const void *a, size_t asz, const void *b, size_t bsz) {
  FILE *f = fopen(fn, "wb");
  if (!f) {
    snprintf(c->err, sizeof(c->err), "cannot write %s", fn);
    return ETD_ERR_IO;
  }
  const int ok = (!asz || fwrite(a, 1, asz, f) == asz) && (!bsz || fwrite(b, 1, bsz, f) == bsz);
  if (fclose(f) || !ok) {
    snprintf(c->err, sizeof(c->err), "short write to %s", fn);
    return ETD_ERR_IO;
  }
  return ETD_OK;
}

// write_chk (vars.c:967-1011): <work_dir>00/eta<v>_chk<chk>_000000.dat =
// int irun, then the n^3 doubles of the field (ghosts included, x fastest);
// the CH files c<q>_chk<chk>_000000.dat in the same layout (the reference
// has their code, disabled: vars.c:990-1010).
extern "C" int etd_write_checkpoint(etd_ctx *c, int irun, int chk) {
  if (!c || (chk != 0 && chk != 1)) return ETD_ERR_ARG;
  if (c->P > 1) {
    snprintf(c->err, sizeof(c->err), "%s: one box only (px,py,pz = 1,1,1)", __func__);
    return ETD_ERR_UNSUPPORTED;
  }
  double *h = (double *)malloc(sizeof(double) * c->n3);
  if (!h) return ETD_ERR_NOMEM;
  int rc = ETD_OK;
  char fn[2200];
  for (int v = 0; v < c->V && rc == ETD_OK; v++) {
    if (v < c->nac)
      snprintf(fn, sizeof fn, "%s%02d/eta%d_chk%d_%02d%02d%02d.dat", c->p.work_dir, 0, v, chk, 0, 0, 0);
    else
      snprintf(fn, sizeof fn, "%s%02d/c%d_chk%d_%02d%02d%02d.dat", c->p.work_dir, 0, v - c->nac, chk, 0, 0, 0);
    rc = etd_get_field(c, v, h);
    if (!rc) rc = put_file(c, fn, &irun, sizeof(int), h, sizeof(double) * c->n3);
  }
  free(h);
  return rc;
}

// read_chk (vars.c:1013-1065): the AC files must exist; the CH files are read
// when present (else the CH fields are left as they are).
extern "C" int etd_read_checkpoint(etd_ctx *c, int chk, int *irun) {
  if (!c || !irun || (chk != 0 && chk != 1)) return ETD_ERR_ARG;
  if (c->P > 1) {
    snprintf(c->err, sizeof(c->err), "%s: one box only (px,py,pz = 1,1,1)", __func__);
    return ETD_ERR_UNSUPPORTED;
  }
  double *h = (double *)malloc(sizeof(double) * c->n3);
  if (!h) return ETD_ERR_NOMEM;
  int rc = ETD_OK;
  char fn[2200];
  for (int v = 0; v < c->V && rc == ETD_OK; v++) {
    const int ch = v >= c->nac;
    if (!ch)
      snprintf(fn, sizeof fn, "%s%02d/eta%d_chk%d_%02d%02d%02d.dat", c->p.work_dir, 0, v, chk, 0, 0, 0);
    else
      snprintf(fn, sizeof fn, "%s%02d/c%d_chk%d_%02d%02d%02d.dat", c->p.work_dir, 0, v - c->nac, chk, 0, 0, 0);
    FILE *f = fopen(fn, "rb");
    if (!f) {
      if (ch) continue;
      snprintf(c->err, sizeof(c->err), "cannot open %s", fn);
      rc = ETD_ERR_IO;
      break;
    }
    int ir = 0;
    const int ok = fread(&ir, sizeof(int), 1, f) == 1 && fread(h, sizeof(double), c->n3, f) == c->n3;
    fclose(f);
    if (!ok) {
      snprintf(c->err, sizeof(c->err), "%s is shorter than int + %zu doubles", fn, c->n3);
      rc = ETD_ERR_IO;
      break;
    }
    if (!ch) *irun = ir;
    rc = etd_set_field(c, v, h);
  }
  free(h);
  return rc;
}

template <typename T>
static int field2b_t(etd_ctx *c, int irun) {
  const T *u = fields<T>(c).u;
  uint16_t *d = (uint16_t *)fields<T>(c).gtmp;  // 3 n^3 T of scratch between steps
  uint16_t *h = (uint16_t *)malloc(sizeof(uint16_t) * c->n3);
  if (!h) return ETD_ERR_NOMEM;
  int rc = ETD_OK;
  char fn[2200];
  const unsigned blocks = (unsigned)((c->n3 + 255) / 256 < 4096 ? (c->n3 + 255) / 256 : 4096);
  for (int v = 0; v < c->nac && rc == ETD_OK; v++) {
    k_field2b<T><<<blocks, 256, 0, c->stream>>>(u + c->n3 * v, d, c->n3);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(uint16_t) * c->n3, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      snprintf(c->err, sizeof(c->err), "field2b: %s", hipGetErrorString(e));
      rc = ETD_ERR_DEVICE;
      break;
    }
    snprintf(fn, sizeof fn, "%s%02d/eta%d_%06d_%02d%02d%02d.dat", c->p.data_dir, 0, v, irun, 0, 0, 0);
    rc = put_file(c, fn, h, sizeof(uint16_t) * c->n3, nullptr, 0);
  }
  free(h);
  return rc;
}

// Dtype2B_E + write_field2B (anisotropic_hip.c:3920-3931, 4290-4299;
// vars.c:1068-1094): every AC field as <data_dir>00/eta<v>_<irun:06d>_000000.dat,
// n^3 uint16 (ghosts included, x fastest).
extern "C" int etd_write_field2b(etd_ctx *c, int irun) {
  if (!c) return ETD_ERR_ARG;
  if (c->P > 1) {
    snprintf(c->err, sizeof(c->err), "%s: one box only (px,py,pz = 1,1,1)", __func__);
    return ETD_ERR_UNSUPPORTED;
  }
  return c->mixed ? field2b_t<float>(c, irun) : field2b_t<double>(c, irun);
}

template <typename T>
static int detector_t(etd_ctx *c, int var, unsigned char *bits, float *binary) {
  const int d = c->N / 2;
  const long cells = (long)d * d * d, nbytes = cells / 8;
  uint8_t *db = (uint8_t *)fields<T>(c).gtmp;
  float *dbin = binary ? (float *)(db + ((nbytes + 255) / 256) * 256) : nullptr;
  k_detector<T><<<(unsigned)((nbytes + 255) / 256), 256, 0, c->stream>>>(fields<T>(c).u + c->n3 * var, c->n, c->g, db,
                                                                         dbin, nbytes);
  AIO_CHK(c, hipGetLastError());
  AIO_CHK(c, hipMemcpyAsync(bits, db, (size_t)nbytes, hipMemcpyDeviceToHost, c->stream));
  if (binary) AIO_CHK(c, hipMemcpyAsync(binary, dbin, sizeof(float) * cells, hipMemcpyDeviceToHost, c->stream));
  AIO_CHK(c, hipStreamSynchronize(c->stream));
  return ETD_OK;
}

extern "C" int etd_detector_feed(etd_ctx *c, int var, unsigned char *bits, float *binary) {
  if (!c || !bits || var < 0 || var >= c->nac) return ETD_ERR_ARG;
  if (c->P > 1) {
    snprintf(c->err, sizeof(c->err), "%s: one box only (px,py,pz = 1,1,1)", __func__);
    return ETD_ERR_UNSUPPORTED;
  }
  const int d = c->N / 2;
  if (c->N % 2 || ((long)d * d * d) % 8) {
    snprintf(c->err, sizeof(c->err), "detector feed: N = %d must be even with (N/2)^3 a multiple of 8", c->N);
    return ETD_ERR_UNSUPPORTED;
  }
  return c->mixed ? detector_t<float>(c, var, bits, binary) : detector_t<double>(c, var, bits, binary);
}
