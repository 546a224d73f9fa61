// etd_mlp.hip -- the learned dt controller's network on the GPU: an MLP
// 16 -> 64 -> 64 -> 1 with GELU, FP32, and its hand-written forward,
// backward and Adam kernels.  No counterpart in the reference.
//
// Training is deterministic: the per-sample gradient of a batch goes to its
// own row of a partial buffer, and a second kernel sums each parameter over
// the rows in a fixed order.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "etd_internal.h"

#define NF ETD_NFEAT
#define NH 64
#define P_W1 0
#define P_B1 (P_W1 + NH * NF)
#define P_W2 (P_B1 + NH)
#define P_B2 (P_W2 + NH * NH)
#define P_W3 (P_B2 + NH)
#define P_B3 (P_W3 + NH)
#define NPAR (P_B3 + 1)
static_assert(NPAR == ETD_MLP_PARAMS, "parameter layout");

struct etd_mlp {
  int device;
  hipStream_t stream;
  float *p, *m, *v;   // parameters and Adam moments (device)
  float *g;           // reduced gradient
  float *part;        // per-sample gradients, cap x NPAR
  float *x, *t, *y;   // batch inputs / targets / outputs
  float *loss_part;   // per-sample squared error
  int cap;
  long long step;     // Adam step count
  float fmean[NF], fstd[NF], ymean, ystd;
};

__device__ __forceinline__ float gelu(float a) { return 0.5f * a * (1.0f + erff(a * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float a) {
  const float Phi = 0.5f * (1.0f + erff(a * 0.70710678118654752f));
  const float phi = 0.39894228040143268f * expf(-0.5f * a * a);
  return Phi + a * phi;
}

__device__ __forceinline__ float block64_sum(float v, float *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if (threadIdx.x == 0) sh[0] = v;
  __syncthreads();
  const float r = sh[0];
  __syncthreads();
  return r;
}

// One 64-thread block per sample: forward; with t != NULL also the sample's
// squared error and gradient (scaled by 2 / batch) into part[s][*].
__global__ __launch_bounds__(64) void k_mlp(const float *__restrict__ P, const float *__restrict__ X,
                                            const float *__restrict__ Tg, float *__restrict__ Y,
                                            float *__restrict__ part, float *__restrict__ lpart, int batch) {
  const int s = blockIdx.x, j = threadIdx.x;
  __shared__ float x[NF], h1[NH], h2[NH], d2[NH], red[1];
  if (j < NF) x[j] = X[(size_t)s * NF + j];
  __syncthreads();
  float a1 = P[P_B1 + j];
#pragma unroll
  for (int k = 0; k < NF; k++) a1 += P[P_W1 + j * NF + k] * x[k];
  h1[j] = gelu(a1);
  __syncthreads();
  float a2 = P[P_B2 + j];
#pragma unroll 8
  for (int i = 0; i < NH; i++) a2 += P[P_W2 + j * NH + i] * h1[i];
  const float hh2 = gelu(a2);
  h2[j] = hh2;
  const float y = block64_sum(P[P_W3 + j] * hh2, red) + P[P_B3];
  if (j == 0) Y[s] = y;
  if (!Tg) return;
  const float r = y - Tg[s];
  const float dy = 2.0f * r / (float)batch;
  float *G = part + (size_t)s * NPAR;
  if (j == 0) {
    lpart[s] = r * r;
    G[P_B3] = dy;
  }
  G[P_W3 + j] = dy * hh2;
  const float da2 = dy * P[P_W3 + j] * gelu_d(a2);
  d2[j] = da2;
  G[P_B2 + j] = da2;
#pragma unroll 8
  for (int i = 0; i < NH; i++) G[P_W2 + j * NH + i] = da2 * h1[i];
  __syncthreads();
  float dh1 = 0.0f;
#pragma unroll 8
  for (int i = 0; i < NH; i++) dh1 += P[P_W2 + i * NH + j] * d2[i];
  const float da1 = dh1 * gelu_d(a1);
  G[P_B1 + j] = da1;
#pragma unroll
  for (int k = 0; k < NF; k++) G[P_W1 + j * NF + k] = da1 * x[k];
}

// g[p] = sum over samples in order; loss = mean squared error
__global__ void k_mlp_reduce(const float *__restrict__ part, const float *__restrict__ lpart, float *__restrict__ g,
                             float *__restrict__ loss, int batch) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < NPAR) {
    float acc = 0.0f;
    for (int s = 0; s < batch; s++) acc += part[(size_t)s * NPAR + p];
    g[p] = acc;
  }
  if (p == 0) {
    float l = 0.0f;
    for (int s = 0; s < batch; s++) l += lpart[s];
    loss[0] = l / (float)batch;
  }
}

__global__ void k_adam(float *__restrict__ P, float *__restrict__ M, float *__restrict__ Vv, const float *__restrict__ g,
                       float lr, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPAR) return;
  const float gp = g[p];
  const float m = 0.9f * M[p] + 0.1f * gp;
  const float v = 0.999f * Vv[p] + 0.001f * gp * gp;
  M[p] = m;
  Vv[p] = v;
  P[p] -= lr * (m / c1) / (sqrtf(v / c2) + 1e-8f);
}

// ------------------------------------------------------------------ host
#define MCHK(call)                                \
  do {                                            \
    if ((call) != hipSuccess) return ETD_ERR_DEVICE; \
  } while (0)

static int ensure_cap(etd_mlp *m, int batch) {
  if (batch <= m->cap) return ETD_OK;
  (void)hipFree(m->part);
  (void)hipFree(m->x);
  (void)hipFree(m->t);
  (void)hipFree(m->y);
  (void)hipFree(m->loss_part);
  m->part = m->x = m->t = m->y = m->loss_part = nullptr;
  m->cap = 0;
  MCHK(hipMalloc(&m->part, sizeof(float) * (size_t)batch * NPAR));
  MCHK(hipMalloc(&m->x, sizeof(float) * (size_t)batch * NF));
  MCHK(hipMalloc(&m->t, sizeof(float) * (size_t)batch));
  MCHK(hipMalloc(&m->y, sizeof(float) * (size_t)batch));
  MCHK(hipMalloc(&m->loss_part, sizeof(float) * (size_t)batch));
  m->cap = batch;
  return ETD_OK;
}

extern "C" int etd_mlp_create(int device, etd_mlp **out) {
  if (!out) return ETD_ERR_ARG;
  *out = nullptr;
  etd_mlp *m = (etd_mlp *)calloc(1, sizeof(etd_mlp));
  if (!m) return ETD_ERR_NOMEM;
  m->device = device;
  *out = m;
  MCHK(hipSetDevice(device));
  MCHK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  MCHK(hipMalloc(&m->p, sizeof(float) * NPAR));
  MCHK(hipMalloc(&m->m, sizeof(float) * NPAR));
  MCHK(hipMalloc(&m->v, sizeof(float) * NPAR));
  MCHK(hipMalloc(&m->g, sizeof(float) * (NPAR + 1)));
  return etd_mlp_init(m, 1);
}

extern "C" void etd_mlp_destroy(etd_mlp *m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  (void)hipFree(m->p);
  (void)hipFree(m->m);
  (void)hipFree(m->v);
  (void)hipFree(m->g);
  (void)hipFree(m->part);
  (void)hipFree(m->x);
  (void)hipFree(m->t);
  (void)hipFree(m->y);
  (void)hipFree(m->loss_part);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  free(m);
}

static unsigned long long splitmix(unsigned long long *s) {
  unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static double unif(unsigned long long *s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

extern "C" int etd_mlp_init(etd_mlp *m, unsigned long long seed) {
  if (!m) return ETD_ERR_ARG;
  float *h = (float *)calloc(NPAR, sizeof(float));
  if (!h) return ETD_ERR_NOMEM;
  unsigned long long st = seed;
  const double l1 = sqrt(6.0 / (NF + NH)), l2 = sqrt(6.0 / (NH + NH)), l3 = sqrt(6.0 / (NH + 1));
  for (int i = 0; i < NH * NF; i++) h[P_W1 + i] = (float)((2.0 * unif(&st) - 1.0) * l1);
  for (int i = 0; i < NH * NH; i++) h[P_W2 + i] = (float)((2.0 * unif(&st) - 1.0) * l2);
  for (int i = 0; i < NH; i++) h[P_W3 + i] = (float)((2.0 * unif(&st) - 1.0) * l3);
  for (int k = 0; k < NF; k++) {
    m->fmean[k] = 0.0f;
    m->fstd[k] = 1.0f;
  }
  m->ymean = 0.0f;
  m->ystd = 1.0f;
  m->step = 0;
  int rc = hipMemcpy(m->p, h, sizeof(float) * NPAR, hipMemcpyHostToDevice) == hipSuccess ? ETD_OK : ETD_ERR_DEVICE;
  free(h);
  if (rc) return rc;
  MCHK(hipMemset(m->m, 0, sizeof(float) * NPAR));
  MCHK(hipMemset(m->v, 0, sizeof(float) * NPAR));
  return ETD_OK;
}

extern "C" int etd_mlp_set(etd_mlp *m, const float *params, const float *fm, const float *fs, float ym, float ys) {
  if (!m || !params || !fm || !fs || !(ys > 0.0f)) return ETD_ERR_ARG;
  for (int k = 0; k < NF; k++)
    if (!(fs[k] > 0.0f)) return ETD_ERR_ARG;
  MCHK(hipMemcpy(m->p, params, sizeof(float) * NPAR, hipMemcpyHostToDevice));
  memcpy(m->fmean, fm, sizeof(m->fmean));
  memcpy(m->fstd, fs, sizeof(m->fstd));
  m->ymean = ym;
  m->ystd = ys;
  MCHK(hipMemset(m->m, 0, sizeof(float) * NPAR));
  MCHK(hipMemset(m->v, 0, sizeof(float) * NPAR));
  m->step = 0;
  return ETD_OK;
}

extern "C" int etd_mlp_get(const etd_mlp *m, float *params, float *fm, float *fs, float *ym, float *ys) {
  if (!m) return ETD_ERR_ARG;
  if (params) MCHK(hipMemcpy(params, m->p, sizeof(float) * NPAR, hipMemcpyDeviceToHost));
  if (fm) memcpy(fm, m->fmean, sizeof(m->fmean));
  if (fs) memcpy(fs, m->fstd, sizeof(m->fstd));
  if (ym) *ym = m->ymean;
  if (ys) *ys = m->ystd;
  return ETD_OK;
}

static int run_batch(etd_mlp *m, const float *xd, const float *td, int batch) {
  k_mlp<<<batch, 64, 0, m->stream>>>(m->p, xd, td, m->y, m->part, m->loss_part, batch);
  if (td) k_mlp_reduce<<<(NPAR + 255) / 256, 256, 0, m->stream>>>(m->part, m->loss_part, m->g, m->g + NPAR, batch);
  return hipGetLastError() == hipSuccess ? ETD_OK : ETD_ERR_DEVICE;
}

extern "C" int etd_mlp_forward(etd_mlp *m, const float *x, int batch, float *y) {
  if (!m || !x || !y || batch < 1) return ETD_ERR_ARG;
  int rc = ensure_cap(m, batch);
  if (rc) return rc;
  MCHK(hipMemcpyAsync(m->x, x, sizeof(float) * (size_t)batch * NF, hipMemcpyHostToDevice, m->stream));
  if ((rc = run_batch(m, m->x, nullptr, batch))) return rc;
  MCHK(hipMemcpyAsync(y, m->y, sizeof(float) * batch, hipMemcpyDeviceToHost, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  return ETD_OK;
}

extern "C" int etd_mlp_loss_grad(etd_mlp *m, const float *x, const float *t, int batch, float *loss, float *grad) {
  if (!m || !x || !t || batch < 1) return ETD_ERR_ARG;
  int rc = ensure_cap(m, batch);
  if (rc) return rc;
  MCHK(hipMemcpyAsync(m->x, x, sizeof(float) * (size_t)batch * NF, hipMemcpyHostToDevice, m->stream));
  MCHK(hipMemcpyAsync(m->t, t, sizeof(float) * batch, hipMemcpyHostToDevice, m->stream));
  if ((rc = run_batch(m, m->x, m->t, batch))) return rc;
  if (grad) MCHK(hipMemcpyAsync(grad, m->g, sizeof(float) * NPAR, hipMemcpyDeviceToHost, m->stream));
  if (loss) MCHK(hipMemcpyAsync(loss, m->g + NPAR, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  return ETD_OK;
}

extern "C" int etd_mlp_train(etd_mlp *m, const float *x, const float *t, int count, int epochs, int batch, float lr,
                             unsigned long long seed, float *epoch_loss) {
  if (!m || !x || !t || count < 1 || epochs < 1 || batch < 1 || !(lr > 0.0f)) return ETD_ERR_ARG;
  if (batch > count) batch = count;
  int rc = ensure_cap(m, batch);
  if (rc) return rc;
  int *idx = (int *)malloc(sizeof(int) * count);
  float *hx = (float *)malloc(sizeof(float) * (size_t)batch * NF), *ht = (float *)malloc(sizeof(float) * batch);
  float *hl = (float *)malloc(sizeof(float) * ((size_t)count / batch + 1));
  if (!idx || !hx || !ht || !hl) {
    free(idx);
    free(hx);
    free(ht);
    free(hl);
    return ETD_ERR_NOMEM;
  }
  for (int i = 0; i < count; i++) idx[i] = i;
  unsigned long long st = seed;
  for (int e = 0; e < epochs && rc == ETD_OK; e++) {
    for (int i = count - 1; i > 0; i--) {  // Fisher-Yates, seeded
      const int j = (int)(splitmix(&st) % (unsigned long long)(i + 1));
      const int tmp = idx[i];
      idx[i] = idx[j];
      idx[j] = tmp;
    }
    const int nb = count / batch;  // full batches only
    for (int b = 0; b < nb && rc == ETD_OK; b++) {
      for (int s = 0; s < batch; s++) {
        const int k = idx[b * batch + s];
        memcpy(hx + (size_t)s * NF, x + (size_t)k * NF, sizeof(float) * NF);
        ht[s] = t[k];
      }
      if (hipMemcpyAsync(m->x, hx, sizeof(float) * (size_t)batch * NF, hipMemcpyHostToDevice, m->stream) ||
          hipMemcpyAsync(m->t, ht, sizeof(float) * batch, hipMemcpyHostToDevice, m->stream)) {
        rc = ETD_ERR_DEVICE;
        break;
      }
      rc = run_batch(m, m->x, m->t, batch);
      if (rc) break;
      m->step++;
      const float c1 = 1.0f - powf(0.9f, (float)m->step), c2 = 1.0f - powf(0.999f, (float)m->step);
      k_adam<<<(NPAR + 255) / 256, 256, 0, m->stream>>>(m->p, m->m, m->v, m->g, lr, c1, c2);
      if (hipMemcpyAsync(hl + b, m->g + NPAR, sizeof(float), hipMemcpyDeviceToHost, m->stream) ||
          hipStreamSynchronize(m->stream)) {  // the host batch buffers are reused next iteration
        rc = ETD_ERR_DEVICE;
        break;
      }
    }
    if (rc == ETD_OK && epoch_loss) {
      double acc = 0.0;
      for (int b = 0; b < nb; b++) acc += hl[b];
      epoch_loss[e] = (float)(acc / nb);
    }
  }
  free(idx);
  free(hx);
  free(ht);
  free(hl);
  return rc;
}

extern "C" int etd_mlp_predict(etd_mlp *m, const double *features, double *y) {
  if (!m || !features || !y) return ETD_ERR_ARG;
  float xn[NF], yn;
  for (int k = 0; k < NF; k++) xn[k] = (float)((features[k] - m->fmean[k]) / m->fstd[k]);
  int rc = etd_mlp_forward(m, xn, 1, &yn);
  if (rc) return rc;
  *y = (double)yn * m->ystd + m->ymean;
  return ETD_OK;
}
