// etd_ensemble.hip -- batched ensembles of independent contexts (one hipGraph
// replayed as one launch) and the kernel-variant autotuner.  No counterpart in
// the reference, which runs one simulation per MPI job.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "etd_internal.h"

struct etd_ensemble {
  int count, steps;
  etd_ctx **m;
  hipStream_t main;
  hipEvent_t fork;
  hipEvent_t *join;
  hipGraph_t graph;
  hipGraphExec_t exec;
  long long *evals_per_launch;  // elastic evaluations each replay adds, per member
  char err[512];
};

static int gcd_i(int a, int b) {
  while (b) {
    const int t = a % b;
    a = b;
    b = t;
  }
  return a;
}

static int enerr(etd_ensemble *e, int code, const char *msg, hipError_t he = hipSuccess) {
  snprintf(e->err, sizeof(e->err), "%s%s%s", msg, he != hipSuccess ? ": " : "",
           he != hipSuccess ? hipGetErrorString(he) : "");
  return code;
}

#define ECHK(call)                                  \
  do {                                              \
    hipError_t he_ = (call);                        \
    if (he_ != hipSuccess) return enerr(e, ETD_ERR_DEVICE, #call, he_); \
  } while (0)

extern "C" const char *etd_ensemble_error(const etd_ensemble *e) { return e ? e->err : "no ensemble"; }

extern "C" int etd_ensemble_create(etd_ctx *const *members, int count, int steps, etd_ensemble **out) {
  if (!members || count < 1 || steps < 1 || !out) return ETD_ERR_ARG;
  *out = nullptr;
  etd_ensemble *e = (etd_ensemble *)calloc(1, sizeof(etd_ensemble));
  if (!e) return ETD_ERR_NOMEM;
  *out = e;
  e->count = count;
  e->steps = steps;
  e->m = (etd_ctx **)calloc(count, sizeof(etd_ctx *));
  e->join = (hipEvent_t *)calloc(count, sizeof(hipEvent_t));
  e->evals_per_launch = (long long *)calloc(count, sizeof(long long));
  if (!e->m || !e->join || !e->evals_per_launch) return ETD_ERR_NOMEM;
  int dev = -1;
  for (int i = 0; i < count; i++) {
    etd_ctx *c = members[i];
    if (c && c->P > 1) return ETD_ERR_UNSUPPORTED;  // members are single boxes
    if (c) {  // a replayed graph transforms u itself every step
      c->no_carry = 1;
      c->carry_ok = 0;
    }
    if (!c) return enerr(e, ETD_ERR_ARG, "null member");
    e->m[i] = c;
    int d;
    ECHK(hipSetDevice(c->device));
    ECHK(hipGetDevice(&d));
    if (dev >= 0 && d != dev) return enerr(e, ETD_ERR_UNSUPPORTED, "members on different devices");
    dev = d;
    for (int j = 0; j < i; j++)
      if (e->m[j] == c) return enerr(e, ETD_ERR_ARG, "a context appears twice");
    // the captured felas ring positions must come back to the same phase
    if (c->nel > 0) {
      const int eps = c->p.etd2 ? 2 : 1, R = c->ring;
      const int need = R / gcd_i((eps * c->nel) % R == 0 ? R : (eps * c->nel) % R, R);
      if (steps % need) {
        char msg[256];
        snprintf(msg, sizeof msg, "member %d (%d elastic fields, %s): steps per launch must be a multiple of %d", i,
                 c->nel, c->p.etd2 ? "ETD2" : "ETD1", need);
        return enerr(e, ETD_ERR_ARG, msg);
      }
      e->evals_per_launch[i] = (long long)steps * eps;
    }
  }
  ECHK(hipStreamCreateWithFlags(&e->main, hipStreamNonBlocking));
  ECHK(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
  for (int i = 0; i < count; i++) ECHK(hipEventCreateWithFlags(&e->join[i], hipEventDisableTiming));
  // every kernel's one-time host set-up happens outside the capture: finish
  // pending work first
  // the CH mean targets are taken outside the capture, so that replays do
  // not retake them
  for (int i = 0; i < count; i++)
    if (e->m[i]->mixed) ch_mean_mark(e->m[i], e->m[i]->f32.u);
  for (int i = 0; i < count; i++) ECHK(hipStreamSynchronize(e->m[i]->stream));
  long long *saved = (long long *)malloc(sizeof(long long) * count);
  for (int i = 0; i < count; i++) saved[i] = e->m[i]->elastic_evals;
  ECHK(hipStreamBeginCapture(e->main, hipStreamCaptureModeRelaxed));
  int rc = ETD_OK;
  hipError_t he = hipEventRecord(e->fork, e->main);
  for (int i = 0; i < count && he == hipSuccess; i++) he = hipStreamWaitEvent(e->m[i]->stream, e->fork, 0);
  for (int s = 0; s < steps && he == hipSuccess && rc == ETD_OK; s++)
    for (int i = 0; i < count && rc == ETD_OK; i++) {
      etd_ctx *c = e->m[i];
      rc = c->mixed ? step_once_t<float>(c) : step_once_t<double>(c);
    }
  for (int i = 0; i < count && he == hipSuccess; i++) {
    he = hipEventRecord(e->join[i], e->m[i]->stream);
    if (he == hipSuccess) he = hipStreamWaitEvent(e->main, e->join[i], 0);
  }
  hipError_t he2 = hipStreamEndCapture(e->main, &e->graph);
  for (int i = 0; i < count; i++) e->m[i]->elastic_evals = saved[i];  // nothing ran yet
  free(saved);
  if (rc) return enerr(e, rc, "capturing a member's step failed");
  if (he != hipSuccess) return enerr(e, ETD_ERR_DEVICE, "capture", he);
  if (he2 != hipSuccess) return enerr(e, ETD_ERR_DEVICE, "hipStreamEndCapture", he2);
  ECHK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
  return ETD_OK;
}

extern "C" int etd_ensemble_run(etd_ensemble *e, int launches) {
  if (!e || launches < 0) return ETD_ERR_ARG;
  for (int i = 0; i < e->count; i++) ECHK(hipStreamSynchronize(e->m[i]->stream));
  for (int l = 0; l < launches; l++) ECHK(hipGraphLaunch(e->exec, e->main));
  ECHK(hipStreamSynchronize(e->main));
  for (int i = 0; i < e->count; i++) {
    etd_ctx *c = e->m[i];
    c->elastic_evals += e->evals_per_launch[i] * launches;
    for (long long k = 0; k < (long long)e->steps * launches; k++) c->t += c->p.dt;  // as etd_step adds it
  }
  return ETD_OK;
}

extern "C" void etd_ensemble_destroy(etd_ensemble *e) {
  if (!e) return;
  if (e->main) (void)hipStreamSynchronize(e->main);
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  if (e->fork) (void)hipEventDestroy(e->fork);
  for (int i = 0; i < e->count && e->join; i++)
    if (e->join[i]) (void)hipEventDestroy(e->join[i]);
  if (e->main) (void)hipStreamDestroy(e->main);
  free(e->m);
  free(e->join);
  free(e->evals_per_launch);
  free(e);
}

// ------------------------------------------------------------------ autotuner
extern "C" int etd_get_tuning(const etd_ctx *c, etd_tuning *t) {
  if (!c || !t) return ETD_ERR_ARG;
  t->mp_variant = c->mp_variant;
  t->split_waves = c->split_waves;
  t->fft_variant = c->fft_variant;
  t->ms_per_step = 0.0;
  return ETD_OK;
}

extern "C" int etd_set_tuning(etd_ctx *c, const etd_tuning *t) {
  if (!c || !t) return ETD_ERR_ARG;
  const int v = t->mp_variant;
  if (!(v == 0 || v == 1 || v == 2 || v == 4 || v == 16)) return ETD_ERR_ARG;
  if (!(t->split_waves == 0 || t->split_waves == 4 || t->split_waves == 8)) return ETD_ERR_ARG;
  if (!(t->fft_variant == 0 || t->fft_variant == 1)) return ETD_ERR_ARG;
  c->mp_variant = v;
  c->split_waves = t->split_waves;
  c->fft_variant = t->fft_variant;
  return ETD_OK;
}

extern "C" int etd_autotune(const etd_params *p, const double *V, const double *Vi, const double *D, int steps,
                            etd_tuning *best) {
  if (!p || !best || steps < 1) return ETD_ERR_ARG;
  etd_ctx *c = nullptr;
  int rc = etd_create(p, V, Vi, D, &c);
  if (rc) {
    etd_close(c);
    return rc;
  }
  // a smooth seeded state so the timed steps do real work
  for (int v = 0; v < c->V && rc == ETD_OK; v++) rc = etd_init_noise(c, v, 7919ULL * (v + 1), v < c->nac ? 0.0 : 0.1, 0.01);
  static const int mp_cands[] = {1, 0, 2, 4, 16};
  static const int sw_cands[] = {8, 4};
  static const int fft_cands[] = {1, 0};
  best->ms_per_step = 1e300;
  double ms[2 + PC_COUNT];
  for (int a = 0; a < (c->mixed ? 1 : 5) && rc == ETD_OK; a++)
    for (int b = 0; b < (c->mixed ? 2 : 1) && rc == ETD_OK; b++)
      for (int f = 0; f < 2 && rc == ETD_OK; f++) {
        etd_tuning t = {mp_cands[a], c->mixed ? sw_cands[b] : 0, fft_cands[f], 0.0};
        if (c->mixed && t.split_waves == 8 && (c->n + 15) / 16 > 14) continue;
        if (t.fft_variant == 1 && (!fft16_supported(c->N) || !(c->nel > 0))) continue;
        if ((rc = etd_set_tuning(c, &t))) break;
        if ((rc = etd_time_steps(c, 1, 0, ms, nullptr))) break;  // warm: first launches
        if ((rc = etd_time_steps(c, steps, 0, ms, nullptr))) break;
        t.ms_per_step = ms[0] / steps;
        if (t.ms_per_step < best->ms_per_step) *best = t;
      }
  etd_close(c);
  return rc;
}
