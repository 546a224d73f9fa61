// etd_api.hip -- the C ABI of libetd: context lifetime, host<->device copies,
// the step orchestration and the per-kernel / timing entry points.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <rocblas/rocblas.h>

#include "etd_internal.h"

#define HCHK(c, call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      if (c) snprintf((c)->err, sizeof((c)->err), "%s: %s", #call, hipGetErrorString(e_)); \
      return ETD_ERR_DEVICE;                                                           \
    }                                                                                  \
  } while (0)

static int seterr(etd_ctx *c, int code, const char *fmt, ...) {
  if (c) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof(c->err), fmt, ap);
    va_end(ap);
  }
  return code;
}

extern "C" const char *etd_error_string(int code) {
  switch (code) {
    case ETD_OK: return "ok";
    case ETD_ERR_ARG: return "invalid argument";
    case ETD_ERR_IO: return "input file missing or malformed";
    case ETD_ERR_UNSUPPORTED: return "unsupported configuration";
    case ETD_ERR_DEVICE: return "HIP device error";
    case ETD_ERR_NOMEM: return "out of memory";
    case ETD_ERR_NONFINITE: return "non-finite state";
    default: return "unknown error";
  }
}

extern "C" const char *etd_last_error(const etd_ctx *c) { return c ? c->err : "no context"; }

// ------------------------------------------------------------------ profiling
void etd_prof_begin(etd_ctx *c, int cls) {
  if (!c->prof || c->nev + 2 > c->cap_ev) return;
  (void)hipEventRecord(c->ev[c->nev], c->stream);
  c->ev_class[c->nev] = cls;
  c->nev++;
}
void etd_prof_end(etd_ctx *c, int cls) {
  if (!c->prof || c->nev + 1 > c->cap_ev) return;
  (void)hipEventRecord(c->ev[c->nev], c->stream);
  c->ev_class[c->nev] = -1 - cls;
  c->nev++;
}

// ------------------------------------------------------------------ helpers
static int fft_plan_radices(int N, int *rad) {
  int np = 0, x = N;
  while (x % 4 == 0) { rad[np++] = 4; x /= 4; }
  while (x % 2 == 0) { rad[np++] = 2; x /= 2; }
  while (x % 3 == 0) { rad[np++] = 3; x /= 3; }
  while (x % 5 == 0) { rad[np++] = 5; x /= 5; }
  if (x != 1 || np > 16) return -1;
  return np;
}

static void c4_from_voigt(const double *C2, double *C4) {
  auto vg = [](int i, int j) {
    if (i == j) return i;
    if ((i == 1 && j == 2) || (i == 2 && j == 1)) return 3;
    if ((i == 0 && j == 2) || (i == 2 && j == 0)) return 4;
    return 5;
  };
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) C4[((i * 3 + j) * 3 + k) * 3 + l] = C2[vg(i, j) * 6 + vg(k, l)];
}

static int validate(etd_ctx *c, const etd_params *p) {
  if (p->nx != p->ny || p->nx != p->nz)
    return seterr(c, ETD_ERR_UNSUPPORTED, "nx, ny, nz must be equal (the axis rotation of calc.c:585-629 needs a cube)");
  if (p->px < 1 || p->py < 1 || p->pz < 1 || (long)p->px * p->py * p->pz > 512)
    return seterr(c, ETD_ERR_ARG, "px,py,pz = %d,%d,%d", p->px, p->py, p->pz);
  if (p->px * p->py * p->pz > 1) {
    // a decomposed run: every box on this GPU, the reference's localized-ETD numerics
    if (p->periodic != 1)
      return seterr(c, ETD_ERR_UNSUPPORTED, "px,py,pz = %d,%d,%d: boxes with physical outer sides are not built "
                    "(periodic = 1 only)", p->px, p->py, p->pz);
    if (p->precision != ETD_PREC_FP64)
      return seterr(c, ETD_ERR_UNSUPPORTED, "px,py,pz > 1 runs in FP64 (precision 0)");
    if (p->variable_mobility || p->free_energy)
      return seterr(c, ETD_ERR_UNSUPPORTED, "px,py,pz > 1 with VariableMobility or the CALPHAD force is not built");
    if (p->elastic && (p->px != p->py || p->px != p->pz))
      return seterr(c, ETD_ERR_UNSUPPORTED, "elasticity over px,py,pz = %d,%d,%d boxes needs px = py = pz (a cubic "
                    "global grid)", p->px, p->py, p->pz);
  }
  if (p->periodic != 0 && p->periodic != 1) return seterr(c, ETD_ERR_ARG, "periodic = %d is not 0 or 1", p->periodic);
  if (p->variable_mobility != 0 && p->variable_mobility != 1)
    return seterr(c, ETD_ERR_ARG, "VariableMobility = %d is not 0 or 1", p->variable_mobility);
  if (p->variable_mobility && (p->nch != 3 || p->nac < 1))
    return seterr(c, ETD_ERR_UNSUPPORTED,
                  "VariableMobility = 1 needs nch = 3 (Cu, Mn, Ni, mobility.c:860-922) and an AC field 0 "
                  "(nac = %d, nch = %d)",
                  p->nac, p->nch);
  if (p->anisotropic < 0 || p->anisotropic > 2)
    return seterr(c, ETD_ERR_ARG, "ANISOTROPIC = %d is not 0, 1 or 2", p->anisotropic);
  if (p->free_energy != 0 && p->free_energy != 1)
    return seterr(c, ETD_ERR_ARG, "free_energy = %d is not 0 (Landau) or 1 (CALPHAD)", p->free_energy);
  if (p->elastic < 0 || p->elastic > 3) return seterr(c, ETD_ERR_ARG, "ELASTIC = %d is not 0..3", p->elastic);
  if (p->elastic == 2 && p->nch > 16) return seterr(c, ETD_ERR_UNSUPPORTED, "elasticity supports nch <= 16");
  if (p->nac < 0 || p->nch < 0 || p->nac + p->nch < 1) return seterr(c, ETD_ERR_ARG, "nac = 0 and nch = 0");
  if (p->nac + p->nch > ETD_MAX_VAR) return seterr(c, ETD_ERR_UNSUPPORTED, "too many fields");
  if (p->nghost < 1 || p->nx < 2 * p->nghost + 2) return seterr(c, ETD_ERR_ARG, "nghost / nx");
  if (p->nx > 16 * 17) return seterr(c, ETD_ERR_UNSUPPORTED, "n > 272");
  if (p->elastic == 1 && p->nac > 16) return seterr(c, ETD_ERR_UNSUPPORTED, "elasticity supports nac <= 16");
  if (p->precision != ETD_PREC_FP64 && p->precision != ETD_PREC_MIXED && p->precision != ETD_PREC_MIXED_F16)
    return seterr(c, ETD_ERR_ARG, "precision %d is not FP64 (0), mixed BF16 (1) or mixed FP16 (2)", p->precision);
  if (p->precision != ETD_PREC_FP64 && p->nx % 4)
    return seterr(c, ETD_ERR_UNSUPPORTED, "the mixed-precision mode products need nx %% 4 == 0 (nx = %d)", p->nx);
  return ETD_OK;
}

extern "C" int etd_create(const etd_params *p, const double *Vh, const double *Vih, const double *Dh, etd_ctx **out) {
  if (!p || !out) return ETD_ERR_ARG;
  *out = nullptr;
  etd_ctx *c = (etd_ctx *)calloc(1, sizeof(etd_ctx));
  if (!c) return ETD_ERR_NOMEM;
  int rc = validate(c, p);
  if (rc) {
    fprintf(stderr, "etd_create: %s\n", c->err);
    free(c);
    return rc;
  }
  c->p = *p;
  c->n = p->nx;
  c->g = p->nghost;
  c->N = c->n - 2 * c->g;
  c->nac = p->nac;
  c->nch = p->nch;
  c->V = p->nac + p->nch;
  const int n = c->n, V = c->V;
  c->n3 = (size_t)n * n * n;
  c->N3 = (size_t)c->N * c->N * c->N;
  c->px = p->px;
  c->py = p->py;
  c->pz = p->pz;
  c->P = p->px * p->py * p->pz;
  c->box = 0;
  c->bs = (size_t)V * c->n3;
  c->Ne = c->px * c->N;  // px = py = pz whenever elastic; the grid edge only matters then
  c->Ne3 = (size_t)c->Ne * c->Ne * c->Ne;
  *out = c;  // from here on etd_close releases partial state
  c->hV = (double *)malloc(sizeof(double) * n * n);
  c->hVi = (double *)malloc(sizeof(double) * n * n);
  c->hD = (double *)malloc(sizeof(double) * n);
  if (Vh && Vih && Dh) {
    memcpy(c->hV, Vh, sizeof(double) * n * n);
    memcpy(c->hVi, Vih, sizeof(double) * n * n);
    memcpy(c->hD, Dh, sizeof(double) * n);
  } else if ((rc = etd_eigbasis(n, p->periodic ? 3 : 0, c->hV, c->hVi, c->hD)) != ETD_OK) {
    return seterr(c, rc, "eigenbasis");
  }
  c->mp_variant = 1;  // direct-to-LDS pipelined 4x4x4 kernel, two workgroups per CU
  const char *mv = getenv("ETD_MFMA");
  if (mv) {
    const int v = atoi(mv);
    if (v == 0 || v == 1 || v == 2 || v == 4 || v == 16) c->mp_variant = v;
  }
  c->mp_nb = 8;
  const char *cv = getenv("ETD_CARRY");
  c->no_carry = cv && atoi(cv) == 0;  // every step transforms u itself
  carry_base(c, nullptr, 1.0, nullptr, 0.0);
  c->fft_variant = 1;
  const char *fv = getenv("ETD_FFT");
  if (fv && atoi(fv) == 0) c->fft_variant = 0;
  const char *nbv = getenv("ETD_MFMA_NB");
  if (nbv && atoi(nbv) == 4) c->mp_nb = 4;
  int dev = 0;
  const char *ev = getenv("ETD_DEVICE");
  if (ev) dev = atoi(ev);
  HCHK(c, hipSetDevice(dev));
  c->device = dev;
  HCHK(c, hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, dev));
  c->split_persistent = 1;
  const char *spv = getenv("ETD_SPLIT_PERSIST");
  if (spv && atoi(spv) == 0) c->split_persistent = 0;
  c->split_kernel = 10;
  const char *skv = getenv("ETD_SPLIT_KERNEL");
  if (skv && atoi(skv) >= 0 && atoi(skv) <= 10) c->split_kernel = atoi(skv);
  c->nl_batch = 0;
  const char *nlb = getenv("ETD_NL_BATCH");
  if (nlb && (atoi(nlb) == 4 || atoi(nlb) == 8 || atoi(nlb) == 12 || atoi(nlb) == 16)) c->nl_batch = atoi(nlb);
  HCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  const size_t n3 = c->n3;
  c->mixed = p->precision == ETD_PREC_MIXED || p->precision == ETD_PREC_MIXED_F16;
  c->split_f16 = p->precision == ETD_PREC_MIXED_F16;
  c->split_terms = 3;
  const char *sv = getenv("ETD_SPLIT");
  if (sv && atoi(sv) == 2) c->split_terms = 2;
  if (c->split_f16) c->split_terms = 2;
  c->split_zs = c->split_f16 && c->split_kernel == 10 && f16_zs_supported(p->nx);
  if (c->split_kernel == 10 && !c->split_zs) c->split_kernel = 0;  // BF16, or n > 256: the default of those
  if (!c->mixed) {
    const size_t PV = (size_t)c->P * V;  // every box's fields
    HCHK(c, hipMalloc(&c->u, sizeof(double) * n3 * 2 * PV));  // [u | N1]
    c->N1 = c->u + n3 * PV;
    HCHK(c, hipMemsetAsync(c->u, 0, sizeof(double) * n3 * 2 * PV, c->stream));
    if (p->etd2) HCHK(c, hipMalloc(&c->N2, sizeof(double) * n3 * PV));
    HCHK(c, hipMalloc(&c->w1, sizeof(double) * n3 * 2 * PV));
    HCHK(c, hipMalloc(&c->w2, sizeof(double) * n3 * 2 * PV));
    HCHK(c, hipMalloc(&c->spec, sizeof(double) * n3 * 2 * PV));
    c->bu = c->u;
    c->bN1 = c->N1;
    c->bN2 = c->N2;
    c->bspec = c->spec;
    HCHK(c, hipMalloc(&c->gtmp, sizeof(double) * n3 * 3));
    // eigenbasis as MFMA A fragments
    const int MT = (n + 15) / 16, nkc = (n + 15) / 16;
    const size_t a4n = arrange_A_chunk(MT) * nkc;
    double *h4 = (double *)malloc(sizeof(double) * a4n);
    HCHK(c, hipMalloc(&c->V4, sizeof(double) * a4n));
    HCHK(c, hipMalloc(&c->Vi4, sizeof(double) * a4n));
    arrange_A(n, c->hVi, h4, MT, nkc);
    HCHK(c, hipMemcpy(c->Vi4, h4, sizeof(double) * a4n, hipMemcpyHostToDevice));
    arrange_A(n, c->hV, h4, MT, nkc);
    HCHK(c, hipMemcpy(c->V4, h4, sizeof(double) * a4n, hipMemcpyHostToDevice));
    free(h4);
  } else {
    etd_fields<float> &F = c->f32;
    HCHK(c, hipMalloc(&F.u, sizeof(float) * n3 * 2 * V));  // [u | N1]
    F.N1 = F.u + n3 * V;
    HCHK(c, hipMemsetAsync(F.u, 0, sizeof(float) * n3 * 2 * V, c->stream));
    if (p->etd2) HCHK(c, hipMalloc(&F.N2, sizeof(float) * n3 * V));
    HCHK(c, hipMalloc(&F.w1, sizeof(float) * n3 * 2 * V));
    HCHK(c, hipMalloc(&F.w2, sizeof(float) * n3 * 2 * V));
    HCHK(c, hipMalloc(&F.spec, sizeof(float) * n3 * 2 * V));
    HCHK(c, hipMalloc(&F.gtmp, sizeof(float) * n3 * 3));
    const size_t sb = split_A_bytes(n, c->split_terms);
    uint16_t *h3 = (uint16_t *)malloc(sb);
    HCHK(c, hipMalloc(&c->V3, sb));
    HCHK(c, hipMalloc(&c->Vi3, sb));
    if (c->split_f16) {
      c->Vi3_inv = arrange_split_A_f16(n, c->hVi, h3);
      HCHK(c, hipMemcpy(c->Vi3, h3, sb, hipMemcpyHostToDevice));
      c->V3_inv = arrange_split_A_f16(n, c->hV, h3);
      HCHK(c, hipMemcpy(c->V3, h3, sb, hipMemcpyHostToDevice));
    } else {
      c->V3_inv = c->Vi3_inv = 1.0f;
      arrange_split_A(n, c->hVi, h3, c->split_terms);
      HCHK(c, hipMemcpy(c->Vi3, h3, sb, hipMemcpyHostToDevice));
      arrange_split_A(n, c->hV, h3, c->split_terms);
      HCHK(c, hipMemcpy(c->V3, h3, sb, hipMemcpyHostToDevice));
    }
    free(h3);
  }
  // D / h^2 per axis (read_matrices, vars.c:512,544,576)
  double *dd = (double *)malloc(sizeof(double) * 3 * n);
  for (int l = 0; l < n; l++) {
    dd[l] = c->hD[l] / p->hx / p->hx;
    dd[n + l] = c->hD[l] / p->hy / p->hy;
    dd[2 * n + l] = c->hD[l] / p->hz / p->hz;
  }
  HCHK(c, hipMalloc(&c->DDX, sizeof(double) * 3 * n));
  c->DDY = c->DDX + n;
  c->DDZ = c->DDX + 2 * n;
  HCHK(c, hipMemcpy(c->DDX, dd, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  free(dd);
  HCHK(c, hipMalloc(&c->lam, sizeof(double) * 9 * (V > 0 ? V : 1)));
  HCHK(c, hipMemcpy(c->lam, p->lam, sizeof(double) * 9 * (p->nac > 0 ? p->nac : 1), hipMemcpyHostToDevice));
  // elasticity: ELASTIC = 1 on the AC fields, 2 on the CH fields (elastic.c:689-762); 3 (BOTH_FUNCTION)
  // matches neither branch of elastic_calculate, so the reference computes no force
  c->nel = p->elastic == 1 ? p->nac : p->elastic == 2 ? p->nch : 0;
  c->el_v0 = p->elastic == 2 ? p->nac : 0;
  if (c->nel > 0) {
    c->fft_npass = fft_plan_radices(c->Ne, c->fft_radix);
    if (c->fft_npass < 0) return seterr(c, ETD_ERR_UNSUPPORTED, "N = %d has prime factors other than 2, 3, 5", c->Ne);
    c->npair = (c->nel + 1) / 2;
    c->ring = c->nel + 2;
    if (!c->mixed) {
      HCHK(c, hipMalloc(&c->Z, sizeof(double) * 2 * c->Ne3 * c->npair));
      HCHK(c, hipMalloc(&c->Y, sizeof(double) * 2 * c->Ne3 * c->npair));
      HCHK(c, hipMalloc(&c->E, sizeof(double) * c->Ne3 * c->ring));
      HCHK(c, hipMemsetAsync(c->E, 0, sizeof(double) * c->Ne3 * c->ring, c->stream));
    } else {
      HCHK(c, hipMalloc(&c->f32.Z, sizeof(float) * 2 * c->N3 * c->npair));
      HCHK(c, hipMalloc(&c->f32.Y, sizeof(float) * 2 * c->N3 * c->npair));
      HCHK(c, hipMalloc(&c->f32.E, sizeof(float) * c->N3 * c->ring));
      HCHK(c, hipMemsetAsync(c->f32.E, 0, sizeof(float) * c->N3 * c->ring, c->stream));
    }
    double *tw = (double *)malloc(sizeof(double) * 2 * c->Ne);
    const double pi = 3.14159265358979323846;
    for (int e = 0; e < c->Ne; e++) {
      tw[2 * e] = cos(2.0 * pi * e / c->Ne);
      tw[2 * e + 1] = -sin(2.0 * pi * e / c->Ne);
    }
    HCHK(c, hipMalloc(&c->tw, sizeof(double) * 2 * c->Ne));
    HCHK(c, hipMemcpy(c->tw, tw, sizeof(double) * 2 * c->Ne, hipMemcpyHostToDevice));
    free(tw);
    etd_dev_consts *k = (etd_dev_consts *)calloc(1, sizeof(etd_dev_consts));
    c4_from_voigt(p->C, k->C4);
    for (int q = 0; q < c->nel; q++) {
      const double *e = p->eps0 + 9 * q;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          double s = 0.0;
          for (int kk = 0; kk < 3; kk++)
            for (int l = 0; l < 3; l++) s += k->C4[((i * 3 + j) * 3 + kk) * 3 + l] * e[kk * 3 + l];
          k->sig[9 * q + i * 3 + j] = s;
        }
    }
    for (int a = 0; a < c->nel; a++)
      for (int b = 0; b < c->nel; b++) {
        double t0 = 0.0;
        const double *ea = p->eps0 + 9 * a, *eb = p->eps0 + 9 * b;
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++)
            for (int kk = 0; kk < 3; kk++)
              for (int l = 0; l < 3; l++) t0 += k->C4[((i * 3 + j) * 3 + kk) * 3 + l] * ea[i * 3 + j] * eb[kk * 3 + l];
        k->t0[a * ETD_MAX_VAR + b] = t0;
      }
    HCHK(c, hipMalloc(&c->dconst, sizeof(etd_dev_consts)));
    HCHK(c, hipMemcpy(c->dconst, k, sizeof(etd_dev_consts), hipMemcpyHostToDevice));
    free(k);
  }
  // energy partials
  c->npart = (int)etd_diag_partials(c);
  HCHK(c, hipMalloc(&c->part, sizeof(double) * c->npart));
  if (p->free_energy == 1) etd_calphad(&c->gbcc, &c->ghcp);
  if (p->variable_mobility) {
    etd_vm_constants(c->vm_X, c->vm_DA, c->vm_DG);
    HCHK(c, hipMalloc(&c->vm_part, sizeof(double) * (3 * 256 + 3)));
  }
  // the mixed path's CH centring relies on the constant-mobility term's conservation
  if (c->mixed && c->nch > 0 && !p->variable_mobility && p->periodic && p->elastic != 2) {
    HCHK(c, hipMalloc(&c->ch_sum, sizeof(double) * c->nch * (1 + CH_BLOCKS)));
    HCHK(c, hipMalloc(&c->ch_const, sizeof(float) * c->nch * c->n3));
  }
  c->ch_stale = 1;
  HCHK(c, hipHostMalloc(&c->hpart, sizeof(double) * c->npart, hipHostMallocDefault));
  HCHK(c, hipStreamSynchronize(c->stream));
  return ETD_OK;
}

extern "C" double etd_get_time(const etd_ctx *c) { return c ? c->t : 0.0; }
extern "C" void etd_set_time(etd_ctx *c, double t) {
  if (c) c->t = t;
}

extern "C" void etd_close(etd_ctx *c) {
  if (!c) return;
  (void)hipFree(c->ad_S0);
  (void)hipFree(c->ad_SB);
  (void)hipFree(c->ad_spec);
  (void)hipFree(c->ad_part);
  (void)hipFree(c->ch_sum);
  (void)hipFree(c->ch_const);
  if (c->ad_hpart) (void)hipHostFree(c->ad_hpart);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->bu) {  // the buffers' starts, whatever box a failed call left in view
    c->u = c->bu;
    c->N2 = c->bN2;
    c->spec = c->bspec;
  }
  (void)hipFree(c->cr_buf);
  (void)hipFree(c->cr_phi);
  if (c->cr_blas) rocblas_destroy_handle((rocblas_handle)c->cr_blas);
  (void)hipFree(c->u);
  (void)hipFree(c->N2);
  (void)hipFree(c->w1);
  (void)hipFree(c->w2);
  (void)hipFree(c->spec);
  (void)hipFree(c->gtmp);
  (void)hipFree(c->V4);
  (void)hipFree(c->Vi4);
  (void)hipFree(c->DDX);
  (void)hipFree(c->lam);
  (void)hipFree(c->Z);
  (void)hipFree(c->Y);
  (void)hipFree(c->E);
  (void)hipFree(c->tw);
  (void)hipFree(c->dconst);
  (void)hipFree(c->part);
  (void)hipFree(c->mask);
  (void)hipFree(c->vm_part);
  if (c->hpart) (void)hipHostFree(c->hpart);
  (void)hipFree(c->f32.u);
  (void)hipFree(c->f32.N2);
  (void)hipFree(c->f32.w1);
  (void)hipFree(c->f32.w2);
  (void)hipFree(c->f32.spec);
  (void)hipFree(c->f32.gtmp);
  (void)hipFree(c->f32.Z);
  (void)hipFree(c->f32.Y);
  (void)hipFree(c->f32.E);
  (void)hipFree(c->V3);
  (void)hipFree(c->Vi3);
  for (int i = 0; i < c->cap_ev; i++) (void)hipEventDestroy(c->ev[i]);
  free(c->ev);
  free(c->ev_class);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  free(c->hV);
  free(c->hVi);
  free(c->hD);
  free(c);
}

extern "C" int etd_get_params(const etd_ctx *c, etd_params *p) {
  if (!c || !p) return ETD_ERR_ARG;
  *p = c->p;
  return ETD_OK;
}

extern "C" int etd_open(const char *dir, const char *suffix, etd_ctx **out) {
  if (!dir || !suffix || !out) return ETD_ERR_ARG;
  etd_params p;
  etd_default_params(&p);
  char fn[2048];
  snprintf(fn, sizeof fn, "%s/input%s.ini", dir, suffix);
  int rc = etd_parse_ini(fn, &p);
  if (rc) {
    fprintf(stderr, "etd_open: cannot parse %s\n", fn);
    return rc;
  }
  if (p.elastic != 0) {
    snprintf(fn, sizeof fn, "%s/elastic_input.txt", dir);
    if ((rc = etd_parse_elastic(fn, &p))) {
      fprintf(stderr, "open elastic_input.txt Fail!");
      return rc;
    }
  }
  if (p.anisotropic == 1) {
    snprintf(fn, sizeof fn, "%s/anisotropic_input.txt", dir);
    if ((rc = etd_parse_anisotropic(fn, &p))) {
      fprintf(stderr, "open anisotropic_input.txt Fail!");
      return rc;
    }
  }
  snprintf(fn, sizeof fn, "%s/landau_input.txt", dir);
  FILE *f = fopen(fn, "r");
  if (f) {
    fclose(f);
    if ((rc = etd_parse_landau(fn, &p))) return rc;
  }
  const int n = p.nx;
  double *V = (double *)malloc(sizeof(double) * n * n), *Vi = (double *)malloc(sizeof(double) * n * n),
         *D = (double *)malloc(sizeof(double) * n);
  // boundary id (read_matrices, vars.c:488-582): 3 with neighbours on both sides, 0 with physical sides
  int have = etd_read_matrices(dir, n, p.periodic ? 3 : 0, V, Vi, D) == ETD_OK;
  rc = etd_create(&p, have ? V : nullptr, have ? Vi : nullptr, have ? D : nullptr, out);
  free(V);
  free(Vi);
  free(D);
  return rc;
}

// ------------------------------------------------------------------ fields
// host FP64 <-> device FP32 through a temporary host buffer (mixed path)
static int put_f32(etd_ctx *c, float *dev, const double *host, size_t cnt) {
  float *t = (float *)malloc(sizeof(float) * cnt);
  if (!t) return ETD_ERR_NOMEM;
  for (size_t i = 0; i < cnt; i++) t[i] = (float)host[i];
  hipError_t e = hipMemcpy(dev, t, sizeof(float) * cnt, hipMemcpyHostToDevice);
  free(t);
  if (e != hipSuccess) return seterr(c, ETD_ERR_DEVICE, "hipMemcpy: %s", hipGetErrorString(e));
  return ETD_OK;
}
static int get_f32(etd_ctx *c, double *host, const float *dev, size_t cnt) {
  float *t = (float *)malloc(sizeof(float) * cnt);
  if (!t) return ETD_ERR_NOMEM;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(t, dev, sizeof(float) * cnt, hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    for (size_t i = 0; i < cnt; i++) host[i] = t[i];
  free(t);
  if (e != hipSuccess) return seterr(c, ETD_ERR_DEVICE, "hipMemcpy: %s", hipGetErrorString(e));
  return ETD_OK;
}

extern "C" int etd_set_field(etd_ctx *c, int var, const double *host) {
  if (!c || !host || var < 0 || var >= c->V) return seterr(c, ETD_ERR_ARG, "set_field: bad variable %d", var);
  c->elastic_fresh = 0;
  c->carry_ok = 0;
  c->ch_stale = 1;
  if (c->mixed) return put_f32(c, c->f32.u + c->n3 * var, host, c->n3);
  for (int b = 0; b < c->P; b++)  // a decomposed run: the boxes one after another
    HCHK(c, hipMemcpyAsync(c->bu + b * c->bs + c->n3 * var, host + b * c->n3, sizeof(double) * c->n3,
                           hipMemcpyHostToDevice, c->stream));
  HCHK(c, hipStreamSynchronize(c->stream));
  return ETD_OK;
}

extern "C" int etd_set_mask(etd_ctx *c, const int *mask) {
  if (!c) return ETD_ERR_ARG;
  if (mask && c->P > 1) return seterr(c, ETD_ERR_UNSUPPORTED, "set_mask: one box only (px,py,pz = 1,1,1)");
  if (!mask) {
    if (c->mask) HCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(c->mask);
    c->mask = nullptr;
    return ETD_OK;
  }
  unsigned char *h = (unsigned char *)malloc(c->n3);
  if (!h) return ETD_ERR_NOMEM;
  for (size_t l = 0; l < c->n3; l++) h[l] = mask[l] != 0;
  hipError_t e = hipSuccess;
  if (!c->mask) e = hipMalloc(&c->mask, c->n3);
  if (e == hipSuccess) e = hipMemcpyAsync(c->mask, h, c->n3, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  free(h);
  if (e != hipSuccess) return seterr(c, ETD_ERR_DEVICE, "set_mask: %s", hipGetErrorString(e));
  c->elastic_fresh = 0;
  return ETD_OK;
}

extern "C" int etd_get_field(etd_ctx *c, int var, double *host) {
  if (!c || !host || var < 0 || var >= c->V) return seterr(c, ETD_ERR_ARG, "get_field: bad variable %d", var);
  if (c->mixed) return get_f32(c, host, c->f32.u + c->n3 * var, c->n3);
  for (int b = 0; b < c->P; b++)
    HCHK(c, hipMemcpyAsync(host + b * c->n3, c->bu + b * c->bs + c->n3 * var, sizeof(double) * c->n3,
                           hipMemcpyDeviceToHost, c->stream));
  HCHK(c, hipStreamSynchronize(c->stream));
  return ETD_OK;
}

extern "C" int etd_init_noise(etd_ctx *c, int var, unsigned long long seed, double mean, double amp) {
  if (!c || var < 0 || var >= c->V) return seterr(c, ETD_ERR_ARG, "init_noise: bad variable %d", var);
  double *h = (double *)malloc(sizeof(double) * c->n3 * c->P);
  if (!h) return ETD_ERR_NOMEM;
  // box b of a decomposed run draws from seed + 1000003 b
  for (int b = 0; b < c->P; b++) etd_host_noise(c->n, c->g, seed + 1000003ULL * b, mean, amp, h + b * c->n3);
  int rc = etd_set_field(c, var, h);
  free(h);
  return rc;
}

// ------------------------------------------------------------------ boxes
// The per-box launches of a decomposed run see box b's fields through u, N1,
// N2 and spec (and make_nl's box coordinates); box 0 is the buffers' start,
// which every other launch uses.
static void box_view(etd_ctx *c, int b) {
  const size_t o = (size_t)b * c->bs;
  c->u = c->bu + o;
  c->N1 = c->bN1 + o;
  c->N2 = c->bN2 ? c->bN2 + o : nullptr;
  c->spec = c->bspec + o;
  c->box = b;
}

// the nonlinear terms of every box (out, sub: box 0's first field, the
// others at stride bs); each box reads its neighbours' fields for the
// outside values, so all boxes are evaluated before any is advanced
template <typename T>
static void nonlinear_boxes(etd_ctx *c, T *out, const T *sub, int ch_centre) {
  if (c->P == 1) {
    launch_nonlinear<T>(c, out, sub, ch_centre);
    return;
  }
  if constexpr (sizeof(T) == 8) {
    for (int b = 0; b < c->P; b++) {
      box_view(c, b);
      launch_nonlinear<T>(c, out + b * c->bs, sub ? sub + b * c->bs : nullptr, ch_centre);
    }
    box_view(c, 0);
  }
}

// phi scaling of every box: u-hat of box b at spec + b bs, N-hat after all
// boxes' u-hat (the forward transform's input was [u of all boxes | N1 of all])
template <typename T>
static void phiscale_boxes(etd_ctx *c, T *spec, const T *nhat0, int order2) {
  const size_t bs = c->bs;
  for (int b = 0; b < c->P; b++)
    launch_phiscale<T>(c, spec + b * bs, nhat0 ? nhat0 + b * bs : nullptr, spec + b * bs, 0, c->V, order2);
}

// ------------------------------------------------------------------ step
// Three rotating mode products by V^-1 (forward) or V (backward) over nf
// fields; FP64 MFMA on the plain path, split-BF16 MFMA on the mixed path.
template <typename T>
void modeprod3(etd_ctx *c, int forward, const T *in, T *out, int nf, int accumulate) {
  const size_t n3 = c->n3;
  const etd_fields<T> F = fields<T>(c);
  if (c->prof) c->mp_fields += 3LL * nf;
  if constexpr (sizeof(T) == 8) {
    const double *A = forward ? c->Vi4 : c->V4;
    launch_modeprod(c, A, in, n3, F.w1, n3, nf, 1, 0);
    launch_modeprod(c, A, F.w1, n3, F.w2, n3, nf, 1, 0);
    launch_modeprod(c, A, F.w2, n3, out, n3, nf, 1, accumulate);
  } else {
    const uint16_t *A = forward ? c->Vi3 : c->V3;
    const float ai = forward ? c->Vi3_inv : c->V3_inv;
    // contracting the slowest axis rotates the other way: z, y, x, back to xyz
    const int rot = c->split_zs ? 2 : 1;
    launch_modeprod_split(c, A, in, n3, F.w1, n3, nf, rot, 0, ai);
    launch_modeprod_split(c, A, F.w1, n3, F.w2, n3, nf, rot, 0, ai);
    launch_modeprod_split(c, A, F.w2, n3, out, n3, nf, rot, accumulate, ai);
  }
}

// ScLETD.c:87-189 for every field at once: u and N(u) share the forward
// transforms (one launch over 2V fields), the phi update writes u-hat in
// place, three inverse products return u in xyz.
// ghost fill, elasticity, N1 = N(u), spec = V^-1 x3 [u | N1]: the part of a
// step that does not depend on dt
//
// With the AC spectrum carried (carry_ok: the AC fields of u were transformed
// back from carry_alpha s1 + carry_beta s2, by default spec[0 .. nac)), only
// [u_CH | N1] is transformed; the AC spectrum of the ghost-filled fields is
// the carried one plus the transform of the fill's changes, which phi_stage
// forms in its pass (carry_pend).
template <typename T>
void stage_forward(etd_ctx *c) {
  const int V = c->V;
  const etd_fields<T> F = fields<T>(c);
  const bool carry = c->carry_ok && carry_possible(c);
  const T *b1 = c->carry_s1 ? (const T *)c->carry_s1 : F.spec, *b2 = (const T *)c->carry_s2;
  const double ba = c->carry_alpha, bb = c->carry_beta;
  c->carry_ok = 0;
  c->carry_pend = 0;
  carry_base(c, nullptr, 1.0, nullptr, 0.0);
  if (carry) {
    T *dz, *dy, *dx;
    carry_deltas<T>(c, &dz, &dy, &dx);
    launch_ghost_fill<T>(c, F.u, c->nac, c->n3, dz, dy, dx);
    if (c->nch) launch_ghost_fill<T>(c, F.u + c->n3 * c->nac, c->nch, c->n3);
  } else {
    launch_ghost_fill<T>(c, F.u, V, c->n3);
  }
  // an adaptive acceptance test may already have evaluated felas of exactly
  // these fields (same ghost fill, same kernels): reuse it
  if (c->nel > 0 && !c->elastic_fresh) launch_elastic<T>(c);
  c->elastic_fresh = 0;
  nonlinear_boxes<T>(c, F.N1, (const T *)nullptr, 1);  // mixed: CH terms centred, N1 - N(c_mean)
  if constexpr (sizeof(T) == 4) ch_center(c, F.u, nullptr, -1);
  if (carry) {
    modeprod3<T>(c, 1, F.u + c->n3 * c->nac, F.spec + c->n3 * c->nac, c->nch + V, 0);
    // the correction joins the phi update (phi_stage)
    c->carry_pend = 1;
    c->pend_s1 = b1;
    c->pend_alpha = ba;
    c->pend_s2 = b2;
    c->pend_beta = bb;
  } else {
    modeprod3<T>(c, 1, F.u, F.spec, 2 * V * c->P, 0);
  }
  // u stays centred: every caller overwrites it with the backward transform;
  // the ETD2 stage needs N1 itself (N(a) - N1)
  if constexpr (sizeof(T) == 4)
    if (c->p.etd2) ch_center(c, nullptr, F.N1, +1);
}

template <typename T>
void phi_stage(etd_ctx *c, T *out, double dt, T *out2, double dt2) {
  const etd_fields<T> F = fields<T>(c);
  const size_t n3 = c->n3;
  const int V = c->V, nac = c->nac;
  const T *nhat = F.spec + n3 * V * c->P;
  if (!c->carry_pend) {
    if (out2)
      launch_phiscale_pair<T>(c, F.spec, nhat, out, dt, out2, dt2, 0, V);
    else
      phiscale_boxes<T>(c, F.spec, nhat, 0);
    return;
  }
  c->carry_pend = 0;  // only single-box contexts carry: N-hat is at spec + n3 V
  carry_correct_phi<T>(c, (const T *)c->pend_s1, c->pend_alpha, (const T *)c->pend_s2, c->pend_beta, nhat, out, dt,
                       out2, dt2);
  if (c->nch) {  // the CH fields were transformed: their plain phi update
    const size_t o = n3 * nac;
    if (out2)
      launch_phiscale_pair<T>(c, F.spec + o, nhat + o, out + o, dt, out2 + o, dt2, nac, c->nch);
    else
      launch_phiscale<T>(c, F.spec + o, nhat + o, out + o, nac, c->nch, 0);
  }
}

// The ETD2 corrector (calc.c:537-583) on the predictor a in u (its spectrum
// a-hat was transformed back into u; N1 = N(u_n) is in place): u = a +
// S phi2 S^-1 (N(a) - N1) at the context's dt.  With spec_out, the fill's
// changes of a are recorded and spec_out = a-hat + S^-1 (those changes) +
// phi2 S^-1 dN, the spectrum of the new u, which a following step carries
// (spec_out may be a-hat).
template <typename T>
void etd2_stage(etd_ctx *c, const T *ahat, T *spec_out) {
  const int V = c->V;
  const size_t n3 = c->n3;
  const etd_fields<T> F = fields<T>(c);
  const int PV = V * c->P;
  if constexpr (sizeof(T) == 4) ch_center(c, F.u, nullptr, +1);
  if (spec_out) {  // the fill's changes of a enter the final spectrum too
    T *dz, *dy, *dx;
    carry_deltas<T>(c, &dz, &dy, &dx);
    launch_ghost_fill<T>(c, F.u, c->nac, n3, dz, dy, dx);
    if (c->nch) launch_ghost_fill<T>(c, F.u + n3 * c->nac, c->nch, n3);
  } else {
    launch_ghost_fill<T>(c, F.u, V, n3);
  }
  if (c->nel > 0) launch_elastic<T>(c);
  nonlinear_boxes<T>(c, F.N2, (const T *)F.N1, 0);  // N(a) - N(u_n)
  // the corrector's spectra in the N-hat half (free after the phi update),
  // so that a-hat stays where it is
  T *dsp = F.spec + n3 * PV;
  modeprod3<T>(c, 1, F.N2, dsp, PV, 0);
  phiscale_boxes<T>(c, dsp, (const T *)nullptr, 1);
  modeprod3<T>(c, 0, dsp, F.u, PV, 1);  // u = a + S phi2 S^-1 dN
  if (spec_out) carry_correct<T>(c, ahat, 1.0, dsp, 1.0, spec_out);
  if constexpr (sizeof(T) == 4) ch_mean_restore(c, F.u);
}

template <typename T>
int step_once_t(etd_ctx *c) {
  const etd_fields<T> F = fields<T>(c);
  if constexpr (sizeof(T) == 4) ch_mean_mark(c, F.u);
  stage_forward<T>(c);
  const int PV = c->V * c->P;
  phi_stage<T>(c, F.spec, c->p.dt, (T *)nullptr, 0.0);  // in place, at the context's dt
  modeprod3<T>(c, 0, F.spec, F.u, PV, 0);
  if (c->p.etd2) {
    etd2_stage<T>(c, F.spec, c->carry_next ? F.spec : (T *)nullptr);
  } else {
    if constexpr (sizeof(T) == 4) ch_mean_restore_centered(c, F.u);  // adds the mean back
  }
  HCHK(c, hipGetLastError());
  return carry_status(c);
}

template void modeprod3<double>(etd_ctx *, int, const double *, double *, int, int);
template void modeprod3<float>(etd_ctx *, int, const float *, float *, int, int);
template void phi_stage<double>(etd_ctx *, double *, double, double *, double);
template void phi_stage<float>(etd_ctx *, float *, double, float *, double);
template void stage_forward<double>(etd_ctx *);
template void stage_forward<float>(etd_ctx *);
template void etd2_stage<double>(etd_ctx *, const double *, double *);
template void etd2_stage<float>(etd_ctx *, const float *, float *);
template int step_once_t<double>(etd_ctx *);
template int step_once_t<float>(etd_ctx *);

// one fixed step; afterwards spec[0 .. nac) is the AC spectrum of u, which
// the next step carries (the adaptive driver and the ensemble call
// step_once_t directly and never set carry_ok)
static int step_once(etd_ctx *c) {
  c->carry_next = carry_ready(c);
  const int rc = c->mixed ? step_once_t<float>(c) : step_once_t<double>(c);
  c->carry_ok = rc == ETD_OK && c->carry_next;
  carry_base(c, nullptr, 1.0, nullptr, 0.0);
  c->carry_next = 0;
  return rc;
}
// check_soln_new of a decomposed run: each box over its own ranges, then the
// reduction of field.c:513-516 (max, min, and the sums in box order)
static int diag_boxes(etd_ctx *c, double *st) {
  const int nv = 4 * c->V + 1;
  double *sb = (double *)malloc(sizeof(double) * nv);
  if (!sb) return ETD_ERR_NOMEM;
  int rc = ETD_OK;
  for (int b = 0; b < c->P && !rc; b++) {
    box_view(c, b);
    rc = launch_diag<double>(c, b == 0 ? st : sb);
    if (rc || b == 0) continue;
    for (int v = 0; v < c->V; v++) {
      if (sb[4 * v] > st[4 * v]) st[4 * v] = sb[4 * v];
      if (sb[4 * v + 1] < st[4 * v + 1]) st[4 * v + 1] = sb[4 * v + 1];
      st[4 * v + 2] += sb[4 * v + 2];
      st[4 * v + 3] += sb[4 * v + 3];
    }
    st[4 * c->V] += sb[4 * c->V];
  }
  box_view(c, 0);
  free(sb);
  return rc;
}
static int diag_any(etd_ctx *c, double *st) {
  if (c->P > 1) return diag_boxes(c, st);
  return c->mixed ? launch_diag<float>(c, st) : launch_diag<double>(c, st);
}

extern "C" int etd_step(etd_ctx *c, int nsteps) {
  if (!c || nsteps < 0) return ETD_ERR_ARG;
  for (int s = 0; s < nsteps; s++) {
    int rc = step_once(c);
    if (rc) return rc;
    c->t += c->p.dt;
  }
  HCHK(c, hipStreamSynchronize(c->stream));
  return ETD_OK;
}

extern "C" int etd_diag(etd_ctx *c, double *stats) {
  if (!c || !stats) return ETD_ERR_ARG;
  int rc = diag_any(c, stats);
  if (rc) return seterr(c, rc, "diag: %s", hipGetErrorString(hipGetLastError()));
  return ETD_OK;
}

extern "C" int etd_write_csv(etd_ctx *c, int iter, const double *stats) {
  if (!c || !stats) return ETD_ERR_ARG;
  int rc = etd_csv_rows(c->p.work_dir, c->nac, c->nch, iter, stats);
  if (rc) return seterr(c, rc, "cannot write CSV under %s", c->p.work_dir);
  return ETD_OK;
}

extern "C" int etd_write_field(etd_ctx *c, int var, int ioutput) {
  if (!c || var < 0 || var >= c->V) return ETD_ERR_ARG;
  double *h = (double *)malloc(sizeof(double) * c->n3 * c->P);
  if (!h) return ETD_ERR_NOMEM;
  int rc = etd_get_field(c, var, h);
  for (int b = 0; b < c->P && !rc; b++)  // one file per box, named by its rank coordinates
    rc = etd_field_file(c->p.data_dir, var, ioutput, c->n, h + b * c->n3, b % c->px, (b / c->px) % c->py,
                        b / (c->px * c->py));
  free(h);
  return rc;
}

// ------------------------------------------------------------------ per-kernel entry points
// On a mixed-precision context these run the FP32 kernels (split-BF16 mode
// products); host buffers stay FP64 and are converted at the copy.
extern "C" int etd_k_transfer(etd_ctx *c) {
  if (!c) return ETD_ERR_ARG;
  c->carry_ok = 0;
  if (c->mixed)
    launch_ghost_fill<float>(c, c->f32.u, c->V, c->n3);
  else
    launch_ghost_fill<double>(c, c->u, c->V, c->n3);
  HCHK(c, hipGetLastError());
  HCHK(c, hipStreamSynchronize(c->stream));
  return ETD_OK;
}

template <typename T>
static int k_elastic_t(etd_ctx *c, double *host_felas) {
  launch_elastic<T>(c);
  if (host_felas) {
    // felas p of box b at (p P + b) n3
    const size_t cnt = c->n3 * c->nel * c->P;
    T *d;
    HCHK(c, hipMalloc(&d, sizeof(T) * cnt));
    if (c->P == 1) {
      launch_felas_materialize<T>(c, d);
    } else if constexpr (sizeof(T) == 8) {
      T *t;
      HCHK(c, hipMalloc(&t, sizeof(T) * c->n3 * c->nel));
      for (int b = 0; b < c->P; b++) {
        box_view(c, b);
        launch_felas_materialize<T>(c, t);
        for (int q = 0; q < c->nel; q++)
          HCHK(c, hipMemcpyAsync(d + ((size_t)q * c->P + b) * c->n3, t + (size_t)q * c->n3, sizeof(T) * c->n3,
                                 hipMemcpyDeviceToDevice, c->stream));
      }
      box_view(c, 0);
      HCHK(c, hipStreamSynchronize(c->stream));
      HCHK(c, hipFree(t));
    }
    int rc = ETD_OK;
    if constexpr (sizeof(T) == 8) {
      HCHK(c, hipMemcpyAsync(host_felas, d, sizeof(double) * cnt, hipMemcpyDeviceToHost, c->stream));
    } else {
      rc = get_f32(c, host_felas, d, cnt);
    }
    HCHK(c, hipStreamSynchronize(c->stream));
    HCHK(c, hipFree(d));
    if (rc) return rc;
  }
  HCHK(c, hipGetLastError());
  HCHK(c, hipStreamSynchronize(c->stream));
  return ETD_OK;
}

extern "C" int etd_k_elastic(etd_ctx *c, double *host_felas) {
  if (!c) return ETD_ERR_ARG;
  if (c->nel == 0) return seterr(c, ETD_ERR_ARG, "elasticity is off");
  return c->mixed ? k_elastic_t<float>(c, host_felas) : k_elastic_t<double>(c, host_felas);
}

extern "C" int etd_k_fft3d(etd_ctx *c, double *data, int dir) {
  if (!c || !data || (dir != 1 && dir != -1)) return ETD_ERR_ARG;
  if (!c->tw) return seterr(c, ETD_ERR_ARG, "fft3d: the context has no elastic FFT plan");
  const size_t cnt = 2 * c->Ne3;
  if (!c->mixed) {
    double *d;
    HCHK(c, hipMalloc(&d, sizeof(double) * cnt));
    HCHK(c, hipMemcpy(d, data, sizeof(double) * cnt, hipMemcpyHostToDevice));
    fft3d<double>(c, d, 1, dir);
    HCHK(c, hipGetLastError());
    HCHK(c, hipStreamSynchronize(c->stream));
    HCHK(c, hipMemcpy(data, d, sizeof(double) * cnt, hipMemcpyDeviceToHost));
    (void)hipFree(d);
    return ETD_OK;
  }
  float *d;
  HCHK(c, hipMalloc(&d, sizeof(float) * cnt));
  int rc = put_f32(c, d, data, cnt);
  if (!rc) {
    fft3d<float>(c, d, 1, dir);
    rc = get_f32(c, data, d, cnt);
  }
  (void)hipFree(d);
  return rc;
}

extern "C" int etd_k_nonlinear(etd_ctx *c, int var, double *host_out) {
  if (!c || !host_out || var < 0 || var >= c->V) return ETD_ERR_ARG;
  if (c->mixed) {
    launch_nonlinear<float>(c, c->f32.N1, nullptr);
    HCHK(c, hipGetLastError());
    return get_f32(c, host_out, c->f32.N1 + c->n3 * var, c->n3);
  }
  nonlinear_boxes<double>(c, c->N1, (const double *)nullptr, 0);
  HCHK(c, hipGetLastError());
  for (int b = 0; b < c->P; b++)  // a decomposed run: the boxes one after another
    HCHK(c, hipMemcpyAsync(host_out + b * c->n3, c->N1 + b * c->bs + c->n3 * var, sizeof(double) * c->n3,
                           hipMemcpyDeviceToHost, c->stream));
  HCHK(c, hipStreamSynchronize(c->stream));
  return ETD_OK;
}

extern "C" int etd_k_modeprod(etd_ctx *c, int n, const double *A, const double *in, double *out, int rotate) {
  if (!c || !A || !in || !out || rotate < 0 || rotate > 3) return ETD_ERR_ARG;
  if (n != c->n) return seterr(c, ETD_ERR_ARG, "modeprod: n must equal the context edge %d", c->n);
  if (rotate >= 2 && !(c->split_f16 && f16_zs_supported(n)))
    return seterr(c, ETD_ERR_UNSUPPORTED, "modeprod: the slowest-axis contraction (rotate %d) is the FP16 split-2 "
                  "kernel's, for n <= 256", rotate);
  if (c->mixed) {
    const size_t sb = split_A_bytes(n, c->split_terms);
    uint16_t *h3 = (uint16_t *)malloc(sb);
    if (!h3) return ETD_ERR_NOMEM;
    float ai = 1.0f;
    if (c->split_f16)
      ai = arrange_split_A_f16(n, A, h3);
    else
      arrange_split_A(n, A, h3, c->split_terms);
    uint16_t *dA;
    float *din, *dout;
    HCHK(c, hipMalloc(&dA, sb));
    HCHK(c, hipMalloc(&din, sizeof(float) * c->n3));
    HCHK(c, hipMalloc(&dout, sizeof(float) * c->n3));
    HCHK(c, hipMemcpy(dA, h3, sb, hipMemcpyHostToDevice));
    free(h3);
    int rc = put_f32(c, din, in, c->n3);
    if (!rc) {
      launch_modeprod_split(c, dA, din, c->n3, dout, c->n3, 1, rotate, 0, ai);
      HCHK(c, hipGetLastError());
      rc = get_f32(c, out, dout, c->n3);
    }
    (void)hipFree(dA);
    (void)hipFree(din);
    (void)hipFree(dout);
    return rc;
  }
  const int MT = (n + 15) / 16, nkc = (n + 15) / 16;
  const size_t a4n = arrange_A_chunk(MT) * nkc;
  double *h4 = (double *)malloc(sizeof(double) * a4n);
  arrange_A(n, A, h4, MT, nkc);
  double *dA, *din, *dout;
  HCHK(c, hipMalloc(&dA, sizeof(double) * a4n));
  HCHK(c, hipMalloc(&din, sizeof(double) * c->n3));
  HCHK(c, hipMalloc(&dout, sizeof(double) * c->n3));
  HCHK(c, hipMemcpy(dA, h4, sizeof(double) * a4n, hipMemcpyHostToDevice));
  HCHK(c, hipMemcpy(din, in, sizeof(double) * c->n3, hipMemcpyHostToDevice));
  free(h4);
  launch_modeprod(c, dA, din, c->n3, dout, c->n3, 1, rotate, 0);
  HCHK(c, hipGetLastError());
  HCHK(c, hipStreamSynchronize(c->stream));
  HCHK(c, hipMemcpy(out, dout, sizeof(double) * c->n3, hipMemcpyDeviceToHost));
  (void)hipFree(dA);
  (void)hipFree(din);
  (void)hipFree(dout);
  return ETD_OK;
}

extern "C" int etd_k_phiscale(etd_ctx *c, int var, double *uhat, const double *nhat) {
  if (!c || !uhat || !nhat || var < 0 || var >= c->V) return ETD_ERR_ARG;
  if (c->mixed) {
    float *d;
    HCHK(c, hipMalloc(&d, sizeof(float) * 2 * c->n3));
    int rc = put_f32(c, d, uhat, c->n3);
    if (!rc) rc = put_f32(c, d + c->n3, nhat, c->n3);
    if (!rc) {
      launch_phiscale<float>(c, d, d + c->n3, d, var, 1, 0);
      HCHK(c, hipGetLastError());
      rc = get_f32(c, uhat, d, c->n3);
    }
    (void)hipFree(d);
    return rc;
  }
  double *d;
  HCHK(c, hipMalloc(&d, sizeof(double) * 2 * c->n3));
  HCHK(c, hipMemcpy(d, uhat, sizeof(double) * c->n3, hipMemcpyHostToDevice));
  HCHK(c, hipMemcpy(d + c->n3, nhat, sizeof(double) * c->n3, hipMemcpyHostToDevice));
  launch_phiscale<double>(c, d, d + c->n3, d, var, 1, 0);
  HCHK(c, hipGetLastError());
  HCHK(c, hipStreamSynchronize(c->stream));
  HCHK(c, hipMemcpy(uhat, d, sizeof(double) * c->n3, hipMemcpyDeviceToHost));
  (void)hipFree(d);
  return ETD_OK;
}

// ------------------------------------------------------------------ counter calibration
extern "C" int etd_calib_copy(etd_ctx *c, size_t bytes) {
  if (!c) return ETD_ERR_ARG;
  const size_t esz = c->mixed ? sizeof(float) : sizeof(double);
  const size_t cap = esz * c->n3 * 2 * c->V;
  if (bytes == 0 || bytes % 16 || bytes > cap)
    return seterr(c, ETD_ERR_ARG, "calibration copy of %zu bytes: must be a positive multiple of 16 up to %zu", bytes,
                  cap);
  if (c->mixed)
    launch_copy16(c, c->f32.spec, c->f32.w1, bytes / 16);
  else
    launch_copy16(c, c->spec, c->w1, bytes / 16);
  HCHK(c, hipGetLastError());
  HCHK(c, hipStreamSynchronize(c->stream));
  return ETD_OK;
}

// ------------------------------------------------------------------ timing
extern "C" int etd_time_steps(etd_ctx *c, int nsteps, int with_diag, double *ms, int *launches) {
  if (!c || nsteps < 1 || !ms) return ETD_ERR_ARG;
  // enough event slots for every bracketed launch group of nsteps steps
  const int per_step = 2 * (2 * 12 + 2 * 4 + 8 + 6);
  const int need = per_step * nsteps + 16;
  if (c->cap_ev < need) {
    for (int i = 0; i < c->cap_ev; i++) (void)hipEventDestroy(c->ev[i]);
    free(c->ev);
    free(c->ev_class);
    c->ev = (hipEvent_t *)calloc(need, sizeof(hipEvent_t));
    c->ev_class = (int *)calloc(need, sizeof(int));
    c->cap_ev = 0;
    for (int i = 0; i < need; i++) {
      HCHK(c, hipEventCreate(&c->ev[i]));
      c->cap_ev++;
    }
  }
  hipEvent_t t0, t1;
  HCHK(c, hipEventCreate(&t0));
  HCHK(c, hipEventCreate(&t1));
  HCHK(c, hipStreamSynchronize(c->stream));
  c->prof = 1;
  c->nev = 0;
  c->mp_fields = 0;
  double *st = (double *)malloc(sizeof(double) * (4 * c->V + 1));
  HCHK(c, hipEventRecord(t0, c->stream));
  int rc = ETD_OK;
  for (int s = 0; s < nsteps && rc == ETD_OK; s++) {
    if (with_diag) rc = diag_any(c, st);
    if (rc == ETD_OK) rc = step_once(c);
  }
  HCHK(c, hipEventRecord(t1, c->stream));
  HCHK(c, hipEventSynchronize(t1));
  c->prof = 0;
  free(st);
  if (rc) return rc;
  float tot;
  HCHK(c, hipEventElapsedTime(&tot, t0, t1));
  for (int k = 0; k < 1 + PC_COUNT; k++) ms[k] = 0.0;
  ms[0] = tot;
  ms[1 + PC_COUNT] = (double)c->mp_fields;
  int cnt[PC_COUNT] = {0};
  for (int i = 0; i < c->nev; i++) {
    if (c->ev_class[i] < 0) continue;
    const int cls = c->ev_class[i];
    int j = i + 1;
    while (j < c->nev && c->ev_class[j] != -1 - cls) j++;
    if (j >= c->nev) break;
    float e;
    HCHK(c, hipEventElapsedTime(&e, c->ev[i], c->ev[j]));
    ms[1 + cls] += e;
    cnt[cls]++;
  }
  if (launches) {
    for (int k = 0; k < PC_COUNT; k++) launches[k] = cnt[k];
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return ETD_OK;
}
