/*
 * etd_internal.h -- device state and kernel launchers of libetd (gfx950).
 *
 * HBM layout (plain FP64 path), n = extended edge, n3 = n^3, V = nac + nch
 * (a decomposed run repeats every field buffer per box, box-major; the
 * elastic buffers then cover the global grid of edge Ne = px N):
 *   u     V x n3   fields, xyz (x fastest), ghosts included
 *   N1    V x n3   N(u_n); N2 V x n3: N(a) - N(u_n) for the ETD2 corrector
 *   w1,w2 2V x n3  mode-product intermediates (u and N batched together)
 *   spec  2V x n3  spectral u-hat / N-hat in the rotated x'y'z' layout
 *   Vi4,V4         V^-1 and V rearranged as MFMA A-fragments (see modeprod)
 *   Z,Y   P x N^3  complex pairs (eta_2r + i eta_2r+1) for the elastic FFTs
 *   E     R x N^3  ring of elastic driving forces (interior), R = nac + 2
 */
#ifndef ETD_INTERNAL_H
#define ETD_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "etd.h"

template <typename T>
struct cplx;
template <>
struct cplx<double> {
  typedef double2 t;
};
template <>
struct cplx<float> {
  typedef float2 t;
};

/* The step's field buffers for one storage type (see the layout above);
 * Z and Y hold interleaved complex values. */
template <typename T>
struct etd_fields {
  T *u, *N1, *N2, *w1, *w2, *spec, *gtmp, *Z, *Y, *E;
};

#define ETD_MAXN 512

struct etd_dev_consts {
  /* elasticity, shared by every k-point */
  double C4[81];
  double sig[ETD_MAX_VAR * 9];
  double t0[ETD_MAX_VAR * ETD_MAX_VAR];
};

struct etd_ctx {
  etd_params p;
  int device;
  int n, g, N, V, nac, nch;
  size_t n3, N3;
  /* decomposed run (procs px, py, pz, periodic): P boxes of edge n, box b at
   * (b % px, (b / px) % py, b / (px py)); every field buffer holds the boxes
   * one after another, box b's V fields at b bs (bs = V n3).  box: the box
   * the per-box launches see (u, N1, N2, spec point at its fields); Ne: edge
   * of the elastic grid (px N; N for one box).  bu, bN1, bN2, bspec: box 0,
   * the buffers' starts. */
  int px, py, pz, P, box;
  int Ne;
  size_t Ne3, bs;
  double *bu, *bN1, *bN2, *bspec;
  hipStream_t stream;
  char err[512];
  /* host copies */
  double *hV, *hVi, *hD;
  /* device */
  double *u, *N1, *N2, *w1, *w2, *spec;
  double *V4, *Vi4;        /* MFMA-arranged A operands */
  double *DDX, *DDY, *DDZ; /* D / h^2 per axis */
  double *lam;             /* nac x 9 */
  double *gtmp;            /* 3 x n3 gradient scratch for F2, or CH t */
  /* elastic */
  double *Z, *Y;           /* npair x N^3 complex */
  double *E;               /* ring x N^3 */
  double *tw;              /* N complex twiddles exp(-2 pi i e / N) */
  etd_dev_consts *dconst;
  int npair, ring;
  /* the fields the elastic force acts on: AC fields (ELASTIC = 1) or CH
   * fields (ELASTIC = 2), nel of them from field el_v0; nel = 0 without */
  int nel, el_v0;
  long long elastic_evals; /* completed elastic evaluations */
  int elastic_fresh;       /* the latest evaluation is of the current fields (adaptive energy test) */
  int fft_radix[16], fft_npass;
  int fft_variant;         /* 1: two-stage register FFT for N % 16 == 0 (default), 0: radix passes in LDS */
  int split_waves;         /* mixed mode products: 8 or 4 waves per workgroup, 0 by size */
  int split_persistent;    /* 1: persistent workgroups looping over tiles (default) */
  int split_kernel;        /* BF16 split-3: 0: field chunks staged through LDS, 8 waves x 32 columns;
                              1: field chunks prefetched into registers; 2: LDS-staged, 4 waves x 64
                              columns; 3: two workgroups per CU; 4: wave grid; 9: the default with
                              32-byte field copies (default: whole 128-byte runs).  FP16 split-2 (n <= 256):
                              0: three-slot field ring, whole-line copies; 5: three-slot ring, 32-byte
                              copies; 7: two-slot ring; 10 (the default where f16_zs_supported): the
                              three-slot ring contracting the slowest axis (whole 1 KB rows per copy) */
  int split_zs;            /* the step's FP16 mode products contract the slowest axis (split_kernel 10) */
  int nl_batch;            /* AC fields whose loads the nonlinear kernel issues together (4, 8, 12, 16);
                              0: by storage type and field count */
  int num_cu;              /* compute units of the device */
  /* mixed precision (ETD_PREC_MIXED): the same buffers in FP32, and V / V^-1
   * split into split_terms BF16 terms in MFMA fragment order */
  int mixed;
  int split_terms;
  etd_fields<float> f32;
  uint16_t *V3, *Vi3;
  int split_f16;           /* ETD_PREC_MIXED_F16: V3 / Vi3 hold two FP16 terms of V / V^-1 scaled by 1 / *_inv */
  float V3_inv, Vi3_inv;   /* the power of two that undoes that scaling (1 for BF16) */
  /* simulated time; adaptive stepping scratch (allocated on first use):
   * saved state, full-step result, phi-scaled spectrum (V x n3 each, in the
   * storage type), reduction partials */
  double t;
  /* mixed path with CH fields: the conserved interior sum of each CH field
   * (nch targets, then nch x CH_BLOCKS per-block partials), restoring the mean
   * that the split-BF16 V V^-1 does not conserve exactly; ch_stale: the
   * fields were set from outside, the targets are taken at the next step */
  double *ch_sum;
  int ch_stale;
  float *ch_const;         /* nch x n3: N(c_mean), the CH nonlinear term of the uniform field at the target mean */
  void *ad_S0, *ad_SB, *ad_spec;
  double *ad_part, *ad_hpart;
  struct etd_mlp *mlp;
  /* beta->alpha grain mask (n^3 bytes, 1 where the grain is, or NULL) */
  unsigned char *mask;
  /* variable-mobility CH: per-block maxima of the three mobilities, then C_q */
  double *vm_part;
  double vm_X[3], vm_DA[3], vm_DG[3];  /* x (1 - x), D_A, D_G of Cu, Mn, Ni */
  double gbcc, ghcp;       /* CALPHAD Gibbs energies (free_energy 1) */
  /* the AC spectrum carried between steps (etd_carry.hip): carry_ok when
   * the AC fields of u were transformed back from the spectrum
   * carry_alpha s1 + carry_beta s2 (s1 null: spec[0 .. nac); s2 null: none)
   * and nothing has changed u since; no_carry turns it off (ETD_CARRY=0,
   * ensemble members); cr_buf the correction's buffers */
  int carry_ok, no_carry;
  const void *carry_s1, *carry_s2;
  double carry_alpha, carry_beta;
  /* set by stage_forward: the AC spectrum still lacks the fill's changes,
   * which phi_stage adds from this base on its way */
  int carry_pend;
  const void *pend_s1, *pend_s2;
  double pend_alpha, pend_beta;
  void *cr_blas;           /* rocBLAS handle of the correction's batched GEMMs */
  int carry_next;          /* inside step_once: leave spec[0 .. nac) as the AC spectrum of u */
  void *cr_buf;
  double *cr_phi;          /* the per-axis factors of exp(-dt H) of the phi-fused correction */
  int blas_status;         /* first failing rocBLAS status of the correction since the last check */
  long long mp_fields;     /* field passes of the mode-product launches while timing */
  long long ad_fetches;    /* host fetches of the adaptive partial sums */
  /* energy partials */
  double *part;            /* device partial sums */
  double *hpart;           /* pinned host copy */
  int npart;
  /* profiling */
  int mp_variant;          /* 0..2: direct-to-LDS 4x4x4 (1 is the default), 4: register-staged 4x4x4, 16: 16x16x4 */
  int mp_nb;               /* column groups per wave of the 4x4x4 kernel (4 or 8) */
  int prof;                /* record per-class event pairs */
  hipEvent_t *ev;
  int *ev_class;
  int nev, cap_ev;
};

/* profile classes; PC_FFT (the 3D FFT passes) is nested inside PC_ELASTIC */
enum { PC_MODEPROD = 0, PC_ELASTIC, PC_NONLIN, PC_PHI, PC_ENERGY, PC_HALO, PC_FFT, PC_CARRY, PC_COUNT };

#ifdef __cplusplus
extern "C" {
#endif
/* host-side helpers in etd_io.c */
void etd_host_noise(int n, int g, unsigned long long seed, double mean, double amp, double *u);
int etd_csv_rows(const char *work_dir, int nac, int nch, int iter, const double *stats);
/* outputfield (field.c:17-56): one box, (cx, cy, cz) its Cartesian rank coordinates */
int etd_field_file(const char *data_dir, int var, int ioutput, int n, const double *u, int cx, int cy, int cz);
/* CALPHAD Gibbs energies (free_energy 1); variable-mobility x(1-x), D_A, D_G of Cu, Mn, Ni */
void etd_calphad(double *gbcc, double *ghcp);
void etd_vm_constants(double X[3], double DA[3], double DG[3]);
#ifdef __cplusplus
}
#endif

/* launchers (etd_kernels.hip) */
void etd_prof_begin(etd_ctx *c, int cls);
void etd_prof_end(etd_ctx *c, int cls);

template <typename T>
inline etd_fields<T> fields(const etd_ctx *c);
template <>
inline etd_fields<double> fields<double>(const etd_ctx *c) {
  return etd_fields<double>{c->u, c->N1, c->N2, c->w1, c->w2, c->spec, c->gtmp, c->Z, c->Y, c->E};
}
template <>
inline etd_fields<float> fields<float>(const etd_ctx *c) {
  return c->f32;
}

/* T = double (plain FP64 path) or float (FP32 storage of the mixed path;
 * arithmetic in registers stays FP64 in these kernels) */
template <typename T>
void launch_ghost_fill(etd_ctx *c, T *base, int nfields, size_t fstride, T *dz = nullptr, T *dy = nullptr,
                       T *dx = nullptr);
/* phi scaling: ac_updateU_new / ch_updateU_new (calc.c:489-535) and
 * prepare_U2_new (553-567) with phi1, phi2 (vars.c:302-359) recomputed
 * per point from D */
struct phi_args {
  int n, nac;
  double dt, epn2, LE, KE, LCI, KCI, kkx, kky, kkz;
  double dt2;  // with a second output: its step
};

// phi1, phi2 and 1 - phi1 H' of one point for one kind of field (AC or CH);
// the formula and the |H| > 1e-8 switch are the reference's (vars.c:315-325)
struct phi_pt {
  double phi1, phi2, lin;
};
__device__ __forceinline__ phi_pt phi_at(double H, double dt) {
  phi_pt r;
  if (fabs(H) > 1.0e-8) {
    const double e = exp(-dt * H);
    r.phi1 = (1.0 - e) / H;
    r.phi2 = (1.0 - r.phi1 / dt) / H;
  } else {
    r.phi1 = dt;
    r.phi2 = dt / 2;
  }
  const double Hp = fabs(H) < 1.0e-8 ? 0.0 : H;
  r.lin = 1.0 - r.phi1 * Hp;
  return r;
}

static inline phi_args phi_args_of(const etd_ctx *c, double dt, double dt2) {
  phi_args a;
  a.n = c->n;
  a.nac = c->nac;
  a.dt = dt;
  a.epn2 = c->p.epn2;
  a.LE = c->p.LE;
  a.KE = c->p.KE;
  a.LCI = c->p.LCI;
  a.KCI = c->p.KCI;
  a.kkx = c->p.kkx;
  a.kky = c->p.kky;
  a.kkz = c->p.kkz;
  a.dt2 = dt2;
  return a;
}
template <typename T>
void launch_phiscale(etd_ctx *c, const T *uhat, const T *nhat, T *out, int first_var, int nvars, int order2);
/* first order over fields first_var .. first_var + nvars (pointers at
 * field first_var): out at step dt, out2 at step dt2 */
template <typename T>
void launch_phiscale_pair(etd_ctx *c, const T *uhat, const T *nhat, T *out, double dt, T *out2, double dt2,
                          int first_var, int nvars);
/* the first-order phi update of the forward stage's spectra (spec, N-hat
 * after all boxes' u-hat) into out (and out2 at step dt2; single: out =
 * spec, dt = the context's dt); a carried AC spectrum that stage_forward
 * left pending is corrected in the same pass */
template <typename T>
void phi_stage(etd_ctx *c, T *out, double dt, T *out2, double dt2);
/* the ETD2 corrector on the predictor in u (its spectrum a-hat); with
 * spec_out, also the spectrum of the result for a carrying step */
template <typename T>
void etd2_stage(etd_ctx *c, const T *ahat, T *spec_out);
template <typename T>
void launch_nonlinear(etd_ctx *c, T *out, const T *sub, int ch_centre = 0);
template <typename T>
void launch_elastic(etd_ctx *c);
template <typename T>
void fft3d(etd_ctx *c, T *data, int npair, int sgn);
template <typename T>
void launch_felas_materialize(etd_ctx *c, T *felas);
template <typename T>
int launch_diag(etd_ctx *c, double *stats);
/* doubles of partial sums one check_soln_new evaluation writes (the size of c->part) */
size_t etd_diag_partials(const etd_ctx *c);
/* the two-stage register FFT (k_fft16) handles this N */
bool fft16_supported(int N);
/* CH mean conservation of the mixed path: ch_mean_mark takes the interior
 * sums of the CH fields of the state u (V fields, stride n3) as the targets
 * if the fields were set since the last mark; ch_mean_restore shifts the CH
 * fields of u so that their interior sums are the targets (a shift below the
 * fields' rounding is carried to a later step, the targets being fixed).
 * No-ops unless the context is mixed and has CH fields. */
#define CH_BLOCKS 256
void ch_mean_mark(etd_ctx *c, const float *u);
void ch_mean_restore(etd_ctx *c, float *u);
/* ch_center(c, u, NULL, +1) followed by ch_mean_restore(c, u) in one pass:
 * ghost cells get the mean added, interior cells the shift that gives the
 * target sum (computed from the centred field). */
void ch_mean_restore_centered(etd_ctx *c, float *u);
/* The mixed path transforms the CH fields' deviation from their conserved
 * mean: a uniform field is a fixed point of the scheme, so the step evolves
 * u - c_mean with N(u) - N(c_mean) and adds c_mean back.  The split-BF16
 * products then round the small deviation instead of the large mean, whose
 * rounding would otherwise leak into the barely damped long-wave CH modes
 * every step.  sign -1 centres (u and/or N, either may be null), +1 undoes.
 * The step leaves u centred after the forward transform (the backward
 * transform overwrites it) and undoes it in ch_mean_restore_centered. */
void ch_center(etd_ctx *c, float *u, float *N, int sign);

/* the carried AC spectrum (etd_carry.hip): carry_possible says whether this
 * context may carry; carry_prepare allocates the buffers; carry_deltas gives
 * the buffers the ghost fill records its changes of the AC fields into;
 * carry_correct writes out = alpha s1 + beta s2 + S^-1 (recorded changes)
 * for the nac AC fields (s2 may be null) */
bool carry_possible(const etd_ctx *c);
/* what the next carried step starts from: alpha s1 + beta s2 (s1 null:
 * spec[0 .. nac); s2 null: nothing) */
static inline void carry_base(etd_ctx *c, const void *s1, double alpha, const void *s2, double beta) {
  c->carry_s1 = s1;
  c->carry_alpha = alpha;
  c->carry_s2 = s2;
  c->carry_beta = beta;
}
template <typename T>
int carry_prepare(etd_ctx *c);
/* ETD_OK, or the error of a failed correction GEMM since the last call
 * (message in c->err) */
int carry_status(etd_ctx *c);
/* carry_possible, with the buffers prepared; where they cannot be
 * allocated the context stops carrying (no_carry) and this is false */
bool carry_ready(etd_ctx *c);
template <typename T>
void carry_deltas(etd_ctx *c, T **dz, T **dy, T **dx);
template <typename T>
void carry_correct(etd_ctx *c, const T *s1, double alpha, const T *s2, double beta, T *out);
/* the same spectrum, phi-updated with N-hat (nac fields from nhat) at step
 * dt into out, and at step dt2 into out2 unless out2 is null */
template <typename T>
void carry_correct_phi(etd_ctx *c, const T *s1, double alpha, const T *s2, double beta, const T *nhat, T *out,
                       double dt, T *out2, double dt2);

/* step pieces (etd_api.hip) */
template <typename T>
void modeprod3(etd_ctx *c, int forward, const T *in, T *out, int nf, int accumulate);
template <typename T>
void stage_forward(etd_ctx *c);
template <typename T>
int step_once_t(etd_ctx *c);

void launch_modeprod(etd_ctx *c, const double *A4, const double *in, size_t in_fs, double *out, size_t out_fs,
                     int nfields, int rotate, int accumulate);
void launch_copy16(etd_ctx *c, const void *in, void *out, size_t n16);
size_t arrange_A_chunk(int MT);
void arrange_A(int n, const double *Acm, double *A4host, int MT, int nkc);

/* mixed precision mode products: FP32 in / out, A split into `terms` BF16
 * terms, FP32 accumulation (etd_kernels_mixed.hip) */
void launch_modeprod_split(etd_ctx *c, const uint16_t *A3, const float *in, size_t in_fs, float *out, size_t out_fs,
                           int nfields, int rotate, int accumulate, float a_inv = 1.0f);
size_t split_A_bytes(int n, int terms);
void arrange_split_A(int n, const double *Acm, uint16_t *A3host, int terms);
/* FP16 split-2 arrangement of A / s (s a power of two bringing max |A| to
 * [2^13, 2^14)); returns 1 / s. */
float arrange_split_A_f16(int n, const double *Acm, uint16_t *A2host);
/* The FP16 split-2 three-slot kernel can contract the slowest axis at this n
 * (launch_modeprod_split with rotate = 2: rotated, out[m + n col]; 3: not
 * rotated, out[col + n^2 m]; col = i + n j, contracted index k of
 * in[col + n^2 k]) */
bool f16_zs_supported(int n);

#endif
