/*
 * etd.h -- C ABI of libetd, the MI355X (gfx950) implementation of the ScLETD
 * exponential-time-differencing step of moxiaoan/Energy-stability.
 *
 * The reference has no plugin or FFI: its time-step loop
 * (source_beta_to_omega/src/ScLETD.c:84-191) calls a fixed sequence of void
 * functions on process-global state.  Each entry point below replaces one of
 * those functions (cited per declaration) or the init chain around them, with
 * explicit state (etd_ctx), plain pointers and sizes, int return codes
 * (0 = OK, negative = error) and etd_last_error() instead of printf+exit.
 *
 * Data layout at the boundary: one field = n^3 doubles, index (k*n + j)*n + i
 * with i (x) fastest, ghost cells included, exactly the reference's
 * ac[v].fieldE / ch[v].fieldCI arrays (source_beta_to_omega/src/vars.c:175).
 * Variables are numbered AC fields first (0..nac-1), then CH (nac..nac+nch-1).
 * Matrices use the reference's in-memory layout: A(r,c) at A[r + n*c].
 *
 * Threading: an etd_ctx owns its device buffers and one HIP stream; it is not
 * thread-safe.  Several contexts may share a device.
 */
#ifndef ETD_H
#define ETD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ETD_MAX_VAR 64

enum {
  ETD_OK = 0,
  ETD_ERR_ARG = -1,      /* invalid argument */
  ETD_ERR_IO = -2,       /* missing / malformed input file */
  ETD_ERR_UNSUPPORTED = -3,
  ETD_ERR_DEVICE = -4,   /* HIP runtime failure */
  ETD_ERR_NOMEM = -5,
  ETD_ERR_NONFINITE = -6 /* a NaN / infinite state or candidate that no step size cures */
};

/* precision of the step: FP64 throughout; or FP32 fields with the mode
 * products as MFMA products of split operands accumulated in FP32 and the
 * energy in FP64 -- BF16 split-3 (six BF16 products, 2^-24 per term) or FP16
 * split-2 (three FP16 products after power-of-two scaling of V / V^-1 and of
 * every 32-long column chunk of the field, 2^-22 per term) */
enum { ETD_PREC_FP64 = 0, ETD_PREC_MIXED = 1, ETD_PREC_MIXED_F16 = 2 };

/* Everything the reference reads from input<s>.ini (ScLETD.c:18-34),
 * elastic_input.txt (elastic.c:375-418), anisotropic_input.txt
 * (anisotropic.c:645-673) and the constants hard-coded in formula.c:9-29. */
typedef struct etd_params {
  int nx, ny, nz;          /* extended edge (must be equal) */
  int px, py, pz;          /* process grid: 1,1,1 one subdomain; otherwise the
                              reference's decomposition into px x py x pz boxes
                              of nx^3, all on this GPU, exchanged as its ranks
                              exchange them (periodic = 1, FP64, Landau force,
                              constant mobility; px = py = pz when elastic).
                              Box b sits at (b % px, (b / px) % py, b / (px py)). */
  int nghost;
  char work_dir[1024];
  char data_dir[1024];
  double dt, t_total;
  double epn2;
  int periodic;            /* 1; or 0: physical boundaries on every side of the one box (boundary
                              id 0 basis, h = L / nx, no ghost fill or outside values) */
  int nac, nch;
  int restart, restart_iter;
  char restart_dir[1024];
  double xmin, ymin, zmin, xmax, ymax, zmax;
  int etd2;
  int elastic;             /* 0 none; 1 on the AC fields; 2 on the CH fields (t += LCI felas, and
                              -LCI felas in the outside chemical potential); 3 (BOTH_FUNCTION)
                              matches no branch of elastic_calculate: no force, as in the reference */
  int anisotropic;         /* 0 none; 1 F2 = div(Lambda grad u), omega (anisotropic.c:164-627);
                              2 the compact Lambda_ij d_i d_j stencil of the beta->alpha driver
                              (anisotropic_hip.c:528-647) */
  double kkx, kky, kkz;
  int variable_mobility;   /* 0; 1 the variable-mobility CH of mobility.c:838-922 (nch == 3:
                              Cu, Mn, Ni mobilities from AC field 0) */
  /* derived / hard-coded in the reference */
  double hx, hy, hz;       /* (max-min)/gnx, vars.c:153-155 */
  double LE, KE;           /* formula.c:9-18: 1, 42 */
  double a2, a3, a4;       /* formula.c:126-128 */
  double LCI, KCI;         /* CH mobility / stabiliser (unset in the reference) */
  double elastic_scale;
  double C[36];            /* Voigt 6x6, row-major */
  double eps0[ETD_MAX_VAR * 9];
  double lam[ETD_MAX_VAR * 9];  /* Lambda = diag(kk) - lambda_in */
  int precision;           /* ETD_PREC_FP64 or ETD_PREC_MIXED */
  /* the beta->alpha ini's output and restart schedule
   * (source_beta_to_alpha/src/ScLETD.cpp:29-31): checkpoint 1 starts fresh,
   * 2 reads checkpoint buffer chk; a checkpoint every nchk steps (alternating
   * buffers), uint16 field dumps every nout steps; approx (the reference's
   * truncated-elasticity mode count) is read and not used */
  int checkpoint, nchk, chk, nout, approx;
  /* AC driving force: 0 the Landau polynomial (omega, formula.c:30-38); 1 the
   * CALPHAD Ti-Al-V driving force of the beta->alpha driver
   * (anisotropic_hip.c:2002-2188, parameters vars.c:144-299): bcc / hcp Gibbs
   * energies at its fixed compositions, f = LE KE u - LE (dG/gnormal + 4 ALPHA
   * u sum_{m != p} u_m^2), with that driver's order of the elastic, gradient
   * and grain-mask terms (ac_calc_F2_dev, 2190-2211) */
  int free_energy;
} etd_params;

typedef struct etd_ctx etd_ctx;

/* Fills *p with the reference's defaults (ω constants of formula.c:9-29). */
void etd_default_params(etd_params *p);

/* Replaces the positional fscanf block of main (ScLETD.c:16-35), including
 * CRLF tolerance.  Derives hx,hy,hz as init_para does (vars.c:58-157). */
int etd_parse_ini(const char *path, etd_params *p);

/* The beta->alpha ini order (source_beta_to_alpha/src/ScLETD.cpp:15-31;
 * epn2 = 0.048 and ETD1 as that driver sets them). */
int etd_parse_ini_alpha(const char *path, etd_params *p);

/* Replaces elastic_input (elastic.c:349-438): ElasticScale, 36 Voigt
 * entries, then nac 3x3 eigenstrains. */
int etd_parse_elastic(const char *path, etd_params *p);

/* Replaces anisotropic_input (anisotropic.c:630-674): nac 3x3 lambda, then
 * Lambda = diag(kk) - lambda. */
int etd_parse_anisotropic(const char *path, etd_params *p);

/* Optional extension file landau_input.txt: "LE KE a2 a3 a4 [LCI KCI]"
 * (the reference hard-codes these in formula.c:9-29). */
int etd_parse_landau(const char *path, etd_params *p);

/* Replaces read_matrices' file reads (vars.c:370-476): d<id><n>.dat (n
 * values), v<id><n>.dat and vi<id><n>.dat (n*n values, row-major file,
 * stored column-major).  Returns ETD_ERR_IO when a file is missing. */
int etd_read_matrices(const char *dir, int n, int id, double *V, double *Vi, double *D);

/* Closed-form eigenbasis of the truncated second-difference matrix for
 * boundary id 0..3 (what the .dat files hold). */
int etd_eigbasis(int n, int id, double *V, double *Vi, double *D);

/* Creates a context: allocates device state, uploads the eigenbasis (V, Vi,
 * D may be NULL: closed form for id 3), builds the phi tables
 * (read_matrices -> ac_init_phi / ch_init_phi, vars.c:302-359,583-590).
 * Replaces alloc_vars + init_vars + init_para + read_matrices +
 * elastic_init (vars.c:166-300, elastic.c:607-613). */
int etd_create(const etd_params *p, const double *V, const double *Vi, const double *D, etd_ctx **out);

/* Parses <dir>/input<suffix>.ini, elastic_input.txt, anisotropic_input.txt
 * (and landau_input.txt when present), reads d/v/vi<3><n>.dat from dir when
 * present, then etd_create. */
int etd_open(const char *dir, const char *suffix, etd_ctx **out);

void etd_close(etd_ctx *c);
const char *etd_last_error(const etd_ctx *c);
const char *etd_error_string(int code);
int etd_get_params(const etd_ctx *c, etd_params *p);

/* Host <-> device copies of one field (n^3 doubles, ghosts included; a
 * decomposed context takes and gives its boxes one after another, P n^3). */
int etd_set_field(etd_ctx *c, int var, const double *host);
int etd_get_field(etd_ctx *c, int var, double *host);

/* The grain mask of the beta->alpha driver (read_ori, field.c:111-129;
 * ac_calc_F2_dev, anisotropic_hip.c:2206-2210): n^3 ints, x fastest; every AC
 * nonlinear term gets -0.5 u where mask == 0.  NULL removes it. */
int etd_set_mask(etd_ctx *c, const int *mask);

/* The Gibbs energies that free_energy = 1 drives with: bcc and hcp Ti-Al-V at
 * the beta->alpha driver's fixed compositions and T = 1093 K (Free_energy,
 * HCP_energy, BCC_energy: source_beta_to_alpha/src/vars.c:144-299; used at
 * anisotropic_hip.c:2029-2048).  Host only, no context needed. */
void etd_calphad(double *gbcc, double *ghcp);

/* Seeded initial state: interior U[mean-amp, mean+amp) from splitmix64
 * ((x >> 11) * 2^-53), ghosts as periodic images (box b of a decomposed
 * context: its own images, seed + 1000003 b).  Replaces init_field
 * (field.c:742-777) for the synthetic workloads. */
int etd_init_noise(etd_ctx *c, int var, unsigned long long seed, double mean, double amp);

/* One or more full time steps: transfer, elastic_calculate, calc_mu,
 * ac_calc_FU / ch_calc_FU, mode products, phi update, and the ETD2
 * corrector (ScLETD.c:87-189).  Does not run the diagnostics. */
int etd_step(etd_ctx *c, int nsteps);

/* check_soln_new (field.c:407-678) without file output:
 * stats[4*v + {0,1,2,3}] = max, min, vol, eng for each variable v, then
 * stats[4*(nac+nch)] = couple energy (0 unless nac > 0 and nch > 0); over
 * the boxes of a decomposed context reduced as its MPI_Reduce calls do
 * (max, min, sums; field.c:513-516). */
int etd_diag(etd_ctx *c, double *stats);

/* The allocation-free variant of etd_diag; it has no counterpart in the
 * reference.  It is for callers that cannot allocate: code under HIP graph
 * capture, tight serving loops.  The same statistics, bit for bit, go to
 * stats_dev, a caller's device buffer of 4 (nac + nch) + 1 doubles, from
 * kernels queued on the context's stream: no host synchronisation, no
 * allocation, no memset.  The partial sums live in a static 512 KiB device
 * array (one per GPU, held until the process ends), which must hold
 * (3 (nac + nch) + 1 + nch + [nac > 0 and nch > 0]) (N - 1) 8 doubles
 * (65536; e.g. 64,176 for 192^3 x (12 + 1) fields); larger contexts get
 * ETD_ERR_UNSUPPORTED.  Calls on one GPU share that array: calls that
 * overlap in time on one GPU (other streams, other threads) race.  The
 * result is ready when the context's stream has reached it. */
int etd_diag_noalloc(etd_ctx *c, double *stats_dev);

/* Appends the CSV rows of check_soln_new (phi_ac_<v>.csv, phi_ch_<v>.csv,
 * phi_couple.csv, format "%d,%+1.15lf,...") under work_dir; iter == 0 also
 * writes the headers. */
int etd_write_csv(etd_ctx *c, int iter, const double *stats);

/* Writes <data_dir>eta<v>_<ioutput:06d>_<cx:02d><cy:02d><cz:02d>.dat per box,
 * text %+1.15lf, full n^3 (outputfield, field.c:17-56). */
int etd_write_field(etd_ctx *c, int var, int ioutput);

/* ---- per-kernel entry points (unit parity; host buffers in/out) ---- */

/* transfer() (trans.c:1939-1996): periodic ghost fill of every field (nothing
 * on a box with physical sides). */
int etd_k_transfer(etd_ctx *c);
/* elastic_calculate (elastic.c:689-762); the force of every field it acts on
 * copied to host_felas when non-NULL: nac * n^3 doubles for ELASTIC = 1 (AC
 * fields), nch * n^3 for ELASTIC = 2 (CH fields); a decomposed context gives
 * P n^3 per field (field p of box b at (p P + b) n^3) and transforms its
 * global (px N)^3 interior, as fftMPI's distributed transform does. */
int etd_k_elastic(etd_ctx *c, double *host_felas);
/* The 3D FFT of the elastic path (replaces fft_forward / fft_backward,
 * fft3d.c:98-138) on N^3 interleaved complex values, N = n - 2 nghost (px N
 * for a decomposed context), in place; dir = +1 forward exp(-i..), -1
 * backward exp(+i..), both unscaled.  Needs a context with elasticity. */
int etd_k_fft3d(etd_ctx *c, double *data, int dir);
/* ac_calc_FU / ch_calc_FU (calc.c:8-68,475-487) for variable var on the
 * current device state (after etd_k_transfer / etd_k_elastic); P n^3 for a
 * decomposed context. */
int etd_k_nonlinear(etd_ctx *c, int var, double *host_out);
/* PUX/PUY/PUZ (calc.c:631-684): out = A x_1 in, optionally rotated
 * (flat (a,b,c) -> (b,c,a)); A column-major n x n; n^3 in/out.
 * rotate 0 / 1: contraction over the fastest axis a, as the reference's
 * products (out[m + n (b + n c)] / out[(b + n c) + n^2 m]); 2 / 3 (FP16
 * split-2 contexts, n <= 256): over the slowest axis c, rotated
 * (out[m + n (a + n b)], (a,b,c) -> (c',a,b)) / not (out[(a + n b) + n^2 m]),
 * the kernel the step uses there. */
int etd_k_modeprod(etd_ctx *c, int n, const double *A, const double *in, double *out, int rotate);
/* ac_updateU_new / ch_updateU_new (calc.c:489-535) on spectral data in
 * the rotated (x'y'z', x' fastest) layout. */
int etd_k_phiscale(etd_ctx *c, int var, double *uhat, const double *nhat);

/* ---- timing: runs nsteps on the context stream between HIP events ---- */
/* ms[0] = total, ms[1] = mode-product kernels, ms[2] = elastic (FFT +
 * contraction), ms[3] = nonlinear, ms[4] = phi scaling, ms[5] = energy,
 * ms[6] = halo, ms[7] = the 3D FFT passes alone (part of ms[2]), ms[8] = the
 * carried spectrum's correction, ms[9] = the number of field passes the
 * mode-product launches made (a count: fields x 3 axes); launches[k] counts
 * the launch groups of class k (8 ints). */
int etd_time_steps(etd_ctx *c, int nsteps, int with_diag, double *ms, int *launches);

/* ---- adaptive ETD and the learned dt controller ----
 * No counterpart in the reference, which steps with a fixed dt
 * (ScLETD.c:84-191).  Each attempt from u_n with step dt computes one step of
 * dt and two of dt/2 (the first half step shares N(u_n) and its forward
 * transforms with the full step); err = max |u_half - u_full| over all fields.
 * An attempt is accepted when err <= tol and, with energy_test, the FP64
 * energy does not increase.  The energy is the scheme's discrete free energy
 * (energy_kind 1: periodic forward-difference gradient energy with kk, the
 * central-difference anisotropic term, Landau bulk, CH double well; the
 * semi-discrete dynamics is its gradient flow) or check_soln_new's AC + CH
 * energy (energy_kind 0, the reference's 8-corner diagnostic, which is not
 * a Lyapunov function of the scheme); with energy_elastic plus the elastic
 * energy ES/2 h^3 sum_p eta_p felas_p (one more elastic evaluation per
 * attempt, reused by the next step when accepted).
 * Slack: 1e-12 (FP64 path) or 1e-7 (mixed path, the FP32 rounding level)
 * times the energy's scale: the larger sum of the absolute values of its
 * terms over the two states (scheme energy), or the largest |E| of the run
 * (check_soln_new energy).  Rejected
 * attempts restore u_n.  A candidate whose error estimate or energy is NaN or
 * infinite is rejected with the factor shrink; at dt_min the run stops with
 * ETD_ERR_NONFINITE and u_n restored (a step at dt_min is accepted only when
 * finite), and a non-finite initial energy is ETD_ERR_NONFINITE before any
 * step.  The plain controller proposes
 * dt * clamp(safety (tol/err)^(1/2), shrink, grow); the learned controller
 * proposes learned_safety * dt_plain * 10^y, where y = MLP(features) over
 * ETD_NFEAT features of the run (below) predicts log10 of the error model's
 * largest step over the plain proposal dt_plain, bounded to
 * [min(dt_plain, last dt), 10 dt_plain] (so the learned step is not held to
 * the plain growth limit but cannot shrink where the plain rule would not),
 * and falls back to the plain proposal after a rejection or for a
 * non-finite output. */
#define ETD_NFEAT 16
enum { ETD_CTRL_PLAIN = 0, ETD_CTRL_LEARNED = 1 };

typedef struct etd_adapt_opts {
  double t_end;        /* integrate from the context's time to t_end (last step clipped) */
  double dt0;          /* first proposal */
  double dt_min, dt_max;
  double tol;          /* bound on the step-doubling estimate */
  double safety, grow, shrink;
  double learned_safety; /* the learned proposal's margin below the error model's largest step (0.97) */
  int controller;      /* ETD_CTRL_PLAIN or ETD_CTRL_LEARNED (needs etd_set_controller) */
  int extrapolate;     /* 1: accept 2 u_half - u_full (local extrapolation) */
  int energy_test;     /* 1: reject an energy increase */
  int energy_elastic;  /* 1: the tested energy adds the elastic energy ES/2 h^3 sum eta_p felas_p */
  int energy_kind;     /* 1: the scheme's discrete free energy (default); 0: check_soln_new's */
  int max_attempts;
  int learned_online;  /* 1 (default): the learned proposal's prediction is lowered by the largest recent
                          over-prediction (decaying by 0.8 per learned step), measured after each attempt
                          against the error model's largest step at the observed error; 0: the network alone */
} etd_adapt_opts;

typedef struct etd_adapt_stats {
  double t, dt_last, energy;
  int accepted, rejected_err, rejected_energy;
  int learned_proposals, learned_fallbacks;
  int step_evals;      /* ETD step evaluations: 3 per attempt (one shares N with another) */
} etd_adapt_stats;

/* Features of the proposal for step k (all from the run so far):
 *  0..3  log10 of the last four accepted dt (newest first)
 *  4     log10(err / tol) of the last accepted step
 *  5     log10 |dE/dt| over the last accepted step
 *  6     log10 |E|
 *  7     log10 max |d eta| over the last accepted step (AC fields)
 *  8     interface-volume fraction: interior cells with max_p |eta_p| > 0.1
 *  9     mean |grad eta| over interior cells and AC fields
 *  10    log10(t + 1e-6)      11  log10(1 + accepted steps)
 *  12    rejections before the last acceptance
 *  13    log10 max |eta|      14  log10(tol)
 *  15    log10 of the plain controller's proposal */
#define ETD_LOG_COLS (4 + ETD_NFEAT)

void etd_adapt_defaults(etd_adapt_opts *o);
/* Runs to o->t_end.  When log != NULL, each accepted step appends one row of
 * ETD_LOG_COLS doubles (t after the step, dt, err, E, then the features seen
 * when dt was proposed) while fewer than log_cap rows are stored. */
int etd_run_adaptive(etd_ctx *c, const etd_adapt_opts *o, etd_adapt_stats *st, double *log, int log_cap,
                     int *log_rows);
/* Simulated time of the context (advanced by etd_step and etd_run_adaptive). */
double etd_get_time(const etd_ctx *c);
void etd_set_time(etd_ctx *c, double t);

/* The controller network: MLP ETD_NFEAT -> 64 -> 64 -> 1, GELU, FP32, with
 * hand-written forward / backward / Adam kernels.  Parameters (ETD_MLP_PARAMS
 * floats): W1 [64][16] row-major, b1 [64], W2 [64][64], b2 [64], w3 [64],
 * b3.  Inputs and targets are normalised by the stored feature mean / std
 * and target mean / std; the controller's network predicts
 * log10(dt_best / dt_plain) for the next step (see etd_run_adaptive). */
#define ETD_MLP_PARAMS (64 * ETD_NFEAT + 64 + 64 * 64 + 64 + 64 + 1)
typedef struct etd_mlp etd_mlp;
int etd_mlp_create(int device, etd_mlp **out);
void etd_mlp_destroy(etd_mlp *m);
/* Seeded init (uniform +-sqrt(6/(fan_in + fan_out)), zero biases) and unit normalisation. */
int etd_mlp_init(etd_mlp *m, unsigned long long seed);
int etd_mlp_set(etd_mlp *m, const float *params, const float *feat_mean, const float *feat_std, float y_mean,
                float y_std);
int etd_mlp_get(const etd_mlp *m, float *params, float *feat_mean, float *feat_std, float *y_mean, float *y_std);
/* Network on normalised inputs x[batch][ETD_NFEAT] -> normalised outputs y[batch]. */
int etd_mlp_forward(etd_mlp *m, const float *x, int batch, float *y);
/* Mean squared error over the batch and its gradient (ETD_MLP_PARAMS floats),
 * normalised inputs and targets. */
int etd_mlp_loss_grad(etd_mlp *m, const float *x, const float *t, int batch, float *loss, float *grad);
/* Adam (beta 0.9 / 0.999, eps 1e-8) over epochs of seeded shuffled batches of
 * the normalised data set; epoch_loss[e] = mean loss of epoch e (may be NULL). */
int etd_mlp_train(etd_mlp *m, const float *x, const float *t, int count, int epochs, int batch, float lr,
                  unsigned long long seed, float *epoch_loss);
/* Raw features -> the network's prediction in target units (de-normalised);
 * what the learned controller calls once per accepted step. */
int etd_mlp_predict(etd_mlp *m, const double *features, double *y);
/* Attach (or detach with NULL) the controller network of a context. */
int etd_set_controller(etd_ctx *c, etd_mlp *m);

/* ---- batched ensembles and the tile autotuner ----
 * No counterpart in the reference (one simulation per MPI job).  An ensemble
 * steps many independent contexts of uneven sizes on one GPU: one hipGraph,
 * built once, holds steps_per_launch fixed-dt steps of every member, each
 * member on its own stream forked from and joined to the ensemble's stream,
 * so members run concurrently; etd_ensemble_run replays it.  A replay is bit
 * for bit the same as stepping the members one by one.  With elasticity the
 * graph repeats the captured felas ring positions, so steps_per_launch *
 * (elastic evaluations per step) * nac must be a multiple of the ring length
 * nac + 2 for every member (etd_ensemble_create says which multiple).
 * Members must live on one device and must not be stepped elsewhere while a
 * replay runs. */
typedef struct etd_ensemble etd_ensemble;
int etd_ensemble_create(etd_ctx *const *members, int count, int steps_per_launch, etd_ensemble **out);
/* launches replays; each advances every member by steps_per_launch steps. */
int etd_ensemble_run(etd_ensemble *e, int launches);
void etd_ensemble_destroy(etd_ensemble *e);
const char *etd_ensemble_error(const etd_ensemble *e);

/* Kernel choices of a context (the autotuner's search space). */
typedef struct etd_tuning {
  int mp_variant;      /* FP64 mode products: 0, 1, 2 (direct-to-LDS), 4 (register-staged), 16 (16x16x4) */
  int split_waves;     /* mixed mode products: 8 or 4 waves per workgroup (0: by size) */
  int fft_variant;     /* 1: two-stage register FFT (16 | N), 0: radix passes */
  double ms_per_step;  /* measured by etd_autotune */
} etd_tuning;
/* Times `steps` fixed-dt steps of every candidate on a scratch context with
 * the same parameters (the caller's contexts are untouched) and returns the
 * fastest choice. */
int etd_autotune(const etd_params *p, const double *V, const double *Vi, const double *D, int steps,
                 etd_tuning *best);
int etd_set_tuning(etd_ctx *c, const etd_tuning *t);
int etd_get_tuning(const etd_ctx *c, etd_tuning *t);

/* ---- the beta->alpha driver's output and restart formats ----
 * Files go under work_dir / data_dir plus the reference's per-rank
 * subdirectory "00/" (one subdomain: cart ids 0,0,0), which must exist. */

/* write_chk (source_beta_to_alpha/src/vars.c:967-1011):
 * <work_dir>00/eta<v>_chk<chk>_000000.dat = int irun, then the n^3 doubles of
 * AC field v (ghosts included, x fastest); CH fields as c<q>_chk<chk>_000000.dat
 * (the reference's disabled CH branch, vars.c:990-1010).  chk is 0 or 1. */
int etd_write_checkpoint(etd_ctx *c, int irun, int chk);
/* read_chk (vars.c:1013-1065): the AC files must exist, CH files are read
 * when present; *irun from the files. */
int etd_read_checkpoint(etd_ctx *c, int chk, int *irun);
/* Dtype2B_E + write_field2B (source_beta_to_alpha/src/anisotropic_hip.c:3920-3931,
 * 4290-4299; vars.c:1068-1094): each AC field as
 * <data_dir>00/eta<v>_<irun:06d>_000000.dat, n^3 uint16 (eta + 0.1) * 30000
 * (truncated, saturated to [0, 2^32 - 1], low 16 bits kept), converted on the GPU. */
int etd_write_field2b(etd_ctx *c, int irun);
/* The detector feed (anisotropic_hip.c:3897-3918 skip_copyin_new_2,
 * 4264-4278 binarisation, 4400-4427 float2bin): the interior of AC field var
 * averaged over 2x2x2 blocks ((N/2)^3 cells, x fastest), each cell 1 if its
 * mean exceeds 0.1; bits[t] packs cells 8t .. 8t+7 (bit e = cell 8t + e);
 * binary (NULL to skip) gets the (N/2)^3 cells as 1.0f / 0.0f.  N/2 even. */
int etd_detector_feed(etd_ctx *c, int var, unsigned char *bits, float *binary);

/* Measurement support, no counterpart in the reference: one plain copy
 * kernel that reads and writes exactly `bytes` (a multiple of 16, at most the
 * context's 2 (nac+nch) n^3 doubles of scratch) between two of the context's
 * own buffers, so HBM counters can be scaled against a known size.  Clobbers
 * the spectral scratch only; the fields are untouched. */
int etd_calib_copy(etd_ctx *c, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
