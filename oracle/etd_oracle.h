/*
 * etd_oracle.h -- FP64 CPU oracle for the ScLETD exponential-time-differencing
 * step (test infrastructure only).
 *
 * It checks results; it is not the product.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it.  The product library
 * (libetd, HIP) never links or calls it.
 *
 * It restates, for one subdomain (procs = 1,1,1) or a periodic decomposition
 * into boxes (orc_group), the algorithm of the reference's beta->omega CPU
 * path:
 *   source_beta_to_omega/src/ScLETD.c:84-191   stage order (ETD1 + optional ETD2)
 *   source_beta_to_omega/src/calc.c:8-68,70-473,489-684
 *   source_beta_to_omega/src/anisotropic.c:9-627
 *   source_beta_to_omega/src/formula.c:27-38   (polynomial form, generalised)
 *   source_beta_to_omega/src/vars.c:58-157,302-359,361-591
 *   source_beta_to_omega/src/trans.c:7-925,927-1432,1433-1937,1939-2013
 *   source_beta_to_omega/src/elastic.c:349-762
 *   source_beta_to_omega/src/fft3d.c:98-138    (fftMPI forward / scaled backward)
 *   source_beta_to_omega/src/field.c:407-678   (diagnostics / energy)
 *   source_beta_to_omega/src/mobility.c:1-944  (variable-mobility CH)
 *   source_beta_to_alpha/src/anisotropic_hip.c:528-647, 2002-2211
 *                                               (full Lambda_ij d_i d_j stencil,
 *                                                CALPHAD driving force, grain mask)
 *   source_beta_to_alpha/src/vars.c:144-299    (CALPHAD parameters)
 * Each function in etd_oracle.c cites the lines it follows, and states what
 * pins it (closed forms, invariants, golden vectors) or "parity unpinned".
 *
 * Layout: field index (k*n + j)*n + i, i (x) fastest, n = N + 2g extended edge.
 * Matrices: A(r,c) stored column-major at A[r + n*c] (as the reference keeps
 * V and V^-1 after read_matrices).
 */
#ifndef ETD_ORACLE_H
#define ETD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
  int n;            /* extended edge (nx = ny = nz) */
  int g;            /* ghost width (nghost); the reference's slabs assume 2 */
  int nac, nch;     /* Allen-Cahn and Cahn-Hilliard field counts */
  int etd2;         /* 0: ETD1, 1: ETD2 corrector */
  int elastic;      /* 0: none, 1: on AC fields */
  int aniso;        /* 1: gradient term F2 = div(Lambda grad u) (omega); 2: the
                       compact Lambda_ij d_i d_j stencil of alpha */
  double dt, epn2;
  double hx, hy, hz;
  double kk[3];     /* kkx, kky, kkz */
  double LE, KE;    /* AC mobility and stabiliser */
  double a2, a3, a4;/* Landau coefficients of f' = a2 u - a3 u^2 + a4 (sum u^2) u */
  double LCI, KCI;  /* CH mobility and stabiliser */
  double elastic_scale;
  double C[36];     /* 6x6 Voigt stiffness, row-major, order 11,22,33,23,13,12 */
  const double *eps0; /* nac x 3x3 eigenstrain, row-major */
  const double *lam;  /* nac x 3x3 gradient tensor Lambda = diag(kk) - lambda_in */
  const double *V;    /* n x n column-major eigenvectors of the 1D operator */
  const double *Vi;   /* n x n column-major inverse */
  const double *D;    /* n eigenvalues (unscaled, for h = 1) */
  /* beta->alpha and variable-mobility physics (appended; 0 / NULL = omega) */
  int free_energy;    /* 0: Landau polynomial (omega), 1: CALPHAD Ti-Al-V (alpha) */
  const int *mask;    /* NULL, or n^3 grain mask: N -= 0.5 u where mask == 0 (alpha) */
  int varmob;         /* 1: variable-mobility CH (omega mobility.c; nch == 3, nac >= 1) */
  int physical;       /* 1: physical boundaries on every side (periodic = 0, one box) */
} orc_params;

typedef struct orc_ctx orc_ctx;

orc_ctx *orc_create(const orc_params *p);
void orc_destroy(orc_ctx *c);
double *orc_field(orc_ctx *c, int v);     /* state of field v (AC first, then CH), n^3 */
double *orc_felas(orc_ctx *c, int p);     /* elastic driving force of AC field p, n^3 */
int orc_threads(void);
void orc_set_threads(int n);
/* test hook: plant a known mistake (1: sign of CH_f, 2: lam for lam^2 in the
 * CH operator) so a test can show its check fails; 0 restores the oracle */
void orc_set_mutation(int m);

/* individual stages (for unit parity) */
void orc_ghost_fill(int n, int g, double *u);
void orc_transfer(orc_ctx *c);
void orc_elastic(orc_ctx *c);
void orc_nonlinear(orc_ctx *c, int v, double *out);
void orc_modeprod(int n, const double *A, const double *in, double *out, int rotate);
void orc_phi_update(orc_ctx *c, int v, double *uhat, const double *nhat);  /* zxy layout */
void orc_step(orc_ctx *c);
/* adaptive stepping support: change dt (rebuilds the phi tables); the
 * step-doubling pair of one ETD1 attempt (ufull: V x n3, u_n + dt; the
 * context's fields: u_n + dt/2; dt left at dt/2) */
void orc_set_dt(orc_ctx *c, double dt);
void orc_step_pair(orc_ctx *c, double dt, double *ufull);
/* the CALPHAD bcc / hcp Gibbs energies at alpha's fixed compositions */
void orc_calphad(double *gbcc, double *ghcp);
/* its lattice stabilities and interaction parameters at T_a (13 values:
 * GHSERAL, GHSERTI, GHSERV, GBCCAL, GBCCTI, GHCPAL, GHCPV, BL13_0, BL12_0,
 * BL12_1, BL32_0, HL13_0, HL32_0) */
void orc_calphad_terms(double *t);
/* the variable-mobility divergence div((M) grad a) of mobility.c:435-722 on
 * one field (out-of-box neighbours 0), for unit checks */
void orc_div_m_grad(int n, double hx, double hy, double hz, const double *M, const double *a, double *out);
/* stats: per field v (nac+nch) 4 doubles max,min,vol,eng; then couple eng */
void orc_diag(orc_ctx *c, double *stats);

/* Decomposed runs (omega's procs px,py,pz with periodic = 1): boxes[b] is
 * the box at (b % px, (b / px) % py, b / (px py)), all created with the same
 * parameters (n, g, field counts, h = L / (p N)); the group borrows them
 * (destroy the group before the boxes).  Each box is then stepped by
 * orc_group_step only; orc_nonlinear on a box reads its neighbours' fields
 * for the outside values. */
typedef struct orc_group orc_group;
orc_group *orc_group_create(orc_ctx **boxes, int px, int py, int pz);
void orc_group_destroy(orc_group *G);
void orc_group_transfer(orc_group *G);
void orc_group_elastic(orc_group *G);
void orc_group_step(orc_group *G);
/* stats as orc_diag, reduced over the boxes */
void orc_group_diag(orc_group *G, double *stats);

/* closed-form eigenbasis of the truncated 1D second-difference operator */
int orc_eigbasis(int n, int id, double *V, double *Vi, double *D);

/* own complex FFT, 3D, in place on interleaved complex (re,im), x fastest.
 * dir = +1: forward exp(-i..), unscaled; dir = -1: backward exp(+i..), scaled 1/N^3 */
int orc_fft3d(int N, double *data, int dir);

/* elastic B_pq(n_hat) at integer wavevector (gx,gy,gz) of an N^3 grid */
double orc_bpq(const orc_params *p, int N, int gx, int gy, int gz, int pp, int qq);

#ifdef __cplusplus
}
#endif
#endif
