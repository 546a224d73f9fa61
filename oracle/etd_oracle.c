/*
 * etd_oracle.c -- FP64 CPU oracle for the ScLETD step (test infrastructure).
 *
 * Not the product: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, to check results.  Function headers cite the
 * reference lines restated (paths relative to the reference root) and what
 * pins each one:
 *
 *   orc_eigbasis     closed forms; pinned by the reference's n=260 files
 *                    (tests/golden/eig260.npz) and V*diag(D)*V^-1 = T.
 *   orc_ghost_fill   pinned by the periodic-image invariant and the fixed-
 *                    point KAT; the edge/corner order is restated from
 *                    trans.c:154-251,1939-1996 (not independently pinned).
 *   outside values   pinned by the uniform-state fixed point (T + outside
 *                    term = periodic Laplacian) and translation equivariance.
 *   orc_nonlinear    f' pinned pointwise (closed form); F2 pinned against an
 *                    analytic Fourier mode away from the box edges.
 *   orc_elastic      pinned by the k=0 KAT, B symmetry / evenness and a
 *                    brute-force DFT convolution at N <= 6.
 *   orc_fft3d        pinned by a brute-force DFT and round trip / Parseval.
 *   orc_modeprod     pinned by a dense Kronecker product at n <= 8.
 *   orc_step (ETD1)  pinned by dense expm(dt*L) at n = 6 with N == 0.
 *   orc_step (ETD2)  pinned by order-2 convergence.
 *   orc_diag         pinned by hand-computed 2^3 / 3^3 grids.
 *   orc_group_*      decomposed runs (procs > 1): a group of one box is the
 *                    single-box oracle bit for bit; check_soln_new of the
 *                    reference's own input64.ini start (2x2x2 boxes) equals
 *                    the first line of its shipped energy trace; face ghosts
 *                    and outside values are global periodic images; the
 *                    force is that of one box holding the global interior
 *                    (tests/test_oracle_decomposed.py).
 *   CH (ch_nonlinear, h_ch)
 *                    the reference's CH_f is broken (formula.c:40-56); the
 *                    double well W = (c^2-1)^2/4 of field.c:575 / trans.c:993
 *                    is used (CH_f = W').  Pinned by: t(c) up to a constant,
 *                    pointwise, read off the Laplacian of a one-cell bump;
 *                    a Fourier-mode closed form of the nonlinear term away
 *                    from the box edges; one ETD1 step against dense
 *                    expm of the truncated CH operator plus the periodic CH
 *                    right-hand side.  Planted mistakes (orc_set_mutation)
 *                    fail them.
 */
#include "etd_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(n, k, j, i) ((((size_t)(k)) * (n) + (j)) * (n) + (i))

struct orc_ctx {
  orc_params p;
  int V;                 /* nac + nch */
  size_t n3;
  double *eps0, *lam, *Vm, *Vim, *D;
  double *DDX, *DDY, *DDZ;
  double *field;         /* V * n3 */
  double *N1, *N2;       /* V * n3: N(u_n), N(a) */
  double *felas;         /* nac * n3 */
  double *fbuf;          /* persistent elastic transfer buffer (elastic.c:228) */
  double *phi1, *phi2;   /* (nac? 1:0 + nch? 1:0) tables, zxy layout */
  double *t0, *t1, *t2, *t3;  /* n3 scratch */
  int *mask;             /* copy of p.mask (or NULL) */
  int nfel;              /* felas arrays allocated */
  /* decomposed run (orc_group): the boxes this one exchanges with, and its
   * coordinates in the process grid (NULL / 0 for a single box) */
  const struct orc_group *grp;
  int bx, by, bz;
};

/* A decomposed run on one periodic domain: px x py x pz boxes of the same
 * extended edge n, box (bx, by, bz) at index (bz py + by) px + bx owning the
 * global interior cells [bx N, (bx + 1) N) x ... (ω's Cartesian ranks,
 * mpi.c / ScLETD.c:20-47, one box per rank). */
struct orc_group {
  int px, py, pz;
  orc_ctx **box;
};

/* the fields the elastic force acts on (elastic.c:689-762): the AC fields for
 * ELASTIC = 1 (AC_FUNCTION), the CH fields for 2 (CH_FUNCTION); 3
 * (BOTH_FUNCTION) matches neither branch, so no force is computed */
static int el_count(const orc_params *p) { return p->elastic == 1 ? p->nac : p->elastic == 2 ? p->nch : 0; }
static int el_first(const orc_params *p) { return p->elastic == 2 ? p->nac : 0; }

int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* thread count of the following parallel regions (CPU baseline timing at 1
 * core and at all cores) */
void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* Eigenbasis: vars.c:361-591 reads d/v/vi<id><n>.dat for the truncated
 * second-difference matrix T of each boundary id (id = (k-1)+(l-1)*2 with
 * k/l = 1 for a physical side, 2 for a neighbour side, vars.c:488-506).
 * Angles are reduced as exact integers before sin/cos.
 *   id 3 : Dirichlet truncation both sides, T = tridiag(1,-2,1):
 *          lambda_m = -4 sin^2(m pi / (2(n+1))), v_i = sin(i m pi/(n+1)) (DST-I)
 *   id 0 : reflecting both sides, rows [-2,2] ... [2,-2]:
 *          lambda_m = -4 sin^2(m pi / (2(n-1))), v_i = cos(i m pi/(n-1)) (DCT-I)
 *   id 1 : neighbour low side, reflecting high side (T[n-1][n-2] = 2):
 *          lambda_m = -4 sin^2((2m+1) pi / (4n)), v_i = sin((i+1)(2m+1) pi/(2n))
 *   id 2 : mirror image of id 1: v_i = cos(i (2m+1) pi / (2n))
 * V columns are eigenvectors (A(r,c) at V[r + n*c]); V^-1 from Gauss-Jordan
 * elimination with partial pivoting (orthonormal for id 3: V^-1 = V^T).
 * Eigenvalue order: decreasing magnitude, as in the shipped d<id>260.dat. */
static int invert(int n, const double *A, double *Ai) {
  double *M = malloc(sizeof(double) * (size_t)n * 2 * n);
  if (!M) return -1;
  /* row-major augmented [A | I], A(r,c) = A[r + n c] */
  for (int r = 0; r < n; r++)
    for (int c = 0; c < 2 * n; c++)
      M[(size_t)r * 2 * n + c] = c < n ? A[r + (size_t)n * c] : (c - n == r ? 1.0 : 0.0);
  for (int col = 0; col < n; col++) {
    int piv = col;
    for (int r = col + 1; r < n; r++)
      if (fabs(M[(size_t)r * 2 * n + col]) > fabs(M[(size_t)piv * 2 * n + col])) piv = r;
    if (M[(size_t)piv * 2 * n + col] == 0.0) { free(M); return -1; }
    if (piv != col)
      for (int c = 0; c < 2 * n; c++) {
        double t = M[(size_t)col * 2 * n + c];
        M[(size_t)col * 2 * n + c] = M[(size_t)piv * 2 * n + c];
        M[(size_t)piv * 2 * n + c] = t;
      }
    double d = M[(size_t)col * 2 * n + col];
    for (int c = 0; c < 2 * n; c++) M[(size_t)col * 2 * n + c] /= d;
    for (int r = 0; r < n; r++) {
      if (r == col) continue;
      double f = M[(size_t)r * 2 * n + col];
      if (f == 0.0) continue;
      for (int c = 0; c < 2 * n; c++) M[(size_t)r * 2 * n + c] -= f * M[(size_t)col * 2 * n + c];
    }
  }
  for (int r = 0; r < n; r++)
    for (int c = 0; c < n; c++) Ai[r + (size_t)n * c] = M[(size_t)r * 2 * n + n + c];
  free(M);
  return 0;
}

int orc_eigbasis(int n, int id, double *V, double *Vi, double *D) {
  if (n < 3 || id < 0 || id > 3) return -1;
  const double pi = 3.14159265358979323846;
  for (int m = 0; m < n; m++) {
    /* column c = m holds the mode of the m-th largest |lambda| */
    double th, lam;
    int mm;
    switch (id) {
      case 3:
        mm = n - m;  /* modes 1..n, decreasing magnitude */
        lam = -4.0 * pow(sin(mm * pi / (2.0 * (n + 1))), 2);
        for (int i = 0; i < n; i++) V[i + (size_t)n * m] = sqrt(2.0 / (n + 1)) * sin((double)(((long)(i + 1) * mm) % (2L * (n + 1))) * pi / (n + 1));
        break;
      case 0:
        mm = n - 1 - m;  /* modes 0..n-1 */
        lam = -4.0 * pow(sin(mm * pi / (2.0 * (n - 1))), 2);
        for (int i = 0; i < n; i++) V[i + (size_t)n * m] = cos((double)(((long)i * mm) % (2L * (n - 1))) * pi / (n - 1));
        break;
      case 1:
        mm = n - 1 - m;
        th = (2.0 * mm + 1.0) * pi / (2.0 * n);
        lam = -4.0 * pow(sin(th / 2.0), 2);
        for (int i = 0; i < n; i++) V[i + (size_t)n * m] = sin((double)(((long)(i + 1) * (2 * mm + 1)) % (4L * n)) * pi / (2.0 * n));
        break;
      default: /* 2 */
        mm = n - 1 - m;
        th = (2.0 * mm + 1.0) * pi / (2.0 * n);
        lam = -4.0 * pow(sin(th / 2.0), 2);
        for (int i = 0; i < n; i++) V[i + (size_t)n * m] = cos((double)(((long)i * (2 * mm + 1)) % (4L * n)) * pi / (2.0 * n));
        break;
    }
    D[m] = lam;
  }
  if (id == 3) {
    for (int r = 0; r < n; r++)
      for (int c = 0; c < n; c++) Vi[r + (size_t)n * c] = V[c + (size_t)n * r];
    return 0;
  }
  return invert(n, V, Vi);
}

/* ------------------------------------------------------------------ */
/* Ghost fill of one field: the periodic self-exchange of transfer()
 * (trans.c:1939-1996 with the packs 7-152 and unpack 154-251, mpi.c:71-82).
 * Every ghost value is read from the pre-transfer state; z faces are written
 * last (unpack front/back), then y faces, then x faces, so the cell (k,j,i)
 * receives
 *   k ghost          : u[wrap(k)][j][i]
 *   else j ghost     : u[k][wrap(j)][i]
 *   else i ghost     : u[k][j][wrap(i)]
 * with wrap(e) = e+N (e<g), e-N (e>=n-g).  In place this is z faces first
 * (they read y/x ghost cells of interior planes before those change), then
 * y faces, then x faces. */
static inline int wrap_e(int e, int n, int g) {
  int N = n - 2 * g;
  return e < g ? e + N : (e >= n - g ? e - N : e);
}

void orc_ghost_fill(int n, int g, double *u) {
  for (int k = 0; k < n; k++) {
    if (!(k < g || k >= n - g)) continue;
    int ks = wrap_e(k, n, g);
    memcpy(u + IDX(n, k, 0, 0), u + IDX(n, ks, 0, 0), sizeof(double) * (size_t)n * n);
  }
  for (int k = g; k < n - g; k++)
    for (int j = 0; j < n; j++) {
      if (!(j < g || j >= n - g)) continue;
      int js = wrap_e(j, n, g);
      memcpy(u + IDX(n, k, j, 0), u + IDX(n, k, js, 0), sizeof(double) * (size_t)n);
    }
  for (int k = g; k < n - g; k++)
    for (int j = g; j < n - g; j++)
      for (int i = 0; i < n; i++) {
        if (!(i < g || i >= n - g)) continue;
        u[IDX(n, k, j, i)] = u[IDX(n, k, j, wrap_e(i, n, g))];
      }
}

void orc_transfer(orc_ctx *c) {
  if (c->p.physical) return;  /* no neighbour: unpack writes nothing (trans.c:154-251, left < 0) */
  for (int v = 0; v < c->V; v++) orc_ghost_fill(c->p.n, c->p.g, c->field + c->n3 * v);
}

/* Periodic interior image P(z,y,x) for -2 <= z,y,x < n+2.  Every entry of the
 * enlarged slabs fielde_* built by enlarge() (trans.c:253-925) from the
 * exchanged slabs equals the pre-transfer value at the interior image of its
 * position; the outside values fieldEu_* (ac_mu, trans.c:927-1432) and the
 * CH chemical potential at the outside cells (ch_mu, trans.c:1433-1937) are
 * read from these slabs.  Interior cells are untouched by the ghost fill, so
 * P can be read from the post-fill field. */
static inline int box_shift(int e, int n, int g) { return e < g ? -1 : (e >= n - g ? 1 : 0); }
static inline const orc_ctx *nb_box(const orc_ctx *c, int sx, int sy, int sz) {
  const struct orc_group *G = c->grp;
  if (!G) return c;
  int x = (c->bx + sx + G->px) % G->px, y = (c->by + sy + G->py) % G->py, z = (c->bz + sz + G->pz) % G->pz;
  return G->box[((size_t)z * G->py + y) * G->px + x];
}
/* the same array of box b as u is of box c (fields, felas, the elastic
 * transfer buffer) */
static const double *same_array(const orc_ctx *c, const orc_ctx *b, const double *u) {
  if (b == c) return u;
  uintptr_t a = (uintptr_t)u;
  size_t nf = (size_t)c->n3 * c->V, ne = (size_t)c->n3 * c->nfel;
  if (a >= (uintptr_t)c->field && a < (uintptr_t)(c->field + nf)) return b->field + (u - c->field);
  if (a >= (uintptr_t)c->felas && a < (uintptr_t)(c->felas + ne)) return b->felas + (u - c->felas);
  if (a >= (uintptr_t)c->fbuf && a < (uintptr_t)(c->fbuf + c->n3)) return b->fbuf + (u - c->fbuf);
  return u;
}
/* P(z, y, x) for -2 <= z, y, x < n + 2: in a decomposed run the interior
 * image lies in the neighbouring box the position points into (a corner
 * image in the diagonal box), as the relayed slabs of left_right_pack /
 * top_bottom_pack carry it (trans.c:7-152, 253-925). */
static inline double PimgC(const orc_ctx *c, const double *u, int z, int y, int x) {
  const int n = c->p.n, g = c->p.g;
  const double *src = u;
  if (c->grp) src = same_array(c, nb_box(c, box_shift(x, n, g), box_shift(y, n, g), box_shift(z, n, g)), u);
  return src[IDX(n, wrap_e(z, n, g), wrap_e(y, n, g), wrap_e(x, n, g))];
}

/* Neighbour rule of omega's stencils (anisotropic.c:177-216, calc.c, mobility.c
 * laplace): the low neighbour of cell 0 is 0 when that side has a neighbour
 * (periodic box) and the mirror cell 1 when it is physical (left < 0, i ==
 * ix1: n_left = ix1 + 1); the high neighbour of cell n - 1 is 0 in both
 * cases (with left < 0 the first branch applies and sets n_right = -1). */
static inline double lo_nb(const double *a, size_t l, size_t step, int i, int phys) {
  return i > 0 ? a[l - step] : (phys ? a[l + step] : 0.0);
}
static inline double hi_nb(const double *a, size_t l, size_t step, int i, int n) {
  return i < n - 1 ? a[l + step] : 0.0;
}

/* ------------------------------------------------------------------ */
/* Nonlinear term of an AC field: ac_calc_FU (calc.c:8-68) ->
 * ac_calc_F1 (anisotropic.c:9-161) with SWTICH_AC_FIELD1 (ScLETD.h:108-120)
 * and AC_f (formula.c:30-38) generalised to nac variants:
 *   f1 = LE*KE*u_p - LE*(a2 u_p - a3 u_p^2 + a4 (sum_m u_m^2) u_p)
 *   f1 += LE*epn2*u_out/h^2 on each of the 6 box faces (no kk factor)
 *   N  = f1 - epn2*LE*F2 (ANISOTROPIC) - LE*ElasticScale*felas (ELASTIC=1)
 * F2 (anisotropic.c:164-627): g = Lambda (gx, gy, gz) with central 2h
 * differences where gy, gz use (u[-1] - u[+1]) (the reference's top = j-1,
 * front = k-1 naming), then F2 = d/dx g_x + (g_y[-1]-g_y[+1])/2h + ...;
 * neighbours outside the extended box read 0 (both sides have neighbours). */
static void ac_f2(const orc_ctx *c, int pv, const double *u, double *gx_, double *gy_, double *gz_, double *f2) {
  const int n = c->p.n;
  const double hx = c->p.hx, hy = c->p.hy, hz = c->p.hz;
  const double *L = c->lam + 9 * pv;
  const int ph = c->p.physical;
  const size_t nn = (size_t)n * n;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        const size_t l = IDX(n, k, j, i);
        double a_left = lo_nb(u, l, 1, i, ph), a_right = hi_nb(u, l, 1, i, n);
        double a_top = lo_nb(u, l, n, j, ph), a_bottom = hi_nb(u, l, n, j, n);
        double a_front = lo_nb(u, l, nn, k, ph), a_back = hi_nb(u, l, nn, k, n);
        double gradx = (a_right - a_left) / (2.0 * hx);
        double grady = (a_top - a_bottom) / (2.0 * hy);
        double gradz = (a_front - a_back) / (2.0 * hz);
        gx_[l] = L[0] * gradx + L[1] * grady + L[2] * gradz;
        gy_[l] = L[3] * gradx + L[4] * grady + L[5] * gradz;
        gz_[l] = L[6] * gradx + L[7] * grady + L[8] * gradz;
      }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        const size_t l = IDX(n, k, j, i);
        double a_left = lo_nb(gx_, l, 1, i, ph), a_right = hi_nb(gx_, l, 1, i, n);
        double a_top = lo_nb(gy_, l, n, j, ph), a_bottom = hi_nb(gy_, l, n, j, n);
        double a_front = lo_nb(gz_, l, nn, k, ph), a_back = hi_nb(gz_, l, nn, k, n);
        double tmp = 0.0;
        tmp += (a_right - a_left) / 2.0 / hx;
        tmp += (a_top - a_bottom) / 2.0 / hy;
        tmp += (a_front - a_back) / 2.0 / hz;
        f2[l] = tmp;
      }
}

static double calphad_f(const orc_ctx *c, size_t l, int pv, double gbcc, double ghcp);
static void ac_f2_full(const orc_ctx *c, int pv, const double *u, double *f2);
static void ac_nonlinear(orc_ctx *c, int pv, double *out) {
  const orc_params *p = &c->p;
  const int n = p->n, nac = p->nac;
  const double LE = p->LE, KE = p->KE, epn2 = p->epn2;
  const double hx = p->hx, hy = p->hy, hz = p->hz;
  const double *u = c->field + c->n3 * pv;
  double gbcc = 0.0, ghcp = 0.0;
  if (p->free_energy == 1) orc_calphad(&gbcc, &ghcp);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        size_t l = IDX(n, k, j, i);
        if (p->free_energy == 1) {
          out[l] = calphad_f(c, l, pv, gbcc, ghcp);
          continue;
        }
        double s = 0.0;
        for (int m = 0; m < nac; m++) {
          double um = c->field[c->n3 * m + l];
          s = s + um * um;
        }
        double up = u[l];
        double fp = p->a2 * up - p->a3 * up * up + p->a4 * s * up;
        double f = LE * KE * up;
        f -= LE * fp;
        out[l] = f;
      }
  /* boundary coupling, order left,right,top,bottom,front,back; only on sides
   * with a neighbour (anisotropic.c:106-158, left >= 0 ...) */
  if (!p->physical) {
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++) out[IDX(n, k, j, 0)] += LE * epn2 * PimgC(c, u, k, j, -1) / hx / hx;
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++) out[IDX(n, k, j, n - 1)] += LE * epn2 * PimgC(c, u, k, j, n) / hx / hx;
  for (int k = 0; k < n; k++)
    for (int i = 0; i < n; i++) out[IDX(n, k, 0, i)] += LE * epn2 * PimgC(c, u, k, -1, i) / hy / hy;
  for (int k = 0; k < n; k++)
    for (int i = 0; i < n; i++) out[IDX(n, k, n - 1, i)] += LE * epn2 * PimgC(c, u, k, n, i) / hy / hy;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) out[IDX(n, 0, j, i)] += LE * epn2 * PimgC(c, u, -1, j, i) / hz / hz;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) out[IDX(n, n - 1, j, i)] += LE * epn2 * PimgC(c, u, n, j, i) / hz / hz;
  }
  if (p->aniso == 1) ac_f2(c, pv, u, c->t0, c->t1, c->t2, c->t3);
  if (p->aniso == 2) ac_f2_full(c, pv, u, c->t3);
  const double *felas = p->elastic == 1 ? c->felas + c->n3 * pv : NULL;
  /* omega's order: F2 then felas (calc.c:20-50); alpha's (ac_calc_F2_dev,
   * anisotropic_hip.c:2190-2211): felas, F2, then the grain mask */
  const int alpha_order = p->free_energy == 1;
#pragma omp parallel for schedule(static)
  for (size_t l = 0; l < c->n3; l++) {
    double f = out[l];
    if (alpha_order && felas) f -= LE * p->elastic_scale * felas[l];
    if (p->aniso) f -= epn2 * LE * c->t3[l];
    if (!alpha_order && felas) f -= LE * p->elastic_scale * felas[l];
    if (p->mask && p->mask[l] == 0) f -= 0.5 * u[l];
    out[l] = f;
  }
}

/* CH nonlinear term: ch_calc_FU_ConstantMobility (calc.c:70-473) with the
 * chemical potential at the outside cells from ch_mu (trans.c:1433-1937):
 *   t   = -LCI*KCI*c + LCI*CH_f(c)          (SWITCH_CH_FIELD1, ScLETD.h:125)
 *   t  -= LCI*epn2*c_out/h^2 on each face
 *   out = Lap_T t  (out-of-box neighbours 0)
 *   out -= mu_out/h^2 on each face, mu_out = LCI*(epn2*Lap c - CH_f(c)) at
 *          the outside cell, evaluated on the periodic image.
 * CH_f(c) = c (c^2 - 1), the derivative of (c^2-1)^2/4 (field.c:575). */
/* Planted mistakes for the tests that show the CH checks can fail
 * (orc_set_mutation): 1 flips the sign of CH_f, 2 uses lam in place of
 * lam^2 in the CH linear operator.  0 (the default) is the restatement. */
static int orc_mutation = 0;
void orc_set_mutation(int m) { orc_mutation = m; }

static inline double ch_f(double cc) { return orc_mutation == 1 ? -cc * (cc * cc - 1.0) : cc * (cc * cc - 1.0); }

static double ch_mu_out(const orc_ctx *c, const double *u, int z, int y, int x) {
  const orc_params *p = &c->p;
  const int q = (int)((u - c->field) / (ptrdiff_t)c->n3) - p->nac;  /* the CH field u is */
  double v0 = PimgC(c, u, z, y, x);
  double mu = (PimgC(c, u, z, y, x - 1) + (-2.0) * v0 + PimgC(c, u, z, y, x + 1)) / p->hx / p->hx;
  mu += (PimgC(c, u, z, y - 1, x) + (-2.0) * v0 + PimgC(c, u, z, y + 1, x)) / p->hy / p->hy;
  mu += (PimgC(c, u, z - 1, y, x) + (-2.0) * v0 + PimgC(c, u, z + 1, y, x)) / p->hz / p->hz;
  mu = p->LCI * p->epn2 * mu;
  mu -= p->LCI * ch_f(v0);
  /* ELASTIC = 2: minus LCI felas at the outside cell (SWITCH_CH_FIELDMU, ScLETD.h:153-156), from
   * the enlarged slabs of the force's own exchange (elastic.c:757-759): its interior image */
  if (p->elastic == 2) mu -= p->LCI * PimgC(c, c->felas + c->n3 * q, z, y, x);
  return mu;
}

static void ch_nonlinear(orc_ctx *c, int v, double *out) {
  const orc_params *p = &c->p;
  const int n = p->n;
  const double LCI = p->LCI, KCI = p->KCI, epn2 = p->epn2;
  const double hx = p->hx, hy = p->hy, hz = p->hz;
  const double *u = c->field + c->n3 * v;
  double *t = c->t0;
  const double *fe = p->elastic == 2 ? c->felas + c->n3 * (v - p->nac) : NULL;
#pragma omp parallel for schedule(static)
  for (size_t l = 0; l < c->n3; l++) {
    double cc = u[l];
    double f = -LCI * KCI * cc;
    f += LCI * ch_f(cc);
    if (fe) f += LCI * fe[l];  /* ELASTIC == CH_FUNCTION (calc.c:90-101) */
    t[l] = f;
  }
  const int ph = p->physical;
  if (!ph) {
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++) t[IDX(n, k, j, 0)] -= LCI * epn2 * PimgC(c, u, k, j, -1) / hx / hx;
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++) t[IDX(n, k, j, n - 1)] -= LCI * epn2 * PimgC(c, u, k, j, n) / hx / hx;
  for (int k = 0; k < n; k++)
    for (int i = 0; i < n; i++) t[IDX(n, k, 0, i)] -= LCI * epn2 * PimgC(c, u, k, -1, i) / hy / hy;
  for (int k = 0; k < n; k++)
    for (int i = 0; i < n; i++) t[IDX(n, k, n - 1, i)] -= LCI * epn2 * PimgC(c, u, k, n, i) / hy / hy;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) t[IDX(n, 0, j, i)] -= LCI * epn2 * PimgC(c, u, -1, j, i) / hz / hz;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) t[IDX(n, n - 1, j, i)] -= LCI * epn2 * PimgC(c, u, n, j, i) / hz / hz;
  }
  const size_t nn = (size_t)n * n;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        const size_t l = IDX(n, k, j, i);
        double tc = t[l];
        double f_left = lo_nb(t, l, 1, i, ph), f_right = hi_nb(t, l, 1, i, n);
        double f_top = lo_nb(t, l, n, j, ph), f_bottom = hi_nb(t, l, n, j, n);
        double f_front = lo_nb(t, l, nn, k, ph), f_back = hi_nb(t, l, nn, k, n);
        double r = (f_left + f_right - 2.0 * tc) / hx / hx;
        r += (f_top + f_bottom - 2.0 * tc) / hy / hy;
        r += (f_front + f_back - 2.0 * tc) / hz / hz;
        out[l] = r;
      }
  if (ph) return;  /* no outside chemical potential on physical sides */
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++) out[IDX(n, k, j, 0)] -= ch_mu_out(c, u, k, j, -1) / hx / hx;
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++) out[IDX(n, k, j, n - 1)] -= ch_mu_out(c, u, k, j, n) / hx / hx;
  for (int k = 0; k < n; k++)
    for (int i = 0; i < n; i++) out[IDX(n, k, 0, i)] -= ch_mu_out(c, u, k, -1, i) / hy / hy;
  for (int k = 0; k < n; k++)
    for (int i = 0; i < n; i++) out[IDX(n, k, n - 1, i)] -= ch_mu_out(c, u, k, n, i) / hy / hy;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) out[IDX(n, 0, j, i)] -= ch_mu_out(c, u, -1, j, i) / hz / hz;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) out[IDX(n, n - 1, j, i)] -= ch_mu_out(c, u, n, j, i) / hz / hz;
}

/* ------------------------------------------------------------------ */
/* beta->alpha physics (source_beta_to_alpha/src).                      */

/* CALPHAD Gibbs energies of the bcc and hcp phases (J/mol) at the fixed
 * compositions of ac_calc_F1_dev (anisotropic_hip.c:2029-2048): x_Al, x_V =
 * 0.102, 0.036 (bcc) and 0.1057, 0.0223 (hcp), x_Ti the rest; lattice
 * stabilities and Redlich-Kister parameters of init_vars (vars.c:146-190) at
 * T_a = 1093 K, R_a = 8.31451 (ScLETD.h:24-26).  Pinned by the values the
 * reference notes beside its parameters (GHSERAL 8.9664e4, GHSERTI 7.2379e3,
 * GHSERV 2.7046e4, GBCCAL 9.4487e4, GBCCTI 7.4701e3, GHCPAL 9.3179e4,
 * GHCPV 3.3669e4, BL13_0 -85873, BL12_0 -63024, BL12_1 -1965, BL32_0 8860.5,
 * HL13_0 -91123, HL32_0 19320) and by the ideal-mixing term recomputed from
 * the compositions. */
#define CAL_T 1093.0
#define CAL_R 8.31451
/* t: GHSERAL, GHSERTI, GHSERV, GBCCAL, GBCCTI, GHCPAL, GHCPV (vars.c:146-160),
 * then BL13_0, BL12_0, BL12_1, BL32_0, HL13_0, HL32_0 at T_a */
static void calphad_terms(double *t) {
  const double T = CAL_T, lT = log(T);
  t[0] = -11277.683 + 188.661987 * T - 31.748192 * T * lT - 1.234e28 * pow(T, -9) + 137529.83;
  t[1] = 908.837 + 67.048538 * T - 14.9466 * T * lT - 0.0081465 * T * T + 2.02715e-7 * T * T * T - 1477660.0 / T +
         58166.677;
  t[2] = -7967.84 + 143.291 * T - 25.9 * T * lT + 6.25e-5 * T * T - 6.8e-7 * T * T * T + 77277.217;
  t[3] = 10083.0 - 4.812 * T + t[0];
  t[4] = 5758.6 + 38.38987 * T - 7.4305 * T * lT + 0.0093636 * T * T - 1.04805e-6 * T * T * T - 525093.0 / T + t[1];
  t[5] = 5481.0 - 1.799 * T + t[0];
  t[6] = 4000.0 + 2.4 * T + t[2];
  t[7] = -128500.0 + 39.0 * T;
  t[8] = -98000.0 + 32.0 * T;
  t[9] = 3500.0 - 5.0 * T;
  t[10] = 10500.0 - 1.5 * T;
  t[11] = -133750.0 + 39.0 * T;
  t[12] = 30250.0 - 10.0 * T;
}
void orc_calphad_terms(double *t) { calphad_terms(t); }

void orc_calphad(double *gbcc, double *ghcp) {
  double t[13];
  calphad_terms(t);
  const double T = CAL_T;
  const double ser_ti = t[1], ser_v = t[2], bcc_al = t[3], bcc_ti = t[4], hcp_al = t[5], hcp_v = t[6];
  /* interaction parameters: 1 = Al, 2 = V, 3 = Ti */
  const double b13_0 = -128500.0 + 39.0 * T, b13_1 = 6000.0, b13_2 = 21200.0, b132_0 = -8000.0;
  const double b12_0 = -98000.0 + 32.0 * T, b12_1 = 3500.0 - 5.0 * T;
  const double b32_0 = 10500.0 - 1.5 * T, b32_1 = 2000.0, b32_2 = 1000.0;
  const double h13_0 = -133750.0 + 39.0 * T, h13_1 = 250.0, h13_2 = 17250.0;
  const double h132_0 = -60000.0, h132_1 = -30000.0, h132_2 = 60000.0;
  const double h12_0 = -98000.0 + 32.0 * T, h12_1 = 3500.0 - 5.0 * T, h32_0 = 30250.0 - 10.0 * T;
  {
    const double al = 0.102, v = 0.036, ti = 1.0 - al - v;
    *gbcc = ti * bcc_ti + al * bcc_al + v * ser_v + CAL_R * T * (ti * log(ti) + al * log(al) + v * log(v)) +
            al * v * (b12_0 + b12_1 * (al - v)) + ti * v * (b32_0 + b32_1 * (ti - v) + b32_2 * (ti - v) * (ti - v)) +
            al * ti * (b13_0 + b13_1 * (al - ti) + b13_2 * (al - ti) * (al - ti)) + al * v * ti * b132_0;
  }
  {
    const double al = 0.1057, v = 0.0223, ti = 1.0 - al - v;
    *ghcp = ti * ser_ti + al * hcp_al + v * hcp_v + CAL_R * T * (ti * log(ti) + al * log(al) + v * log(v)) +
            al * v * (h12_0 + h12_1 * (al - v)) + ti * v * h32_0 +
            al * ti * (h13_0 + h13_1 * (al - ti) + h13_2 * (al - ti) * (al - ti)) +
            al * v * ti * (al * h132_0 + ti * h132_1 + v * h132_2);
  }
}

/* alpha's driving force (ac_calc_F1_dev, anisotropic_hip.c:2075-2140) for
 * variant p, generalised from 6 variants to nac:
 *   dp   = 30 u^2 (1 - u)^2                 (derivative of u^3 (10 - 15u + 6u^2))
 *   dq   = 4 wmega u (1 - u)(2 - u),        wmega = 100 (vars.c:144)
 *   f    = LE KE u - LE (( -(Gbcc dp) + (Ghcp dp) + dq) / gnormal
 *                        + 4 ALPHA sum_{m != p} u u_m u_m)
 * gnormal = 50000, ALPHA = 0.02 (ScLETD.h:27-29); the sum in field order.
 * Pinned: dp and dq as derivatives of their polynomials (finite differences
 * of independently written energies), the coupling as the derivative of
 * ALPHA sum_{m<m'} u_m^2 u_m'^2 * 4/2 ... (see tests/test_oracle.py). */
#define CAL_WMEGA 100.0
#define CAL_GNORMAL 50000.0
#define CAL_ALPHA 0.02
static double calphad_f(const orc_ctx *c, size_t l, int pv, double gbcc, double ghcp) {
  const orc_params *p = &c->p;
  const double up = c->field[c->n3 * pv + l];
  double q = 0.0;
  for (int m = 0; m < p->nac; m++) {
    if (m == pv) continue;
    const double um = c->field[c->n3 * m + l];
    q += up * um * um;
  }
  const double dp = 30.0 * up * up * (1 - up) * (1 - up);
  const double dq = 4.0 * CAL_WMEGA * up * (1 - up) * (2 - up);
  const double dfeta = -(gbcc * dp) + (ghcp * dp) + dq;
  double f = p->LE * p->KE * up;
  f -= p->LE * (dfeta / CAL_GNORMAL + 4.0 * CAL_ALPHA * q);
  return f;
}

/* alpha's full gradient term (anisotropic_calc_dev_2, anisotropic_hip.c:
 * 528-647): f2 = sum_ij Lambda_ij d_i d_j u with three-point second
 * differences and four-corner cross differences,
 *   fxx = (u[i+1] - 2u + u[i-1]) / hx^2
 *   fxy = ((u[i+1,j+1] - u[i-1,j+1]) - (u[i+1,j-1] - u[i-1,j-1])) / (2hx 2hy)
 * on every extended cell; the neighbours outside the box come from the
 * enlarged slabs, i.e. the periodic image (one box, periodic).  Pinned: exact
 * on quadratic fields away from the box faces (any Lambda, cross terms
 * included), and the periodic image against a shifted field. */
static inline double pimg_or_in(const orc_ctx *c, const double *u, int n, int z, int y, int x);
static inline int mirror_in(int x, int n);
static void ac_f2_full(const orc_ctx *c, int pv, const double *u, double *f2) {
  const int n = c->p.n;
  const double hx = c->p.hx, hy = c->p.hy, hz = c->p.hz;
  const double *L = c->lam + 9 * pv;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
#define U(dz, dy, dx)                                                                                          \
  (c->p.physical ? u[IDX(n, mirror_in(k + (dz), n), mirror_in(j + (dy), n), mirror_in(i + (dx), n))]            \
                 : pimg_or_in(c, u, n, k + (dz), j + (dy), i + (dx)))
        const double m = U(0, 0, 0);
        const double fxx = (U(0, 0, 1) - 2 * m + U(0, 0, -1)) / hx / hx;
        const double fyy = (U(0, 1, 0) - 2 * m + U(0, -1, 0)) / hy / hy;
        const double fzz = (U(1, 0, 0) - 2 * m + U(-1, 0, 0)) / hz / hz;
        const double fxy = ((U(0, 1, 1) - U(0, 1, -1)) - (U(0, -1, 1) - U(0, -1, -1))) / 2.0 / hx / 2.0 / hy;
        const double fxz = ((U(1, 0, 1) - U(1, 0, -1)) - (U(-1, 0, 1) - U(-1, 0, -1))) / 2.0 / hx / 2.0 / hz;
        const double fyz = ((U(1, 1, 0) - U(1, -1, 0)) - (U(-1, 1, 0) - U(-1, -1, 0))) / 2.0 / hy / 2.0 / hz;
#undef U
        double t = 0.0;
        t += L[0] * fxx;
        t += L[1] * fxy;
        t += L[2] * fxz;
        t += L[3] * fxy;
        t += L[4] * fyy;
        t += L[5] * fyz;
        t += L[6] * fxz;
        t += L[7] * fyz;
        t += L[8] * fzz;
        f2[IDX(n, k, j, i)] = t;
      }
}
/* a cell inside the box reads the field, one outside its periodic image */
static inline double pimg_or_in(const orc_ctx *c, const double *u, int n, int z, int y, int x) {
  if (z >= 0 && z < n && y >= 0 && y < n && x >= 0 && x < n) return u[IDX(n, z, y, x)];
  return PimgC(c, u, z, y, x);
}
/* physical sides (anisotropic_hip.c:546-591): n_left = i + 1 at i == ix1,
 * n_right = i - 1 at i == ix4 - 1, per axis; the mirror stays in the box */
static inline int mirror_in(int x, int n) { return x < 0 ? -x : (x >= n ? 2 * (n - 1) - x : x); }

/* ------------------------------------------------------------------ */
/* Variable-mobility CH: ch_calc_FU_VariableMobility (mobility.c:838-858)
 * with update_M_C (860-922) and the pieces of 12-836, one periodic box:
 *   M_q   = x_q (1 - x_q)(D_qA (1 - eta_0) + D_qG eta_0) on every cell,
 *           (x, D) of Cu, Mn, Ni for q = 0, 1, 2 (mobility.c:864-906),
 *           C_q = max over the box of M_q
 *   f0    = -LCI KCI c + LCI CH_f(c)           (SWITCH_CH_FIELD1)
 *   ft    = -f0 + KCI c, + epn2 c_out / h^2 on the faces
 *   f1    = -C Lap_T ft - C mu_out / h^2 on the faces
 *   ft'   = epn2 Lap_T c - f0, + epn2 c_out / h^2 on the faces
 *   f2    = -div((M - C) grad ft')             (ch_divergence_M_gradient)
 *   N     = LCI (f1 + f2)
 * Lap_T and div read 0 outside the box (mobility.c:12-244, 435-722, the
 * periodic branch).  mu_out as in the constant-mobility term.  Pinned: at
 * eta = 0 (M = C) N is C times the constant-mobility term for LCI = 1, and
 * the divergence is checked on its own against a closed form. */
static void vm_mobility(const orc_ctx *c, int q, double *M, double *Cmax) {
  const double gconst = 8.314472, tempr = 823.0, RT = gconst * tempr;
  const double x0[3] = {0.15, 0.01, 0.01};
  const double DCuA0 = 4.7e-5 * exp(-2.44e5 / RT);
  double DA, DG;
  if (q == 0) {
    DA = 1.0;
    DG = (4.3e-5 * exp(-2.80e5 / RT)) / DCuA0;
  } else if (q == 1) {
    DA = (1.49e-4 * exp(-2.63e5 / RT)) / DCuA0;
    DG = (2.78e-5 * exp(-2.64e5 / RT)) / DCuA0;
  } else {
    DA = (1.4e-4 * exp(-2.56e5 / RT)) / DCuA0;
    DG = (1.08e-5 * exp(-2.73e5 / RT)) / DCuA0;
  }
  const double *eta = c->field;
  double cm = -1.0e30;
  for (size_t l = 0; l < c->n3; l++) {
    M[l] = x0[q] * (1.0 - x0[q]) * (DA * (1.0 - eta[l]) + DG * eta[l]);
    cm = cm > M[l] ? cm : M[l];
  }
  *Cmax = cm;
}

static void lap_zero(int n, double hx, double hy, double hz, const double *a, double *out, int ph) {
  const size_t nn = (size_t)n * n;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        const size_t l = IDX(n, k, j, i);
        const double ac = a[l];
        const double al = lo_nb(a, l, 1, i, ph), ar = hi_nb(a, l, 1, i, n);
        const double at = lo_nb(a, l, n, j, ph), ab = hi_nb(a, l, n, j, n);
        const double af = lo_nb(a, l, nn, k, ph), ak = hi_nb(a, l, nn, k, n);
        out[l] = (al + ar - 2 * ac) / hx / hx + (at + ab - 2 * ac) / hy / hy + (af + ak - 2 * ac) / hz / hz;
      }
}

static void div_m_grad(int n, double hx, double hy, double hz, const double *M, const double *a, double *out, int ph) {
  const size_t nn = (size_t)n * n;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        const size_t l = IDX(n, k, j, i);
        const double am = a[l], mm = M[l];
        const double al = lo_nb(a, l, 1, i, ph), ml = lo_nb(M, l, 1, i, ph);
        const double ar = hi_nb(a, l, 1, i, n), mr = hi_nb(M, l, 1, i, n);
        const double at = lo_nb(a, l, n, j, ph), mt = lo_nb(M, l, n, j, ph);
        const double ab = hi_nb(a, l, n, j, n), mb = hi_nb(M, l, n, j, n);
        const double af = lo_nb(a, l, nn, k, ph), mf = lo_nb(M, l, nn, k, ph);
        const double ak = hi_nb(a, l, nn, k, n), mk = hi_nb(M, l, nn, k, n);
        double t = 0.0;
        t += (ml + mm) * (al - am) / 2.0 / hx / hx;
        t += (mr + mm) * (ar - am) / 2.0 / hx / hx;
        t += (mt + mm) * (at - am) / 2.0 / hy / hy;
        t += (mb + mm) * (ab - am) / 2.0 / hy / hy;
        t += (mf + mm) * (af - am) / 2.0 / hz / hz;
        t += (mk + mm) * (ak - am) / 2.0 / hz / hz;
        out[l] = t;
      }
}
void orc_div_m_grad(int n, double hx, double hy, double hz, const double *M, const double *a, double *out) {
  div_m_grad(n, hx, hy, hz, M, a, out, 0);
}

static void ch_nonlinear_vm(orc_ctx *c, int v, double *out) {
  const orc_params *p = &c->p;
  const int n = p->n, q = v - p->nac;
  const double LCI = p->LCI, KCI = p->KCI, epn2 = p->epn2, hx = p->hx, hy = p->hy, hz = p->hz;
  const double *u = c->field + c->n3 * v;
  double *M = c->t0, *f0 = c->t1, *ft = c->t2, *w = c->t3;
  double *f1 = malloc(sizeof(double) * c->n3);
  double Cq;
  vm_mobility(c, q, M, &Cq);
  for (size_t l = 0; l < c->n3; l++) f0[l] = -LCI * KCI * u[l] + LCI * ch_f(u[l]);
  /* f1 = -C Lap_T(-f0 + KCI c + faces) - C mu_out / h^2 */
  for (size_t l = 0; l < c->n3; l++) ft[l] = -f0[l] + KCI * u[l];
#define FACES(arr, coef)                                                                                 \
  for (int k = 0; k < n; k++)                                                                            \
    for (int j = 0; j < n; j++) arr[IDX(n, k, j, 0)] += coef * PimgC(c, u, k, j, -1) / hx / hx;       \
  for (int k = 0; k < n; k++)                                                                            \
    for (int j = 0; j < n; j++) arr[IDX(n, k, j, n - 1)] += coef * PimgC(c, u, k, j, n) / hx / hx;    \
  for (int k = 0; k < n; k++)                                                                            \
    for (int i = 0; i < n; i++) arr[IDX(n, k, 0, i)] += coef * PimgC(c, u, k, -1, i) / hy / hy;       \
  for (int k = 0; k < n; k++)                                                                            \
    for (int i = 0; i < n; i++) arr[IDX(n, k, n - 1, i)] += coef * PimgC(c, u, k, n, i) / hy / hy;    \
  for (int j = 0; j < n; j++)                                                                            \
    for (int i = 0; i < n; i++) arr[IDX(n, 0, j, i)] += coef * PimgC(c, u, -1, j, i) / hz / hz;       \
  for (int j = 0; j < n; j++)                                                                            \
    for (int i = 0; i < n; i++) arr[IDX(n, n - 1, j, i)] += coef * PimgC(c, u, n, j, i) / hz / hz;
  const int ph = p->physical;
  if (!ph) { FACES(ft, epn2) }
  lap_zero(n, hx, hy, hz, ft, f1, ph);
  for (size_t l = 0; l < c->n3; l++) f1[l] = -Cq * f1[l];
  if (!ph) {
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++) f1[IDX(n, k, j, 0)] -= Cq * ch_mu_out(c, u, k, j, -1) / hx / hx;
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++) f1[IDX(n, k, j, n - 1)] -= Cq * ch_mu_out(c, u, k, j, n) / hx / hx;
  for (int k = 0; k < n; k++)
    for (int i = 0; i < n; i++) f1[IDX(n, k, 0, i)] -= Cq * ch_mu_out(c, u, k, -1, i) / hy / hy;
  for (int k = 0; k < n; k++)
    for (int i = 0; i < n; i++) f1[IDX(n, k, n - 1, i)] -= Cq * ch_mu_out(c, u, k, n, i) / hy / hy;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) f1[IDX(n, 0, j, i)] -= Cq * ch_mu_out(c, u, -1, j, i) / hz / hz;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) f1[IDX(n, n - 1, j, i)] -= Cq * ch_mu_out(c, u, n, j, i) / hz / hz;
  }
  /* f2 = -div((M - C) grad(epn2 Lap_T c - f0 + faces)) */
  lap_zero(n, hx, hy, hz, u, ft, ph);
  for (size_t l = 0; l < c->n3; l++) ft[l] = epn2 * ft[l] - f0[l];
  if (!ph) { FACES(ft, epn2) }
#undef FACES
  for (size_t l = 0; l < c->n3; l++) M[l] = M[l] - Cq;
  div_m_grad(n, hx, hy, hz, M, ft, w, ph);
  for (size_t l = 0; l < c->n3; l++) out[l] = LCI * (f1[l] + -w[l]);
  free(f1);
}

void orc_nonlinear(orc_ctx *c, int v, double *out) {
  if (v < c->p.nac) ac_nonlinear(c, v, out);
  else if (c->p.varmob) ch_nonlinear_vm(c, v, out);
  else ch_nonlinear(c, v, out);
}

/* ------------------------------------------------------------------ */
/* Own mixed-radix complex FFT (the reference links fftMPI, fft3d.c:63-131,
 * which is not vendored).  Recursive decimation in time over the prime
 * factors of N with exact cos/sin twiddles.  Forward: exp(-2 pi i jk/N),
 * unscaled; backward: exp(+...), scaled by 1/N^3 over the 3D transform
 * (fftMPI default scaled = 1; fixed by the k = 0 KAT). */
typedef struct { double re, im; } cplx;

/* work: N scratch entries shared down the recursion (a parent only uses it
 * after all its children have returned) */
static void fft1d_rec(int N, const cplx *in, int stride, cplx *out, int sign, const int *fac, const cplx *tw, int twN, cplx *res) {
  if (N == 1) { out[0] = in[0]; return; }
  int r = fac[0], m = N / r;
  /* sub-transforms of the r decimated sequences */
  for (int q = 0; q < r; q++) fft1d_rec(m, in + (size_t)q * stride, stride * r, out + (size_t)q * m, sign, fac + 1, tw, twN, res);
  for (int k = 0; k < m; k++)
    for (int s = 0; s < r; s++) {
      int kk = k + s * m; /* output index */
      double sr = 0.0, si = 0.0;
      for (int q = 0; q < r; q++) {
        /* twiddle exp(sign*2 pi i q kk / N) */
        size_t e = ((size_t)q * kk) % N;
        size_t ti = e * (size_t)(twN / N);
        double wr = tw[ti].re, wi = sign * tw[ti].im;
        cplx x = out[(size_t)q * m + k];
        sr += x.re * wr - x.im * wi;
        si += x.re * wi + x.im * wr;
      }
      res[kk].re = sr;
      res[kk].im = si;
    }
  memcpy(out, res, sizeof(cplx) * (size_t)N);
}

static int factorize(int N, int *fac) {
  int nf = 0;
  int x = N;
  for (int pr = 2; pr * pr <= x; pr++)
    while (x % pr == 0) { fac[nf++] = pr; x /= pr; }
  if (x > 1) fac[nf++] = x;
  return nf;
}

int orc_fft3d(int N, double *data, int dir) {
  if (N < 1) return -1;
  int fac[64];
  factorize(N, fac);
  cplx *tw = malloc(sizeof(cplx) * (size_t)N);
  const double pi = 3.14159265358979323846;
  for (int e = 0; e < N; e++) { /* exp(-2 pi i e / N) */
    tw[e].re = cos(2.0 * pi * e / N);
    tw[e].im = -sin(2.0 * pi * e / N);
  }
  int sign = dir > 0 ? 1 : -1;
  cplx *d = (cplx *)data;
  size_t NN = (size_t)N * N;
  /* lines along x, y, z */
  for (int axis = 0; axis < 3; axis++) {
#pragma omp parallel
    {
      cplx *line = malloc(sizeof(cplx) * (size_t)N), *res = malloc(sizeof(cplx) * (size_t)N);
      cplx *work = malloc(sizeof(cplx) * (size_t)N);
#pragma omp for schedule(static)
      for (size_t l = 0; l < NN; l++) {
        size_t base, st;
        if (axis == 0) { base = l * N; st = 1; }
        else if (axis == 1) { base = (l / N) * NN + (l % N); st = N; }
        else { base = l; st = NN; }
        for (int t = 0; t < N; t++) line[t] = d[base + t * st];
        fft1d_rec(N, line, 1, res, sign, fac, tw, N, work);
        for (int t = 0; t < N; t++) d[base + t * st] = res[t];
      }
      free(line);
      free(res);
      free(work);
    }
  }
  if (dir < 0) {
    double s = 1.0 / ((double)N * N * N);
    for (size_t l = 0; l < NN * N; l++) { d[l].re *= s; d[l].im *= s; }
  }
  free(tw);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Khachaturyan B_pq(n) (elastic.c:460-605, inverse_3x3 442-457, C_ijkl and
 * sigma0 = C:eps0 from elastic_input 349-438):
 *   B = C_ijkl e_p,ij e_q,kl - n_i s_p,ij Omega_jk s_q,kl n_l,
 *   Omega = (C_ijkl n_j n_l)^-1, n from signed wavenumbers (g < N/2 ? g : g-N),
 *   k = 0 uses n = (0,1,0). */
static int voigt(int i, int j) {
  if (i == j) return i;
  if ((i == 1 && j == 2) || (i == 2 && j == 1)) return 3;
  if ((i == 0 && j == 2) || (i == 2 && j == 0)) return 4;
  return 5;
}

static void c4_from_voigt(const double *C2, double *C4) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) C4[((i * 3 + j) * 3 + k) * 3 + l] = C2[voigt(i, j) * 6 + voigt(k, l)];
}

static void inverse3(double *om, const double *oi) {
  double nr = 0.0;
  nr += oi[0] * (oi[4] * oi[8] - oi[5] * oi[7]);
  nr -= oi[1] * (oi[3] * oi[8] - oi[5] * oi[6]);
  nr += oi[2] * (oi[3] * oi[7] - oi[4] * oi[6]);
  om[0] = (+1) * (oi[4] * oi[8] - oi[5] * oi[7]) / nr;
  om[1] = (-1) * (oi[3] * oi[8] - oi[6] * oi[5]) / nr;
  om[2] = (+1) * (oi[3] * oi[7] - oi[6] * oi[4]) / nr;
  om[3] = (-1) * (oi[1] * oi[8] - oi[7] * oi[2]) / nr;
  om[4] = (+1) * (oi[0] * oi[8] - oi[2] * oi[6]) / nr;
  om[5] = (-1) * (oi[0] * oi[7] - oi[6] * oi[1]) / nr;
  om[6] = (+1) * (oi[1] * oi[5] - oi[4] * oi[2]) / nr;
  om[7] = (-1) * (oi[0] * oi[5] - oi[3] * oi[2]) / nr;
  om[8] = (+1) * (oi[0] * oi[4] - oi[1] * oi[3]) / nr;
}

double orc_bpq(const orc_params *p, int N, int gx, int gy, int gz, int pp, int qq) {
  double C4[81];
  c4_from_voigt(p->C, C4);
  const double *ep = p->eps0 + 9 * pp, *eq = p->eps0 + 9 * qq;
  double sp[9], sq[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      sp[i * 3 + j] = 0;
      sq[i * 3 + j] = 0;
      for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) {
          sp[i * 3 + j] += C4[((i * 3 + j) * 3 + k) * 3 + l] * ep[k * 3 + l];
          sq[i * 3 + j] += C4[((i * 3 + j) * 3 + k) * 3 + l] * eq[k * 3 + l];
        }
    }
  int cnt = N / 2;
  double n1 = gx < cnt ? 1.0 * gx : -1.0 * (N - gx);
  double n2 = gy < cnt ? 1.0 * gy : -1.0 * (N - gy);
  double n3 = gz < cnt ? 1.0 * gz : -1.0 * (N - gz);
  double nv[3];
  if (gx == 0 && gy == 0 && gz == 0) {
    nv[0] = 0; nv[1] = 1; nv[2] = 0;
  } else {
    double r = pow((pow(n1, 2) + pow(n2, 2) + pow(n3, 2)), 0.5);
    nv[0] = n1 / r; nv[1] = n2 / r; nv[2] = n3 / r;
  }
  double oi[9], om[9];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) {
      oi[i * 3 + k] = 0.0;
      for (int j = 0; j < 3; ++j)
        for (int l = 0; l < 3; ++l) oi[i * 3 + k] += C4[((i * 3 + j) * 3 + k) * 3 + l] * nv[j] * nv[l];
    }
  inverse3(om, oi);
  double t0 = 0, t1 = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) t0 += C4[((i * 3 + j) * 3 + k) * 3 + l] * ep[i * 3 + j] * eq[k * 3 + l];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) t1 += nv[i] * sp[i * 3 + j] * om[j * 3 + k] * sq[k * 3 + l] * nv[l];
  return t0 - t1;
}

/* elastic_calculate (elastic.c:689-762) -> elastic_calculate_ElasDri
 * (617-687): felas_p = Re IFFT[sum_q B_pq FFT(eta_q)] on the interior N^3,
 * written into the interior of the persistent buffer f, then f takes the
 * same self-exchange as the fields (its ghost cells start from the previous
 * variant's values, as the single buffer f does), then felas_p = f.
 * The reference runs nac^2 forward/backward pairs; by linearity nac forward
 * transforms and one backward per p give the same values to roundoff. */
/* B_pq for all q at one wavevector, factored as
 *   tmp0_pq - (sigma_p^T n) . Omega . (sigma_q n)
 * (the same sums as orc_bpq regrouped; equal to roundoff). */
typedef struct {
  double C4[81];
  double sig[9 * 64];
  double t0[64 * 64];
} bfast;

static void bfast_init(const orc_params *p, bfast *b) {
  c4_from_voigt(p->C, b->C4);
  for (int q = 0; q < el_count(p); q++) {
    const double *e = p->eps0 + 9 * q;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double s = 0.0;
        for (int k = 0; k < 3; k++)
          for (int l = 0; l < 3; l++) s += b->C4[((i * 3 + j) * 3 + k) * 3 + l] * e[k * 3 + l];
        b->sig[9 * q + i * 3 + j] = s;
      }
  }
  for (int pp = 0; pp < el_count(p); pp++)
    for (int qq = 0; qq < el_count(p); qq++) {
      double t0 = 0.0;
      const double *ep = p->eps0 + 9 * pp, *eq = p->eps0 + 9 * qq;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++) t0 += b->C4[((i * 3 + j) * 3 + k) * 3 + l] * ep[i * 3 + j] * eq[k * 3 + l];
      b->t0[pp * 64 + qq] = t0;
    }
}

static void bfast_row(const orc_params *p, const bfast *b, int N, int gx, int gy, int gz, int pp, double *row) {
  int cnt = N / 2;
  double n1 = gx < cnt ? 1.0 * gx : -1.0 * (N - gx);
  double n2 = gy < cnt ? 1.0 * gy : -1.0 * (N - gy);
  double n3 = gz < cnt ? 1.0 * gz : -1.0 * (N - gz);
  double nv[3];
  if (gx == 0 && gy == 0 && gz == 0) {
    nv[0] = 0; nv[1] = 1; nv[2] = 0;
  } else {
    double r = pow((pow(n1, 2) + pow(n2, 2) + pow(n3, 2)), 0.5);
    nv[0] = n1 / r; nv[1] = n2 / r; nv[2] = n3 / r;
  }
  double oi[9], om[9];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) {
      oi[i * 3 + k] = 0.0;
      for (int j = 0; j < 3; ++j)
        for (int l = 0; l < 3; ++l) oi[i * 3 + k] += b->C4[((i * 3 + j) * 3 + k) * 3 + l] * nv[j] * nv[l];
    }
  inverse3(om, oi);
  const double *sp = b->sig + 9 * pp;
  double a[3], w[3];
  for (int j = 0; j < 3; j++) a[j] = nv[0] * sp[0 * 3 + j] + nv[1] * sp[1 * 3 + j] + nv[2] * sp[2 * 3 + j];
  for (int k = 0; k < 3; k++) w[k] = a[0] * om[0 * 3 + k] + a[1] * om[1 * 3 + k] + a[2] * om[2 * 3 + k];
  for (int qq = 0; qq < el_count(p); qq++) {
    const double *sq = b->sig + 9 * qq;
    double t1 = 0.0;
    for (int k = 0; k < 3; k++) t1 += w[k] * (sq[k * 3 + 0] * nv[0] + sq[k * 3 + 1] * nv[1] + sq[k * 3 + 2] * nv[2]);
    row[qq] = b->t0[pp * 64 + qq] - t1;
  }
}

void orc_elastic(orc_ctx *c) {
  const orc_params *p = &c->p;
  const int n = p->n, g = p->g, N = n - 2 * g, nac = el_count(p), v0 = el_first(p);
  if (nac == 0 || nac > 64) return;
  size_t N3 = (size_t)N * N * N;
  double *spec = malloc(sizeof(double) * 2 * N3 * nac);
  double *acc = malloc(sizeof(double) * 2 * N3);
  bfast *bf = malloc(sizeof(bfast));
  bfast_init(p, bf);
  for (int q = 0; q < nac; q++) {
    double *s = spec + 2 * N3 * q;
    const double *u = c->field + c->n3 * (v0 + q);
    for (int z = 0; z < N; z++)
      for (int y = 0; y < N; y++)
        for (int x = 0; x < N; x++) {
          size_t m = ((size_t)z * N + y) * N + x;
          s[2 * m] = u[IDX(n, z + g, y + g, x + g)];
          s[2 * m + 1] = 0.0;
        }
    orc_fft3d(N, s, +1);
  }
  for (int pp = 0; pp < nac; pp++) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; z++) {
      double row[64];
      for (int y = 0; y < N; y++)
        for (int x = 0; x < N; x++) {
          size_t m = ((size_t)z * N + y) * N + x;
          bfast_row(p, bf, N, x, y, z, pp, row);
          double re = 0.0, im = 0.0;
          for (int qq = 0; qq < nac; qq++) {
            re += spec[2 * N3 * qq + 2 * m] * row[qq];
            im += spec[2 * N3 * qq + 2 * m + 1] * row[qq];
          }
          acc[2 * m] = re;
          acc[2 * m + 1] = im;
        }
    }
    orc_fft3d(N, acc, -1);
    double *f = c->fbuf;
    for (int z = 0; z < N; z++)
      for (int y = 0; y < N; y++)
        for (int x = 0; x < N; x++) f[IDX(n, z + g, y + g, x + g)] = acc[2 * (((size_t)z * N + y) * N + x)];
    if (!p->physical) orc_ghost_fill(n, g, f);  /* physical sides: f keeps 0 outside the interior */
    memcpy(c->felas + c->n3 * pp, f, sizeof(double) * c->n3);
  }
  free(spec);
  free(acc);
  free(bf);
}

/* ------------------------------------------------------------------ */
/* Mode product: PUX/PUY/PUZ (calc.c:631-684) = cblas_dgemm(ColMajor, NN,
 * m=n, n=n^2, k=n) followed by the cyclic rotation of xyz_yzx/yzx_zxy/
 * zxy_xyz (calc.c:585-629); flat (a,b,c) -> (b,c,a).  rotate = 0 keeps the
 * GEMM layout (PUZ stage 0/1). */
void orc_modeprod(int n, const double *A, const double *in, double *out, int rotate) {
  const size_t n2 = (size_t)n * n;
  double *At = malloc(sizeof(double) * n2); /* At[m][a] = A(m,a) */
  for (int m = 0; m < n; m++)
    for (int a = 0; a < n; a++) At[(size_t)m * n + a] = A[m + (size_t)n * a];
  const int CB = 64;
  size_t nblk = (n2 + CB - 1) / CB;
#pragma omp parallel
  {
    double *Bt = calloc((size_t)n * CB, sizeof(double));
    double *Cb = malloc(sizeof(double) * 4 * CB);
#pragma omp for schedule(static)
    for (size_t b = 0; b < nblk; b++) {
      size_t c0 = b * CB;
      int cw = (int)((c0 + CB <= n2) ? CB : n2 - c0);
      for (int cc = 0; cc < cw; cc++)
        for (int a = 0; a < n; a++) Bt[(size_t)a * CB + cc] = in[a + (size_t)n * (c0 + cc)];
      for (int m0 = 0; m0 < n; m0 += 4) {
        int mw = n - m0 < 4 ? n - m0 : 4;
        for (int t = 0; t < 4 * CB; t++) Cb[t] = 0.0;
        for (int a = 0; a < n; a++) {
          const double *br = Bt + (size_t)a * CB;
          for (int r = 0; r < mw; r++) {
            double av = At[(size_t)(m0 + r) * n + a];
            double *cr = Cb + r * CB;
            for (int cc = 0; cc < CB; cc++) cr[cc] += av * br[cc];
          }
        }
        for (int r = 0; r < mw; r++) {
          int m = m0 + r;
          if (rotate)
            for (int cc = 0; cc < cw; cc++) out[(c0 + cc) + n2 * m] = Cb[r * CB + cc];
          else
            for (int cc = 0; cc < cw; cc++) out[m + (size_t)n * (c0 + cc)] = Cb[r * CB + cc];
        }
      }
    }
    free(Bt);
    free(Cb);
  }
  free(At);
}

/* flat (a,b,c) -> (b,c,a): zxy_xyz (calc.c:615-629) */
static void rotate3(int n, const double *f, double *ft) {
  const size_t n2 = (size_t)n * n;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < n; c++)
    for (int b = 0; b < n; b++)
      for (int a = 0; a < n; a++) ft[b + (size_t)n * c + n2 * a] = f[a + (size_t)n * b + n2 * c];
}

/* ------------------------------------------------------------------ */
/* phi tables: ac_init_phi / ch_init_phi (vars.c:302-359), zxy layout
 * l = j n^2 + i n + k:
 *   AC: H = -LE (epn2 (kkz DDZ_k + kky DDY_j + kkx DDX_i) - KE)
 *   CH: H = LCI (epn2 lam^2 - KCI lam), lam = DDZ_k + DDY_j + DDX_i
 *   |H| > 1e-8: phi1 = (1 - exp(-dt H))/H, phi2 = (1 - phi1/dt)/H
 *   else phi1 = dt, phi2 = dt/2 */
static double h_ac(const orc_ctx *c, int i, int j, int k) {
  const orc_params *p = &c->p;
  double tmp = p->kk[2] * c->DDZ[k] + p->kk[1] * c->DDY[j] + p->kk[0] * c->DDX[i];
  return -p->LE * (tmp * p->epn2 - p->KE);
}
static double h_ch(const orc_ctx *c, int i, int j, int k) {
  const orc_params *p = &c->p;
  double tmp = c->DDZ[k] + c->DDY[j] + c->DDX[i];
  if (orc_mutation == 2) return p->LCI * (-tmp * p->epn2 - p->KCI * tmp);
  return p->LCI * (tmp * tmp * p->epn2 - p->KCI * tmp);
}

static void init_phi(orc_ctx *c) {
  const int n = c->p.n;
  const double dt = c->p.dt;
  for (int t = 0; t < 2; t++) {
    double *phi = c->phi1 + c->n3 * t, *phi2 = c->phi2 + c->n3 * t;
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) {
          size_t l = (size_t)j * n * n + (size_t)i * n + k;
          double H = t == 0 ? h_ac(c, i, j, k) : h_ch(c, i, j, k);
          if (fabs(H) > 1.0e-8) {
            double e = exp(-dt * H);
            phi[l] = (1.0 - e) / H;
            phi2[l] = (1.0 - phi[l] / dt) / H;
          } else {
            phi[l] = dt;
            phi2[l] = dt / 2;
          }
        }
  }
}

/* ac_updateU_new / ch_updateU_new (calc.c:489-535):
 *   uhat <- (1 - phi1 H') uhat + phi1 nhat, H' = 0 where |H| < 1e-8 */
void orc_phi_update(orc_ctx *c, int v, double *uhat, const double *nhat) {
  const int n = c->p.n;
  const int isch = v >= c->p.nac;
  const double *phi = c->phi1 + c->n3 * isch;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++)
      for (int k = 0; k < n; k++) {
        size_t l = (size_t)j * n * n + (size_t)i * n + k;
        double H = isch ? h_ch(c, i, j, k) : h_ac(c, i, j, k);
        if (fabs(H) < 1.0e-8) H = 0.0;
        double tmp = 1.0 - phi[l] * H;
        uhat[l] = tmp * uhat[l] + phi[l] * nhat[l];
      }
}

/* ------------------------------------------------------------------ */
orc_ctx *orc_create(const orc_params *p) {
  if (!p || p->n < 2 * p->g + 2 || p->g < 1 || p->nac < 0 || p->nch < 0 || p->nac + p->nch < 1) return NULL;
  orc_ctx *c = calloc(1, sizeof(orc_ctx));
  c->p = *p;
  c->V = p->nac + p->nch;
  const int n = p->n;
  c->n3 = (size_t)n * n * n;
  size_t nn = (size_t)n * n;
  c->Vm = malloc(sizeof(double) * nn);
  c->Vim = malloc(sizeof(double) * nn);
  c->D = malloc(sizeof(double) * n);
  memcpy(c->Vm, p->V, sizeof(double) * nn);
  memcpy(c->Vim, p->Vi, sizeof(double) * nn);
  memcpy(c->D, p->D, sizeof(double) * n);
  const int neps = el_count(p) > p->nac ? el_count(p) : p->nac;
  c->eps0 = calloc((size_t)9 * (neps > 0 ? neps : 1), sizeof(double));
  c->lam = calloc((size_t)9 * (p->nac > 0 ? p->nac : 1), sizeof(double));
  if (p->eps0) memcpy(c->eps0, p->eps0, sizeof(double) * 9 * neps);
  if (p->lam) memcpy(c->lam, p->lam, sizeof(double) * 9 * p->nac);
  c->p.eps0 = c->eps0;
  c->p.lam = c->lam;
  c->p.V = c->Vm;
  c->p.Vi = c->Vim;
  c->p.D = c->D;
  if (p->mask) {
    c->mask = malloc(sizeof(int) * (size_t)n * n * n);
    memcpy(c->mask, p->mask, sizeof(int) * (size_t)n * n * n);
  }
  c->p.mask = c->mask;
  if (p->varmob && (p->nch != 3 || p->nac < 1)) {
    free(c->Vm); free(c->Vim); free(c->D); free(c->eps0); free(c->lam); free(c->mask); free(c);
    return NULL;
  }
  c->DDX = malloc(sizeof(double) * n);
  c->DDY = malloc(sizeof(double) * n);
  c->DDZ = malloc(sizeof(double) * n);
  for (int l = 0; l < n; l++) {
    c->DDX[l] = c->D[l] / p->hx / p->hx;
    c->DDY[l] = c->D[l] / p->hy / p->hy;
    c->DDZ[l] = c->D[l] / p->hz / p->hz;
  }
  c->field = calloc(c->n3 * c->V, sizeof(double));
  c->N1 = calloc(c->n3 * c->V, sizeof(double));
  c->N2 = calloc(c->n3 * c->V, sizeof(double));
  c->nfel = el_count(p) > 0 ? el_count(p) : 1;
  c->felas = calloc(c->n3 * c->nfel, sizeof(double));
  c->fbuf = calloc(c->n3, sizeof(double));
  c->phi1 = calloc(c->n3 * 2, sizeof(double));
  c->phi2 = calloc(c->n3 * 2, sizeof(double));
  c->t0 = calloc(c->n3, sizeof(double));
  c->t1 = calloc(c->n3, sizeof(double));
  c->t2 = calloc(c->n3, sizeof(double));
  c->t3 = calloc(c->n3, sizeof(double));
  init_phi(c);
  return c;
}

void orc_destroy(orc_ctx *c) {
  if (!c) return;
  free(c->Vm); free(c->Vim); free(c->D); free(c->eps0); free(c->lam);
  free(c->DDX); free(c->DDY); free(c->DDZ);
  free(c->field); free(c->N1); free(c->N2); free(c->felas); free(c->fbuf);
  free(c->phi1); free(c->phi2); free(c->t0); free(c->t1); free(c->t2); free(c->t3);
  free(c->mask);
  free(c);
}

double *orc_field(orc_ctx *c, int v) { return c->field + c->n3 * v; }
double *orc_felas(orc_ctx *c, int p) { return c->felas + c->n3 * p; }

/* S^-1 along x, y, z: PUX, PUY, PUZ(stage 0/1) -> zxy spectral layout */
static void fwd3(orc_ctx *c, const double *in, double *out, double *w1, double *w2) {
  const int n = c->p.n;
  orc_modeprod(n, c->Vim, in, w1, 1);
  orc_modeprod(n, c->Vim, w1, w2, 1);
  orc_modeprod(n, c->Vim, w2, out, 0);
}
/* zxy_xyz then S along x, y, z (stage 2), result in xyz */
static void bwd3(orc_ctx *c, const double *spec_zxy, double *out, double *w1, double *w2) {
  const int n = c->p.n;
  rotate3(n, spec_zxy, w1);
  orc_modeprod(n, c->Vm, w1, w2, 1);
  orc_modeprod(n, c->Vm, w2, w1, 1);
  orc_modeprod(n, c->Vm, w1, out, 1);
}

/* One time step: ScLETD.c:86-189 (check_soln_new is separate: orc_diag). */
void orc_step(orc_ctx *c) {
  const orc_params *p = &c->p;
  const size_t n3 = c->n3;
  double *uh = malloc(sizeof(double) * n3), *nh = malloc(sizeof(double) * n3);
  double *w1 = malloc(sizeof(double) * n3), *w2 = malloc(sizeof(double) * n3);
  orc_transfer(c);
  if (el_count(p) > 0) orc_elastic(c);
  for (int v = 0; v < c->V; v++) orc_nonlinear(c, v, c->N1 + n3 * v);
  for (int v = 0; v < c->V; v++) {
    double *u = c->field + n3 * v;
    fwd3(c, u, uh, w1, w2);
    fwd3(c, c->N1 + n3 * v, nh, w1, w2);
    orc_phi_update(c, v, uh, nh);
    bwd3(c, uh, u, w1, w2);
  }
  if (p->etd2) {
    orc_transfer(c);
    if (el_count(p) > 0) orc_elastic(c);
    for (int v = 0; v < c->V; v++) orc_nonlinear(c, v, c->N2 + n3 * v);
    for (int v = 0; v < c->V; v++) {
      double *dn = c->N2 + n3 * v;
      const double *n1 = c->N1 + n3 * v;
      for (size_t l = 0; l < n3; l++) dn[l] = dn[l] - n1[l];       /* prepare_U1_new */
      fwd3(c, dn, nh, w1, w2);
      const double *ph2 = c->phi2 + n3 * (v >= p->nac);
      for (size_t l = 0; l < n3; l++) nh[l] = ph2[l] * nh[l];      /* prepare_U2_new */
      bwd3(c, nh, uh, w1, w2);
      double *u = c->field + n3 * v;
      for (size_t l = 0; l < n3; l++) u[l] += uh[l];               /* correct_U_new */
    }
  }
  free(uh); free(nh); free(w1); free(w2);
}

/* ------------------------------------------------------------------ */
/* Adaptive stepping support (no counterpart in the reference; restates the
 * GPU driver's step-doubling attempt so its decisions can be checked).  */
void orc_set_dt(orc_ctx *c, double dt) {
  c->p.dt = dt;
  init_phi(c);
}

/* One ETD1 attempt's shared part: N(u_n) and the forward transforms once;
 * ufull = u_n advanced by dt, field = u_n advanced by dt/2 (the first half
 * step).  Leaves the context's dt at dt/2.  Pinned by the identity with
 * orc_step at dt and at dt/2 (tests). */
void orc_step_pair(orc_ctx *c, double dt, double *ufull) {
  const size_t n3 = c->n3;
  double *uh = malloc(sizeof(double) * n3), *nh = malloc(sizeof(double) * n3);
  double *u0h = malloc(sizeof(double) * n3 * c->V), *n0h = malloc(sizeof(double) * n3 * c->V);
  double *w1 = malloc(sizeof(double) * n3), *w2 = malloc(sizeof(double) * n3);
  orc_transfer(c);
  if (el_count(&c->p) > 0) orc_elastic(c);
  for (int v = 0; v < c->V; v++) orc_nonlinear(c, v, c->N1 + n3 * v);
  for (int v = 0; v < c->V; v++) {
    fwd3(c, c->field + n3 * v, u0h + n3 * v, w1, w2);
    fwd3(c, c->N1 + n3 * v, n0h + n3 * v, w1, w2);
  }
  const double dts[2] = {dt, 0.5 * dt};
  for (int s = 0; s < 2; s++) {
    orc_set_dt(c, dts[s]);
    for (int v = 0; v < c->V; v++) {
      memcpy(uh, u0h + n3 * v, sizeof(double) * n3);
      memcpy(nh, n0h + n3 * v, sizeof(double) * n3);
      orc_phi_update(c, v, uh, nh);
      bwd3(c, uh, s == 0 ? ufull + n3 * v : c->field + n3 * v, w1, w2);
    }
  }
  free(uh); free(nh); free(u0h); free(n0h); free(w1); free(w2);
}

/* ------------------------------------------------------------------ */
/* check_soln_new (field.c:407-678).  Per AC field: max, min, vol over the
 * interior [g, n-g)^3; eng over cells [g, n-g-1)^3 with the 8-corner mean
 * and 4-edge-averaged gradient, summed over all AC fields m:
 *   epn2/2 grad.(Lambda^T grad) + a2/2 sum u^2 - a3/3 sum u^3 + a4/4 (sum u^2)^2
 * (the reference hard-codes 4 variants and a2,a3,a4 = 2376/2372, ...).
 * The same eng is reported for every AC field.  CH: epn2/2 |grad c|^2 +
 * (c^2-1)^2/4.  Couple (nac>0 and nch>0): continues the last CH accumulator
 * (not reset, field.c:597-644) with gradient terms of AC fields 0..nch-1 and
 * CH fields, then scales by h^3 again.  Partial sums per plane make the
 * result independent of the thread count. */
static void cell_grad(const double *u, int n, int k, int j, int i, double hx, double hy, double hz, double *m, double *ux, double *uy, double *uz) {
  double f[8];
  f[0] = u[IDX(n, k, j, i)];
  f[1] = u[IDX(n, k, j, i + 1)];
  f[2] = u[IDX(n, k, j + 1, i)];
  f[3] = u[IDX(n, k, j + 1, i + 1)];
  f[4] = u[IDX(n, k + 1, j, i)];
  f[5] = u[IDX(n, k + 1, j, i + 1)];
  f[6] = u[IDX(n, k + 1, j + 1, i)];
  f[7] = u[IDX(n, k + 1, j + 1, i + 1)];
  *m = (f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7]) / 8;
  *ux = ((f[1] - f[0]) + (f[3] - f[2]) + (f[5] - f[4]) + (f[7] - f[6])) / 4 / hx;
  *uy = ((f[2] - f[0]) + (f[3] - f[1]) + (f[6] - f[4]) + (f[7] - f[5])) / 4 / hy;
  *uz = ((f[4] - f[0]) + (f[5] - f[1]) + (f[6] - f[2]) + (f[7] - f[3])) / 4 / hz;
}

void orc_diag(orc_ctx *c, double *stats) {
  const orc_params *p = &c->p;
  const int n = p->n, g = p->g, nac = p->nac, nch = p->nch;
  /* the loops run over [ix2, ix3) (field.c:457-527): [g, n - g) with neighbours on
   * both sides, [0, n - g) when the low side is physical (init_para, vars.c:58-157:
   * left < 0 gives ix2 = 0, ix3 = nx - nghost) */
  const int lo = p->physical ? 0 : g;
  const double hx = p->hx, hy = p->hy, hz = p->hz, epn2 = p->epn2;
  const int K = n - 2 * g;
  double *plane = calloc((size_t)n, sizeof(double));
  for (int v = 0; v < c->V; v++) {
    const double *u = c->field + c->n3 * v;
    double mx = -1.0e30, mn = 1.0e30, vol = 0.0;
    for (int k = lo; k < n - g; k++)
      for (int j = lo; j < n - g; j++)
        for (int i = lo; i < n - g; i++) {
          double t = u[IDX(n, k, j, i)];
          if (t > mx) mx = t;
          if (t < mn) mn = t;
        }
#pragma omp parallel for schedule(static)
    for (int k = lo; k < n - g; k++) {
      double s = 0.0;
      for (int j = lo; j < n - g; j++)
        for (int i = lo; i < n - g; i++) s += u[IDX(n, k, j, i)] + 1.0;
      plane[k] = s;
    }
    for (int k = lo; k < n - g; k++) vol += plane[k];
    vol *= hx * hy * hz;
    stats[4 * v + 0] = mx;
    stats[4 * v + 1] = mn;
    stats[4 * v + 2] = vol;
  }
  (void)K;
  /* AC energy (common to all AC fields) */
  if (nac > 0) {
#pragma omp parallel for schedule(static)
    for (int k = lo; k < n - g - 1; k++) {
      double e = 0.0;
      double *um = malloc(sizeof(double) * nac);
      for (int j = lo; j < n - g - 1; j++)
        for (int i = lo; i < n - g - 1; i++) {
          for (int m = 0; m < nac; m++) {
            double ux, uy, uz;
            cell_grad(c->field + c->n3 * m, n, k, j, i, hx, hy, hz, &um[m], &ux, &uy, &uz);
            const double *L = c->lam + 9 * m;
            double ffx = ux * L[0] + uy * L[3] + uz * L[6];
            double ffy = ux * L[1] + uy * L[4] + uz * L[7];
            double ffz = ux * L[2] + uy * L[5] + uz * L[8];
            e += epn2 * (ffx * ux + ffy * uy + ffz * uz) / 2;
          }
          double s2 = 0.0, s3 = 0.0;
          for (int m = 0; m < nac; m++) { s2 += um[m] * um[m]; s3 += um[m] * um[m] * um[m]; }
          e += (p->a2 / 2.0) * s2 - (p->a3 / 3.0) * s3 + (p->a4 / 4.0) * s2 * s2;
        }
      free(um);
      plane[k] = e;
    }
    double eng = 0.0;
    for (int k = lo; k < n - g - 1; k++) eng += plane[k];
    eng = eng * hx * hy * hz;
    for (int v = 0; v < nac; v++) stats[4 * v + 3] = eng;
  }
  double engtmp = 0.0;
  for (int q = 0; q < nch; q++) {
    const double *u = c->field + c->n3 * (nac + q);
#pragma omp parallel for schedule(static)
    for (int k = lo; k < n - g - 1; k++) {
      double e = 0.0;
      for (int j = lo; j < n - g - 1; j++)
        for (int i = lo; i < n - g - 1; i++) {
          double cm, ux, uy, uz;
          cell_grad(u, n, k, j, i, hx, hy, hz, &cm, &ux, &uy, &uz);
          e += epn2 * (ux * ux + uy * uy + uz * uz) / 2;
          e += ((cm * cm - 1) * (cm * cm - 1) / 4);
        }
      plane[k] = e;
    }
    engtmp = 0.0;
    for (int k = lo; k < n - g - 1; k++) engtmp += plane[k];
    engtmp = engtmp * hx * hy * hz;
    stats[4 * (nac + q) + 3] = engtmp;
  }
  double couple = 0.0;
  if (nac > 0 && nch > 0) {
#pragma omp parallel for schedule(static)
    for (int k = lo; k < n - g - 1; k++) {
      double e = 0.0;
      for (int j = lo; j < n - g - 1; j++)
        for (int i = lo; i < n - g - 1; i++) {
          for (int q = 0; q < nch && q < nac; q++) {
            double cm, ux, uy, uz;
            cell_grad(c->field + c->n3 * q, n, k, j, i, hx, hy, hz, &cm, &ux, &uy, &uz);
            e += epn2 * (ux * ux + uy * uy + uz * uz) / 2;
          }
          for (int q = 0; q < nch; q++) {
            double cm, ux, uy, uz;
            cell_grad(c->field + c->n3 * (nac + q), n, k, j, i, hx, hy, hz, &cm, &ux, &uy, &uz);
            e += epn2 * (ux * ux + uy * uy + uz * uz) / 2;
          }
        }
      plane[k] = e;
    }
    couple = engtmp;
    for (int k = lo; k < n - g - 1; k++) couple += plane[k];
    couple = couple * hx * hy * hz;
  }
  stats[4 * c->V] = couple;
  free(plane);
}

/* ------------------------------------------------------------------ */
/* Decomposed runs: px x py x pz periodic boxes (omega's procs line with
 * periodic = 1), every box with the same extended edge n and the same
 * boundary id 3 eigenbasis (each side has a neighbour, vars.c:488-582).
 * h = L / (p N) comes from init_para (vars.c:58-157) through the caller's
 * hx, hy, hz. */

orc_group *orc_group_create(orc_ctx **boxes, int px, int py, int pz) {
  if (!boxes || px < 1 || py < 1 || pz < 1) return NULL;
  const int P = px * py * pz;
  const orc_params *p0 = &boxes[0]->p;
  for (int b = 0; b < P; b++) {
    const orc_params *p = &boxes[b]->p;
    if (p->n != p0->n || p->g != p0->g || p->nac != p0->nac || p->nch != p0->nch || p->physical || p->varmob)
      return NULL;
  }
  orc_group *G = calloc(1, sizeof(orc_group));
  G->px = px;
  G->py = py;
  G->pz = pz;
  G->box = malloc(sizeof(orc_ctx *) * P);
  for (int b = 0; b < P; b++) {
    G->box[b] = boxes[b];
    boxes[b]->grp = G;
    boxes[b]->bx = b % px;
    boxes[b]->by = (b / px) % py;
    boxes[b]->bz = b / (px * py);
  }
  return G;
}

void orc_group_destroy(orc_group *G) {
  if (!G) return;
  for (int b = 0; b < G->px * G->py * G->pz; b++) G->box[b]->grp = NULL;
  free(G->box);
  free(G);
}

/* transfer() across the boxes (trans.c:1939-1996 with the slab packing of
 * 7-251): every send reads the sender's state before any unpack, and unpack
 * writes x ghosts, then y ghosts over every plane, then z ghosts, so the last
 * writer of an edge or corner cell is the z (then y) exchange, which carries
 * whole planes (rows) of the neighbour including its own ghost cells.  That
 * is the single-box sweep of orc_ghost_fill with the source box being the
 * neighbour in the sweep's direction; running each sweep over all boxes
 * before the next keeps every read on pre-transfer data (a z sweep reads
 * interior planes, a y sweep interior rows whose x ghosts only the x sweep
 * writes). */
static void group_fill(orc_group *G, double *(*arr)(orc_ctx *, int), int v) {
  const int P = G->px * G->py * G->pz;
  const int n = G->box[0]->p.n, g = G->box[0]->p.g;
  for (int b = 0; b < P; b++) {
    orc_ctx *c = G->box[b];
    double *u = arr(c, v);
    for (int k = 0; k < n; k++) {
      if (!(k < g || k >= n - g)) continue;
      const double *src = arr((orc_ctx *)nb_box(c, 0, 0, box_shift(k, n, g)), v);
      memcpy(u + IDX(n, k, 0, 0), src + IDX(n, wrap_e(k, n, g), 0, 0), sizeof(double) * (size_t)n * n);
    }
  }
  for (int b = 0; b < P; b++) {
    orc_ctx *c = G->box[b];
    double *u = arr(c, v);
    for (int k = g; k < n - g; k++)
      for (int j = 0; j < n; j++) {
        if (!(j < g || j >= n - g)) continue;
        const double *src = arr((orc_ctx *)nb_box(c, 0, box_shift(j, n, g), 0), v);
        memcpy(u + IDX(n, k, j, 0), src + IDX(n, k, wrap_e(j, n, g), 0), sizeof(double) * (size_t)n);
      }
  }
  for (int b = 0; b < P; b++) {
    orc_ctx *c = G->box[b];
    double *u = arr(c, v);
    const double *lo = arr((orc_ctx *)nb_box(c, -1, 0, 0), v), *hi = arr((orc_ctx *)nb_box(c, 1, 0, 0), v);
    for (int k = g; k < n - g; k++)
      for (int j = g; j < n - g; j++)
        for (int i = 0; i < n; i++) {
          if (i < g) u[IDX(n, k, j, i)] = lo[IDX(n, k, j, wrap_e(i, n, g))];
          else if (i >= n - g) u[IDX(n, k, j, i)] = hi[IDX(n, k, j, wrap_e(i, n, g))];
        }
  }
}
static double *arr_field(orc_ctx *c, int v) { return c->field + c->n3 * v; }
static double *arr_fbuf(orc_ctx *c, int v) { (void)v; return c->fbuf; }

void orc_group_transfer(orc_group *G) {
  const int V = G->box[0]->V;
  for (int v = 0; v < V; v++) group_fill(G, arr_field, v);
}

/* elastic_calculate (elastic.c:689-762) on the global interior of
 * (px N)^3 cells, the distributed FFT of fftMPI (fft3d.c:98-138) taken as
 * one transform: B_pq at the global wavenumbers, felas_p written into each
 * box's interior of its persistent transfer buffer f, which the exchange of
 * elastic_calculate_ElasDri (its own persistent requests on f, 669-686)
 * then fills as transfer() fills a field; felas_p = f. */
void orc_group_elastic(orc_group *G) {
  orc_ctx *c0 = G->box[0];
  const orc_params *p = &c0->p;
  const int n = p->n, g = p->g, N = n - 2 * g, nel = el_count(p), v0 = el_first(p);
  if (nel == 0 || nel > 64 || G->px != G->py || G->px != G->pz) return;
  const int P = G->px * G->py * G->pz, Ng = G->px * N;
  const size_t N3 = (size_t)Ng * Ng * Ng;
  double *spec = malloc(sizeof(double) * 2 * N3 * nel);
  double *acc = malloc(sizeof(double) * 2 * N3);
  bfast *bf = malloc(sizeof(bfast));
  bfast_init(p, bf);
  for (int q = 0; q < nel; q++) {
    double *s = spec + 2 * N3 * q;
    for (int b = 0; b < P; b++) {
      const orc_ctx *c = G->box[b];
      const double *u = c->field + c->n3 * (v0 + q);
      for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
          for (int x = 0; x < N; x++) {
            size_t m = (((size_t)(c->bz * N + z)) * Ng + (c->by * N + y)) * Ng + (c->bx * N + x);
            s[2 * m] = u[IDX(n, z + g, y + g, x + g)];
            s[2 * m + 1] = 0.0;
          }
    }
    orc_fft3d(Ng, s, +1);
  }
  for (int pp = 0; pp < nel; pp++) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < Ng; z++) {
      double row[64];
      for (int y = 0; y < Ng; y++)
        for (int x = 0; x < Ng; x++) {
          size_t m = ((size_t)z * Ng + y) * Ng + x;
          bfast_row(p, bf, Ng, x, y, z, pp, row);
          double re = 0.0, im = 0.0;
          for (int qq = 0; qq < nel; qq++) {
            re += spec[2 * N3 * qq + 2 * m] * row[qq];
            im += spec[2 * N3 * qq + 2 * m + 1] * row[qq];
          }
          acc[2 * m] = re;
          acc[2 * m + 1] = im;
        }
    }
    orc_fft3d(Ng, acc, -1);
    for (int b = 0; b < P; b++) {
      orc_ctx *c = G->box[b];
      for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
          for (int x = 0; x < N; x++) {
            size_t m = (((size_t)(c->bz * N + z)) * Ng + (c->by * N + y)) * Ng + (c->bx * N + x);
            c->fbuf[IDX(n, z + g, y + g, x + g)] = acc[2 * m];
          }
    }
    group_fill(G, arr_fbuf, 0);
    for (int b = 0; b < P; b++) memcpy(G->box[b]->felas + G->box[b]->n3 * pp, G->box[b]->fbuf, sizeof(double) * c0->n3);
  }
  free(spec);
  free(acc);
  free(bf);
}

/* One time step of every box: ScLETD.c:86-189 with transfer / elastic across
 * the boxes; each box's nonlinear terms read its neighbours (outside values)
 * before any box is advanced. */
void orc_group_step(orc_group *G) {
  const int P = G->px * G->py * G->pz;
  orc_ctx *c0 = G->box[0];
  const orc_params *p = &c0->p;
  const size_t n3 = c0->n3;
  const int V = c0->V;
  double *uh = malloc(sizeof(double) * n3), *nh = malloc(sizeof(double) * n3);
  double *w1 = malloc(sizeof(double) * n3), *w2 = malloc(sizeof(double) * n3);
  orc_group_transfer(G);
  if (el_count(p) > 0) orc_group_elastic(G);
  for (int b = 0; b < P; b++)
    for (int v = 0; v < V; v++) orc_nonlinear(G->box[b], v, G->box[b]->N1 + n3 * v);
  for (int b = 0; b < P; b++) {
    orc_ctx *c = G->box[b];
    for (int v = 0; v < V; v++) {
      double *u = c->field + n3 * v;
      fwd3(c, u, uh, w1, w2);
      fwd3(c, c->N1 + n3 * v, nh, w1, w2);
      orc_phi_update(c, v, uh, nh);
      bwd3(c, uh, u, w1, w2);
    }
  }
  if (p->etd2) {
    orc_group_transfer(G);
    if (el_count(p) > 0) orc_group_elastic(G);
    for (int b = 0; b < P; b++)
      for (int v = 0; v < V; v++) orc_nonlinear(G->box[b], v, G->box[b]->N2 + n3 * v);
    for (int b = 0; b < P; b++) {
      orc_ctx *c = G->box[b];
      for (int v = 0; v < V; v++) {
        double *dn = c->N2 + n3 * v;
        const double *n1 = c->N1 + n3 * v;
        for (size_t l = 0; l < n3; l++) dn[l] = dn[l] - n1[l];
        fwd3(c, dn, nh, w1, w2);
        const double *ph2 = c->phi2 + n3 * (v >= p->nac);
        for (size_t l = 0; l < n3; l++) nh[l] = ph2[l] * nh[l];
        bwd3(c, nh, uh, w1, w2);
        double *u = c->field + n3 * v;
        for (size_t l = 0; l < n3; l++) u[l] += uh[l];
      }
    }
  }
  free(uh); free(nh); free(w1); free(w2);
}

/* check_soln_new over the boxes: each box's sums over its own ranges, then
 * MPI_Reduce (field.c:529-532): max, min, and the sums of vol and eng in box
 * order. */
void orc_group_diag(orc_group *G, double *stats) {
  const int P = G->px * G->py * G->pz, V = G->box[0]->V;
  double *s = malloc(sizeof(double) * (4 * V + 1));
  for (int b = 0; b < P; b++) {
    orc_diag(G->box[b], s);
    if (b == 0) {
      memcpy(stats, s, sizeof(double) * (4 * V + 1));
      continue;
    }
    for (int v = 0; v < V; v++) {
      if (s[4 * v] > stats[4 * v]) stats[4 * v] = s[4 * v];
      if (s[4 * v + 1] < stats[4 * v + 1]) stats[4 * v + 1] = s[4 * v + 1];
      stats[4 * v + 2] += s[4 * v + 2];
      stats[4 * v + 3] += s[4 * v + 3];
    }
    stats[4 * V] += s[4 * V];
  }
  free(s);
}
