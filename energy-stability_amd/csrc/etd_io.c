/*
 * etd_io.c -- host-side C: the reference's input formats, eigenbasis files,
 * closed-form eigenbasis and the CSV / field outputs.
 */
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "etd.h"

void etd_default_params(etd_params *p) {
  memset(p, 0, sizeof(*p));
  p->px = p->py = p->pz = 1;
  p->nghost = 2;
  strcpy(p->work_dir, "./");
  strcpy(p->data_dir, "./");
  p->periodic = 1;
  p->kkx = p->kky = p->kkz = 1.0;
  /* source_beta_to_omega/src/formula.c:9-18 (LE = 1, KE = 42) and the
   * Landau constants of formula.c:126-128 */
  p->LE = 1.0;
  p->KE = 42.0;
  p->a2 = 2376.0 / 2372.0;
  p->a3 = (7128.0 + 12.0 * 2372.0) / 2372.0;
  p->a4 = (4752.0 + 12.0 * 2372.0) / 2372.0;
  /* CH constants are never set by the reference (ch[n].LCI/KCI); L = 1 and
   * a stabiliser >= max f''(c)/... of the double well c^3 - c */
  p->LCI = 1.0;
  p->KCI = 2.0;
  p->elastic_scale = 1.0;
  p->precision = ETD_PREC_FP64;
}

/* CALPHAD Gibbs energies of bcc and hcp Ti-Al-V at the beta->alpha driver's
 * fixed compositions (anisotropic_hip.c:2029-2048): lattice stabilities and
 * Redlich-Kister parameters of init_vars (vars.c:146-190) at T = 1093 K,
 * R = 8.31451 (ScLETD.h:24-26). */
void etd_calphad(double *gbcc, double *ghcp) {
  const double T = 1093.0, R = 8.31451, lT = log(T);
  const double ser_al = -11277.683 + 188.661987 * T - 31.748192 * T * lT - 1.234e28 * pow(T, -9) + 137529.83;
  const double ser_ti = 908.837 + 67.048538 * T - 14.9466 * T * lT - 0.0081465 * T * T + 2.02715e-7 * T * T * T -
                        1477660.0 / T + 58166.677;
  const double ser_v = -7967.84 + 143.291 * T - 25.9 * T * lT + 6.25e-5 * T * T - 6.8e-7 * T * T * T + 77277.217;
  const double g_bcc[3] = {/* Al, V, Ti */ 10083.0 - 4.812 * T + ser_al, ser_v,
                           5758.6 + 38.38987 * T - 7.4305 * T * lT + 0.0093636 * T * T - 1.04805e-6 * T * T * T -
                               525093.0 / T + ser_ti};
  const double g_hcp[3] = {5481.0 - 1.799 * T + ser_al, 4000.0 + 2.4 * T + ser_v, ser_ti};
  /* x = (Al, V, Ti); binary terms x_a x_b sum_k L_k (x_a - x_b)^k for (Al,V), (Ti,V), (Al,Ti) */
  const double xb[3] = {0.102, 0.036, 1.0 - 0.102 - 0.036}, xh[3] = {0.1057, 0.0223, 1.0 - 0.1057 - 0.0223};
  const double Lb[3][3] = {{-98000.0 + 32.0 * T, 3500.0 - 5.0 * T, 0.0},
                           {10500.0 - 1.5 * T, 2000.0, 1000.0},
                           {-128500.0 + 39.0 * T, 6000.0, 21200.0}};
  const double Lh[3][3] = {{-98000.0 + 32.0 * T, 3500.0 - 5.0 * T, 0.0},
                           {30250.0 - 10.0 * T, 0.0, 0.0},
                           {-133750.0 + 39.0 * T, 250.0, 17250.0}};
  const int pa[3] = {0, 2, 0}, pb[3] = {1, 1, 2};
  for (int ph = 0; ph < 2; ph++) {
    const double *x = ph ? xh : xb, *G = ph ? g_hcp : g_bcc;
    const double(*L)[3] = ph ? Lh : Lb;
    double s = x[2] * G[2] + x[0] * G[0] + x[1] * G[1];
    s += R * T * (x[2] * log(x[2]) + x[0] * log(x[0]) + x[1] * log(x[1]));
    for (int q = 0; q < 3; q++) {
      const double d = x[pa[q]] - x[pb[q]];
      s += x[pa[q]] * x[pb[q]] * (L[q][0] + L[q][1] * d + L[q][2] * d * d);
    }
    /* ternary: bcc x_Al x_V x_Ti (-8000); hcp x_Al x_V x_Ti (x_Al L0 + x_Ti L1 + x_V L2) */
    if (ph == 0)
      s += x[0] * x[1] * x[2] * -8000.0;
    else
      s += x[0] * x[1] * x[2] * (x[0] * -60000.0 + x[2] * -30000.0 + x[1] * 60000.0);
    if (ph)
      *ghcp = s;
    else
      *gbcc = s;
  }
}

/* Mobility constants of update_M_C (mobility.c:864-906) for Cu, Mn, Ni:
 * x (1 - x), and D_A, D_G of D = D0 exp(-Q / RT) relative to D_Cu in the A
 * phase, T = 823 K. */
void etd_vm_constants(double X[3], double DA[3], double DG[3]) {
  const double RT = 8.314472 * 823.0;
  const double x0[3] = {0.15, 0.01, 0.01};
  const double dcu = 4.7e-5 * exp(-2.44e5 / RT);
  DA[0] = 1.0;
  DG[0] = (4.3e-5 * exp(-2.80e5 / RT)) / dcu;
  DA[1] = (1.49e-4 * exp(-2.63e5 / RT)) / dcu;
  DG[1] = (2.78e-5 * exp(-2.64e5 / RT)) / dcu;
  DA[2] = (1.4e-4 * exp(-2.56e5 / RT)) / dcu;
  DG[2] = (1.08e-5 * exp(-2.73e5 / RT)) / dcu;
  for (int q = 0; q < 3; q++) X[q] = x0[q] * (1.0 - x0[q]);
}

static void set_lambda_iso(etd_params *p) {
  for (int v = 0; v < ETD_MAX_VAR; v++) {
    double *L = p->lam + 9 * v;
    memset(L, 0, 9 * sizeof(double));
    L[0] = p->kkx;
    L[4] = p->kky;
    L[8] = p->kkz;
  }
}

static void derive(etd_params *p) {
  /* init_para (vars.c:58-157), periodic single subdomain:
   * gnx = lnx*px - 2*nghost*px */
  int gnx = p->nx * p->px - 2 * p->nghost * p->px;
  int gny = p->ny * p->py - 2 * p->nghost * p->py;
  int gnz = p->nz * p->pz - 2 * p->nghost * p->pz;
  if (!p->periodic) {
    gnx = p->nx * p->px - 2 * p->nghost * (p->px - 1);
    gny = p->ny * p->py - 2 * p->nghost * (p->py - 1);
    gnz = p->nz * p->pz - 2 * p->nghost * (p->pz - 1);
  }
  p->hx = (p->xmax - p->xmin) / gnx;
  p->hy = (p->ymax - p->ymin) / gny;
  p->hz = (p->zmax - p->zmin) / gnz;
}

/* ScLETD.c:16-35, positional; fscanf skips the CR of CRLF files as space */
int etd_parse_ini(const char *path, etd_params *p) {
  if (!path || !p) return ETD_ERR_ARG;
  FILE *f = fopen(path, "r");
  if (!f) return ETD_ERR_IO;
  int ok = 1;
  ok &= fscanf(f, "%d,%d,%d", &p->nx, &p->ny, &p->nz) == 3;
  ok &= fscanf(f, "%d,%d,%d", &p->px, &p->py, &p->pz) == 3;
  ok &= fscanf(f, "%d", &p->nghost) == 1;
  ok &= fscanf(f, "%1023s", p->work_dir) == 1;
  ok &= fscanf(f, "%1023s", p->data_dir) == 1;
  ok &= fscanf(f, "%lf,%lf", &p->dt, &p->t_total) == 2;
  ok &= fscanf(f, "%lf", &p->epn2) == 1;
  ok &= fscanf(f, "%d", &p->periodic) == 1;
  ok &= fscanf(f, "%d, %d", &p->nac, &p->nch) == 2;
  ok &= fscanf(f, "%d, %d, %1023s", &p->restart, &p->restart_iter, p->restart_dir) == 3;
  ok &= fscanf(f, "%lf,%lf,%lf", &p->xmin, &p->ymin, &p->zmin) == 3;
  ok &= fscanf(f, "%lf,%lf,%lf", &p->xmax, &p->ymax, &p->zmax) == 3;
  ok &= fscanf(f, "%d", &p->etd2) == 1;
  ok &= fscanf(f, "%d", &p->elastic) == 1;
  ok &= fscanf(f, "%d", &p->anisotropic) == 1;
  ok &= fscanf(f, "%lf,%lf,%lf", &p->kkx, &p->kky, &p->kkz) == 3;
  ok &= fscanf(f, "%d", &p->variable_mobility) == 1;
  fclose(f);
  if (!ok) return ETD_ERR_IO;
  derive(p);
  set_lambda_iso(p);
  return ETD_OK;
}

/* source_beta_to_alpha/src/ScLETD.cpp:15-31 order; epn2 = 0.048 and ETD1
 * as set in source_beta_to_alpha/src/vars.c */
int etd_parse_ini_alpha(const char *path, etd_params *p) {
  if (!path || !p) return ETD_ERR_ARG;
  FILE *f = fopen(path, "r");
  if (!f) return ETD_ERR_IO;
  int ok = 1;
  ok &= fscanf(f, "%d,%d,%d", &p->nx, &p->ny, &p->nz) == 3;
  ok &= fscanf(f, "%d,%d,%d", &p->px, &p->py, &p->pz) == 3;
  ok &= fscanf(f, "%d", &p->nghost) == 1;
  ok &= fscanf(f, "%1023s", p->work_dir) == 1;
  ok &= fscanf(f, "%1023s", p->data_dir) == 1;
  ok &= fscanf(f, "%lf,%lf", &p->dt, &p->t_total) == 2;
  ok &= fscanf(f, "%d", &p->restart) == 1;
  ok &= fscanf(f, "%d", &p->periodic) == 1;
  ok &= fscanf(f, "%d, %d", &p->nac, &p->nch) == 2;
  ok &= fscanf(f, "%d", &p->elastic) == 1;
  ok &= fscanf(f, "%d", &p->anisotropic) == 1;
  ok &= fscanf(f, "%lf,%lf,%lf", &p->xmin, &p->ymin, &p->zmin) == 3;
  ok &= fscanf(f, "%lf,%lf,%lf", &p->xmax, &p->ymax, &p->zmax) == 3;
  ok &= fscanf(f, "%lf,%lf,%lf", &p->kkx, &p->kky, &p->kkz) == 3;
  ok &= fscanf(f, "%d,%d,%d", &p->checkpoint, &p->nchk, &p->chk) == 3;
  ok &= fscanf(f, "%d", &p->nout) == 1;
  ok &= fscanf(f, "%d", &p->approx) == 1;
  fclose(f);
  if (!ok) return ETD_ERR_IO;
  /* the beta->alpha driver's physics: ac[n].LE = 2, KE = 3, epn2 = 0.048
   * (vars.c:114-139), the CALPHAD driving force (anisotropic_hip.c:2002-2188),
   * ETD1; its ANISOTROPIC = 1 is the compact Lambda_ij d_i d_j stencil
   * (anisotropic_calc_dev_2, launched at anisotropic_hip.c:4090) */
  p->epn2 = 0.048;
  p->LE = 2.0;
  p->KE = 3.0;
  p->free_energy = 1;
  if (p->anisotropic == 1) p->anisotropic = 2;
  p->etd2 = 0;
  derive(p);
  set_lambda_iso(p);
  return ETD_OK;
}

static int voigt_ok(const etd_params *p) { return p->nac >= 0 && p->nac <= ETD_MAX_VAR; }

/* elastic.c:375-418 */
int etd_parse_elastic(const char *path, etd_params *p) {
  if (!path || !p || !voigt_ok(p)) return ETD_ERR_ARG;
  FILE *f = fopen(path, "r");
  if (!f) return ETD_ERR_IO;
  int ok = fscanf(f, "%lf", &p->elastic_scale) == 1;
  for (int i = 0; i < 36 && ok; i++) ok &= fscanf(f, "%lf", &p->C[i]) == 1;
  /* one eigenstrain per variant the force acts on: the AC fields, or the CH fields for ELASTIC = 2 */
  const int nvar = p->elastic == 2 ? p->nch : p->nac;
  if (nvar > ETD_MAX_VAR) ok = 0;
  for (int v = 0; v < nvar && ok; v++)
    for (int t = 0; t < 9 && ok; t++) ok &= fscanf(f, "%lf", &p->eps0[9 * v + t]) == 1;
  fclose(f);
  return ok ? ETD_OK : ETD_ERR_IO;
}

/* anisotropic.c:645-673: lambda read per variant, then diag(kk) - lambda */
int etd_parse_anisotropic(const char *path, etd_params *p) {
  if (!path || !p || !voigt_ok(p)) return ETD_ERR_ARG;
  FILE *f = fopen(path, "r");
  if (!f) return ETD_ERR_IO;
  int ok = 1;
  double D[9] = {p->kkx, 0, 0, 0, p->kky, 0, 0, 0, p->kkz};
  for (int v = 0; v < p->nac && ok; v++)
    for (int t = 0; t < 9 && ok; t++) {
      double x;
      ok &= fscanf(f, "%lf", &x) == 1;
      p->lam[9 * v + t] = D[t] - x;
    }
  fclose(f);
  return ok ? ETD_OK : ETD_ERR_IO;
}

int etd_parse_landau(const char *path, etd_params *p) {
  if (!path || !p) return ETD_ERR_ARG;
  FILE *f = fopen(path, "r");
  if (!f) return ETD_ERR_IO;
  int ok = fscanf(f, "%lf %lf %lf %lf %lf", &p->LE, &p->KE, &p->a2, &p->a3, &p->a4) == 5;
  double l, k;
  if (ok && fscanf(f, "%lf %lf", &l, &k) == 2) {
    p->LCI = l;
    p->KCI = k;
  }
  fclose(f);
  return ok ? ETD_OK : ETD_ERR_IO;
}

/* vars.c:370-476: file element (row j, col i) -> A[i*n + j] i.e. A(j,i)
 * column-major */
static int read_vec(const char *fn, int cnt, double *out) {
  FILE *f = fopen(fn, "r");
  if (!f) return ETD_ERR_IO;
  for (int l = 0; l < cnt; l++)
    if (fscanf(f, "%lf", out + l) != 1) { fclose(f); return ETD_ERR_IO; }
  fclose(f);
  return ETD_OK;
}

int etd_read_matrices(const char *dir, int n, int id, double *V, double *Vi, double *D) {
  if (!dir || n < 1 || id < 0 || id > 3 || !V || !Vi || !D) return ETD_ERR_ARG;
  char fn[2048];
  double *tmp = malloc(sizeof(double) * (size_t)n * n);
  if (!tmp) return ETD_ERR_NOMEM;
  int rc;
  snprintf(fn, sizeof fn, "%s/d%d%d.dat", dir, id, n);
  if ((rc = read_vec(fn, n, D)) != ETD_OK) goto out;
  snprintf(fn, sizeof fn, "%s/v%d%d.dat", dir, id, n);
  if ((rc = read_vec(fn, n * n, tmp)) != ETD_OK) goto out;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) V[(size_t)i * n + j] = tmp[(size_t)j * n + i];
  snprintf(fn, sizeof fn, "%s/vi%d%d.dat", dir, id, n);
  if ((rc = read_vec(fn, n * n, tmp)) != ETD_OK) goto out;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) Vi[(size_t)i * n + j] = tmp[(size_t)j * n + i];
out:
  free(tmp);
  return rc;
}

/* Closed forms of the four truncated operators (what the .dat files hold):
 *   id 3  tridiag(1,-2,1):          -4 sin^2(m pi/(2(n+1))), sin(i m pi/(n+1))
 *   id 0  reflecting both ends:      -4 sin^2(m pi/(2(n-1))), cos(i m pi/(n-1))
 *   id 1  open low / reflecting high: -4 sin^2((2m+1) pi/(4n)), sin((i+1)(2m+1) pi/(2n))
 *   id 2  reflecting low / open high: same eigenvalues, cos(i (2m+1) pi/(2n))
 * Integer angle reduction keeps sin/cos arguments below 2 pi.  V^-1 by
 * Gauss-Jordan with partial pivoting (V^T for the orthonormal id 3). */
static int invert_cm(int n, const double *A, double *Ai) {
  size_t w = 2 * (size_t)n;
  double *M = malloc(sizeof(double) * n * w);
  if (!M) return ETD_ERR_NOMEM;
  for (int r = 0; r < n; r++)
    for (size_t c = 0; c < w; c++) M[r * w + c] = c < (size_t)n ? A[r + (size_t)n * c] : ((int)(c - n) == r ? 1.0 : 0.0);
  for (int col = 0; col < n; col++) {
    int piv = col;
    for (int r = col + 1; r < n; r++)
      if (fabs(M[r * w + col]) > fabs(M[piv * w + col])) piv = r;
    if (M[piv * w + col] == 0.0) { free(M); return ETD_ERR_ARG; }
    if (piv != col)
      for (size_t c = 0; c < w; c++) { double t = M[col * w + c]; M[col * w + c] = M[piv * w + c]; M[piv * w + c] = t; }
    double d = M[col * w + col];
    for (size_t c = 0; c < w; c++) M[col * w + c] /= d;
    for (int r = 0; r < n; r++) {
      if (r == col) continue;
      double f = M[r * w + col];
      if (f != 0.0)
        for (size_t c = 0; c < w; c++) M[r * w + c] -= f * M[col * w + c];
    }
  }
  for (int r = 0; r < n; r++)
    for (int c = 0; c < n; c++) Ai[r + (size_t)n * c] = M[r * w + n + c];
  free(M);
  return ETD_OK;
}

int etd_eigbasis(int n, int id, double *V, double *Vi, double *D) {
  if (n < 3 || id < 0 || id > 3 || !V || !Vi || !D) return ETD_ERR_ARG;
  const double pi = 3.14159265358979323846;
  for (int m = 0; m < n; m++) {
    if (id == 3) {
      long mm = n - m;
      D[m] = -4.0 * pow(sin(mm * pi / (2.0 * (n + 1))), 2);
      for (int i = 0; i < n; i++)
        V[i + (size_t)n * m] = sqrt(2.0 / (n + 1)) * sin((double)(((i + 1) * mm) % (2L * (n + 1))) * pi / (n + 1));
    } else if (id == 0) {
      long mm = n - 1 - m;
      D[m] = -4.0 * pow(sin(mm * pi / (2.0 * (n - 1))), 2);
      for (int i = 0; i < n; i++) V[i + (size_t)n * m] = cos((double)((i * mm) % (2L * (n - 1))) * pi / (n - 1));
    } else {
      long mm = n - 1 - m;
      D[m] = -4.0 * pow(sin((2.0 * mm + 1.0) * pi / (4.0 * n)), 2);
      for (int i = 0; i < n; i++) {
        long a = id == 1 ? (long)(i + 1) * (2 * mm + 1) : (long)i * (2 * mm + 1);
        double ang = (double)(a % (4L * n)) * pi / (2.0 * n);
        V[i + (size_t)n * m] = id == 1 ? sin(ang) : cos(ang);
      }
    }
  }
  if (id == 3) {
    for (int r = 0; r < n; r++)
      for (int c = 0; c < n; c++) Vi[r + (size_t)n * c] = V[c + (size_t)n * r];
    return ETD_OK;
  }
  return invert_cm(n, V, Vi);
}

/* splitmix64 -> [0,1) */
static unsigned long long sm64(unsigned long long *x) {
  unsigned long long z = (*x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__attribute__((visibility("hidden"))) void etd_host_noise(int n, int g, unsigned long long seed, double mean, double amp, double *u) {
  int N = n - 2 * g;
  unsigned long long x = seed;
  double *P = malloc(sizeof(double) * (size_t)N * N * N);
  for (size_t l = 0; l < (size_t)N * N * N; l++) {
    double r = (double)(sm64(&x) >> 11) * (1.0 / 9007199254740992.0);
    P[l] = mean + amp * (2.0 * r - 1.0);
  }
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        int kk = ((k - g) % N + N) % N, jj = ((j - g) % N + N) % N, ii = ((i - g) % N + N) % N;
        u[((size_t)k * n + j) * n + i] = P[((size_t)kk * N + jj) * N + ii];
      }
  free(P);
}

/* check_soln_new's CSV rows (field.c:432-455, 524, 591, 647) */
__attribute__((visibility("hidden"))) int etd_csv_rows(const char *work_dir, int nac, int nch, int iter, const double *stats) {
  char fn[2048];
  FILE *f;
  if (iter == 0) {
    if (nac > 0 && nch > 0) {
      snprintf(fn, sizeof fn, "%sphi_couple.csv", work_dir);
      if (!(f = fopen(fn, "w"))) return ETD_ERR_IO;
      fprintf(f, "iter,eng\n");
      fclose(f);
    }
    for (int v = 0; v < nac; v++) {
      snprintf(fn, sizeof fn, "%sphi_ac_%d.csv", work_dir, v);
      if (!(f = fopen(fn, "w"))) return ETD_ERR_IO;
      fprintf(f, "iter,maxphiE,minphiE,volE,engE\n");
      fclose(f);
    }
    for (int v = 0; v < nch; v++) {
      snprintf(fn, sizeof fn, "%sphi_ch_%d.csv", work_dir, v);
      if (!(f = fopen(fn, "w"))) return ETD_ERR_IO;
      fprintf(f, "iter,maxphiCI,minphiCI,volCI,engCI\n");
      fclose(f);
    }
  }
  for (int v = 0; v < nac + nch; v++) {
    if (v < nac) snprintf(fn, sizeof fn, "%sphi_ac_%d.csv", work_dir, v);
    else snprintf(fn, sizeof fn, "%sphi_ch_%d.csv", work_dir, v - nac);
    if (!(f = fopen(fn, "a"))) return ETD_ERR_IO;
    const double *s = stats + 4 * v;
    fprintf(f, "%d,%+1.15lf,%+1.15lf,%+1.15lf,%+1.15lf\n", iter, s[0], s[1], s[2], s[3]);
    fclose(f);
  }
  if (nac > 0 && nch > 0) {
    snprintf(fn, sizeof fn, "%sphi_couple.csv", work_dir);
    if (!(f = fopen(fn, "a"))) return ETD_ERR_IO;
    fprintf(f, "%d,%+1.15lf\n", iter, stats[4 * (nac + nch)]);
    fclose(f);
  }
  return ETD_OK;
}

/* outputfield (field.c:17-37): text %+1.15lf, one value per line, full n^3 */
__attribute__((visibility("hidden"))) int etd_field_file(const char *data_dir, int var, int ioutput, int n, const double *u,
                                                       int cx, int cy, int cz) {
  char fn[2048];
  snprintf(fn, sizeof fn, "%seta%d_%06d_%02d%02d%02d.dat", data_dir, var, ioutput, cx, cy, cz);
  FILE *f = fopen(fn, "w");
  if (!f) return ETD_ERR_IO;
  for (size_t l = 0; l < (size_t)n * n * n; l++) fprintf(f, "%+1.15lf\n", u[l]);
  fclose(f);
  return ETD_OK;
}
