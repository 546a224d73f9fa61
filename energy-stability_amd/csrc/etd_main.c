/*
 * etd_main.c -- drop-in driver for the reference's time-step loop
 * (source_beta_to_omega/src/ScLETD.c:9-201) on one MI355X:
 *
 *   etd_run <suffix> [dir]
 *
 * reads <dir>/input<suffix>.ini (omega order), elastic_input.txt,
 * anisotropic_input.txt, optional landau_input.txt and d/v/vi3<n>.dat, then
 * runs  while (tt < t_total) { check_soln_new(tt); iter++; tt += dt; step; }
 * writing phi_ac_<n>.csv / phi_ch_<n>.csv / phi_couple.csv under work_dir.
 * Errors print the reference's message and exit(1).
 *
 * A procs line other than 1,1,1 runs the reference's decomposition on this
 * GPU: px x py x pz boxes of nx^3, exchanged as its ranks would.
 *
 * Initial state: restart = 1 reads <dir>/input/field_iter<it>_ac<n>_<cx><cy><cz>.txt
 * per box (interior, one value per line, field.c:137-154,746-754);
 * ETD_INIT=sphere the reference's own start for nch = 0 (init_field,
 * field.c:757-766: the sphere of init_field_sphere_center in AC field 3, the
 * other fields 0); otherwise the seeded noise of the synthetic workloads
 * (ETD_SEED, default 1): AC fields U[-0.01, 0.01], CH fields
 * 0.1 + U[-0.005, 0.005].
 *
 * Field dumps: every 6500 iterations (ETD_OUTPUT_EVERY overrides; 0 none), inside the
 * diagnostics as in check_soln_new (field.c:667-668), outputfield writes
 * <data_dir>eta<v>_<ioutput:06d>_<cx:02d><cy:02d><cz:02d>.dat per box with ioutput starting at iter/100
 * (ScLETD.c:83) and counting up per dump (field.c:55).  The reference's loop
 * starts at variant 3 (field.c:23); every AC variant is written here, a
 * superset of its files.
 *
 *   etd_run --alpha <suffix> [dir]
 *
 * the beta -> alpha driver's input and output contract
 * (source_beta_to_alpha/src/ScLETD.cpp:5-180, anisotropic_hip.c:3953-4355) on
 * this library's physics: the alpha-order ini, elastic_input.txt (ELASTIC = 1)
 * and anisotropic_input.txt (ANISOTROPIC = 1); checkpoint 2 restarts from
 * checkpoint buffer chk and continues in the other buffer, 1 starts from the
 * seeded noise; then floor(t_total / dt) fixed steps rounded up to whole
 * blocks of 100 (ScLETD.cpp:131, the while (gyq < 100) loop), and after each
 * step: every nout steps irun += 1 and the uint16 dumps
 * <data_dir>00/eta<v>_<irun:06d>_000000.dat; every nchk steps the binary
 * checkpoint into buffer chk, then chk = 1 - chk (anisotropic_hip.c:4290-4316).
 * After every block of 100 steps whose iteration count is a multiple of
 * detect_iter = 10000 (ETD_DETECT_EVERY overrides), the detector feed of each
 * AC field: ./downsample_output_<nx>_<px><py><pz>/eta<v>_<iter>_0_0_0_ds.txt,
 * one byte of 8 binarised downsampled cells per line as "%d"
 * (ScLETD.cpp:139-160, compress_save_output).  The CenterNet3D detector the
 * reference then runs on the feed is not part of this library.  At the end
 * the check_soln_new rows of the last state are appended to the CSV files.
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "etd.h"

/* restart files (field.c:137-154, 746-754): one per box, named by its rank
 * coordinates, interior values */
static int read_restart(etd_ctx *c, const etd_params *p, const char *dir, const char *kind, int var, int idx,
                        int iter) {
  const int n = p->nx, g = p->nghost, P = p->px * p->py * p->pz;
  const size_t n3 = (size_t)n * n * n;
  double *u = calloc(n3 * P, sizeof(double));
  if (!u) return -1;
  for (int b = 0; b < P; b++) {
    char fn[2048];
    snprintf(fn, sizeof fn, "%s/input/field_iter%d_%s%d_%d%d%d.txt", dir, iter, kind, idx, b % p->px,
             (b / p->px) % p->py, b / (p->px * p->py));
    FILE *f = fopen(fn, "r");
    if (!f) {
      printf("fopen error %s!\n", fn);
      free(u);
      return -1;
    }
    double *ub = u + b * n3;
    for (int k = g; k < n - g; k++)
      for (int j = g; j < n - g; j++)
        for (int i = g; i < n - g; i++)
          if (fscanf(f, "%lf", &ub[((size_t)k * n + j) * n + i]) != 1) {
            fclose(f);
            free(u);
            return -1;
          }
    fclose(f);
  }
  int rc = etd_set_field(c, var, u);
  free(u);
  return rc;
}

/* init_field_sphere_center (field.c:379-405), the reference's start for
 * nch = 0 (init_field, field.c:757-766): AC field 3 is 1 where a box's own
 * coordinate (cx nx + i, cy ny + j, cz nz + k), ghost cells included, lies
 * within 8 of nx px / 2 on every axis, and 0 elsewhere; the other fields 0. */
static int init_sphere(etd_ctx *c, const etd_params *p) {
  const int n = p->nx, P = p->px * p->py * p->pz;
  const size_t n3 = (size_t)n * n * n;
  double *u = calloc(n3 * P, sizeof(double));
  if (!u) return -1;
  int rc = 0;
  for (int v = 0; v < p->nac + p->nch && !rc; v++) rc = etd_set_field(c, v, u);
  const double cnt = n * p->px / 2;
  for (int b = 0; b < P; b++) {
    const int cx = b % p->px, cy = (b / p->px) % p->py, cz = b / (p->px * p->py);
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          const double x = cx * n + i - cnt, y = cy * n + j - cnt, z = cz * n + k - cnt;
          u[b * n3 + ((size_t)k * n + j) * n + i] = x * x + y * y + z * z < 8.0 * 8.0 ? 1.0 : 0.0;
        }
  }
  if (!rc) rc = etd_set_field(c, 3, u);
  free(u);
  return rc;
}

static int write_feed(etd_ctx *c, const etd_params *p, int v, long iter) {
  const int d = (p->nx - 2 * p->nghost) / 2;
  const long nb = (long)d * d * d / 8;
  unsigned char *bits = malloc((size_t)nb);
  if (!bits) return -1;
  int rc = etd_detector_feed(c, v, bits, NULL);
  if (!rc) {
    char fn[2048];
    snprintf(fn, sizeof fn, "./downsample_output_%d_%d%d%d/eta%d_%ld_%d_%d_%d_ds.txt", p->nx, p->px, p->py, p->pz, v,
             iter, 0, 0, 0);
    FILE *f = fopen(fn, "w");
    if (!f) {
      printf("Fail to save\n");
      rc = -1;
    } else {
      for (long t = 0; t < nb; t++) fprintf(f, "%d\n", bits[t]);
      fclose(f);
    }
  }
  free(bits);
  return rc;
}

static int run_alpha(const char *suffix, const char *dir) {
  etd_params p;
  etd_default_params(&p);
  char fn[2048];
  snprintf(fn, sizeof fn, "%s/input%s.ini", dir, suffix);
  if (etd_parse_ini_alpha(fn, &p)) {
    fprintf(stderr, "cannot parse %s\n", fn);
    return 1;
  }
  p.px = p.py = p.pz = 1;  /* one subdomain (the reference's files name ranks by cart id) */
  p.periodic = 1;
  if (p.elastic == 1) {
    snprintf(fn, sizeof fn, "%s/elastic_input.txt", dir);
    if (etd_parse_elastic(fn, &p)) {
      fprintf(stderr, "open elastic_input.txt Fail!");
      return 1;
    }
  }
  if (p.anisotropic) {
    snprintf(fn, sizeof fn, "%s/anisotropic_input.txt", dir);
    if (etd_parse_anisotropic(fn, &p)) {
      fprintf(stderr, "open anisotropic_input.txt Fail!");
      return 1;
    }
  }
  snprintf(fn, sizeof fn, "%s/landau_input.txt", dir);
  FILE *lf = fopen(fn, "r");
  if (lf) {
    fclose(lf);
    if (etd_parse_landau(fn, &p)) return 1;
  }
  etd_ctx *c = NULL;
  int rc = etd_create(&p, NULL, NULL, NULL, &c);
  if (rc) {
    fprintf(stderr, "etd_create: %s%s%s\n", etd_error_string(rc), c ? ": " : "", c ? etd_last_error(c) : "");
    if (c) etd_close(c);
    return 1;
  }
  /* the grain mask: n^3 native ints, x fastest, as read_ori reads its f file (field.c:111-129) */
  const char *mpath = getenv("ETD_ALPHA_MASK");
  if (mpath) {
    const size_t n3 = (size_t)p.nx * p.ny * p.nz;
    int *m = (int *)malloc(sizeof(int) * n3);
    FILE *mf = fopen(mpath, "rb");
    const int ok = m && mf && fread(m, sizeof(int), n3, mf) == n3;
    if (mf) fclose(mf);
    if (!ok || etd_set_mask(c, m)) {
      fprintf(stderr, "fopen error %s!\n", mpath);
      free(m);
      etd_close(c);
      return 1;
    }
    free(m);
  }
  printf("restart\t\t%d\n", p.restart);
  printf("nx,ny,nz\t%d\t%d\t%d\n", p.nx, p.ny, p.nz);
  printf("dt,t_total\t%lf,%lf\n", p.dt, p.t_total);
  printf("nac,nch\t\t%d,%d\n", p.nac, p.nch);
  printf("checkpoint_,nchk,chk\t%d,%d,%d\n", p.checkpoint, p.nchk, p.chk);
  printf("nout\t\t%d\n", p.nout);
  const int V = p.nac + p.nch;
  int irun = 0, chk = p.chk;
  if (p.checkpoint == 2) {
    if (etd_read_checkpoint(c, chk, &irun)) {
      fprintf(stderr, "read_chk: %s\n", etd_last_error(c));
      return 1;
    }
    chk = (chk + 1) % 2;
  } else {
    chk = 0;
    const char *s = getenv("ETD_SEED");
    unsigned long long seed = s ? strtoull(s, NULL, 10) : 1ULL;
    for (int v = 0; v < V; v++) {
      int ch = v >= p.nac;
      if (etd_init_noise(c, v, seed + 7919ULL * v, ch ? 0.1 : 0.0, ch ? 0.005 : 0.01)) return 1;
    }
  }
  const char *de = getenv("ETD_DETECT_EVERY");
  const long detect_iter = de && atol(de) > 0 ? atol(de) : 10000;
  const long blocks = (long)ceil(floor(p.t_total / p.dt) / 100.0);
  long iter = 0;
  int ot1 = 0, ot2 = 0;
  for (long b = 0; b < blocks; b++) {
    for (int s = 0; s < 100; s++) {
      if (etd_step(c, 1)) {
        fprintf(stderr, "etd_step: %s\n", etd_last_error(c));
        return 1;
      }
      iter++;
      if (++ot1 >= p.nout) {
        ot1 = 0;
        irun += 1;
        if (etd_write_field2b(c, irun)) {
          printf("write_field2B: %s\n", etd_last_error(c));
          return 1;
        }
      }
      if (++ot2 >= p.nchk) {
        ot2 = 0;
        if (etd_write_checkpoint(c, irun, chk)) {
          printf("write_chk: %s\n", etd_last_error(c));
          return 1;
        }
        chk = (chk + 1) % 2;
      }
    }
    if (iter % detect_iter == 0)
      for (int v = 0; v < p.nac; v++)
        if (write_feed(c, &p, v, iter)) return 1;
  }
  double *stats = malloc(sizeof(double) * (4 * V + 1));
  if (etd_diag(c, stats) || etd_write_csv(c, (int)iter, stats)) {
    fprintf(stderr, "check_soln_new: %s\n", etd_last_error(c));
    return 1;
  }
  printf("time\t\t%lf\n", iter * p.dt);
  free(stats);
  etd_close(c);
  return 0;
}

int main(int argc, char **argv) {
  if (argc >= 3 && strcmp(argv[1], "--alpha") == 0) return run_alpha(argv[2], argc > 3 ? argv[3] : ".");
  if (argc < 2) {
    fprintf(stderr, "usage: %s [--alpha] <ini suffix> [dir]\n", argv[0]);
    return 1;
  }
  const char *dir = argc > 2 ? argv[2] : ".";
  etd_ctx *c = NULL;
  int rc = etd_open(dir, argv[1], &c);
  if (rc) {
    fprintf(stderr, "etd_open: %s%s%s\n", etd_error_string(rc), c ? ": " : "", c ? etd_last_error(c) : "");
    if (c) etd_close(c);
    exit(1);
  }
  etd_params p;
  etd_get_params(c, &p);
  printf("running   on\t%d   processors\n", 1);
  printf("nx,ny,nz\t%d\t%d\t%d\n", p.nx, p.ny, p.nz);
  printf("nghost\t\t%d\n", p.nghost);
  printf("epn2\t\t%lf\n", p.epn2);
  printf("dt\t\t%lf\n", p.dt);
  printf("nac,nch\t\t%d,%d\n", p.nac, p.nch);
  printf("ETD2\t\t%d\n", p.etd2);
  printf("ELASTIC\t\t%d\n", p.elastic);
  printf("ANISOTROPIC\t\t%d\n", p.anisotropic);
  const int V = p.nac + p.nch;
  const char *init = getenv("ETD_INIT");
  if (p.restart == 1) {
    for (int v = 0; v < p.nac; v++)
      if (read_restart(c, &p, dir, "ac", v, v, p.restart_iter)) exit(1);
    for (int v = 0; v < p.nch; v++)
      if (read_restart(c, &p, dir, "ch", p.nac + v, v, p.restart_iter)) exit(1);
  } else if (init && !strcmp(init, "sphere")) {
    if (p.nch != 0 || p.nac < 4) {
      printf("ETD_INIT=sphere is the reference's nch = 0 start in AC field 3 (nac >= 4)\n");
      exit(1);
    }
    if (init_sphere(c, &p)) exit(1);
  } else {
    const char *s = getenv("ETD_SEED");
    unsigned long long seed = s ? strtoull(s, NULL, 10) : 1ULL;
    for (int v = 0; v < V; v++) {
      int ch = v >= p.nac;
      if (etd_init_noise(c, v, seed + 7919ULL * v, ch ? 0.1 : 0.0, ch ? 0.005 : 0.01)) exit(1);
    }
  }
  int iter = p.restart == 1 ? p.restart_iter : 0;
  double tt = iter * p.dt;
  int ioutput = iter / 100;
  const char *oe = getenv("ETD_OUTPUT_EVERY");
  const int out_every = oe ? atoi(oe) : 6500;  // 0 or less: no field dumps
  double *stats = malloc(sizeof(double) * (4 * V + 1));
  struct timespec a, b;
  while (tt < p.t_total) {
    if (etd_diag(c, stats) || etd_write_csv(c, iter, stats)) {
      fprintf(stderr, "check_soln_new: %s\n", etd_last_error(c));
      exit(1);
    }
    if (out_every > 0 && iter % out_every == 0) {
      for (int v = 0; v < p.nac; v++)
        if (etd_write_field(c, v, ioutput)) {
          printf("fopen error writing eta%d under %s\n", v, p.data_dir);
          exit(1);
        }
      ioutput++;
    }
    iter++;
    tt += p.dt;
    clock_gettime(CLOCK_MONOTONIC, &a);
    if (etd_step(c, 1)) {
      fprintf(stderr, "etd_step: %s\n", etd_last_error(c));
      exit(1);
    }
    clock_gettime(CLOCK_MONOTONIC, &b);
    printf("iter %d runtime %.6f s\n", iter, (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec));
  }
  free(stats);
  etd_close(c);
  return 0;
}
