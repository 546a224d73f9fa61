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
 * Initial state: restart = 1 reads <dir>/input/field_iter<it>_ac<n>_000.txt
 * (interior, one value per line, field.c:137-154,746-754); otherwise the
 * seeded noise of the synthetic workloads (ETD_SEED, default 1): AC fields
 * U[-0.01, 0.01], CH fields 0.1 + U[-0.005, 0.005].
 */
#define _POSIX_C_SOURCE 199309L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "etd.h"

static int read_restart(etd_ctx *c, const char *dir, const char *kind, int var, int idx, int iter, int n, int g) {
  char fn[2048];
  snprintf(fn, sizeof fn, "%s/input/field_iter%d_%s%d_%d%d%d.txt", dir, iter, kind, idx, 0, 0, 0);
  FILE *f = fopen(fn, "r");
  if (!f) {
    printf("fopen error %s!\n", fn);
    return -1;
  }
  double *u = calloc((size_t)n * n * n, sizeof(double));
  for (int k = g; k < n - g; k++)
    for (int j = g; j < n - g; j++)
      for (int i = g; i < n - g; i++)
        if (fscanf(f, "%lf", &u[((size_t)k * n + j) * n + i]) != 1) {
          fclose(f);
          free(u);
          return -1;
        }
  fclose(f);
  int rc = etd_set_field(c, var, u);
  free(u);
  return rc;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s <ini suffix> [dir]\n", argv[0]);
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
  if (p.restart == 1) {
    for (int v = 0; v < p.nac; v++)
      if (read_restart(c, dir, "ac", v, v, p.restart_iter, p.nx, p.nghost)) exit(1);
    for (int v = 0; v < p.nch; v++)
      if (read_restart(c, dir, "ch", p.nac + v, v, p.restart_iter, p.nx, p.nghost)) exit(1);
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
  double *stats = malloc(sizeof(double) * (4 * V + 1));
  struct timespec a, b;
  while (tt < p.t_total) {
    if (etd_diag(c, stats) || etd_write_csv(c, iter, stats)) {
      fprintf(stderr, "check_soln_new: %s\n", etd_last_error(c));
      exit(1);
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
