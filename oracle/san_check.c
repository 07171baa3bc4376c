/*
 * Sanitizer harness for the CPU oracle (test infrastructure only): calls
 * every orc_* function on small valid inputs, in both physics modes, so
 * that a build with -fsanitize=address,undefined reports any out-of-bounds
 * access, use of uninitialised heap memory through UB, signed overflow or
 * misaligned access in the oracle itself.  Light invariants guard against
 * the harness silently doing nothing; the oracle's numbers are pinned
 * elsewhere (golden fixtures, closed forms).
 *
 * Prints one line per case and "san_check: ok" at the end; exits non-zero
 * on the first failed invariant (the sanitizers abort on their own).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lbm_oracle.h"

#define CHECK(cond)                                                    \
  do {                                                                 \
    if (!(cond)) {                                                     \
      fprintf(stderr, "san_check: %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      exit(1);                                                         \
    }                                                                  \
  } while (0)

static double* alloc(size_t n) {
  double* p = (double*)malloc(n * sizeof(double));
  CHECK(p != NULL);
  return p;
}

static int all_finite(const double* a, size_t n) {
  for (size_t i = 0; i < n; i++)
    if (!isfinite(a[i])) return 0;
  return 1;
}

/* the reference driver's REF box of square side L */
static orc_geom ref_geom(int L) {
  orc_geom g;
  memset(&g, 0, sizeof g);
  g.mode = ORC_MODE_REF;
  g.Lx = g.Ly = 4 * L + 1;
  g.xmin = (g.Lx - 1) / 2;
  g.xmax = g.xmin + L;
  g.ymin = (g.Ly - 1) / 2 - L / 2;
  g.ymax = g.ymin + L;
  g.tau = 0.53;
  g.beta = (1.0 / (g.Lx - 1)) * ((double)L / (g.Ly - 1)) * 0.05 * 0.05;
  return g;
}

static orc_geom channel_geom(void) {
  orc_geom g;
  memset(&g, 0, sizeof g);
  g.mode = ORC_MODE_CHANNEL;
  g.Lx = 48;
  g.Ly = 24;
  g.xmin = 9;
  g.xmax = 15;
  g.ymin = 9;
  g.ymax = 15;
  g.tau = 0.55;
  g.u_in = 0.05;
  return g;
}

static void fill_feq(const orc_geom* g, double* f, unsigned seed) {
  srand(seed);
  for (int x = 0; x < g->Lx; x++)
    for (int y = 0; y < g->Ly; y++) {
      const double r1 = rand() / (double)RAND_MAX - 0.5;
      const double r2 = rand() / (double)RAND_MAX - 0.5;
      orc_feq(1.0 + 1e-3 * r1, 0.05 + 1e-3 * r2, 1e-3 * r1 * r2,
              f + 9 * ((size_t)y + (size_t)g->Ly * x));
    }
}

static void steps_and_forces(const orc_geom* g, const char* name) {
  const size_t N = 9 * (size_t)g->Lx * g->Ly;
  double *a = alloc(N), *b = alloc(N), *A = alloc(12);
  fill_feq(g, a, 1);
  const double F = orc_step(g, a, b);
  CHECK(isfinite(F) && all_finite(b, N));
  CHECK(isfinite(orc_force_stress(g, b)) && isfinite(orc_force_mea(g, b)));
  const double acc = orc_run_steps(g, b, a, 12, 1.0, 0.0, 1.0, A);
  CHECK(isfinite(acc) && all_finite(A, 12) && all_finite(b, N));
  double *rho = alloc(N / 9), *ux = alloc(N / 9), *uy = alloc(N / 9);
  orc_macro(g, b, NULL, 0.0, rho, ux, uy);
  CHECK(all_finite(rho, N / 9) && all_finite(ux, N / 9) &&
        all_finite(uy, N / 9));
  for (int x = 0; x < g->Lx; x++)
    for (int y = 0; y < g->Ly; y++) (void)orc_bb_mask(g, x, y);
  if (g->mode == ORC_MODE_REF) {
    double* map = alloc(N / 9);
    for (size_t i = 0; i < N / 9; i++) map[i] = sin(0.37 * (double)i);
    CHECK(isfinite(orc_step_map(g, b, a, map, 0.5)) && all_finite(a, N));
    const double accf =
        orc_run_steps_forced(g, a, b, 5, 9, map, 0.5, 1.0, 0.0, 1.0, A);
    CHECK(isfinite(accf) && all_finite(A, 9));
    orc_macro(g, a, map, 0.5, rho, ux, uy);
    CHECK(all_finite(ux, N / 9));
    /* snapshot perturbation over three snapshots */
    double* snaps = alloc(3 * N);
    for (int n = 0; n < 3; n++) fill_feq(g, snaps + n * N, 10 + n);
    const double pre[3] = {0.2, 0.5, 0.9};
    orc_perturb(a, snaps, 3, (long)N, 0.3, pre);
    CHECK(all_finite(a, N));
    free(snaps);
    free(map);
  } else {
    orc_init_channel(g, a, 1e-3, 7, 0, 3, 1);
    orc_noise(g, a, 1e-4, 7, 0, 3, 2, 1);
    CHECK(all_finite(a, N));
  }
  /* several clones back to back, as the CPU baseline runs them */
  double *st = alloc(3 * N), *sc = alloc(N), acc3[3];
  for (int j = 0; j < 3; j++) fill_feq(g, st + j * N, 20 + j);
  CHECK(isfinite(orc_run_clones(g, st, sc, 3, 4, 1.0, 0.0, 1.0, acc3)));
  printf("san_check: %s %dx%d steps, forces, macro ok\n", name, g->Lx, g->Ly);
  free(st);
  free(sc);
  free(rho);
  free(ux);
  free(uy);
  free(a);
  free(b);
  free(A);
}

static void selection(void) {
  enum { N = 64 };
  double s[N], u[N], R = 0.0;
  int32_t rint[N], nb0[N], nb1[N], pairs[2 * N], labels[N];
  for (int p = 1; p <= 4; p *= 2) {
    for (int j = 0; j < N; j++) {
      s[j] = exp(2.0 * sin(0.7 * j + p));
      u[j] = fmod(0.6180339887 * (j + 1) * p, 1.0);
      rint[j] = (int32_t)(7919 * j + 13 * p);
    }
    int used = -1;
    const int np =
        orc_select(N, s, p, u, rint, nb0, nb1, pairs, labels, &R, &used);
    CHECK(np >= 0 && np <= N && used >= 0 && used <= N && isfinite(R));
    int tot = 0;
    for (int j = 0; j < N; j++) tot += nb1[j];
    CHECK(tot == N);
  }
  /* degenerate: one clone holds all the weight */
  for (int j = 0; j < N; j++) {
    s[j] = j == 5 ? 1.0 : 0.0;
    u[j] = 0.5;
    rint[j] = j;
  }
  int used = -1;
  CHECK(orc_select(N, s, 1, u, rint, nb0, nb1, pairs, labels, &R, &used) ==
        N - 1);
  printf("san_check: selection ok\n");
}

static void generators(void) {
  uint32_t ctr[4] = {1, 2, 3, 4}, key[2] = {5, 6}, out[4];
  orc_philox4x32_10(ctr, key, out);
  double n[4];
  orc_box_muller4(out, n);
  CHECK(all_finite(n, 4));
  double sn[10];
  orc_site_normals(out, sn);
  CHECK(all_finite(sn, 10));
  double z[64];
  orc_field_normals(1, 2, 3, 4, 5, 32, z);
  CHECK(all_finite(z, 64));
  int32_t r[16];
  double d[16];
  orc_glibc_srand(1);
  orc_glibc_rand_n(16, r);
  orc_glibc_seed48_default();
  orc_glibc_drand48_n(16, d);
  for (int i = 0; i < 16; i++) CHECK(r[i] >= 0 && d[i] >= 0.0 && d[i] < 1.0);
  double logR[10], lam[10];
  for (int i = 0; i < 10; i++) logR[i] = 0.01 * i;
  orc_scgf(logR, 10, 0.3, lam);
  CHECK(lam[0] == 0.0 && all_finite(lam, 10));
  printf("san_check: generators and scgf ok\n");
}

int main(void) {
  const orc_geom r4 = ref_geom(4), r5 = ref_geom(5), ch = channel_geom();
  steps_and_forces(&r4, "REF");
  steps_and_forces(&r5, "REF");
  steps_and_forces(&ch, "CHANNEL");
  selection();
  generators();
  printf("san_check: ok\n");
  return 0;
}
