/*
 * CPU FP64 oracle for the cloning-sampler inner loop.
 *
 * TEST INFRASTRUCTURE ONLY: linked by nothing in the product.  Used by tests/,
 * by the smoke test and by bench.py's CPU-baseline leg, as the checker.
 *
 * Every routine restates the reference's algorithm in plain C with the
 * reference's floating-point operation order, so that on x86-64 (SSE2, no FMA
 * contraction) its results are bit-identical to the compiled reference where
 * the reference has the routine.  Citations are to tlestang/cloning-lbm-MPI.
 */
#define _DEFAULT_SOURCE
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* src/mpi_LBM_GK_pops_perturb.cpp:31-32 (velocity set and weights). */
const int orc_c[9][2] = {{0, 0}, {1, 0}, {0, 1}, {-1, 0}, {0, -1},
                         {1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
const double orc_w[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,
                         1.0 / 9.0,  1.0 / 9.0,  1.0 / 36.0,
                         1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
const int orc_opp[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

/* src/lbm_src/global.h:4 */
static inline long IDX(const orc_geom* g, int x, int y, int k) {
  return (long)k + 9L * ((long)y + (long)g->Ly * (long)x);
}

static int is_interior(const orc_geom* g, int x, int y) {
  return x > g->xmin && x < g->xmax && y > g->ymin && y < g->ymax;
}

/* Square part of the bounce-back set, following the writes of
 * computeSquareBounceBack_TEST (src/lbm_src/square.cpp:24-50): three
 * directions per face node, one per corner. */
static int square_mask(const orc_geom* g, int x, int y) {
  const int x0 = g->xmin, x1 = g->xmax, y0 = g->ymin, y1 = g->ymax;
  if (x < x0 || x > x1 || y < y0 || y > y1) return 0;
  if (is_interior(g, x, y)) return 0;
  const int inx = (x > x0 && x < x1), iny = (y > y0 && y < y1);
  if (y == y0 && inx) return (1 << 4) | (1 << 7) | (1 << 8);
  if (y == y1 && inx) return (1 << 2) | (1 << 5) | (1 << 6);
  if (x == x0 && iny) return (1 << 3) | (1 << 6) | (1 << 7);
  if (x == x1 && iny) return (1 << 1) | (1 << 5) | (1 << 8);
  if (x == x0 && y == y0) return 1 << 7;
  if (x == x0 && y == y1) return 1 << 6;
  if (x == x1 && y == y1) return 1 << 5;
  return 1 << 8; /* (x1, y0) */
}

/* src/lbm_src/domain_noSlipWalls.cpp:10-21 */
static int wall_mask(const orc_geom* g, int y) {
  int m = 0;
  if (y == 0) m |= (1 << 2) | (1 << 5) | (1 << 6);
  if (y == g->Ly - 1) m |= (1 << 4) | (1 << 7) | (1 << 8);
  return m;
}

int orc_bb_mask(const orc_geom* g, int x, int y) {
  return wall_mask(g, y) | square_mask(g, x, y);
}

/* Equilibrium in the arithmetic order of src/lbm_src/streamCollCompute.cpp:44-49. */
void orc_feq(double rho, double ux, double uy, double feq[9]) {
  double u2 = -1.5 * (ux * ux + uy * uy);
  for (int k = 0; k < 9; k++) {
    double eu = orc_c[k][0] * ux + orc_c[k][1] * uy;
    double eueu = 4.5 * eu * eu;
    feq[k] = orc_w[k] * rho * (1.0 + 3.0 * eu + eueu + u2);
  }
}

/* BGK collision with Guo forcing at one node, in place
 * (src/lbm_src/streamCollCompute.cpp:30-52). */
static void collide_node(double* f, double beta, double tau) {
  double omega = 1. / tau;
  double omega1 = 1.0 - omega;
  double coeff_forcing = 1.0 - 0.5 * omega;
  double rho = 0.0, ux = 0.0, uy = 0.0;
  for (int k = 0; k < 9; k++) {
    double ft = f[k];
    rho += ft;
    ux += ft * orc_c[k][0];
    uy += ft * orc_c[k][1];
  }
  ux = ux / rho + 0.5 * beta / rho;
  uy /= rho;
  double u2 = -1.5 * (ux * ux + uy * uy);
  for (int k = 0; k < 9; k++) {
    double eu = orc_c[k][0] * ux + orc_c[k][1] * uy;
    double eueu = 4.5 * eu * eu;
    double feq = orc_w[k] * rho * (1.0 + 3.0 * eu + eueu + u2);
    double fd = orc_w[k] * coeff_forcing * beta *
                (3. * (orc_c[k][0] - ux) + 9. * orc_c[k][0] * eu);
    f[k] = f[k] * omega1 + feq * omega + fd;
  }
}

/* One step; map (REF only, may be NULL) gives the forcing mode's per-site
 * body force beta(x,y) = (1.0 + eps*map[y + Dy*x])*beta
 * (src/lbm_src/streamCollCompute.cpp:63, idx(x,y) = y + Dy*x). */
static double step_impl(const orc_geom* g, double* fin, double* fout,
                        const double* map, double eps) {
  const int Dx = g->Lx, Dy = g->Ly;
  const int periodic = (g->mode == ORC_MODE_REF);
  const double beta = periodic ? g->beta : 0.0;
  /* collide + push (src/lbm_src/streamCollCompute.cpp:30-56) */
  for (int x = 0; x < Dx; x++) {
    for (int y = 0; y < Dy; y++) {
      double* f = fin + IDX(g, x, y, 0);
      const double b =
          (periodic && map) ? (1.0 + eps * map[y + Dy * x]) * beta : beta;
      collide_node(f, b, g->tau);
      for (int k = 0; k < 9; k++) {
        int nx = x + orc_c[k][0], ny = y + orc_c[k][1];
        if (periodic) {
          nx = (nx + Dx) % Dx;
          ny = (ny + Dy) % Dy;
        } else if (nx < 0 || nx >= Dx || ny < 0 || ny >= Dy) {
          continue;
        }
        fout[IDX(g, nx, ny, k)] = f[k];
      }
    }
  }
  if (!periodic) {
    /* zero-gradient outlet: ghost column x=Dx is a copy of column Dx-1 */
    for (int y = 0; y < Dy; y++) {
      const int ks[3] = {3, 6, 7};
      for (int i = 0; i < 3; i++) {
        int k = ks[i];
        int ys = y - orc_c[k][1];
        if (ys < 0 || ys >= Dy) continue;
        fout[IDX(g, Dx - 1, y, k)] = fin[IDX(g, Dx - 1, ys, k)];
      }
    }
  }
  /* walls (src/lbm_src/domain_noSlipWalls.cpp:10-21) */
  for (int x = 0; x < Dx; x++) {
    fout[IDX(g, x, Dy - 1, 4)] = fin[IDX(g, x, Dy - 1, 2)];
    fout[IDX(g, x, Dy - 1, 8)] = fin[IDX(g, x, Dy - 1, 6)];
    fout[IDX(g, x, Dy - 1, 7)] = fin[IDX(g, x, Dy - 1, 5)];
    fout[IDX(g, x, 0, 2)] = fin[IDX(g, x, 0, 4)];
    fout[IDX(g, x, 0, 5)] = fin[IDX(g, x, 0, 7)];
    fout[IDX(g, x, 0, 6)] = fin[IDX(g, x, 0, 8)];
  }
  /* square ring bounce-back (src/lbm_src/square.cpp:24-50) */
  for (int x = g->xmin; x <= g->xmax; x++) {
    for (int y = g->ymin; y <= g->ymax; y++) {
      int m = square_mask(g, x, y);
      for (int k = 0; k < 9; k++)
        if (m & (1 << k))
          fout[IDX(g, x, y, k)] = fin[IDX(g, x, y, orc_opp[k])];
    }
  }
  if (!periodic) {
    /* outlet: the zero-gradient column x=Dx-1 is completed to equilibrium at
       the reference density, feq(1, u) with u its own velocity */
    for (int y = 0; y < Dy; y++) {
      double* f = fout + IDX(g, Dx - 1, y, 0);
      double rho = 0.0, ux = 0.0, uy = 0.0;
      for (int k = 0; k < 9; k++) {
        rho += f[k];
        ux += f[k] * orc_c[k][0];
        uy += f[k] * orc_c[k][1];
      }
      ux /= rho;
      uy /= rho;
      double eq[9];
      orc_feq(1.0, ux, uy, eq);
      for (int k = 0; k < 9; k++) f[k] = eq[k];
    }
    /* equilibrium velocity inlet: column x=0 becomes feq(rho_in, (u_in, 0)),
       rho_in = (f0 + f2 + f4 + 2 (f3 + f6 + f7)) / (1 - u_in) from the
       populations that streamed in from the domain (walls included) */
    for (int y = 0; y < Dy; y++) {
      double* f = fout + IDX(g, 0, y, 0);
      double rho_in = (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7])) /
                      (1.0 - g->u_in);
      double eq[9];
      orc_feq(rho_in, g->u_in, 0.0, eq);
      for (int k = 0; k < 9; k++) f[k] = eq[k];
    }
  }
  /* interior reset (src/mpi_LBM_GK_pops_perturb.cpp:430-439) */
  for (int x = g->xmin + 1; x < g->xmax; x++)
    for (int y = g->ymin + 1; y < g->ymax; y++)
      for (int k = 0; k < 9; k++) fout[IDX(g, x, y, k)] = orc_w[k];

  return (g->mode == ORC_MODE_REF) ? orc_force_stress(g, fout)
                                   : orc_force_mea(g, fout);
}

double orc_step(const orc_geom* g, double* fin, double* fout) {
  return step_impl(g, fin, fout, NULL, 0.0);
}

double orc_step_map(const orc_geom* g, double* fin, double* fout,
                    const double* map, double eps) {
  return step_impl(g, fin, fout, map, eps);
}

/* Moments and non-equilibrium stress at one node, in the order of
 * src/lbm_src/force.cpp:26-46. */
static void node_stress(const double* f, double* rho_o, double* pxx,
                        double* pxy) {
  double rho = 0.0, ux = 0.0, uy = 0.0;
  for (int k = 0; k < 9; k++) {
    double ft = f[k];
    rho += ft;
    ux += ft * orc_c[k][0];
    uy += ft * orc_c[k][1];
  }
  uy /= rho;
  ux /= rho;
  double Pi_xx = 0.0, Pi_xy = 0.0;
  double u2 = -1.5 * (ux * ux + uy * uy);
  for (int k = 0; k < 9; k++) {
    double eu = orc_c[k][0] * ux + orc_c[k][1] * uy;
    double eueu = 4.5 * eu * eu;
    double feq = orc_w[k] * rho * (1.0 + 3.0 * eu + eueu + u2);
    double fneq = f[k] - feq;
    Pi_xx += fneq * orc_c[k][0] * orc_c[k][0];
    Pi_xy += fneq * orc_c[k][0] * orc_c[k][1];
  }
  *rho_o = rho;
  *pxx = Pi_xx;
  *pxy = Pi_xy;
}

/* Stress-integral drag (src/lbm_src/force.cpp:8-142). */
double orc_force_stress(const orc_geom* g, const double* f) {
  const double ot = 1. / 3.;
  const double omega = 1. / g->tau;
  const double cf = 1.0 - .5 * omega;
  double rho, pxx, pxy;
  double fW = 0.0, fE = 0.0, fN = 0.0, fS = 0.0;
  for (int y = g->ymin; y < g->ymax + 1; y++) {
    node_stress(f + IDX(g, g->xmin, y, 0), &rho, &pxx, &pxy);
    fW += rho * ot + cf * pxx;
  }
  for (int y = g->ymin; y < g->ymax + 1; y++) {
    node_stress(f + IDX(g, g->xmax, y, 0), &rho, &pxx, &pxy);
    fE += -rho * ot - cf * pxx;
  }
  for (int x = g->xmin; x < g->xmax + 1; x++) {
    node_stress(f + IDX(g, x, g->ymax, 0), &rho, &pxx, &pxy);
    fN += -cf * pxy;
  }
  for (int x = g->xmin; x < g->xmax + 1; x++) {
    node_stress(f + IDX(g, x, g->ymin, 0), &rho, &pxx, &pxy);
    fS += +cf * pxy;
  }
  return fW + fE + fN + fS;
}

/* Momentum-exchange drag on a post-stream state: each population q that
 * arrived at a ring node by bounce-back from the square carries momentum
 * c_q; the square received -2 c_q f_q.  Order: west face (y ascending,
 * corners included), east face, north face (corners excluded), south face. */
static double mea_node(const orc_geom* g, const double* f, int x, int y) {
  int m = square_mask(g, x, y);
  double t = 0.0;
  for (int k = 0; k < 9; k++)
    if (m & (1 << k)) t += (double)(-2 * orc_c[k][0]) * f[IDX(g, x, y, k)];
  return t;
}

double orc_force_mea(const orc_geom* g, const double* f) {
  double fW = 0.0, fE = 0.0, fN = 0.0, fS = 0.0;
  for (int y = g->ymin; y <= g->ymax; y++) fW += mea_node(g, f, g->xmin, y);
  for (int y = g->ymin; y <= g->ymax; y++) fE += mea_node(g, f, g->xmax, y);
  for (int x = g->xmin + 1; x < g->xmax; x++) fN += mea_node(g, f, x, g->ymax);
  for (int x = g->xmin + 1; x < g->xmax; x++) fS += mea_node(g, f, x, g->ymin);
  return fW + fE + fN + fS;
}

/* Interval loop body (src/mpi_LBM_GK_pops_perturb.cpp:424-465). */
double orc_run_steps(const orc_geom* g, double* state, double* scratch, int M,
                     double scale, double shift, double dt, double* A_out) {
  long N = 9L * g->Lx * g->Ly;
  double* cur = state;
  double* oth = scratch;
  double acc = 0.0;
  for (int n = 0; n < M; n++) {
    double F = orc_step(g, cur, oth);
    double* t = cur;
    cur = oth;
    oth = t;
    F = F * scale;
    F = F - shift;
    if (A_out) A_out[n] = F;
    acc += F;
  }
  if (cur != state) memcpy(state, cur, (size_t)N * sizeof(double));
  acc *= dt;
  return acc;
}

/* Fields of a state as the step's collision computes them
 * (src/lbm_src/streamCollCompute.cpp:17-26,44-45). */
void orc_macro(const orc_geom* g, const double* f, const double* map,
               double eps, double* rho, double* ux, double* uy) {
  const int Dx = g->Lx, Dy = g->Ly;
  const int periodic = (g->mode == ORC_MODE_REF);
  for (int x = 0; x < Dx; x++) {
    for (int y = 0; y < Dy; y++) {
      const double* p = f + IDX(g, x, y, 0);
      const double beta =
          !periodic ? 0.0
                    : (map ? (1.0 + eps * map[y + Dy * x]) * g->beta : g->beta);
      double r = 0.0, u = 0.0, v = 0.0;
      for (int k = 0; k < 9; k++) {
        double ft = p[k];
        r += ft;
        u += ft * orc_c[k][0];
        v += ft * orc_c[k][1];
      }
      u = u / r + 0.5 * beta / r;
      v /= r;
      rho[y + Dy * x] = r;
      ux[y + Dy * x] = u;
      uy[y + Dy * x] = v;
    }
  }
}

/* Forcing-mode interval (src/mpi_LBM_GK_forcing.cpp:335-400): M1 steps
 * with the per-site force of map, then M - M1 with the uniform one; the
 * drag of every step is scaled, shifted and summed as in orc_run_steps. */
double orc_run_steps_forced(const orc_geom* g, double* state, double* scratch,
                            int M1, int M, const double* map, double eps,
                            double scale, double shift, double dt,
                            double* A_out) {
  long N = 9L * g->Lx * g->Ly;
  double* cur = state;
  double* oth = scratch;
  double acc = 0.0;
  for (int n = 0; n < M; n++) {
    double F = step_impl(g, cur, oth, n < M1 ? map : NULL, eps);
    double* t = cur;
    cur = oth;
    oth = t;
    F = F * scale;
    F = F - shift;
    if (A_out) A_out[n] = F;
    acc += F;
  }
  if (cur != state) memcpy(state, cur, (size_t)N * sizeof(double));
  acc *= dt;
  return acc;
}

/* Forcing-mode field draw: complex standard normals Z_m = z[2m] + i z[2m+1]
 * for m < n, the pair m = 2p, 2p+1 from one Philox block with counter
 * (p, clone, interval, 0xF1E1D000 + instance) and key (seed_lo, seed_hi)
 * through orc_box_muller4. */
void orc_field_normals(uint32_t seed_lo, uint32_t seed_hi, uint32_t clone,
                       uint32_t interval, uint32_t instance, int n, double* z) {
  const uint32_t key[2] = {seed_lo, seed_hi};
  for (int p = 0; 2 * p < n; p++) {
    uint32_t ctr[4] = {(uint32_t)p, clone, interval, 0xF1E1D000u + instance};
    uint32_t r[4];
    double v[4];
    orc_philox4x32_10(ctr, key, r);
    orc_box_muller4(r, v);
    z[4 * p] = v[0];
    z[4 * p + 1] = v[1];
    if (2 * p + 1 < n) {
      z[4 * p + 2] = v[2];
      z[4 * p + 3] = v[3];
    }
  }
}

double orc_run_clones(const orc_geom* g, double* states, double* scratch,
                      int nclones, int M, double scale, double shift,
                      double dt, double* acc_out) {
  long N = 9L * g->Lx * g->Ly;
  double tot = 0.0;
  for (int j = 0; j < nclones; j++) {
    double a = orc_run_steps(g, states + N * j, scratch, M, scale, shift, dt,
                             NULL);
    if (acc_out) acc_out[j] = a;
    tot += a;
  }
  return tot;
}

/* src/perturbation_routines/generatePerturbedState.cpp:7-36, with the NN
 * drand48() prefactors passed in. */
void orc_perturb(double* fin, const double* snaps, int NN, long N, double eps,
                 const double* a) {
  double* fprim = (double*)calloc((size_t)N, sizeof(double));
  double sum_prefactors = 0.0;
  for (int nn = 0; nn < NN; nn++) {
    double rp = a[nn];
    sum_prefactors += rp;
    const double* P = snaps + (long)nn * N;
    for (long i = 0; i < N; i++) fprim[i] += rp * P[i];
  }
  double norm_0 = 0.0, norm_prim = 0.0;
  for (long i = 0; i < N; i++) {
    norm_0 += fin[i] * fin[i];
    norm_prim += fprim[i] * fprim[i];
  }
  norm_0 = sqrt(norm_0);
  norm_prim = sqrt(norm_prim);
  double weight = norm_0 / norm_prim;
  double A = 1.0 + sum_prefactors * eps * weight;
  double oneOvA = 1. / A;
  for (long i = 0; i < N; i++) {
    fin[i] += eps * fprim[i] * weight;
    fin[i] = fin[i] * oneOvA;
  }
  free(fprim);
}

/* Selection (src/mpi_LBM_GK_pops_perturb.cpp:489-497, 512-571, 590-621). */
int orc_select(int Nc, const double* s, int p_blocks, const double* u,
               const int32_t* rint, int32_t* nb_before, int32_t* nb_after,
               int32_t* pairs, int32_t* labels, double* total_R_out,
               int* n_rint_used) {
  if (Nc <= 0 || p_blocks <= 0 || Nc % p_blocks != 0) return -1;
  const int local = Nc / p_blocks;
  double total_R = 0.0;
  for (int r = 0; r < p_blocks; r++) {
    double R = 0.0;
    for (int j = r * local; j < (r + 1) * local; j++) R += s[j];
    if (r == 0)
      total_R = R;
    else
      total_R += R;
  }
  total_R /= Nc;
  int32_t* nb = (int32_t*)malloc(sizeof(int32_t) * (size_t)Nc);
  int32_t* temp = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)Nc);
  long NcPrime = 0;
  for (int j = 0; j < Nc; j++) {
    double v = floor(s[j] / total_R + u[j]);
    if (!(v >= 0.0 && v <= 2.0 * Nc)) {
      free(nb);
      free(temp);
      return -1;
    }
    nb[j] = (int32_t)v;
    NcPrime += nb[j];
  }
  if (nb_before) memcpy(nb_before, nb, sizeof(int32_t) * (size_t)Nc);
  long kk = 0;
  for (int j = 0; j < Nc; j++)
    for (int i = 0; i < nb[j]; i++) temp[kk++] = j;
  long deltaN = NcPrime - Nc;
  int used = 0;
  if (deltaN > 0) {
    for (long i = 0; i < deltaN; i++) {
      long idx = rint[used++] % (NcPrime - i);
      int c = temp[idx];
      temp[idx] = temp[NcPrime - i - 1];
      nb[c]--;
    }
  } else if (deltaN < 0) {
    for (long i = 0; i < -deltaN; i++) {
      long idx = rint[used++] % NcPrime;
      int c = temp[idx];
      temp[NcPrime + i] = temp[idx];
      nb[c]++;
    }
  }
  if (n_rint_used) *n_rint_used = used;
  if (nb_after) memcpy(nb_after, nb, sizeof(int32_t) * (size_t)Nc);
  int nbComm = 0;
  for (int j = 0; j < Nc; j++) labels[j] = j;
  for (int i = 0; i < Nc; i++) {
    while (nb[i] > 1) {
      int k = 0;
      while (nb[k] != 0) k++;
      pairs[2 * nbComm] = i;
      pairs[2 * nbComm + 1] = k;
      nb[k] = -1;
      nbComm++;
      labels[k] = i;
      nb[i]--;
    }
  }
  *total_R_out = total_R;
  free(nb);
  free(temp);
  return nbComm;
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, "Parallel random
 * numbers: as easy as 1, 2, 3"): 10 rounds of the 4x32 S-box with Weyl key
 * schedule. */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                       uint32_t out[4]) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; r++) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    uint64_t p0 = (uint64_t)M0 * c0;
    uint64_t p1 = (uint64_t)M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

void orc_box_muller4(const uint32_t r[4], double n[4]) {
  const double two32inv = 2.3283064365386963e-10; /* 2^-32 */
  const double twopi = 6.283185307179586;
  for (int h = 0; h < 2; h++) {
    double u1 = ((double)r[2 * h] + 1.0) * two32inv;
    double u2 = (double)r[2 * h + 1] * two32inv;
    double rad = sqrt(-2.0 * log(u1));
    double ang = twopi * u2;
    n[2 * h] = rad * cos(ang);
    n[2 * h + 1] = rad * sin(ang);
  }
}

/* Bits [off, off + len) of a Philox block read as one 128-bit number, r[0]
 * the least significant word. */
static uint32_t draw_bits(const uint32_t r[4], int off, int len) {
  const int i = off / 32, sh = off % 32;
  const uint64_t lo = r[i], hi = i < 3 ? r[i + 1] : 0u;
  const uint64_t two = (hi << 32) | lo;
  return (uint32_t)(two >> sh) & ((1u << len) - 1u);
}

void orc_site_normals(const uint32_t r[4], double n[10]) {
  const double twopi = 6.283185307179586;
  for (int p = 0; p < 5; p++) {
    double u1 = ((double)draw_bits(r, 25 * p, 13) + 1.0) * (1.0 / 8192.0);
    double u2 = (double)draw_bits(r, 25 * p + 13, 12) * (1.0 / 4096.0);
    double rad = sqrt(-2.0 * log(u1));
    double ang = twopi * u2;
    n[2 * p] = rad * cos(ang);
    n[2 * p + 1] = rad * sin(ang);
  }
}

void orc_noise(const orc_geom* g, double* f, double sigma, uint32_t seed_lo,
               uint32_t seed_hi, uint32_t clone, uint32_t interval,
               uint32_t instance) {
  const uint32_t key[2] = {seed_lo, seed_hi};
  for (int x = 0; x < g->Lx; x++) {
    for (int y = 0; y < g->Ly; y++) {
      if (is_interior(g, x, y)) continue;
      uint32_t site = (uint32_t)(y * g->Lx + x);
      uint32_t ctr[4] = {site, clone, interval, instance};
      uint32_t r[4];
      double n[10];
      orc_philox4x32_10(ctr, key, r);
      orc_site_normals(r, n);
      for (int q = 0; q < 9; q++) f[IDX(g, x, y, q)] += sigma * n[q];
    }
  }
}

void orc_init_channel(const orc_geom* g, double* f, double amp, uint32_t seed_lo,
                      uint32_t seed_hi, uint32_t clone, uint32_t instance) {
  const uint32_t key[2] = {seed_lo, seed_hi};
  const double two32inv = 2.3283064365386963e-10;
  for (int x = 0; x < g->Lx; x++) {
    for (int y = 0; y < g->Ly; y++) {
      double feq[9];
      if (is_interior(g, x, y)) {
        for (int k = 0; k < 9; k++) f[IDX(g, x, y, k)] = orc_w[k];
        continue;
      }
      uint32_t site = (uint32_t)(y * g->Lx + x);
      uint32_t ctr[4] = {site, clone, 0xFFFFFFFFu, instance};
      uint32_t r[4];
      orc_philox4x32_10(ctr, key, r);
      double ux = g->u_in + amp * (2.0 * ((double)r[0] * two32inv) - 1.0);
      double uy = amp * (2.0 * ((double)r[1] * two32inv) - 1.0);
      orc_feq(1.0, ux, uy, feq);
      for (int k = 0; k < 9; k++) f[IDX(g, x, y, k)] = feq[k];
    }
  }
}

/* post_processing/read_rvalue.py:14-21: the inner sum restarts from zero
 * for every i and adds log r[1..i-1] in order, which is the running sum
 * below; T += 0.3 per i there, dT here; lam = (1/T) * s. */
void orc_scgf(const double* logR, int n, double dT, double* lam) {
  double s = 0.0, T = 0.0;
  if (n > 0) lam[0] = 0.0;
  for (int i = 1; i < n; i++) {
    if (i >= 2) s = s + logR[i - 1];
    T = T + dT;
    const double oneOvT = 1. / T;
    lam[i] = oneOvT * s;
  }
}

void orc_glibc_srand(unsigned seed) { srand(seed); }
void orc_glibc_rand_n(int n, int32_t* out) {
  for (int i = 0; i < n; i++) out[i] = rand();
}
/* drand48's state in a process that never seeds it, as the reference's ranks
 * do (src/mpi_LBM_GK_pops_perturb.cpp:517, generatePerturbedState.cpp:15):
 * glibc zero-fills the static state and its first use installs only the
 * multiplier and addend, so X starts at 0 (the first draw is 11 / 2^48), not
 * at the 0x1234ABCD330E that srand48-style seeding would give. */
void orc_glibc_seed48_default(void) {
  unsigned short s[3] = {0, 0, 0};
  seed48(s);
}
void orc_glibc_drand48_n(int n, double* out) {
  for (int i = 0; i < n; i++) out[i] = drand48();
}
