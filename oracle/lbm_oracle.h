/*
 * CPU FP64 oracle for the cloning-sampler inner loop (D2Q9 LBM clones, drag,
 * weights, selection, perturbation).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library or driver links,
 * loads or calls this code; it is the checker used by tests/, by the smoke
 * test and by bench.py's CPU-baseline leg.
 *
 * Layout: the reference's AoS population layout, IDX(x,y,k) = k + 9*(y + Dy*x)
 * (src/lbm_src/global.h:4), FP64.
 *
 * Two physics modes share geometry and the cloning machinery:
 *   ORC_MODE_REF     periodic x, uniform body force beta with Guo forcing,
 *                    half-way bounce-back walls and square, interior reset,
 *                    stress-integral drag (the reference's production path);
 *   ORC_MODE_CHANNEL equilibrium velocity inlet (column x=0 is reset to
 *                    feq(rho_in, (u_in,0)) after every stream, rho_in from the
 *                    populations streamed in from the domain), zero-gradient
 *                    outlet (ghost column x=Lx copies column Lx-1), same walls
 *                    and square, momentum-exchange drag.  No reference
 *                    counterpart.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MODE_REF 0
#define ORC_MODE_CHANNEL 1

typedef struct {
  int mode;
  int Lx, Ly;                    /* Dx, Dy of the reference */
  int xmin, xmax, ymin, ymax;    /* square ring; interior is open box */
  double tau;
  double beta;                   /* REF body force (Guo) */
  double u_in;                   /* CHANNEL inlet velocity */
} orc_geom;

/* D2Q9 velocity set and weights, in the reference's ordering. */
extern const int orc_c[9][2];
extern const double orc_w[9];
extern const int orc_opp[9];

/* Bounce-back mask of site (x,y): bit q set when population q arriving at
 * (x,y) is reflected from (x,y)'s own opposite population. */
int orc_bb_mask(const orc_geom* g, int x, int y);

/* One LBM step in push form with explicit boundary passes.  fin holds the
 * post-stream state S_n on entry and the post-collision state P_n on exit;
 * fout receives S_{n+1}.  Returns the raw drag on S_{n+1} (stress integral in
 * REF mode, momentum exchange in CHANNEL mode). */
double orc_step(const orc_geom* g, double* fin, double* fout);

/* The same step with the forcing mode's per-site body force
 * beta(x,y) = (1.0 + eps*map[y + Ly*x])*beta (REF mode;
 * src/lbm_src/streamCollCompute.cpp:50-93); map NULL = orc_step. */
double orc_step_map(const orc_geom* g, double* fin, double* fout,
                    const double* map, double eps);

/* Forcing-mode interval (src/mpi_LBM_GK_forcing.cpp:335-400): M1 steps with
 * map, then M - M1 uniform ones; outputs as orc_run_steps. */
double orc_run_steps_forced(const orc_geom* g, double* state, double* scratch,
                            int M1, int M, const double* map, double eps,
                            double scale, double shift, double dt,
                            double* A_out);

/* Complex normals of one clone's random-field draw: z[2m] + i z[2m+1],
 * m < n; pair p = (2p, 2p+1) from Philox counter
 * (p, clone, interval, 0xF1E1D000 + instance), key (seed_lo, seed_hi),
 * orc_box_muller4. */
void orc_field_normals(uint32_t seed_lo, uint32_t seed_hi, uint32_t clone,
                       uint32_t interval, uint32_t instance, int n, double* z);

/* Macroscopic fields of state f as the step's collision computes them
 * (src/lbm_src/streamCollCompute.cpp:17-26,44-45): rho, ux/rho + beta/(2 rho)
 * (REF; beta per site with map, NULL = uniform; CHANNEL beta = 0), uy/rho,
 * each Lx*Ly doubles at idx(x,y) = y + Ly*x. */
void orc_macro(const orc_geom* g, const double* f, const double* map,
               double eps, double* rho, double* ux, double* uy);

/* Drag functionals on a post-stream state. */
double orc_force_stress(const orc_geom* g, const double* f);
double orc_force_mea(const orc_geom* g, const double* f);

/* M steps on `state` (S form in and out; scratch of the same size).
 * A_out[n] = F_{n+1}*scale - shift (may be NULL); returns
 * dt * sum_n A_n, summed sequentially. */
double orc_run_steps(const orc_geom* g, double* state, double* scratch, int M,
                     double scale, double shift, double dt, double* A_out);

/* Snapshot-mix perturbation with injected prefactors a[NN]:
 * f' = sum_n a_n P_n; r = |fin|/|f'|; fin <- (fin + eps*r*f') / (1 + eps*r*sum a).
 * snaps is NN contiguous states of N doubles. */
void orc_perturb(double* fin, const double* snaps, int NN, long N, double eps,
                 const double* a);

/* Selection with injected draws.  s[Nc] weights, u[Nc] uniforms for the floor
 * step, rint[] non-negative ints consumed one per prune/add iteration
 * (at most Nc are needed).  p_blocks: number of equal rank blocks for the
 * total_R summation order (1 = sequential).  Outputs: nb_before[Nc] (after
 * the floor step), nb_after[Nc] (after prune/add), pairs[2*Nc] (src,dst),
 * labels[Nc], *total_R.  Returns the number of pairs, or -1 on bad input.
 * *n_rint_used receives the number of rint entries consumed. */
int orc_select(int Nc, const double* s, int p_blocks, const double* u,
               const int32_t* rint, int32_t* nb_before, int32_t* nb_after,
               int32_t* pairs, int32_t* labels, double* total_R,
               int* n_rint_used);

/* Philox4x32-10 block (Salmon et al., SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);

/* Standard normals from one Philox block via Box-Muller:
 * u1 = (r0+1)*2^-32, u2 = r1*2^-32 -> (n0, n1); same for (r2, r3). */
void orc_box_muller4(const uint32_t r[4], double n[4]);

/* Nine normals (n[0..8]; n[9] unused) of one site's copy noise from one
 * Philox block r, as five Box-Muller pairs: pair p takes bits
 * [25p, 25p + 25) of r read as one 128-bit number (r[0] least significant):
 * a 13-bit a, u1 = (a + 1) 2^-13, then a 12-bit b, u2 = b 2^-12;
 * n[2p] = sqrt(-2 ln u1) cos(2 pi u2), n[2p+1] = ... sin(2 pi u2). */
void orc_site_normals(const uint32_t r[4], double n[10]);

/* CHANNEL population noise of one clone: every fluid population q of site
 * y*Lx + x gets sigma * n[q] of orc_site_normals on the Philox block with
 * counter (site, clone, interval, instance), key (seed_lo, seed_hi). */
void orc_noise(const orc_geom* g, double* f, double sigma, uint32_t seed_lo,
               uint32_t seed_hi, uint32_t clone, uint32_t interval,
               uint32_t instance);

/* CHANNEL initial state of one clone: feq(1, (u_in + a*(2U-1), a*(2U-1)))
 * on fluid sites, w on the solid interior; counter (site, clone, 0xFFFFFFFF,
 * instance). */
void orc_init_channel(const orc_geom* g, double* f, double amp, uint32_t seed_lo,
                      uint32_t seed_hi, uint32_t clone, uint32_t instance);

/* Equilibrium of (rho, ux, uy) in the reference's arithmetic order. */
void orc_feq(double rho, double ux, double uy, double feq[9]);

/* Running scaled-cumulant estimate from the per-interval log mean weights,
 * as post_processing/read_rvalue.py:14-21 forms it: lam[0] = 0 and, for
 * i >= 1, T accumulates dT by repeated addition and
 * lam[i] = (1/T) * sum_{j=1}^{i-1} logR[j] (logR[0] never enters). */
void orc_scgf(const double* logR, int n, double dT, double* lam);

/* glibc RNG access (the reference's own generators). */
void orc_glibc_srand(unsigned seed);
void orc_glibc_rand_n(int n, int32_t* out);
void orc_glibc_seed48_default(void);
void orc_glibc_drand48_n(int n, double* out);

/* Time of an OpenMP-free, single-thread AoS run: clone-LUPS measurement
 * helper for the CPU baseline (runs `nclones` independent clones for M
 * steps each, returns the accumulated sum of A over all clones). */
double orc_run_clones(const orc_geom* g, double* states, double* scratch,
                      int nclones, int M, double scale, double shift,
                      double dt, double* acc_out);

#ifdef __cplusplus
}
#endif
#endif
