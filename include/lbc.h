/*
 * lbc.h -- C ABI of the MI355X cloning-sampler inner loop.
 *
 * N D2Q9 lattice-Boltzmann clones of flow past a square obstacle advance
 * together on one GPU; their drag is time-integrated into FP64 log-weights
 * and the population is resampled, copied and perturbed on the device at
 * every cloning step.  One context may hold several independent sampler
 * instances (lattice size, bias, clone count, steps per interval) that share
 * one physics mode and one storage precision; all of them advance in one
 * persistent kernel launch per cloning interval.
 *
 * In the reference (tlestang/cloning-lbm-MPI) this path is a set of free C++
 * functions over caller-owned double arrays plus inline code in the MPI
 * driver's main(); each entry point below names what it replaces.
 *
 * Conventions
 *  - Every int-returning function returns LBC_OK (0) or a negative LBC_E_*
 *    code; lbc_last_error() gives the message of the last failure.
 *  - Host arrays are caller-owned and only read/written during the call.
 *    Device arrays are library-owned.
 *  - Population arrays crossing this boundary are FP64 in the reference's
 *    AoS layout IDX(x,y,k) = k + 9*(y + Ly*x) (src/lbm_src/global.h:4),
 *    one lattice after another.
 *  - A context is single-threaded.  All device work is ordered on the
 *    context's own stream; functions that return host data synchronise it.
 *  - Every call makes the context's device current and restores the
 *    caller's current device before it returns.
 */
#ifndef LBC_H
#define LBC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBC_OK 0
#define LBC_E_ARG -1     /* invalid argument or configuration */
#define LBC_E_HIP -2     /* HIP runtime error */
#define LBC_E_NOMEM -3   /* device allocation failed */
#define LBC_E_STATE -4   /* call not valid in the current state */
#define LBC_E_NUMERIC -5 /* non-finite weight or population detected */

/* Physics modes. */
#define LBC_MODE_REF 0     /* periodic x, body force, stress-integral drag */
#define LBC_MODE_CHANNEL 1 /* equilibrium inlet, zero-gradient outlet,
                              momentum-exchange drag */

/* Population storage.  F64 stores f; the others store f - w_q (shifted by
 * the rest weights) and collide in FP32. */
#define LBC_STORE_F64 0
#define LBC_STORE_F32 1
#define LBC_STORE_F16 2
#define LBC_STORE_BF16 3

typedef struct lbc_instance {
  int Lx, Ly;                 /* lattice (Dx, Dy in the reference); one
                                 clone's populations must stay below 2 GiB
                                 in the chosen storage (FP64: about 29.8M
                                 sites) */
  int xmin, xmax, ymin, ymax; /* square ring nodes; solid = open interior */
  int n_clones;               /* Nc */
  int steps;                  /* LBM steps per cloning interval (M) */
  double tau;                 /* BGK relaxation time */
  double beta;                /* REF: uniform body force (beta0) */
  double u_in;                /* CHANNEL: inlet velocity */
  double force_scale;         /* A_n = F_n*force_scale - force_shift */
  double force_shift;
  double dt;                  /* interval observable = dt * sum_n A_n */
  double bias;                /* log-weight = bias * observable (alpha, k) */
  double eps;                 /* REF: snapshot-perturbation amplitude */
  double noise_sigma;         /* CHANNEL: population noise after a copy */
  int rank_blocks;            /* total_R summed in this many equal blocks (>=1) */
  int n_snapshots;            /* REF: snapshots mixed by the perturbation (NN) */
  /* Forcing-mode perturbation (src/mpi_LBM_GK_forcing.cpp:333-395; REF
   * mode): the first forced_steps steps of every interval (lbmTimeSteps1)
   * run with the body force (1 + field_eps*map(x,y))*beta, map a fresh
   * Gaussian random field per clone and interval drawn on the device by
   * circulant embedding of the covariance exp(-r^2/(2 field_sigma^2)) in a
   * (2Lx-1) x (2Ly-1) torus, negative eigenvalues set to zero
   * (src/perturbation_routines/generate_random_field.cpp:9-181;
   * field_sigma 0 = the reference's 32).  Clones 2k and 2k+1 take the real
   * and imaginary parts of one transform: two independent fields of the
   * same law.  The shifted storages transform in single precision and
   * collide with FP32 maps.
   * forced_steps = 0: off. */
  int forced_steps;
  double field_eps;
  double field_sigma;
  /* Macroscopic output (src/mpi_LBM_GK_pops_perturb.cpp:450-455 with
   * src/lbm_src/write_vtk.cpp): with macro_every = E > 0 every interval
   * stores rho, ux and uy of every clone at steps 0, E, 2E, ... < steps,
   * the fields the collision of that step computes
   * (src/lbm_src/streamCollCompute.cpp:17-26,44-45: ux includes the
   * beta/(2 rho) shift; the solid interior holds the rest state), for
   * lbc_get_macro.  0 = off. */
  int macro_every;
} lbc_instance;

typedef struct lbc_config {
  int mode;                   /* LBC_MODE_* */
  int storage;                /* LBC_STORE_* */
  int n_instances;
  int device;                 /* HIP device ordinal */
  uint64_t seed;              /* Philox key */
  int threads;                /* workgroup size of the step kernel; 0 = tuned */
  int flags;                  /* step-kernel shape, 0 = tuned with threads:
                                 bits 0-2 sites per thread per iteration (1),
                                 bits 3-4 non-temporal population stores
                                 (1 off, 2 on), bits 5-6 LBM passes per HBM
                                 sweep, chained through LDS rings (1, 2 or
                                 3), bits 7-8 clones per workgroup (1, or 2
                                 with one pass per sweep; 0: tuned where
                                 the tuner runs, else 1); other bits 0.
                                 Every shape gives the same results. */
  const lbc_instance* instances;
} lbc_config;

/* Host-visible result of one selection (all arrays caller-owned, sized for
 * the instance's n_clones; any pointer may be NULL). */
typedef struct lbc_selection {
  int32_t* nb_before;  /* copies per clone after the floor step */
  int32_t* nb_after;   /* after prune/add (sums to n_clones) */
  int32_t* pairs;      /* [2*n_clones]: (source, destination) per copy */
  int32_t* labels;     /* parent of every slot */
  int32_t n_pairs;
  int32_t n_rint_used; /* prune/add draws consumed */
  double total_R;      /* mean weight of the interval */
  double log_total_R;  /* log of the mean weight */
} lbc_selection;

typedef struct lbc_ctx lbc_ctx;

/* Create a context: allocates every device buffer and, if threads == 0,
 * picks the workgroup size by timing.  Replaces the per-rank allocations of
 * src/mpi_LBM_GK_pops_perturb.cpp:180-199 and the geometry globals of
 * src/lbm_src/global.h:11-15. */
int lbc_create(const lbc_config* cfg, lbc_ctx** out);
void lbc_destroy(lbc_ctx* ctx);
const char* lbc_last_error(const lbc_ctx* ctx);

/* Copy n lattices (AoS FP64) into clones clone0..clone0+n-1.
 * Replaces the pops-file reads of src/mpi_LBM_GK_pops_perturb.cpp:291,305. */
int lbc_load_states(lbc_ctx* ctx, int inst, int clone0, int n,
                    const double* aos);
/* Read n lattices back (AoS FP64, unshifted): the layout of the pops files
 * the reference's warm-up program writes (src/lbm_src/main_bodyforce.cpp:
 * 179-182) and its cloning driver reads. */
int lbc_store_states(lbc_ctx* ctx, int inst, int clone0, int n, double* aos);

/* Equilibrium initial state feq(rho=1, u=(u_in + a(2U-1), a(2U-1))) per
 * fluid site, Philox-seeded, rest weights on the solid interior.  No
 * reference counterpart (channel configuration); the rest-state case a=0,
 * u_in=0 is src/lbm_src/initialize_lattice_arrays.cpp:7-20. */
int lbc_init_equilibrium(lbc_ctx* ctx, int inst, double amp);

/* Snapshot states mixed by the REF perturbation (n_snapshots AoS lattices).
 * Replaces popsForPerturb (src/mpi_LBM_GK_pops_perturb.cpp:321-364). */
int lbc_set_snapshots(lbc_ctx* ctx, int inst, const double* aos);

/* Clones marked for perturbation at the start of the next interval
 * (mark_perturb, src/mpi_LBM_GK_pops_perturb.cpp:204-211,642-648,676,693).
 * Selection sets them for every overwritten slot. */
int lbc_set_marks(lbc_ctx* ctx, int inst, const uint8_t* marks);
int lbc_get_marks(lbc_ctx* ctx, int inst, uint8_t* marks);

/* One cloning interval of every instance: perturb marked clones (REF), then
 * `steps` LBM steps of every clone with the drag integrated into FP64
 * log-weights.  prefactors: NULL (device Philox draws) or, per instance in
 * order, n_clones*n_snapshots uniforms (rows of unmarked clones unused).
 * Replaces src/mpi_LBM_GK_pops_perturb.cpp:399-470: generatePerturbedState
 * (src/perturbation_routines/generatePerturbedState.cpp:4-38),
 * streamingAndCollisionComputeMacroBodyForce
 * (src/lbm_src/streamCollCompute.cpp:7-48), computeDomainNoSlipWalls_BB
 * (src/lbm_src/domain_noSlipWalls.cpp:6-22), computeSquareBounceBack_TEST
 * (src/lbm_src/square.cpp:24-75), the interior reset (:430-439),
 * computeForceOnSquare (src/lbm_src/force.cpp:8-144) and the weight
 * accumulation (:445-469). */
int lbc_run_interval(lbc_ctx* ctx, const double* prefactors);

/* Per-step observable A_n = F_n*scale - shift of the last interval,
 * [n_clones][steps] (what the reference appends to data_force.datout,
 * src/mpi_LBM_GK_pops_perturb.cpp:456). */
int lbc_get_drag(lbc_ctx* ctx, int inst, double* A);
/* Interval observable dt*sum A (s_ at :465) and log-weight bias*that. */
int lbc_get_logweights(lbc_ctx* ctx, int inst, double* obs, double* logw);

/* Selection and copy table for one instance.
 * weights: NULL (exp(logw) on device) or n_clones host weights.
 * u: NULL (Philox) or n_clones uniforms in [0,1) for the floor step.
 * rint: NULL (Philox) or n_clones non-negative ints for prune/add.
 * Replaces the master's gather and selection,
 * src/mpi_LBM_GK_pops_perturb.cpp:487-632, and the copy-table broadcast
 * :636-640. */
int lbc_select(lbc_ctx* ctx, int inst, const double* weights, const double* u,
               const int32_t* rint, lbc_selection* out);

/* Copy every selected parent into its destination slot (with CHANNEL noise)
 * and mark the destinations.  Replaces the memcpy / MPI_Send / MPI_Recv
 * loop of src/mpi_LBM_GK_pops_perturb.cpp:658-701. */
int lbc_gather(lbc_ctx* ctx);

/* The same copies, in CHANNEL mode with noise on the shifted storages taken
 * straight into the next interval: each noisy copy is collided into that
 * interval's collide-only pass 0 instead of being stored, which saves the
 * copy's own read and write of the destination's populations (FP64, other
 * modes and instances with macroscopic output copy as lbc_gather).  No counterpart in the reference; for drivers that run the
 * interval, the selection and the copy one call at a time and read no
 * states between a copy and the next interval (lbc_advance does the same
 * internally).  Until the next lbc_run_interval or lbc_advance, the calls
 * that read, replace or reselect states (lbc_store_states,
 * lbc_load_states, lbc_init_equilibrium, lbc_select, lbc_gather,
 * lbc_gather_fused) return LBC_E_STATE. */
int lbc_gather_fused(lbc_ctx* ctx);

/* n full cloning steps (interval, selection, copy) of every instance with
 * device Philox draws, replayed from one captured HIP graph. */
int lbc_advance(lbc_ctx* ctx, int n_intervals);

/* Mean weight of every interval so far (R_record,
 * src/mpi_LBM_GK_pops_perturb.cpp:621, written to rvalues.dat at :705-714). */
int lbc_get_R_history(lbc_ctx* ctx, int inst, double* R, int max_n,
                      int* n_out);

/* Running large-deviation (scaled cumulant generating function) estimate
 * after every interval so far, computed in FP64 on the device from the log
 * mean weights the selection kept, in the form of
 * post_processing/read_rvalue.py:14-21: lam[0] = 0 and, for i >= 1,
 * lam[i] = (1/T_i) * sum_{j=1}^{i-1} log R_j with T_i = dT + ... + dT
 * (i terms; the script hard-codes dT = 0.3).  *n_out receives the number of
 * intervals; at most max_n values are written. */
int lbc_get_scgf(lbc_ctx* ctx, int inst, double dT, double* lam, int max_n,
                 int* n_out);

/* Forcing mode: the random fields of the next interval of instance inst,
 * instead of drawing them on the device: n_clones maps of Lx*Ly doubles in
 * the reference's field layout map[y + Ly*x] (the map argument of
 * src/lbm_src/streamCollCompute.cpp:50-93).  Used once. */
int lbc_set_field_maps(lbc_ctx* ctx, int inst, const double* maps);

/* The random fields the last interval of instance inst ran with, in the same
 * layout, in FP64 (FP32 values widened for the shifted storages). */
int lbc_get_field_maps(lbc_ctx* ctx, int inst, double* maps);

/* Macroscopic fields of the last interval of clones [clone0, clone0 + n) of
 * instance inst (macro_every > 0): per clone, frames = ceil(steps /
 * macro_every) frames (frame k at step k*macro_every), each three planes
 * rho, ux, uy of Lx*Ly doubles in row order y*Lx + x (the point order of
 * the reference's VTK writer, src/lbm_src/write_vtk.cpp:38-49). */
int lbc_get_macro(lbc_ctx* ctx, int inst, int clone0, int n, double* out);

/* Wait for the context's stream (the reference's MPI_Barrier points, e.g.
 * src/mpi_LBM_GK_pops_perturb.cpp:471,507, have no other counterpart here). */
int lbc_sync(lbc_ctx* ctx);

/* Average device time (ms) of the step kernel over the last
 * lbc_run_interval / lbc_advance call, from HIP events on the context's
 * stream, and the number of launches it averaged. */
int lbc_step_kernel_ms(lbc_ctx* ctx, double* avg_ms, int* n_launches);

/* Device time (ms) of the copy kernel of the last lbc_gather call, from HIP
 * events on the context's stream, and the population bytes it moved: each
 * copy reads the parent's 9*Lx*Ly stored populations and writes the
 * destination's, with the copies counted by every instance's last
 * lbc_select.  Instrumentation for the copy's HBM roofline, like
 * lbc_step_kernel_ms; the reference times nothing. */
int lbc_gather_kernel_ms(lbc_ctx* ctx, double* ms, double* bytes);

/* Workgroup size and sites per thread the context uses for the step
 * kernel. */
int lbc_threads(const lbc_ctx* ctx);
int lbc_sites_per_thread(const lbc_ctx* ctx);

/* The step-kernel shape the context settled on (tuned or pinned), in the
 * encoding of lbc_config.flags with every field resolved: sites per thread
 * in bits 0-2, non-temporal stores in bits 3-4 (1 off, 2 on), LBM passes
 * per HBM sweep in bits 5-6 (1, 2 or 3), clones per workgroup in bits 7-8
 * (1 or 2).  0 for a null context. */
int lbc_step_shape(const lbc_ctx* ctx);

/* Largest population of lbc_select_static. */
#define LBC_STATIC_MAX_CLONES 8192

/* The selection of lbc_select (weights NULL, caller's draws) for one
 * population of n clones whose inputs and outputs all live in device memory
 * of the current device, without a context and without allocating: it calls
 * no hipMalloc, hipFree or other allocating routine.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight sampling loops that hold their own
 * buffers).
 *   logw:   n log-weights; the weights are exp(logw), shifted by the largest
 *           log-weight only where exp would leave its range (as lbc_select)
 *   u:      n floor-step uniforms in [0,1);  rint: n non-negative draws for
 *           the prune/add steps (as lbc_select's u and rint)
 *   rank_blocks: total_R summed in that many sequential blocks
 *           (lbc_instance.rank_blocks)
 *   outputs: nb_before[n], nb_after[n], pairs[2n] (the first info[0] pairs
 *           are set), labels[n], marks[n] (1 for an overwritten slot),
 *           info[4] = {n_pairs, n_rint_used, status (0, or LBC_E_NUMERIC
 *           for non-finite weights: every clone kept), 0 (unused)} and
 *           R[2] = {total_R, log total_R}.
 * n is at most LBC_STATIC_MAX_CLONES (larger n: LBC_E_ARG, nothing is
 * launched).  The kernel keeps the FP64 weights in the library's static
 * device scratch (LBC_STATIC_MAX_CLONES doubles, 64 KiB, held on every
 * device the library runs on until the process ends), which every call on
 * a device shares: calls that overlap on one device race, so order them on
 * one stream.  The work is enqueued on `stream` (a hipStream_t; NULL for
 * the null stream) and the call returns without waiting; errors of argument
 * checks and launch come back as the return code (text from
 * lbc_last_error(NULL)), numeric ones in info[2] once the stream has run. */
int lbc_select_static(int n, int rank_blocks, const double* logw,
                      const double* u, const int32_t* rint, int32_t* nb_before,
                      int32_t* nb_after, int32_t* pairs, int32_t* labels,
                      uint8_t* marks, int32_t* info, double* R, void* stream);

/* One reference LBM step on one lattice held in host memory (FP64):
 * fin (post-stream state) becomes the post-collision state, fout receives
 * the next post-stream state, *force the raw drag on it.  Parity hook for
 * the sequence at src/mpi_LBM_GK_pops_perturb.cpp:426-445. */
int lbc_step_host(int mode, const lbc_instance* geom, double* fin,
                  double* fout, double* force);

#ifdef __cplusplus
}
#endif
#endif
