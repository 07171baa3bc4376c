// Harness that links the reference's own LBM kernel translation units
// (compiled from /root/reference by oracle/Makefile into oracle/_ref/) and
// exposes them with C linkage for the tests.
//
// TEST INFRASTRUCTURE ONLY: the product never loads oracle/_ref.
//
// The reference kernels read their geometry and lattice constants from
// globals that the reference's driver defines
// (src/mpi_LBM_GK_pops_perturb.cpp:31-33); this harness plays the driver's
// role and defines them with the same values.
#include <cstdlib>

void streamingAndCollisionComputeMacroBodyForce(double*, double*, double*,
                                                double*, double*, double,
                                                double);
void streamingAndCollisionComputeMacroBodyForceSpatial(double*, double*,
                                                       double*, double*,
                                                       double*, double,
                                                       double*, double,
                                                       double);
void computeDomainNoSlipWalls_BB(double*, double*);
void computeSquareBounceBack_TEST(double*, double*);
double computeForceOnSquare(double*, double);
int generatePerturbedState(double*, double**, double, int, int);

int c[9][2] = {{0, 0}, {1, 0}, {0, 1}, {-1, 0}, {0, -1},
               {1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
double w[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
               1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
int Dx, Dy, xmin, xmax, ymin, ymax;

extern "C" {

void ref_set_geometry(int dx, int dy, int x0, int x1, int y0, int y1) {
  Dx = dx;
  Dy = dy;
  xmin = x0;
  xmax = x1;
  ymin = y0;
  ymax = y1;
}

// One production step as called at src/mpi_LBM_GK_pops_perturb.cpp:426-439:
// stream-collide, walls, square, interior reset.  The caller swaps buffers.
void ref_step(double* fin, double* fout, double* rho, double* ux, double* uy,
              double beta, double tau) {
  streamingAndCollisionComputeMacroBodyForce(fin, fout, rho, ux, uy, beta, tau);
  computeDomainNoSlipWalls_BB(fout, fin);
  computeSquareBounceBack_TEST(fout, fin);
  for (int x = xmin + 1; x < xmax; x++)
    for (int y = ymin + 1; y < ymax; y++)
      for (int k = 0; k < 9; k++) fout[k + 9 * (y + Dy * x)] = w[k];
}

// One forced step as called at src/mpi_LBM_GK_forcing.cpp:337-351: the
// spatial-force stream-collide, walls, square, interior reset.
void ref_step_spatial(double* fin, double* fout, double* rho, double* ux,
                      double* uy, double beta0, double* map, double tau,
                      double eps) {
  streamingAndCollisionComputeMacroBodyForceSpatial(fin, fout, rho, ux, uy,
                                                    beta0, map, tau, eps);
  computeDomainNoSlipWalls_BB(fout, fin);
  computeSquareBounceBack_TEST(fout, fin);
  for (int x = xmin + 1; x < xmax; x++)
    for (int y = ymin + 1; y < ymax; y++)
      for (int k = 0; k < 9; k++) fout[k + 9 * (y + Dy * x)] = w[k];
}

double ref_force(double* f, double omega) {
  return computeForceOnSquare(f, omega);
}

// The driver's per-clone interval loop (src/mpi_LBM_GK_pops_perturb.cpp:
// 399-449) over n clones held back to back in `states` (each Dx*Dy*9
// doubles, post-stream form in and out): M production steps and the force
// on every new state, summed into acc[j] (raw F; the driver's 1/F0 and
// AVG are two more flops per step).  scratch holds one state.  Used to time
// the reference's own kernels; the geometry globals must be set first and
// stay fixed while several threads run this on their own clones.
void ref_run_clones(double* states, double* scratch, int n, int M,
                    double beta, double tau, double* acc) {
  const long N = 9L * Dx * Dy;
  const long S = (long)Dx * Dy;
  double* fields = (double*)std::malloc(3 * S * sizeof(double));
  if (!fields) return;
  for (int j = 0; j < n; j++) {
    double* fin = states + j * N;
    double* fout = scratch;
    double sum = 0.0;
    for (int m = 0; m < M; m++) {
      ref_step(fin, fout, fields, fields + S, fields + 2 * S, beta, tau);
      sum += computeForceOnSquare(fout, 1.0 / tau);
      double* t = fin;
      fin = fout;
      fout = t;
    }
    if (fin != states + j * N)
      for (long i = 0; i < N; i++) states[j * N + i] = fin[i];
    acc[j] = sum;
  }
  std::free(fields);
}

// generatePerturbedState draws its NN prefactors from drand48(); the caller
// controls that stream with ref_seed48 / ref_drand48.
void ref_perturb(double* fin, double* snaps, double eps, int NN, int N) {
  double* ptrs[64];
  for (int i = 0; i < NN && i < 64; i++) ptrs[i] = snaps + (long)i * N;
  generatePerturbedState(fin, ptrs, eps, NN, N);
}

void ref_seed48(unsigned short s0, unsigned short s1, unsigned short s2) {
  unsigned short s[3] = {s0, s1, s2};
  seed48(s);
}

double ref_drand48(void) { return drand48(); }
}
