/*
 * oracle.h -- CPU restatement of the reference's 2D electrostatic PIC step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (libpicamd.so) never links or calls it.
 *
 * Two flavours of every particle stage are provided:
 *   - "f64": the reference algorithm (thesisCode/collisionless/double) with the
 *     intended physics (periodic CIC, spectral Poisson, central-difference E,
 *     leapfrog), continuous CIC weights, plain double accumulation;
 *   - "mirror": the exact operation sequence the GPU kernels use (FP32 or FP64
 *     particles, per-axis fixed-point CIC weights, int64 accumulation), so
 *     integer outputs and fixed-point charge can be compared bit-for-bit.
 *
 * Pinning of each function (tests/test_oracle_*.py):
 *   cell keys / stable sort ........ brute force std::stable_sort, bijection
 *   deposit ........................ partition of unity, single-particle
 *                                    stencils, reference grid.cc via oracle/_ref
 *                                    (reflected-weight bridge)
 *   FFT / Poisson .................. brute-force DFT, single-mode closed form,
 *                                    numpy fixtures (tests/golden)
 *   E-field ........................ closed form sin(k dx)/dx factor
 *   gather / push .................. constant-field ballistic orbit, zero
 *                                    self-force, reference particle.cc via _ref
 *   whole step ..................... Landau damping rate, two-stream growth
 *   Philox4x32-10 .................. Random123 known-answer vectors
 *   Philox2x32-10 .................. Random123 known-answer vectors
 *   std engines .................... ISO C++ [rand.predef] 10000th values
 *   MCC / DSMC ..................... closed-form rates, conservation laws
 *   adjoint ........................ dot-product tests, central FD
 */
#ifndef PIC_ORACLE_H
#define PIC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- random numbers ---------------- */
/* Philox4x32-10 (Salmon et al., SC'11). */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Philox2x32-10 (same paper; key[0] is its single 32-bit key word). */
void or_philox2x32_10(const uint32_t ctr[2], const uint32_t key[1], uint32_t out[2]);
/* n-th output (1-based) of std::minstd_rand0 / std::mt19937 seeded with `seed`. */
uint64_t or_minstd_rand0_nth(uint32_t seed, int n);
uint64_t or_mt19937_nth(uint32_t seed, int n);

/* ---------------- loaders ---------------- */
/* Reference lattice + Maxwellian loader: thesisCode/collisionless/double/src/main.cc:201-244
 * (std::default_random_engine(1234) + std::normal_distribution<double>, vx then vy,
 * row-major; integer-division offset kept as in the reference). */
void or_init_reference(int npx, int npy, int xdim, int ydim, int nx, int ny,
                       double vtx, double vty, double vdx, double vdy,
                       double* x, double* y, double* vx, double* vy);
/* Quiet-lattice loaders with Philox Box-Muller velocities (no reference counterpart). */
void or_init_two_stream(int nx, int ny, double lx, double ly, int ppc, double v0,
                        double vth, double amp, int mode, uint64_t seed,
                        double* x, double* y, double* vx, double* vy);
void or_init_landau(int nx, int ny, double lx, double ly, int ppc, double vth,
                    double alpha, int mode, uint64_t seed,
                    double* x, double* y, double* vx, double* vy);
/* Lattice factorisation used by the loaders: npx = nx*a, npy = ny*b, a*b = ppc. */
void or_lattice_shape(int nx, int ny, int ppc, int* npx, int* npy);

/* ---------------- stages ---------------- */
/* Cell key cx*ny+cy with cx=(int)(x/dx) (thesisCode/collisionless/double/src/pic.cu:69-87). */
void or_cell_keys_f64(int64_t n, const double* x, const double* y, int nx, int ny,
                      double lx, double ly, int32_t* keys);
void or_cell_keys_f32(int64_t n, const float* x, const float* y, int nx, int ny,
                      double lx, double ly, int32_t* keys);
/* Stable counting sort by key (ascending particle index inside a cell),
 * thesisCode/dsmcCollisions/cellTrackerAdvanced/src/pic.cu:419-445. */
void or_stable_sort(int64_t n, const int32_t* keys, int ncells,
                    int32_t* perm, int32_t* counts, int32_t* offsets);

/* Continuous-weight CIC charge density (charge per area, FP64), no background.
 * thesisCode/collisionless/double/src/grid.cc:46-76 with intended CIC weights. */
void or_deposit_f64(int64_t n, const double* x, const double* y, int nx, int ny,
                    double lx, double ly, double qp, double* rho);
/* Fixed-point mirrors of the GPU deposit. f32: one int64 word, 2^18 per axis.
 * f64: two int64 words (hi = W>>30, lo = W & (2^30-1)), 2^30 per axis. */
void or_deposit_fixed_f32(int64_t n, const float* x, const float* y, int nx, int ny,
                          double lx, double ly, int64_t* acc);
void or_deposit_fixed_f64(int64_t n, const double* x, const double* y, int nx, int ny,
                          double lx, double ly, int64_t* hi, int64_t* lo);
/* Convert fixed-point accumulators to charge density (charge per area). */
void or_fixed_to_rho_f32(int nx, int ny, double lx, double ly, double qp,
                         const int64_t* acc, double* rho);
void or_fixed_to_rho_f64(int nx, int ny, double lx, double ly, double qp,
                         const int64_t* hi, const int64_t* lo, double* rho);

/* 2D complex FFT helpers (radix-2, FP64), in place, sign=-1 forward, +1 inverse
 * (unnormalised).  Data is row-major [nx][ny] interleaved (re,im). */
void or_fft2(int nx, int ny, double* data, int sign);
void or_dft2_brute(int nx, int ny, const double* in, double* out, int sign);

/* Spectral Poisson: laplacian(phi) = -rho, k = 2*pi*m/L, k=0 -> 0
 * (thesisCode/collisionless/double/src/fourier.cc:15-126 with intended k, sign). */
void or_poisson(int nx, int ny, double lx, double ly, const double* rho, double* phi);
/* E = -grad(phi), periodic central difference / (2 dx)
 * (thesisCode/collisionless/double/src/grid.cc:85-96, axes as intended). */
void or_efield(int nx, int ny, double lx, double ly, const double* phi,
               double* ex, double* ey);

/* Gather + leapfrog push.  kick = c*dt (c = 0.5 for the first step from synchronous
 * velocities, 1 afterwards).  ex/ey are node fields [nx*ny].  quantized=1 uses the
 * GPU's fixed-point gather weights and FP operation order (mirror); 0 uses
 * continuous weights.  ke_out (may be NULL) receives sum of |v_n|^2 where v_n is
 * the synchronous velocity at the start of the step.
 * thesisCode/collisionless/double/src/particle.cc:37-106 (intended KDK). */
void or_push_f64(int64_t n, double* x, double* y, double* vx, double* vy,
                 const double* ex, const double* ey, int nx, int ny,
                 double lx, double ly, double dt, double qm, double c, int quantized,
                 double* ke_out, double* px_out, double* py_out);
void or_push_f32(int64_t n, float* x, float* y, float* vx, float* vy,
                 const float* ex, const float* ey, int nx, int ny,
                 double lx, double ly, double dt, double qm, double c,
                 double* ke_out, double* px_out, double* py_out);
/* Interpolated field at particle positions (continuous weights, FP64). */
void or_gather_f64(int64_t n, const double* x, const double* y, const double* ex,
                   const double* ey, int nx, int ny, double lx, double ly,
                   double* epx, double* epy);

/* Whole collisionless run in FP64 (continuous weights).  diag (may be NULL) gets
 * nsteps rows of {field_energy, kinetic_energy, px, py, |phi_hat(1,0)|}.
 * Velocities on return are synchronous (final half kick applied).
 * Step order: thesisCode/collisionless/double/src/main.cc:137-172. */
void or_run_f64(int64_t n, double* x, double* y, double* vx, double* vy,
                int nx, int ny, double lx, double ly, double dt, double qm,
                double qp, int nsteps, double* diag);
/* Same in the GPU mirror arithmetic (FP32 particles, fixed-point deposit, FP64 field). */
void or_run_f32(int64_t n, float* x, float* y, float* vx, float* vy,
                int nx, int ny, double lx, double ly, double dt, double qm,
                double qp, int nsteps, double* diag);
/* or_run_f32 on nthreads OpenMP threads (the all-cores CPU baseline): particles in
 * contiguous chunks, per-thread int64 charge grids summed per node, FFT rows and columns
 * split across threads.  Positions and velocities are bit-identical to or_run_f32; the
 * diagnostics' particle sums are added per chunk, so they may differ in the last bits. */
void or_run_f32_threads(int64_t n, float* x, float* y, float* vx, float* vy,
                        int nx, int ny, double lx, double ly, double dt, double qm,
                        double qp, int nsteps, int nthreads, double* diag);

/* ---------------- collisions ---------------- */
/* Analytic cross sections (m^2) at energy E (eV). */
void or_mcc_cross_sections(double e_ev, double* s_el, double* s_ex, double* s_iz);
/* Null-collision MCC pass (mirror of the GPU operator, FP32 particles).  e_per_v2: energy
 * [eV] per |v|^2; nu_dt_per_sigma_v: nu*dt per (sigma [m^2] * |v|); p_coll = 1 - exp(-nu_max
 * dt); mass_ratio2 = 2 m_e / M.  Particle p collides when word 0 of Philox2x32-10 {id[p],
 * step} (key seed_lo ^ seed_hi ^ 'MCC1') gives a uniform below p_coll; a colliding particle
 * then draws Philox4x32-10 {id[p], step, 0|1, 'MCC1'} (u[1] process, u[2] angle, u[3]
 * energy split; the second block for the secondary's direction).  Births
 * (one per ionisation) are appended after n in ascending parent id with ids next_id, ...
 * Returns the new particle count; counts[4] += {elastic, excitation, ionisation, null}. */
int64_t or_mcc_step_f32(int64_t n, int64_t cap, float* x, float* y, float* vx, float* vy,
                        int32_t* id, int64_t next_id, double e_per_v2,
                        double nu_dt_per_sigma_v, double p_coll, double nu_max_dt,
                        double mass_ratio2, uint64_t seed, int64_t step, int64_t* counts);
/* NTC hard-sphere DSMC pass over cell-sorted particles (mirror of the GPU operator), in
 * rounds of disjoint pairs: cell c holds slots [offsets[c], offsets[c+1]); gmax[c] is the
 * running maximum relative speed; candidate k is the pair (pi(2m), pi(2m+1)), m = k % (n/2),
 * of round k / (n/2) under that round's slot permutation pi (Feistel network keyed by
 * Philox {c, step, 0x80000000 | round, 'DSMP'}); Philox {c, step, k, 'DSMC'} gives its
 * acceptance and direction uniforms and {c, step, 0xffffffff, 'DSMC'} the fractional
 * candidate.  Returns the number of accepted collisions. */
int64_t or_dsmc_step_f32(int ncells, const int32_t* offsets, float* vx, float* vy,
                         float* gmax, double coeff, uint64_t seed, int64_t step);
/* out[t] = pi(t), t < n: the slot permutation of one DSMC round (n >= 2). */
void or_dsmc_pair_perm(uint32_t n, int64_t cell, int64_t step, int64_t round, uint64_t seed,
                       uint32_t* out);

/* The reference's own collide arithmetic on an injected uniform stream u[nu] (KEEP list):
 *   or_std_uniform_stream: std::uniform_real_distribution<double>{0,1} over std::mt19937{seed};
 *   or_mcc_reference: thesisCode/mccCollisions/src/particle.cc:112-127 for particles 0..n-1;
 *   or_dsmc_reference: thesisCode/dsmcCollisions/cellTrackerAdvanced/src/main.cc:288-350
 *     (nearest partner in the cell, t > i, +-(Bird normal) pair assignment).
 * Both return the uniforms consumed, -1 when the stream is too short. */
void or_std_uniform_stream(uint32_t seed, int64_t count, double* out);
int64_t or_mcc_reference(int64_t n, double* vx, double* vy, double prob, const double* u,
                         int64_t nu);
int64_t or_dsmc_reference(int64_t n, const double* x, const double* y, double* vx,
                          double* vy, int nx, int ny, double lx, double ly, double prob,
                          const double* u, int64_t nu, int64_t* ncoll);

/* The bit-reproducible elementary functions used by the collision stages. */
void or_sincos2pi(double u, double* s, double* c);
double or_log(double x);

/* ---------------- adjoint ---------------- */
/* Transpose of the field map rho -> (Ex, Ey) (spectral Poisson + central difference):
 * lrho = M^T (lex, ley). */
void or_adjoint_mt(int nx, int ny, double lx, double ly, const double* lex, const double* ley,
                   double* lrho);
/* Loss = sum(Ex^2+Ey^2)*dx*dy after nsteps of the FP64 collisionless run starting from
 * x0 = xl + sum_m a_m cos(kx_m xl + ky_m yl), v0.  Returns loss; if grad_a / grad_vx /
 * grad_vy are non-NULL the reverse-mode gradient is written there (nmodes, n, n). */
double or_adjoint_f64(int64_t n, const double* xl, const double* yl, const double* vx0,
                      const double* vy0, int nmodes, const int32_t* mx, const int32_t* my,
                      const double* amp, int nx, int ny, double lx, double ly, double dt,
                      double qm, double qp, int nsteps,
                      double* grad_a, double* grad_vx, double* grad_vy);

#ifdef __cplusplus
}
#endif
#endif
