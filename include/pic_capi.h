/*
 * pic_capi.h -- C ABI of the MI355X 2D electrostatic PIC engine (libpicamd.so).
 *
 * Plain C: POD structs, pointers and sizes, int status codes.  No C++ or torch types.
 *
 * Boundary it replaces:
 *   extern "C" double picFloatGpu(double *positionX, double *positionY,
 *       double *velocityX, double *velocityY, double *xResult, double *yResult,
 *       std::shared_ptr<Params> p);
 *   declared at thesisCode/collisionless/double/inc/pic.h:5-7 (FP32 twin
 *   thesisCode/collisionless/single/inc/pic.h), defined at
 *   thesisCode/collisionless/double/src/pic.cu:429-611, called at
 *   thesisCode/collisionless/double/src/main.cc:182 (and mccCollisions/src/main.cc:196,
 *   dsmcCollisions/cellTrackerAdvanced/src/main.cc:197).  The reference's Params
 *   (thesisCode/collisionless/double/inc/params.h:21-47) becomes pic_config.
 *
 * Conventions
 *   - Every function returns PIC_OK (0) or a negative PIC_ERR_* code; the message of the
 *     last failure on the calling thread is returned by pic_last_error().  Nothing calls
 *     exit(); invalid input is reported, never trapped.
 *   - Units are normalised: omega_p = 1, lambda_D = 1, eps0 = 1.  Grid node (i, j) is at
 *     (i*dx, j*dy) and is stored at index i*ny + j (x is the slow index).
 *   - A context is not thread-safe; distinct contexts may be used from distinct threads.
 *   - Device pointers passed to pic_op_* and *_device functions must live on the device
 *     of the stream / context; `stream` is a hipStream_t (NULL = default stream).
 */
#ifndef PIC_CAPI_H
#define PIC_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------------- status codes */
#define PIC_OK                  0
#define PIC_ERR_INVALID_CONFIG (-1)  /* bad sizes, non-power-of-two grid, dt <= 0, ... */
#define PIC_ERR_OOM            (-2)  /* device allocation failed */
#define PIC_ERR_HIP            (-3)  /* a HIP runtime call failed */
#define PIC_ERR_OVERFLOW       (-4)  /* particle capacity exceeded */
#define PIC_ERR_NONFINITE      (-5)  /* NaN/Inf found in the state */
#define PIC_ERR_STATE          (-6)  /* call out of order (e.g. step before particles) */
#define PIC_ERR_UNSUPPORTED    (-7)  /* valid request this build does not support */
#define PIC_ERR_ARG            (-8)  /* NULL pointer or out-of-range argument */

/* ------------------------------------------------------------------- enums */
#define PIC_FP32 0                   /* FP32 particles, int64 fixed-point charge, FP64 field */
#define PIC_FP64 1                   /* FP64 particles, two-word fixed-point charge, FP64 field */

#define PIC_COLLIDE_NONE 0
#define PIC_COLLIDE_MCC  1           /* electron-neutral null-collision Monte Carlo */
#define PIC_COLLIDE_DSMC 2           /* like-particle NTC hard-sphere DSMC */

#define PIC_GRID_RHO 0               /* charge density incl. neutralising background */
#define PIC_GRID_PHI 1               /* potential */
#define PIC_GRID_EX  2
#define PIC_GRID_EY  3

/* Diagnostics row (PIC_DIAG_COLS doubles per step, time t_n = n*dt). */
#define PIC_DIAG_COLS 8
#define PIC_DIAG_FIELD_ENERGY 0      /* 1/2 sum |E|^2 dx dy */
#define PIC_DIAG_KINETIC      1      /* 1/2 m sum |v_n|^2 (synchronous v) */
#define PIC_DIAG_PX           2      /* m sum v_x */
#define PIC_DIAG_PY           3      /* m sum v_y */
#define PIC_DIAG_MODE1        4      /* |rho_hat(kx = 2 pi / lx, ky = 0)| / (nx ny) */
#define PIC_DIAG_NPART        5      /* particle count */
#define PIC_DIAG_NCOLL        6      /* collisions in the step (MCC or DSMC) */

/* ----------------------------------------------------------------- config */
/* Replaces thesisCode/collisionless/double/inc/params.h:21-47 (Params).  The reference's
 * integer domain (xDim, yDim), lattice (numParticlesX/Y), tmax and blockSize map to
 * lx/ly, the particle arrays handed to pic_set_particles, nsteps and the tiling fields. */
typedef struct pic_config {
  int32_t nx, ny;            /* cells per axis: powers of two in [4, 2048] */
  double  lx, ly;            /* domain lengths */
  double  dt;                /* time step */
  double  qm;                /* charge-to-mass ratio of the macro-particles (-1: electrons) */
  double  qp;                /* macro-particle charge; 0 => -(lx*ly)/n (omega_p = 1) */
  int64_t capacity;          /* particle capacity per run (>= n); 0 => n at set time */
  int32_t precision;         /* PIC_FP32 | PIC_FP64 */
  int32_t device;            /* HIP device ordinal */
  int32_t sort_every;        /* steps between cell sorts; 0 => autotuner default */
  int32_t tile;              /* deposit tile side in cells (power of two); 0 => default */
  int32_t halo;              /* deposit LDS window halo in cells; 0 => default */
  int32_t splits;            /* workgroups per tile; 0 => default (sized by the particles
                                per tile; in DSMC runs per tile, from each tile's count) */
  int32_t collide;           /* PIC_COLLIDE_* */
  int32_t reserved0;
  uint64_t seed;             /* Philox4x32-10 key for every random stream */
  /* MCC parameters (physical units; see pic_mcc_preset) */
  double  mcc_ng;            /* neutral density [m^-3] */
  double  mcc_v_unit;        /* [m/s] per unit of normalised velocity */
  double  mcc_dt_phys;       /* [s] per unit of normalised time */
  double  mcc_neutral_mass;  /* [kg] */
  /* DSMC parameters (normalised units) */
  double  dsmc_coeff;        /* F_N * sigma * dt / cell_area (candidate-pair coefficient) */
  double  dsmc_gmax0;        /* initial per-cell maximum relative speed */
} pic_config;

typedef struct pic_ctx pic_ctx;          /* opaque engine context (one or many runs) */

/* ----------------------------------------------------------------- general */
const char* pic_last_error(void);
int  pic_version(void);                  /* ABI version, currently 1 */
void pic_config_default(pic_config* cfg);
/* Fill cfg for the MCC preset: neutral density ng [m^-3] at 300 K, electron temperature
 * te_ev, plasma density ne [m^-3]; sets mcc_* and dt so that nu_max*dt_phys ~= 0.01. */
int  pic_mcc_preset(pic_config* cfg, double ng, double te_ev, double ne);
/* nu_max*dt_phys used by the MCC stage for cfg (scan of sigma_T(E) v(E), 0-1000 eV). */
double pic_mcc_nu_max(const pic_config* cfg);

/* ----------------------------------------------------------------- context API */
int pic_create(const pic_config* cfg, pic_ctx** out);
/* Ragged ensemble: nruns independent runs (own grid, particles, dt) stepped together,
 * one launch per stage for all runs.  All runs share precision and device. */
int pic_batch_create(const pic_config* cfgs, int nruns, pic_ctx** out);
int pic_destroy(pic_ctx* ctx);
int pic_num_runs(pic_ctx* ctx, int* nruns);

/* Load particles of run `run` (host FP64 arrays, synchronous velocities).  Particle k gets
 * id k; outputs are returned in id order.  Sorts by cell and deposits the initial charge. */
int pic_set_particles(pic_ctx* ctx, int run, const double* x, const double* y,
                      const double* vx, const double* vy, int64_t n);
/* Same from device arrays in the context precision (float or double). */
int pic_set_particles_device(pic_ctx* ctx, int run, const void* x, const void* y,
                             const void* vx, const void* vy, int64_t n);
/* Copy particles out in id order (host FP64).  sync=1 returns synchronous velocities
 * v_n (the pending half kick is applied to the copy, not to the state). */
int pic_get_particles(pic_ctx* ctx, int run, double* x, double* y, double* vx, double* vy,
                      int64_t n, int sync);
/* Device variant: copy out in id order into device arrays of the context precision. */
int pic_get_particles_device(pic_ctx* ctx, int run, void* x, void* y, void* vx, void* vy,
                             int64_t n, int sync);
int pic_num_particles(pic_ctx* ctx, int run, int64_t* n);

/* Advance every run by nsteps (asynchronous on the context stream; returns after launch). */
int pic_step(pic_ctx* ctx, int nsteps);
/* Wait for the context stream and check the device error flags (capacity overflow). */
int pic_sync(pic_ctx* ctx);
/* Scan every particle for NaN/Inf (synchronises); PIC_ERR_NONFINITE if any. */
int pic_check_finite(pic_ctx* ctx);
/* Cumulative collision counts of a run: MCC {elastic, excitation, ionisation, null};
 * DSMC {accepted pairs, 0, 0, 0}. */
int pic_collision_counts(pic_ctx* ctx, int run, int64_t* out4);
/* Grid of the current positions (host FP64, nx*ny values, node (i,j) at i*ny+j). */
int pic_get_grid(pic_ctx* ctx, int run, int which, double* out);
/* Diagnostics rows [first, first+count) (count*PIC_DIAG_COLS doubles).  The device keeps
 * the last 65536 steps. */
int pic_get_diagnostics(pic_ctx* ctx, int run, int64_t first, int count, double* out);
int pic_current_step(pic_ctx* ctx, int run, int64_t* step);
int pic_stream(pic_ctx* ctx, void** stream);  /* hipStream_t of the context */
/* Use (1) or bypass (0) captured hipGraphs for the step loop (default 1). */
int pic_use_graphs(pic_ctx* ctx, int enable);
/* Tiling / sort cadence in use (after defaults or autotuning); splits reads 0 when the
 * splits are sized per tile (DSMC runs asked for the default). */
int pic_get_tuning(pic_ctx* ctx, int* tile, int* halo, int* splits, int* sort_every);
int pic_set_tuning(pic_ctx* ctx, int tile, int halo, int splits, int sort_every);
/* 1 when the tile sort runs inside the push launches (the step before a sort step counts
 * the tiles of the new positions, the sort step scatters while it pushes; MCC births
 * included); 0 when it runs as separate kernels (DSMC runs, sort_every < 2, or too many
 * tiles for the LDS). */
int pic_sort_fused(pic_ctx* ctx, int* fused);
/* Time candidate tilings for `trial_steps` each and keep the fastest (restores state). */
int pic_autotune(pic_ctx* ctx, int trial_steps, double* best_ms_per_step);
/* Run nsteps eagerly with hipEvents around every kernel; out_ms[PIC_NSTAGES] receives the
 * mean duration per step of each stage (ms). */
#define PIC_NSTAGES 6
#define PIC_STAGE_PUSH_DEPOSIT 0      /* fused gather + push + deposit (the k_push launch) */
#define PIC_STAGE_FFT_ROWS     1      /* rho -> row spectra */
#define PIC_STAGE_FFT_COLS     2      /* column FFT + Green's function + inverse columns */
#define PIC_STAGE_FFT_FIELD    3      /* inverse rows -> E (+ field energy) */
#define PIC_STAGE_SORT         4      /* sorts (amortised per step) and the step's
                                         diagnostics reduction */
#define PIC_STAGE_COLLIDE      5      /* MCC or DSMC */
int pic_stage_times(pic_ctx* ctx, int nsteps, double* out_ms);

/* Whole run in one call, the semantics of picFloatGpu
 * (thesisCode/collisionless/double/src/pic.cu:429-611): host FP64 SoA in, final
 * positions out in input order, GPU seconds in *elapsed_s (steady-state step loop only;
 * setup excluded).  nsteps replaces the reference's floating-point tmax loop. */
int pic_run_compat(const double* x, const double* y, const double* vx, const double* vy,
                   double* x_out, double* y_out, int64_t n, int nsteps,
                   const pic_config* cfg, double* elapsed_s);

/* ----------------------------------------------------------------- loaders (host)
 * Fill host FP64 arrays (length n = *_count()).  Particle p = i*npy + j sits on a
 * quiet-start lattice of npx x npy points (npx = nx*a, npy = ny*b, a*b = ppc, a the
 * smallest divisor of ppc that is >= sqrt(ppc)); velocities are Box-Muller normals from
 * Philox4x32-10 with counter {p, 0, 0, 0x4c4f4144} and key = seed. */
int64_t pic_lattice_count(int nx, int ny, int ppc);
/* Landau damping: x from the inverse CDF of 1 + alpha cos(2 pi mode x / lx), Maxwellian
 * with thermal speed vth. */
int pic_load_landau(int nx, int ny, double lx, double ly, int ppc, double vth, double alpha,
                    int mode, uint64_t seed, double* x, double* y, double* vx, double* vy);
/* Two-stream: cold beams +-v0 (alternating p) with thermal spread vth, positions displaced
 * by amp*cos(2 pi mode x / lx). */
int pic_load_two_stream(int nx, int ny, double lx, double ly, int ppc, double v0, double vth,
                        double amp, int mode, uint64_t seed, double* x, double* y, double* vx,
                        double* vy);
/* The reference's loader (thesisCode/collisionless/double/src/main.cc:201-244):
 * npx x npy lattice over an integer domain, std::default_random_engine(1234) +
 * std::normal_distribution<double>, drift removed then (vdx, vdy) added. */
int pic_load_reference(int npx, int npy, int xdim, int ydim, int nx, int ny, double vtx,
                       double vty, double vdx, double vdy, double* x, double* y, double* vx,
                       double* vy);

/* ----------------------------------------------------------------- stage operators
 * Drop-in per-stage operators on device memory (the reference's file-static kernels,
 * thesisCode/collisionless/double/src/pic.cu:69-420, made callable).  `precision`
 * selects float or double particle arrays. */
typedef struct pic_grid {
  int32_t nx, ny;
  double  lx, ly;
} pic_grid;

/* keys[p] = cx*ny + cy with cx = (int)(x/dx): thesisCode/collisionless/double/src/pic.cu:69-87 */
int pic_op_cell_keys(const void* x, const void* y, int64_t n, int precision,
                     const pic_grid* g, int32_t* keys, void* stream);
/* Stable counting sort by key (ties in ascending index): perm[n], counts[ncells],
 * offsets[ncells+1].  thesisCode/dsmcCollisions/cellTrackerAdvanced/src/pic.cu:419-445 */
int pic_op_sort(const int32_t* keys, int64_t n, int ncells, int32_t* perm,
                int32_t* counts, int32_t* offsets, void* stream);
/* CIC charge density (charge per unit area, no background) into rho[nx*ny] (FP64),
 * deterministic fixed-point accumulation.  thesisCode/collisionless/double/src/pic.cu:99-138 */
int pic_op_deposit(const void* x, const void* y, int64_t n, int precision,
                   const pic_grid* g, double qp, double* rho, void* stream);
/* Allocation-free variant of pic_op_deposit (same arguments, same result bit for bit).
 * No counterpart in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The fixed-point accumulators live in a static device array of
 * 2 * PIC_DEPOSIT_NOALLOC_MAX_NODES int64 words (16 MiB per device, held until the process
 * ends); grids with nx*ny above PIC_DEPOSIT_NOALLOC_MAX_NODES are refused with
 * PIC_ERR_UNSUPPORTED.  Every call on one device shares that array: calls that overlap in
 * time on one device (other streams, other threads) race and must be serialised by the
 * caller.  The array is cleared at the start of each call by a kernel (no memset node).
 * The call synchronises `stream` before returning, like pic_op_deposit. */
#define PIC_DEPOSIT_NOALLOC_MAX_NODES 1048576
int pic_op_deposit_noalloc(const void* x, const void* y, int64_t n, int precision,
                           const pic_grid* g, double qp, double* rho, void* stream);
/* Spectral Poisson, laplacian(phi) = -rho, mean removed (FP64).
 * thesisCode/collisionless/double/src/pic.cu:147-232 */
int pic_op_poisson(const double* rho, double* phi, const pic_grid* g, void* stream);
/* E = -grad(phi) by periodic central differences (FP64 in, FP64 out).
 * thesisCode/collisionless/double/src/pic.cu:242-283 */
int pic_op_efield(const double* phi, double* ex, double* ey, const pic_grid* g, void* stream);
/* Gather E at the particles and leapfrog: v += c*dt*qm*E(x); x += dt*v (wrapped).
 * ex/ey are FP64 node fields.  ke (device, may be NULL) receives sum |v_n|^2 of the
 * synchronous velocity v_n = v + (c - 1/2)*dt*qm*E(x).
 * thesisCode/collisionless/double/src/pic.cu:298-420 */
int pic_op_gather_push(void* x, void* y, void* vx, void* vy, int64_t n, int precision,
                       const double* ex, const double* ey, const pic_grid* g, double dt,
                       double qm, double c, double* ke, void* stream);
/* One null-collision MCC pass (FP32 particles, ids in `id`).  Ionisation births are
 * appended after n (up to cap); *n_io is updated on the host (synchronises).
 * counts (device int64[4], may be NULL) accumulate {elastic, excitation, ionisation, null}.
 * Slot of thesisCode/mccCollisions/src/pic.cu:435-463. */
int pic_op_mcc(float* x, float* y, float* vx, float* vy, int32_t* id, int64_t* n_io,
               int64_t cap, const pic_config* cfg, int64_t step, int64_t* counts,
               void* stream);
/* One NTC hard-sphere DSMC pass over cell-sorted particles: offsets[ncells+1] delimit the
 * cells (row-major cell numbering, as from pic_op_cell_keys + pic_op_sort), and the order
 * of the particles inside a cell is the canonical one (ascending id, as pic_op_sort leaves
 * it for arrays in id order); `id` is not read.  A cell of n particles draws
 * M ~ 1/2 n (n-1) coeff gmax candidate pairs, dealt to rounds of n/2 disjoint pairs under a
 * pseudo-random slot permutation per round; a candidate with relative speed g is accepted
 * with probability g / max(gmax, g).  gmax[ncells] is the running per-cell maximum
 * relative speed (read and updated).  ncoll (device int64, may be NULL) accumulates
 * accepted collisions.
 * Slot of thesisCode/dsmcCollisions/cellTrackerAdvanced/src/pic.cu:791-864. */
int pic_op_dsmc(float* vx, float* vy, const int32_t* id, const int32_t* offsets,
                int ncells, float* gmax, double coeff, uint64_t seed, int64_t step,
                int64_t* ncoll, void* stream);

/* The reference's own collide arithmetic (FP64 device arrays), driven by an injected
 * uniform stream u[nu] (device) in the order the reference draws it: one std::mt19937 over
 * the particles in index order (std::uniform_real_distribution<double>{0, 1} reproduces
 * the reference's host stream).  A particle that draws takes r = u[k] and, when prob > r,
 * six more uniforms: v = (2(u1 + u2 + u3 - 1.5), 2(u4 + u5 + u6 - 1.5)).  The operators
 * work out on the device which uniforms each particle takes (a chain over the stream,
 * resolved in parallel), write *consumed (host) = uniforms used, so the next call goes on
 * at u + consumed, and fail with PIC_ERR_ARG when the stream is too short.  Synchronous.
 *
 * MCC: every particle draws.  thesisCode/mccCollisions/src/particle.cc:112-127 (prob 0.8
 * there; the reference GPU kernel mccCollisions/src/pic.cu:435-463 uses 0.22). */
int pic_op_mcc_reference(double* vx, double* vy, int64_t n, double prob, const double* u,
                         int64_t nu, int64_t* consumed, void* stream);
/* DSMC: particle i draws when the nearest other particle t of its cell ((int)(x/dx),
 * (int)(y/dy); strictly nearer wins, the lower index on ties) has t > i; a collision sets
 * v_i = (m1, m2) and v_t = (-m1, -m2), later particles overwriting earlier ones as in the
 * reference's loop.  Positions must lie in [0, lx) x [0, ly).  *ncoll (host, may be NULL)
 * = collisions.  thesisCode/dsmcCollisions/cellTrackerAdvanced/src/main.cc:269-350 (q = 0.2
 * there). */
int pic_op_dsmc_reference(const double* x, const double* y, double* vx, double* vy,
                          int64_t n, const pic_grid* g, double prob, const double* u,
                          int64_t nu, int64_t* consumed, int64_t* ncoll, void* stream);

/* ----------------------------------------------------------------- checkpoint / restart
 * No reference counterpart (its -f flag only truncates a file, CLd/src/main.cc:101-108).
 * A versioned binary snapshot of one run: its config, step, velocity phase, id counter,
 * macro-particle charge, collision counters, DSMC per-cell maxima and the particles in id
 * order (run precision), followed by an FNV-1a checksum.  Restarting continues the run bit
 * for bit (particle state; the diagnostics history before the restart is not stored). */
int pic_save_checkpoint(pic_ctx* ctx, int run, const char* path);
/* Config, particle count and completed steps stored in a checkpoint (to create a matching
 * context: set capacity >= n). */
int pic_checkpoint_info(const char* path, pic_config* cfg, int64_t* n, int64_t* step);
/* Load a checkpoint into a run whose grid, domain, dt, qm, precision, collisions and seed
 * match (PIC_ERR_INVALID_CONFIG otherwise); replaces the run's particles and state. */
int pic_load_checkpoint(pic_ctx* ctx, int run, const char* path);

/* ----------------------------------------------------------------- adjoint
 * Differentiable collisionless run (no reference counterpart).  The initial state is
 * x0 = xl + sum_m a_m cos(kx_m xl + ky_m yl), y0 = yl, v0 given; the loss is
 * sum(Ex^2 + Ey^2) dx dy after nsteps.  Reverse mode by hand-written adjoint kernels with
 * particle checkpoints every `checkpoint_every` steps and recomputation in between; the
 * field of every step is kept (nsteps x (nx+1)(ny+1) x 2 reals), so the recomputation is
 * particle pushes only. */
typedef struct pic_adjoint pic_adjoint;
int pic_adjoint_create(const pic_config* cfg, const double* xl, const double* yl, int64_t n,
                       int nmodes, const int32_t* mx, const int32_t* my, int nsteps,
                       int checkpoint_every, pic_adjoint** out);
int pic_adjoint_destroy(pic_adjoint* adj);
/* amps: device double[nmodes]; vx0/vy0: device arrays of the config precision.
 * Writes the loss (host double).  Records checkpoints for the backward pass. */
int pic_adjoint_forward(pic_adjoint* adj, const double* amps, const void* vx0,
                        const void* vy0, double* loss);
/* Reverse sweep for dL = dloss: device outputs grad_amps[nmodes] (double) and
 * grad_vx/grad_vy (config precision).  Must follow pic_adjoint_forward. */
int pic_adjoint_backward(pic_adjoint* adj, double dloss, double* grad_amps, void* grad_vx,
                         void* grad_vy);
int pic_adjoint_stream(pic_adjoint* adj, void** stream);

#ifdef __cplusplus
}
#endif
#endif /* PIC_CAPI_H */
