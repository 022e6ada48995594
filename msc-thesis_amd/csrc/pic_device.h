// pic_device.h -- device-side data layout and inline helpers shared by every kernel.
//
// Layout in HBM (all runs of a context are concatenated; a run is addressed by the base
// offsets in its RunDesc):
//   particles  SoA x, y, vx, vy (float or double) + id (int32), [pbase, pbase + pcap)
//   rho        two int64 buffers (double-buffered by step parity) of nx*ny nodes; FP64
//              contexts keep a second word pair (hi, lo) per node
//   spectra    complex double [ny/2+1][nx] per field (transposed half spectrum)
//   E          interleaved (Ex, Ey) on an (nx+1) x (ny+1) node array: row nx repeats row 0
//              and column ny repeats column 0, so a cell's four nodes need no wrap
//   cells      counts / offsets (nx*ny + 1) per run for the cell sort
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PIC_WG 256                   // threads per workgroup of every kernel
#define PIC_FFT_MAX 2048             // largest grid side
#define PIC_FFT_LOG 11               // log2(PIC_FFT_MAX)
#define PIC_TW_LOG 11                // twiddle table: exp(-2 pi i k / 2048), k < 1024
#define PIC_CHUNK 4096               // particles per chunk item
#define PIC_TAIL_CHUNK 65536         // particle range of a push tail item (births)
#define PIC_DIAG_CAP 65536           // diagnostics ring depth (steps)

struct RunDesc {
  int nx, ny, lognx, logny;
  int tile, tiles_x, tiles_y, halo;
  int splits, win;                   // win = tile + 2*halo + 1 (window nodes per axis)
  int push_item_begin, push_item_end;
  int row_item_begin, row_item_end;  // field-kernel row items of this run
  int cell_item_begin, cell_item_end;
  int pcap, fft_pp;                  // fft_pp: row pairs per workgroup of the forward rows
  int nbuckets, nchunks;             // tiles (sort buckets) and PIC_CHUNK particle chunks
  int tscan_item_begin, tscan_item_end;
  int key_tile_major;                // 1: cell keys ordered tile-major (engine sorts)
  int fft_rf;                        // rows per workgroup of the inverse (field) rows
  long long tbase;                   // bucket x chunk count table base (nbuckets*nchunks)
  long long toffbase;                // tile offsets base (nbuckets + 1)
  // sort fused into the push (no collisions, sort cadence >= 2): the push before a sort step
  // counts, per push item ("chunk"), the tiles of the particles it leaves behind; the push
  // of the sort step scatters them tile-major into the other buffer
  int fchunks;                       // push items (0: fused sort off)
  int fscan_item_begin, fscan_item_end;
  int schunk;                        // particles per chunk of the separate tile sort
  long long ftbase;                  // base of the run's bucket-major count table
  double lx, ly, dx, dy;
  double dt, qm, qp;
  double sort_lead;                  // fused sort: keys are tiles of x + sort_lead * v
  double rho_scale;                  // charge per area of one fixed-point unit
  long long pbase;                   // particle base
  long long gbase;                   // node base (nx*ny)
  long long ebase;                   // padded E base, in (Ex,Ey) pairs ((nx+1)*(ny+1))
  long long sbase;                   // spectral base ((ny/2+1)*nx complex)
  long long cbase;                   // cell-array base (nx*ny + 1)
  long long dbase;                   // diagnostics base (rows)
};

struct RunState {
  int np;                            // live particles
  int nsorted;                       // particles covered by the cell offsets
  int vsync;                         // 1: velocities are synchronous (next kick is a half kick)
  int err;                           // device-side error flags (bit 0: capacity exceeded)
  long long step;                    // completed steps
  long long next_id;                 // next id handed to a birth
};

// ------------------------------------------------------------- particle helpers
template <typename R> struct Fix;
template <> struct Fix<float> { static constexpr int QB = 18; };
template <> struct Fix<double> { static constexpr int QB = 30; };

// cx = (int)(x/dx) clamped to [0, n-1]; fraction in [0, 1]
template <typename R>
__device__ __forceinline__ void locate(R x, R dx, int n, int& c, R& f) {
  R t = x / dx;
  int ci = (int)t;
  ci = ci > n - 1 ? n - 1 : ci;
  ci = ci < 0 ? 0 : ci;
  R fr = t - (R)ci;
  fr = fr > (R)1 ? (R)1 : fr;
  fr = fr < (R)0 ? (R)0 : fr;
  c = ci;
  f = fr;
}

// Weight evaluation (gather and deposit): t = x * inv with inv = 1/dx rounded once from the
// FP64 cell size, one multiply instead of a division.  At a cell edge it may pick the
// neighbouring cell with fraction 1 instead of fraction 0, which gives the same CIC
// weights; the cell key used for sorting stays (int)(x/dx) (locate above).
template <typename R>
__device__ __forceinline__ void locate_w(R x, R inv, int n, int& c, R& f) {
  R t = x * inv;
  int ci = (int)t;
  ci = ci > n - 1 ? n - 1 : ci;
  ci = ci < 0 ? 0 : ci;
  R fr = t - (R)ci;
  fr = fr > (R)1 ? (R)1 : fr;
  fr = fr < (R)0 ? (R)0 : fr;
  c = ci;
  f = fr;
}

// fused multiply-add with one rounding (v_fma_f32 / v_fma_f64); the particle arithmetic
// spells out every fma it uses (FP contraction is off), the CPU mirror uses std::fma
template <typename R>
__device__ __forceinline__ R fmar(R a, R b, R c) {
  if constexpr (sizeof(R) == 4) return __builtin_fmaf(a, b, c);
  else return __builtin_fma(a, b, c);
}

// periodic index on a power-of-two axis (every grid size is a power of two): v mod n for
// any int v, one AND
__device__ __forceinline__ int pwrap(int v, int n) { return v & (n - 1); }

// Item handled by workgroup b of an n-workgroup launch, such that each XCD gets a
// contiguous range of items: the hardware dispatches workgroup b to XCD b % 8, and items
// next to each other share cache lines (the transposed spectra), which each XCD's own L2
// would otherwise fetch separately.  A bijection of [0, n).
__device__ __forceinline__ int xcd_item(int b, int n) {
  const int x = b & 7, q = n >> 3, r = n & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

template <typename R>
__device__ __forceinline__ R wrap_pos(R x, R L) {
  if (x >= L) x -= L;
  else if (x < (R)0) x += L;
  if (!(x >= (R)0 && x < L)) {
    x = x - L * floor(x / L);
    if (!(x >= (R)0 && x < L)) x = (R)0;
  }
  return x;
}

// per-axis fixed-point CIC weight of the upper node, in units of 2^-QB
template <typename R>
__device__ __forceinline__ int qweight(R f) {
  return (int)rint(f * (R)(double)(1 << Fix<R>::QB));
}

// ------------------------------------------------------------- Philox4x32-10
// Each round's two 32 x 32 -> 64-bit products are one v_mad_u64_u32 each (the high and low
// words written separately cost a v_mul_hi_u32 and a v_mul_lo_u32 apiece).
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)c[0] * 0xD2511F53u;
    const uint64_t p1 = (uint64_t)c[2] * 0xCD9E8D57u;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
  }
}

// Philox2x32-10: one product per round (a single v_mad_u64_u32), half the work of the
// four-word generator; the MCC null-collision test needs one uniform per particle.
__device__ __forceinline__ void philox2x32_10(uint32_t c[2], uint32_t k) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) k += 0x9E3779B9u;
    const uint64_t p = (uint64_t)c[0] * 0xD256D193u;
    const uint32_t n0 = (uint32_t)(p >> 32) ^ k ^ c[1];
    c[1] = (uint32_t)p;
    c[0] = n0;
  }
}

#define PIC_TAG_LOAD 0x4c4f4144u
#define PIC_TAG_MCC  0x4d434331u
#define PIC_TAG_DSMC 0x44534d43u
#define PIC_TAG_DSMP 0x44534d50u   // DSMC round permutations

__device__ __forceinline__ double u01(uint32_t r) {
  return ((double)r + 0.5) * 2.3283064365386963e-10;
}

__device__ __forceinline__ void draw4(uint32_t a, uint32_t b, uint32_t c, uint32_t tag,
                                      uint64_t seed, double u[4]) {
  uint32_t ctr[4] = {a, b, c, tag};
  philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
  u[0] = u01(ctr[0]); u[1] = u01(ctr[1]); u[2] = u01(ctr[2]); u[3] = u01(ctr[3]);
}

// ------------------------------------------------------------- own sincos / log
// Only +,-,*,/ and fma: identical bits on every IEEE-754 machine.
__device__ __forceinline__ void sincos2pi_own(double u, double& s, double& c) {
  const double kPi = 3.14159265358979323846;
  double q = floor(u * 8.0);
  double r = u * 8.0 - q;
  int oct = (int)q & 7;
  double a = ((oct & 1) ? (1.0 - r) : r) * (kPi / 4.0);
  double a2 = a * a;
  double ps = 1.0 / 1307674368000.0;
  ps = fma(ps, -a2, 1.0 / 6227020800.0);
  ps = fma(ps, -a2, 1.0 / 39916800.0);
  ps = fma(ps, -a2, 1.0 / 362880.0);
  ps = fma(ps, -a2, 1.0 / 5040.0);
  ps = fma(ps, -a2, 1.0 / 120.0);
  ps = fma(ps, -a2, 1.0 / 6.0);
  double sa = fma(-a * a2, ps, a);
  double pc = 1.0 / 20922789888000.0;
  pc = fma(pc, -a2, 1.0 / 87178291200.0);
  pc = fma(pc, -a2, 1.0 / 479001600.0);
  pc = fma(pc, -a2, 1.0 / 3628800.0);
  pc = fma(pc, -a2, 1.0 / 40320.0);
  pc = fma(pc, -a2, 1.0 / 720.0);
  pc = fma(pc, -a2, 1.0 / 24.0);
  pc = fma(pc, -a2, 0.5);
  double ca = fma(-a2, pc, 1.0);
  double sn, cs;
  switch (oct) {
    case 0: sn = sa;  cs = ca;  break;
    case 1: sn = ca;  cs = sa;  break;
    case 2: sn = ca;  cs = -sa; break;
    case 3: sn = sa;  cs = -ca; break;
    case 4: sn = -sa; cs = -ca; break;
    case 5: sn = -ca; cs = -sa; break;
    case 6: sn = -ca; cs = sa;  break;
    default: sn = -sa; cs = ca; break;
  }
  s = sn; c = cs;
}

__device__ __forceinline__ double log_own(double x) {
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752) { m *= 2.0; e -= 1; }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double p = 1.0 / 25.0;
  p = fma(p, s2, 1.0 / 23.0);
  p = fma(p, s2, 1.0 / 21.0);
  p = fma(p, s2, 1.0 / 19.0);
  p = fma(p, s2, 1.0 / 17.0);
  p = fma(p, s2, 1.0 / 15.0);
  p = fma(p, s2, 1.0 / 13.0);
  p = fma(p, s2, 1.0 / 11.0);
  p = fma(p, s2, 1.0 / 9.0);
  p = fma(p, s2, 1.0 / 7.0);
  p = fma(p, s2, 1.0 / 5.0);
  p = fma(p, s2, 1.0 / 3.0);
  double lm = 2.0 * fma(s * s2, p, s);
  return fma((double)e, 0.69314718055994531, lm);
}

// ------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Sum of v over the 256-thread workgroup, fixed order; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* scratch4) {
  v = wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch4[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = ((scratch4[0] + scratch4[1]) + scratch4[2]) + scratch4[3];
  return r;
}

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}
