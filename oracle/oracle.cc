// oracle.cc -- CPU restatement of the reference PIC step (test infrastructure only;
// see oracle.h for the contract and for how each function is pinned).
//
// Reference files followed (paths relative to the reference checkout):
//   stage order        thesisCode/collisionless/double/src/main.cc:137-172
//   loader             thesisCode/collisionless/double/src/main.cc:201-244
//   cell index         thesisCode/collisionless/double/src/pic.cu:69-87
//   deposit            thesisCode/collisionless/double/src/grid.cc:46-76
//   Poisson            thesisCode/collisionless/double/src/fourier.cc:15-126
//                      thesisCode/collisionless/double/src/pic.cu:147-232
//   E-field            thesisCode/collisionless/double/src/grid.cc:85-96
//   gather             thesisCode/collisionless/double/src/particle.cc:37-60
//   push               thesisCode/collisionless/double/src/particle.cc:69-106
//   cell lists         thesisCode/dsmcCollisions/cellTrackerAdvanced/src/pic.cu:419-445
//   MCC slot           thesisCode/mccCollisions/src/particle.cc:112-127
//   DSMC slot          thesisCode/dsmcCollisions/cellTrackerAdvanced/src/main.cc:288-350
// Deviations from the reference code follow its intended physics (periodic CIC with
// standard weights, k = 2 pi m / L, E = -grad phi along the matching axes, KDK leapfrog,
// x in [0, L)); the MCC and DSMC physics are the analytic cross-section null-collision
// and NTC hard-sphere models, which have no reference code.
#include "oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <random>
#include <vector>

namespace {

typedef std::complex<double> cd;
const double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- Philox
inline void mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

void philox(const uint32_t in[4], const uint32_t k[2], uint32_t out[4]) {
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
  uint32_t k0 = k[0], k1 = k[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(0xD2511F53u, c0, &hi0, &lo0);
    mulhilo(0xCD9E8D57u, c2, &hi1, &lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Philox2x32-10 (same paper): one 32 x 32 product per round, multiplier 0xD256D193, the
// key bumped by 0x9E3779B9 between rounds.
void philox2(const uint32_t in[2], uint32_t k, uint32_t out[2]) {
  uint32_t c0 = in[0], c1 = in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) k += 0x9E3779B9u;
    uint32_t hi, lo;
    mulhilo(0xD256D193u, c0, &hi, &lo);
    c0 = hi ^ k ^ c1;
    c1 = lo;
  }
  out[0] = c0; out[1] = c1;
}

// Stream tags (4th counter word).
enum : uint32_t { TAG_LOAD = 0x4c4f4144u, TAG_MCC = 0x4d434331u, TAG_DSMC = 0x44534d43u };

inline double u01(uint32_t r) { return ((double)r + 0.5) * 2.3283064365386963e-10; }  // (r+1/2)/2^32

void draw4(uint32_t a, uint32_t b, uint32_t c, uint32_t tag, uint64_t seed, double u[4]) {
  uint32_t ctr[4] = {a, b, c, tag};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  philox(ctr, key, o);
  for (int i = 0; i < 4; ++i) u[i] = u01(o[i]);
}

// ------------------------------------------------- own elementary functions
// Written from the mathematics with only +,-,*,/ and fma so that they evaluate to the
// same bits on any IEEE-754 machine (the GPU kernels carry an identical restatement).
// sincos(2*pi*u) for u in [0,1): octant reduction + Taylor series on [0, pi/4].
void sincos2pi(double u, double* s, double* c) {
  double q = std::floor(u * 8.0);            // octant 0..7
  double r = u * 8.0 - q;                    // [0,1) inside the octant (exact: power-of-2 scale)
  int oct = (int)q & 7;
  // angle inside octant, measured so that the polynomial argument is in [0, pi/4]
  double a = ((oct & 1) ? (1.0 - r) : r) * (kPi / 4.0);
  double a2 = a * a;
  // sin(a) = a - a^3/3! + ... ; cos(a) = 1 - a^2/2! + ...   (|a| <= pi/4, to a^15, a^16)
  double ps = 1.0 / 1307674368000.0;  // 1/15!
  ps = std::fma(ps, -a2, 1.0 / 6227020800.0);
  ps = std::fma(ps, -a2, 1.0 / 39916800.0);
  ps = std::fma(ps, -a2, 1.0 / 362880.0);
  ps = std::fma(ps, -a2, 1.0 / 5040.0);
  ps = std::fma(ps, -a2, 1.0 / 120.0);
  ps = std::fma(ps, -a2, 1.0 / 6.0);
  double sa = std::fma(-a * a2, ps, a);
  double pc = 1.0 / 20922789888000.0;  // 1/16!
  pc = std::fma(pc, -a2, 1.0 / 87178291200.0);
  pc = std::fma(pc, -a2, 1.0 / 479001600.0);
  pc = std::fma(pc, -a2, 1.0 / 3628800.0);
  pc = std::fma(pc, -a2, 1.0 / 40320.0);
  pc = std::fma(pc, -a2, 1.0 / 720.0);
  pc = std::fma(pc, -a2, 1.0 / 24.0);
  pc = std::fma(pc, -a2, 0.5);
  double ca = std::fma(-a2, pc, 1.0);
  // map octant back: angle = oct*pi/4 + (odd ? pi/4 - a : a)
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
  *s = sn; *c = cs;
}

// natural log for x > 0: x = m * 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s).
double log_own(double x) {
  int e;
  double m = std::frexp(x, &e);  // m in [0.5, 1)
  if (m < 0.70710678118654752) { m *= 2.0; e -= 1; }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double p = 1.0 / 25.0;
  p = std::fma(p, s2, 1.0 / 23.0);
  p = std::fma(p, s2, 1.0 / 21.0);
  p = std::fma(p, s2, 1.0 / 19.0);
  p = std::fma(p, s2, 1.0 / 17.0);
  p = std::fma(p, s2, 1.0 / 15.0);
  p = std::fma(p, s2, 1.0 / 13.0);
  p = std::fma(p, s2, 1.0 / 11.0);
  p = std::fma(p, s2, 1.0 / 9.0);
  p = std::fma(p, s2, 1.0 / 7.0);
  p = std::fma(p, s2, 1.0 / 5.0);
  p = std::fma(p, s2, 1.0 / 3.0);
  double lm = 2.0 * std::fma(s * s2, p, s);
  return std::fma((double)e, 0.69314718055994531, lm);
}

// ---------------------------------------------------------------- helpers
inline int wrapi(int i, int n) { return i >= n ? i - n : (i < 0 ? i + n : i); }

// Weight evaluation of the GPU mirror: t = x * inv with inv = 1/dx rounded once from the
// FP64 cell size (one multiply instead of a division); cell (int)t and fraction t - cell,
// both clamped.  The cell key used for sorting stays (int)(x/dx) (locate below).
template <typename R>
inline void locate_w(R x, R inv, int n, int* c, R* f) {
  R t = x * inv;
  int ci = (int)t;
  if (ci > n - 1) ci = n - 1;
  if (ci < 0) ci = 0;
  R fr = t - (R)ci;
  if (fr > (R)1) fr = (R)1;
  if (fr < (R)0) fr = (R)0;
  *c = ci;
  *f = fr;
}

template <typename R>
inline R wrap_pos(R x, R L) {
  if (x >= L) x -= L;
  else if (x < (R)0) x += L;
  if (!(x >= (R)0 && x < L)) {
    x = x - L * std::floor(x / L);
    if (!(x >= (R)0 && x < L)) x = (R)0;
  }
  return x;
}

// cell index + fractional offset, cx = (int)(x/dx) clamped to [0,n-1]
template <typename R>
inline void locate(R x, R dx, int n, int* c, R* f) {
  R t = x / dx;
  int ci = (int)t;
  if (ci > n - 1) ci = n - 1;
  if (ci < 0) ci = 0;
  R fr = t - (R)ci;
  if (fr > (R)1) fr = (R)1;
  if (fr < (R)0) fr = (R)0;
  *c = ci;
  *f = fr;
}

// --------------------------------------------------------------- FFT core
void fft1(cd* a, int n, int sign) {
  // iterative radix-2 DIT with bit-reversal
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (int len = 2; len <= n; len <<= 1) {
    for (int k = 0; k < len / 2; ++k) {
      double ang = sign * 2.0 * kPi * k / len;
      cd w(std::cos(ang), std::sin(ang));
      for (int i = 0; i < n; i += len) {
        cd u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
    }
  }
}

void fft2d(std::vector<cd>& a, int nx, int ny, int sign) {
  for (int i = 0; i < nx; ++i) fft1(&a[(size_t)i * ny], ny, sign);
  std::vector<cd> col(nx);
  for (int j = 0; j < ny; ++j) {
    for (int i = 0; i < nx; ++i) col[i] = a[(size_t)i * ny + j];
    fft1(col.data(), nx, sign);
    for (int i = 0; i < nx; ++i) a[(size_t)i * ny + j] = col[i];
  }
}

inline double kwave(int m, int n, double L) {
  int mm = (2 * m < n) ? m : m - n;
  return 2.0 * kPi * mm / L;
}

// spectral Poisson; also returns |rho_hat(1,0)|/(nx ny) via mode_amp
void poisson_impl(int nx, int ny, double lx, double ly, const double* rho, double* phi,
                  double* mode_amp) {
  std::vector<cd> a((size_t)nx * ny);
  for (size_t i = 0; i < a.size(); ++i) a[i] = cd(rho[i], 0.0);
  fft2d(a, nx, ny, -1);
  if (mode_amp) *mode_amp = std::abs(a[(size_t)1 * ny + 0]) / ((double)nx * ny);
  for (int i = 0; i < nx; ++i) {
    double kx = kwave(i, nx, lx);
    for (int j = 0; j < ny; ++j) {
      double ky = kwave(j, ny, ly);
      double k2 = kx * kx + ky * ky;
      size_t id = (size_t)i * ny + j;
      a[id] = (i == 0 && j == 0) ? cd(0.0, 0.0) : a[id] / k2;
    }
  }
  fft2d(a, nx, ny, +1);
  double norm = 1.0 / ((double)nx * ny);
  for (size_t i = 0; i < a.size(); ++i) phi[i] = a[i].real() * norm;
}

void efield_impl(int nx, int ny, double lx, double ly, const double* phi, double* ex,
                 double* ey) {
  double dx = lx / nx, dy = ly / ny;
  for (int i = 0; i < nx; ++i) {
    int im = wrapi(i - 1, nx), ip = wrapi(i + 1, nx);
    for (int j = 0; j < ny; ++j) {
      int jm = wrapi(j - 1, ny), jp = wrapi(j + 1, ny);
      ex[(size_t)i * ny + j] = (phi[(size_t)im * ny + j] - phi[(size_t)ip * ny + j]) / (2.0 * dx);
      ey[(size_t)i * ny + j] = (phi[(size_t)i * ny + jm] - phi[(size_t)i * ny + jp]) / (2.0 * dy);
    }
  }
}

double field_energy(int nx, int ny, double lx, double ly, const double* ex, const double* ey) {
  double s = 0.0;
  for (size_t i = 0; i < (size_t)nx * ny; ++i) s += ex[i] * ex[i] + ey[i] * ey[i];
  return 0.5 * s * (lx / nx) * (ly / ny);
}

// continuous-weight FP64 push (drift optional); returns sync-velocity sums
void push_cont(int64_t n, double* x, double* y, double* vx, double* vy, const double* ex,
               const double* ey, int nx, int ny, double lx, double ly, double dt, double qm,
               double c, int drift, double* ke, double* px, double* py) {
  double dx = lx / nx, dy = ly / ny;
  double sk = 0, sx = 0, sy = 0;
  for (int64_t p = 0; p < n; ++p) {
    int cx, cy;
    double fx, fy;
    locate(x[p], dx, nx, &cx, &fx);
    locate(y[p], dy, ny, &cy, &fy);
    int cxp = wrapi(cx + 1, nx), cyp = wrapi(cy + 1, ny);
    double w00 = (1 - fx) * (1 - fy), w10 = fx * (1 - fy), w01 = (1 - fx) * fy, w11 = fx * fy;
    size_t n00 = (size_t)cx * ny + cy, n10 = (size_t)cxp * ny + cy, n01 = (size_t)cx * ny + cyp,
           n11 = (size_t)cxp * ny + cyp;
    double ax = qm * (w00 * ex[n00] + w10 * ex[n10] + w01 * ex[n01] + w11 * ex[n11]);
    double ay = qm * (w00 * ey[n00] + w10 * ey[n10] + w01 * ey[n01] + w11 * ey[n11]);
    double vsx = vx[p] + (c - 0.5) * dt * ax, vsy = vy[p] + (c - 0.5) * dt * ay;
    sk += vsx * vsx + vsy * vsy;
    sx += vsx;
    sy += vsy;
    vx[p] += c * dt * ax;
    vy[p] += c * dt * ay;
    if (drift) {
      x[p] = wrap_pos(x[p] + dt * vx[p], lx);
      y[p] = wrap_pos(y[p] + dt * vy[p], ly);
    }
  }
  if (ke) *ke = sk;
  if (px) *px = sx;
  if (py) *py = sy;
}

// GPU-mirror push: per-axis fixed-point gather weights, fixed evaluation order, the same
// fused multiply-adds as the device code (std::fma: one rounding, as v_fma_f32/f64).
template <typename R, int QB>
void push_mirror(int64_t n, R* x, R* y, R* vx, R* vy, const R* ex, const R* ey, int nx,
                 int ny, double lx, double ly, double dt, double qm, double c, double* ke,
                 double* px, double* py) {
  const R idx = (R)(1.0 / (lx / nx)), idy = (R)(1.0 / (ly / ny)), Lx = (R)lx, Ly = (R)ly;
  const R scale = (R)(1.0 / (double)(1 << QB));
  const R fq = (R)(double)(1 << QB);
  const R hk = (R)(c * dt * qm), dtr = (R)dt;
  const double sk_c = (c - 0.5) * dt * qm;
  double sk = 0, sx = 0, sy = 0;
  for (int64_t p = 0; p < n; ++p) {
    int cx, cy;
    R fx, fy;
    locate_w(x[p], idx, nx, &cx, &fx);
    locate_w(y[p], idy, ny, &cy, &fy);
    int ix1 = (int)std::rint(fx * fq), iy1 = (int)std::rint(fy * fq);
    R gx1 = (R)ix1 * scale, gy1 = (R)iy1 * scale;
    R gx0 = (R)1 - gx1, gy0 = (R)1 - gy1;
    R g00 = gx0 * gy0, g10 = gx1 * gy0, g01 = gx0 * gy1, g11 = gx1 * gy1;
    int cxp = wrapi(cx + 1, nx), cyp = wrapi(cy + 1, ny);
    size_t n00 = (size_t)cx * ny + cy, n10 = (size_t)cxp * ny + cy, n01 = (size_t)cx * ny + cyp,
           n11 = (size_t)cxp * ny + cyp;
    R epx = g00 * ex[n00];
    epx = std::fma(g10, ex[n10], epx);
    epx = std::fma(g01, ex[n01], epx);
    epx = std::fma(g11, ex[n11], epx);
    R epy = g00 * ey[n00];
    epy = std::fma(g10, ey[n10], epy);
    epy = std::fma(g01, ey[n01], epy);
    epy = std::fma(g11, ey[n11], epy);
    double vsx = (double)vx[p] + sk_c * (double)epx, vsy = (double)vy[p] + sk_c * (double)epy;
    sk += vsx * vsx + vsy * vsy;
    sx += vsx;
    sy += vsy;
    R nvx = std::fma(hk, epx, vx[p]), nvy = std::fma(hk, epy, vy[p]);
    vx[p] = nvx;
    vy[p] = nvy;
    x[p] = wrap_pos(std::fma(dtr, nvx, x[p]), Lx);
    y[p] = wrap_pos(std::fma(dtr, nvy, y[p]), Ly);
  }
  if (ke) *ke = sk;
  if (px) *px = sx;
  if (py) *py = sy;
}

template <typename R, int QB>
void fixed_weights(R x, R y, R idx, R idy, int nx, int ny, int* cx, int* cy, int64_t w[4]) {
  R fx, fy;
  locate_w(x, idx, nx, cx, &fx);
  locate_w(y, idy, ny, cy, &fy);
  const R fq = (R)(double)(1 << QB);
  int64_t ix1 = (int64_t)std::rint(fx * fq), iy1 = (int64_t)std::rint(fy * fq);
  int64_t ix0 = ((int64_t)1 << QB) - ix1, iy0 = ((int64_t)1 << QB) - iy1;
  w[0] = ix0 * iy0;
  w[1] = ix1 * iy0;
  w[2] = ix0 * iy1;
  w[3] = ix1 * iy1;
}

// final half kick of a run (mirror arithmetic, no drift)
void half_kick_f32(int64_t n, const float* x, const float* y, float* vx, float* vy,
                   const float* exf, const float* eyf, int nx, int ny, double lx, double ly,
                   double dt, double qm) {
  const float idx = (float)(1.0 / (lx / nx)), idy = (float)(1.0 / (ly / ny));
  const float hk = (float)(0.5 * dt * qm);
  for (int64_t p = 0; p < n; ++p) {
    int cx, cy;
    float fx, fy;
    locate_w(x[p], idx, nx, &cx, &fx);
    locate_w(y[p], idy, ny, &cy, &fy);
    int ix1 = (int)std::rint(fx * 262144.0f), iy1 = (int)std::rint(fy * 262144.0f);
    float gx1 = (float)ix1 * (1.0f / 262144.0f), gy1 = (float)iy1 * (1.0f / 262144.0f);
    float gx0 = 1.0f - gx1, gy0 = 1.0f - gy1;
    float g00 = gx0 * gy0, g10 = gx1 * gy0, g01 = gx0 * gy1, g11 = gx1 * gy1;
    int cxp = wrapi(cx + 1, nx), cyp = wrapi(cy + 1, ny);
    size_t n00 = (size_t)cx * ny + cy, n10 = (size_t)cxp * ny + cy,
           n01 = (size_t)cx * ny + cyp, n11 = (size_t)cxp * ny + cyp;
    float epx = g00 * exf[n00];
    epx = std::fma(g10, exf[n10], epx);
    epx = std::fma(g01, exf[n01], epx);
    epx = std::fma(g11, exf[n11], epx);
    float epy = g00 * eyf[n00];
    epy = std::fma(g10, eyf[n10], epy);
    epy = std::fma(g01, eyf[n01], epy);
    epy = std::fma(g11, eyf[n11], epy);
    vx[p] = std::fma(hk, epx, vx[p]);
    vy[p] = std::fma(hk, epy, vy[p]);
  }
}

// ---- threaded pieces of the FP32 mirror run (the all-cores CPU baseline).  Each does the
// serial arithmetic on its share: FFT rows / columns are independent transforms, particles
// are split in contiguous chunks, and the int64 charge of the chunks is summed per node
// (integer, so identical to the serial deposit).
void fft2d_threads(std::vector<cd>& a, int nx, int ny, int sign) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) fft1(&a[(size_t)i * ny], ny, sign);
#pragma omp parallel
  {
    std::vector<cd> col(nx);
#pragma omp for schedule(static)
    for (int j = 0; j < ny; ++j) {
      for (int i = 0; i < nx; ++i) col[i] = a[(size_t)i * ny + j];
      fft1(col.data(), nx, sign);
      for (int i = 0; i < nx; ++i) a[(size_t)i * ny + j] = col[i];
    }
  }
}

void poisson_threads(int nx, int ny, double lx, double ly, const double* rho, double* phi,
                     double* mode_amp) {
  std::vector<cd> a((size_t)nx * ny);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)a.size(); ++i) a[i] = cd(rho[i], 0.0);
  fft2d_threads(a, nx, ny, -1);
  if (mode_amp) *mode_amp = std::abs(a[(size_t)1 * ny + 0]) / ((double)nx * ny);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) {
    double kx = kwave(i, nx, lx);
    for (int j = 0; j < ny; ++j) {
      double ky = kwave(j, ny, ly);
      double k2 = kx * kx + ky * ky;
      size_t id = (size_t)i * ny + j;
      a[id] = (i == 0 && j == 0) ? cd(0.0, 0.0) : a[id] / k2;
    }
  }
  fft2d_threads(a, nx, ny, +1);
  double norm = 1.0 / ((double)nx * ny);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)a.size(); ++i) phi[i] = a[i].real() * norm;
}

// ---------------------------------------------------------------- DSMC pairing
const uint32_t TAG_DSMP = 0x44534d50u;

uint32_t feistel_f(uint32_t r, uint32_t k, int half) {
  uint32_t h = (r ^ k) * 0x9E3779B1u;
  h ^= h >> 15;
  h *= 0x85EBCA77u;
  h ^= h >> 13;
  return h >> (32 - half);
}

// slot t of [0, n) through a 12-round balanced Feistel network on 2*half bits (the
// smallest even width with 2^(2 half) >= n; round i keyed by key[i % 4] ^ i*0x632BE5AB),
// repeated until the value is below n (cycle walking: the walk stays on the cycle through
// t, which holds t itself).  Twelve rounds: with four, slot pairs that differ only in the
// low half were paired measurably more often than others on domains of 4-6 bits.
uint32_t pair_perm(uint32_t t, uint32_t n, int half, const uint32_t key[4]) {
  const uint32_t mask = (1u << half) - 1u;
  uint32_t x = t;
  do {
    uint32_t l = x >> half, r = x & mask;
    for (uint32_t i = 0; i < 12; ++i) {
      const uint32_t f = feistel_f(r, key[i & 3] ^ (i * 0x632BE5ABu), half);
      const uint32_t nl = r;
      r = l ^ f;
      l = nl;
    }
    x = (l << half) | r;
  } while (x >= n);
  return x;
}

int pair_half_bits(uint32_t n) {
  int b = 32 - __builtin_clz(n - 1u);  // n >= 2
  b += b & 1;
  return b / 2;
}

}  // namespace

// =================================================================== C API
extern "C" {

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox(ctr, key, out);
}
void or_philox2x32_10(const uint32_t ctr[2], const uint32_t key[1], uint32_t out[2]) {
  philox2(ctr, key[0], out);
}

uint64_t or_minstd_rand0_nth(uint32_t seed, int n) {
  std::minstd_rand0 g(seed);
  uint64_t v = 0;
  for (int i = 0; i < n; ++i) v = g();
  return v;
}

uint64_t or_mt19937_nth(uint32_t seed, int n) {
  std::mt19937 g(seed);
  uint64_t v = 0;
  for (int i = 0; i < n; ++i) v = g();
  return v;
}

void or_init_reference(int npx, int npy, int xdim, int ydim, int nx, int ny, double vtx,
                       double vty, double vdx, double vdy, double* x, double* y, double* vx,
                       double* vy) {
  // thesisCode/collisionless/double/src/main.cc:201-244 (integer-division offset kept)
  std::default_random_engine gen(1234);
  std::normal_distribution<double> nd{0.0, 1.0};
  double xoff = 0.5 * (double)(xdim / nx), yoff = 0.5 * (double)(ydim / ny);
  for (int i = 0; i < npy; ++i) {
    double yc = ((double)ydim / npy) * (double)i + yoff;
    for (int j = 0; j < npx; ++j) {
      size_t p = (size_t)i * npx + j;
      x[p] = ((double)xdim / npx) * (double)j + xoff;
      y[p] = yc;
      vx[p] = vtx * nd(gen);
      vy[p] = vty * nd(gen);
    }
  }
  size_t n = (size_t)npx * npy;
  double sx = 0, sy = 0;
  for (size_t p = 0; p < n; ++p) { sx += vx[p]; sy += vy[p]; }
  double inv = 1.0 / (double)n;
  sx = inv * sx - vdx;
  sy = inv * sy - vdy;
  for (size_t p = 0; p < n; ++p) { vx[p] -= sx; vy[p] -= sy; }
}

void or_lattice_shape(int nx, int ny, int ppc, int* npx, int* npy) {
  int a = 1;
  for (int d = 1; d <= ppc; ++d)
    if (ppc % d == 0 && (double)d >= std::sqrt((double)ppc) - 1e-9) { a = d; break; }
  *npx = nx * a;
  *npy = ny * (ppc / a);
}

static void normals_pair(uint64_t seed, uint32_t p, double* n1, double* n2) {
  double u[4];
  draw4(p, 0u, 0u, TAG_LOAD, seed, u);
  double r = std::sqrt(-2.0 * std::log(u[0]));
  double th = 2.0 * kPi * u[1];
  *n1 = r * std::cos(th);
  *n2 = r * std::sin(th);
}

void or_init_two_stream(int nx, int ny, double lx, double ly, int ppc, double v0, double vth,
                        double amp, int mode, uint64_t seed, double* x, double* y, double* vx,
                        double* vy) {
  int npx, npy;
  or_lattice_shape(nx, ny, ppc, &npx, &npy);
  double k = 2.0 * kPi * mode / lx;
  for (int i = 0; i < npx; ++i) {
    double xl = (i + 0.5) * lx / npx;
    double xp = wrap_pos(xl + amp * std::cos(k * xl), lx);
    for (int j = 0; j < npy; ++j) {
      size_t p = (size_t)i * npy + j;
      x[p] = xp;
      y[p] = (j + 0.5) * ly / npy;
      double n1, n2;
      normals_pair(seed, (uint32_t)p, &n1, &n2);
      vx[p] = ((p & 1) ? -v0 : v0) + vth * n1;
      vy[p] = vth * n2;
    }
  }
}

void or_init_landau(int nx, int ny, double lx, double ly, int ppc, double vth, double alpha,
                    int mode, uint64_t seed, double* x, double* y, double* vx, double* vy) {
  int npx, npy;
  or_lattice_shape(nx, ny, ppc, &npx, &npy);
  double k = 2.0 * kPi * mode / lx;
  for (int i = 0; i < npx; ++i) {
    // inverse CDF of 1 + alpha cos(k x) on [0, lx): F(x) = x + (alpha/k) sin(kx)
    double target = (i + 0.5) * lx / npx;
    double xs = target;
    for (int it = 0; it < 60; ++it) {
      double f = xs + (alpha / k) * std::sin(k * xs) - target;
      double d = 1.0 + alpha * std::cos(k * xs);
      double step = f / d;
      xs -= step;
      if (std::fabs(step) < 1e-15 * lx) break;
    }
    xs = wrap_pos(xs, lx);
    for (int j = 0; j < npy; ++j) {
      size_t p = (size_t)i * npy + j;
      x[p] = xs;
      y[p] = (j + 0.5) * ly / npy;
      double n1, n2;
      normals_pair(seed, (uint32_t)p, &n1, &n2);
      vx[p] = vth * n1;
      vy[p] = vth * n2;
    }
  }
}

void or_cell_keys_f64(int64_t n, const double* x, const double* y, int nx, int ny, double lx,
                      double ly, int32_t* keys) {
  double dx = lx / nx, dy = ly / ny;
  for (int64_t p = 0; p < n; ++p) {
    int cx, cy;
    double f;
    locate(x[p], dx, nx, &cx, &f);
    locate(y[p], dy, ny, &cy, &f);
    keys[p] = cx * ny + cy;
  }
}

void or_cell_keys_f32(int64_t n, const float* x, const float* y, int nx, int ny, double lx,
                      double ly, int32_t* keys) {
  float dx = (float)(lx / nx), dy = (float)(ly / ny);
  for (int64_t p = 0; p < n; ++p) {
    int cx, cy;
    float f;
    locate(x[p], dx, nx, &cx, &f);
    locate(y[p], dy, ny, &cy, &f);
    keys[p] = cx * ny + cy;
  }
}

void or_stable_sort(int64_t n, const int32_t* keys, int ncells, int32_t* perm, int32_t* counts,
                    int32_t* offsets) {
  std::vector<int32_t> cnt(ncells, 0);
  for (int64_t p = 0; p < n; ++p) cnt[keys[p]]++;
  std::vector<int32_t> off(ncells + 1, 0);
  for (int c = 0; c < ncells; ++c) off[c + 1] = off[c] + cnt[c];
  std::vector<int32_t> cur(off.begin(), off.end() - 1);
  for (int64_t p = 0; p < n; ++p) perm[cur[keys[p]]++] = (int32_t)p;
  if (counts) std::memcpy(counts, cnt.data(), sizeof(int32_t) * ncells);
  if (offsets) std::memcpy(offsets, off.data(), sizeof(int32_t) * (ncells + 1));
}

void or_deposit_f64(int64_t n, const double* x, const double* y, int nx, int ny, double lx,
                    double ly, double qp, double* rho) {
  double dx = lx / nx, dy = ly / ny, s = qp / (dx * dy);
  std::fill(rho, rho + (size_t)nx * ny, 0.0);
  for (int64_t p = 0; p < n; ++p) {
    int cx, cy;
    double fx, fy;
    locate(x[p], dx, nx, &cx, &fx);
    locate(y[p], dy, ny, &cy, &fy);
    int cxp = wrapi(cx + 1, nx), cyp = wrapi(cy + 1, ny);
    rho[(size_t)cx * ny + cy] += s * (1 - fx) * (1 - fy);
    rho[(size_t)cxp * ny + cy] += s * fx * (1 - fy);
    rho[(size_t)cx * ny + cyp] += s * (1 - fx) * fy;
    rho[(size_t)cxp * ny + cyp] += s * fx * fy;
  }
}

void or_deposit_fixed_f32(int64_t n, const float* x, const float* y, int nx, int ny, double lx,
                          double ly, int64_t* acc) {
  const float idx = (float)(1.0 / (lx / nx)), idy = (float)(1.0 / (ly / ny));
  std::fill(acc, acc + (size_t)nx * ny, (int64_t)0);
  for (int64_t p = 0; p < n; ++p) {
    int cx, cy;
    int64_t w[4];
    fixed_weights<float, 18>(x[p], y[p], idx, idy, nx, ny, &cx, &cy, w);
    int cxp = wrapi(cx + 1, nx), cyp = wrapi(cy + 1, ny);
    acc[(size_t)cx * ny + cy] += w[0];
    acc[(size_t)cxp * ny + cy] += w[1];
    acc[(size_t)cx * ny + cyp] += w[2];
    acc[(size_t)cxp * ny + cyp] += w[3];
  }
}

void or_deposit_fixed_f64(int64_t n, const double* x, const double* y, int nx, int ny,
                          double lx, double ly, int64_t* hi, int64_t* lo) {
  const double idx = 1.0 / (lx / nx), idy = 1.0 / (ly / ny);
  std::fill(hi, hi + (size_t)nx * ny, (int64_t)0);
  std::fill(lo, lo + (size_t)nx * ny, (int64_t)0);
  const int64_t mask = ((int64_t)1 << 30) - 1;
  for (int64_t p = 0; p < n; ++p) {
    int cx, cy;
    int64_t w[4];
    fixed_weights<double, 30>(x[p], y[p], idx, idy, nx, ny, &cx, &cy, w);
    int cxp = wrapi(cx + 1, nx), cyp = wrapi(cy + 1, ny);
    size_t nd[4] = {(size_t)cx * ny + cy, (size_t)cxp * ny + cy, (size_t)cx * ny + cyp,
                    (size_t)cxp * ny + cyp};
    for (int k = 0; k < 4; ++k) {
      hi[nd[k]] += w[k] >> 30;
      lo[nd[k]] += w[k] & mask;
    }
  }
}

void or_fixed_to_rho_f32(int nx, int ny, double lx, double ly, double qp, const int64_t* acc,
                         double* rho) {
  double scale = qp / ((lx / nx) * (ly / ny)) * std::ldexp(1.0, -36);
  for (size_t i = 0; i < (size_t)nx * ny; ++i) rho[i] = (double)acc[i] * scale;
}

void or_fixed_to_rho_f64(int nx, int ny, double lx, double ly, double qp, const int64_t* hi,
                         const int64_t* lo, double* rho) {
  double scale = qp / ((lx / nx) * (ly / ny));
  for (size_t i = 0; i < (size_t)nx * ny; ++i)
    rho[i] = ((double)hi[i] * std::ldexp(1.0, -30) + (double)lo[i] * std::ldexp(1.0, -60)) * scale;
}

void or_fft2(int nx, int ny, double* data, int sign) {
  std::vector<cd> a((size_t)nx * ny);
  for (size_t i = 0; i < a.size(); ++i) a[i] = cd(data[2 * i], data[2 * i + 1]);
  fft2d(a, nx, ny, sign);
  for (size_t i = 0; i < a.size(); ++i) { data[2 * i] = a[i].real(); data[2 * i + 1] = a[i].imag(); }
}

void or_dft2_brute(int nx, int ny, const double* in, double* out, int sign) {
  for (int k = 0; k < nx; ++k)
    for (int l = 0; l < ny; ++l) {
      cd s(0, 0);
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
          double ang = sign * 2.0 * kPi * ((double)((int64_t)k * i % nx) / nx +
                                           (double)((int64_t)l * j % ny) / ny);
          s += cd(in[2 * ((size_t)i * ny + j)], in[2 * ((size_t)i * ny + j) + 1]) *
               cd(std::cos(ang), std::sin(ang));
        }
      out[2 * ((size_t)k * ny + l)] = s.real();
      out[2 * ((size_t)k * ny + l) + 1] = s.imag();
    }
}

void or_poisson(int nx, int ny, double lx, double ly, const double* rho, double* phi) {
  poisson_impl(nx, ny, lx, ly, rho, phi, nullptr);
}

void or_efield(int nx, int ny, double lx, double ly, const double* phi, double* ex, double* ey) {
  efield_impl(nx, ny, lx, ly, phi, ex, ey);
}

void or_push_f64(int64_t n, double* x, double* y, double* vx, double* vy, const double* ex,
                 const double* ey, int nx, int ny, double lx, double ly, double dt, double qm,
                 double c, int quantized, double* ke_out, double* px_out, double* py_out) {
  if (quantized)
    push_mirror<double, 30>(n, x, y, vx, vy, ex, ey, nx, ny, lx, ly, dt, qm, c, ke_out, px_out,
                            py_out);
  else
    push_cont(n, x, y, vx, vy, ex, ey, nx, ny, lx, ly, dt, qm, c, 1, ke_out, px_out, py_out);
}

void or_push_f32(int64_t n, float* x, float* y, float* vx, float* vy, const float* ex,
                 const float* ey, int nx, int ny, double lx, double ly, double dt, double qm,
                 double c, double* ke_out, double* px_out, double* py_out) {
  push_mirror<float, 18>(n, x, y, vx, vy, ex, ey, nx, ny, lx, ly, dt, qm, c, ke_out, px_out,
                         py_out);
}

void or_gather_f64(int64_t n, const double* x, const double* y, const double* ex,
                   const double* ey, int nx, int ny, double lx, double ly, double* epx,
                   double* epy) {
  double dx = lx / nx, dy = ly / ny;
  for (int64_t p = 0; p < n; ++p) {
    int cx, cy;
    double fx, fy;
    locate(x[p], dx, nx, &cx, &fx);
    locate(y[p], dy, ny, &cy, &fy);
    int cxp = wrapi(cx + 1, nx), cyp = wrapi(cy + 1, ny);
    double w00 = (1 - fx) * (1 - fy), w10 = fx * (1 - fy), w01 = (1 - fx) * fy, w11 = fx * fy;
    size_t n00 = (size_t)cx * ny + cy, n10 = (size_t)cxp * ny + cy, n01 = (size_t)cx * ny + cyp,
           n11 = (size_t)cxp * ny + cyp;
    epx[p] = w00 * ex[n00] + w10 * ex[n10] + w01 * ex[n01] + w11 * ex[n11];
    epy[p] = w00 * ey[n00] + w10 * ey[n10] + w01 * ey[n01] + w11 * ey[n11];
  }
}

void or_run_f64(int64_t n, double* x, double* y, double* vx, double* vy, int nx, int ny,
                double lx, double ly, double dt, double qm, double qp, int nsteps,
                double* diag) {
  size_t G = (size_t)nx * ny;
  std::vector<double> rho(G), phi(G), ex(G), ey(G);
  double mp = qp / qm;
  for (int s = 0; s <= nsteps; ++s) {
    or_deposit_f64(n, x, y, nx, ny, lx, ly, qp, rho.data());
    double amp;
    poisson_impl(nx, ny, lx, ly, rho.data(), phi.data(), &amp);
    efield_impl(nx, ny, lx, ly, phi.data(), ex.data(), ey.data());
    if (s == nsteps) {
      // final half kick to synchronous velocities (no drift)
      push_cont(n, x, y, vx, vy, ex.data(), ey.data(), nx, ny, lx, ly, dt, qm, 0.5, 0, nullptr,
                nullptr, nullptr);
      break;
    }
    double ke, px, py;
    push_cont(n, x, y, vx, vy, ex.data(), ey.data(), nx, ny, lx, ly, dt, qm, s == 0 ? 0.5 : 1.0,
              1, &ke, &px, &py);
    if (diag) {
      double* d = diag + (size_t)s * 5;
      d[0] = field_energy(nx, ny, lx, ly, ex.data(), ey.data());
      d[1] = 0.5 * mp * ke;
      d[2] = mp * px;
      d[3] = mp * py;
      d[4] = amp;
    }
  }
}

void or_run_f32(int64_t n, float* x, float* y, float* vx, float* vy, int nx, int ny, double lx,
                double ly, double dt, double qm, double qp, int nsteps, double* diag) {
  size_t G = (size_t)nx * ny;
  std::vector<int64_t> acc(G);
  std::vector<double> rho(G), phi(G), ex(G), ey(G);
  std::vector<float> exf(G), eyf(G);
  double mp = qp / qm;
  for (int s = 0; s <= nsteps; ++s) {
    or_deposit_fixed_f32(n, x, y, nx, ny, lx, ly, acc.data());
    or_fixed_to_rho_f32(nx, ny, lx, ly, qp, acc.data(), rho.data());
    double amp;
    poisson_impl(nx, ny, lx, ly, rho.data(), phi.data(), &amp);
    efield_impl(nx, ny, lx, ly, phi.data(), ex.data(), ey.data());
    for (size_t i = 0; i < G; ++i) { exf[i] = (float)ex[i]; eyf[i] = (float)ey[i]; }
    if (s == nsteps) {
      half_kick_f32(n, x, y, vx, vy, exf.data(), eyf.data(), nx, ny, lx, ly, dt, qm);
      break;
    }
    double ke, px, py;
    push_mirror<float, 18>(n, x, y, vx, vy, exf.data(), eyf.data(), nx, ny, lx, ly, dt, qm,
                           s == 0 ? 0.5 : 1.0, &ke, &px, &py);
    if (diag) {
      double* d = diag + (size_t)s * 5;
      d[0] = field_energy(nx, ny, lx, ly, ex.data(), ey.data());
      d[1] = 0.5 * mp * ke;
      d[2] = mp * px;
      d[3] = mp * py;
      d[4] = amp;
    }
  }
}

void or_run_f32_threads(int64_t n, float* x, float* y, float* vx, float* vy, int nx, int ny,
                        double lx, double ly, double dt, double qm, double qp, int nsteps,
                        int nthreads, double* diag) {
  const int nt = nthreads > 0 ? nthreads : 1;
  const size_t G = (size_t)nx * ny;
  std::vector<int64_t> acc(G), part((size_t)nt * G);
  std::vector<double> rho(G), phi(G), ex(G), ey(G), sums((size_t)nt * 3);
  std::vector<float> exf(G), eyf(G);
  const double mp = qp / qm;
  auto chunk = [&](int t, int64_t* b, int64_t* e) {
    *b = n * t / nt;
    *e = n * (t + 1) / nt;
  };
#ifdef _OPENMP
  const int saved = omp_get_max_threads();
  omp_set_num_threads(nt);
#endif
  for (int s = 0; s <= nsteps; ++s) {
#pragma omp parallel for schedule(static, 1)
    for (int t = 0; t < nt; ++t) {
      int64_t b, e;
      chunk(t, &b, &e);
      or_deposit_fixed_f32(e - b, x + b, y + b, nx, ny, lx, ly, part.data() + (size_t)t * G);
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)G; ++i) {
      int64_t v = 0;
      for (int t = 0; t < nt; ++t) v += part[(size_t)t * G + i];
      acc[i] = v;
    }
    or_fixed_to_rho_f32(nx, ny, lx, ly, qp, acc.data(), rho.data());
    double amp;
    poisson_threads(nx, ny, lx, ly, rho.data(), phi.data(), &amp);
    efield_impl(nx, ny, lx, ly, phi.data(), ex.data(), ey.data());
    for (size_t i = 0; i < G; ++i) { exf[i] = (float)ex[i]; eyf[i] = (float)ey[i]; }
    if (s == nsteps) {
#pragma omp parallel for schedule(static, 1)
      for (int t = 0; t < nt; ++t) {
        int64_t b, e;
        chunk(t, &b, &e);
        half_kick_f32(e - b, x + b, y + b, vx + b, vy + b, exf.data(), eyf.data(), nx, ny, lx,
                      ly, dt, qm);
      }
      break;
    }
#pragma omp parallel for schedule(static, 1)
    for (int t = 0; t < nt; ++t) {
      int64_t b, e;
      chunk(t, &b, &e);
      push_mirror<float, 18>(e - b, x + b, y + b, vx + b, vy + b, exf.data(), eyf.data(), nx,
                             ny, lx, ly, dt, qm, s == 0 ? 0.5 : 1.0, &sums[3 * t],
                             &sums[3 * t + 1], &sums[3 * t + 2]);
    }
    if (diag) {
      double ke = 0, px = 0, py = 0;
      for (int t = 0; t < nt; ++t) { ke += sums[3 * t]; px += sums[3 * t + 1]; py += sums[3 * t + 2]; }
      double* d = diag + (size_t)s * 5;
      d[0] = field_energy(nx, ny, lx, ly, ex.data(), ey.data());
      d[1] = 0.5 * mp * ke;
      d[2] = mp * px;
      d[3] = mp * py;
      d[4] = amp;
    }
  }
#ifdef _OPENMP
  omp_set_num_threads(saved);
#endif
}

// ------------------------------------------------------------ collisions
void or_mcc_cross_sections(double e, double* s_el, double* s_ex, double* s_iz) {
  *s_el = 1e-19 / (1.0 + e / 10.0);
  *s_ex = (e > 11.5) ? 2e-21 * (1.0 - 11.5 / e) : 0.0;
  *s_iz = (e > 15.8) ? 3e-20 * log_own(e / 15.8) / e : 0.0;
}

// MCC slot of thesisCode/mccCollisions/src/particle.cc:112-127 (collision after the push),
// with null-collision physics: process chosen by u * nu_max against cumulative nu_k(E).
int64_t or_mcc_step_f32(int64_t n, int64_t cap, float* x, float* y, float* vx, float* vy,
                        int32_t* id, int64_t next_id, double e_per_v2,
                        double nu_dt_per_sigma_v, double p_coll, double nu_max_dt,
                        double mass_ratio2, uint64_t seed, int64_t step, int64_t* counts) {
  struct Birth { int32_t parent; float x, y, vx, vy; };
  std::vector<Birth> births;
  const uint32_t s32 = (uint32_t)step;
  for (int64_t p = 0; p < n; ++p) {
    // null test: word 0 of Philox2x32-10 {id, step}, key seed_lo ^ seed_hi ^ 'MCC1'
    const uint32_t nc[2] = {(uint32_t)id[p], s32};
    uint32_t nu[2];
    philox2(nc, (uint32_t)seed ^ (uint32_t)(seed >> 32) ^ TAG_MCC, nu);
    if (u01(nu[0]) >= p_coll) continue;
    double u[4];  // the collision's uniforms: u[1] process, u[2] angle, u[3] energy split
    draw4((uint32_t)id[p], s32, 0u, TAG_MCC, seed, u);
    double vxd = (double)vx[p], vyd = (double)vy[p];
    const double v2 = vxd * vxd + vyd * vyd;
    const double sp = std::sqrt(v2);
    const double e = e_per_v2 * v2;
    double sel, sex, siz;
    or_mcc_cross_sections(e, &sel, &sex, &siz);
    const double nel = nu_dt_per_sigma_v * sel * sp;
    const double nex = nu_dt_per_sigma_v * sex * sp;
    const double niz = nu_dt_per_sigma_v * siz * sp;
    const double r = u[1] * nu_max_dt;
    double sn, cs;
    sincos2pi(u[2], &sn, &cs);
    if (r < nel) {
      const double f = std::sqrt(1.0 - mass_ratio2 * (1.0 - cs));
      const double nvx = f * (vxd * cs - vyd * sn), nvy = f * (vxd * sn + vyd * cs);
      vxd = nvx; vyd = nvy;
      counts[0]++;
    } else if (r < nel + nex) {
      const double f = std::sqrt((e - 11.5) / e);
      const double nvx = f * (vxd * cs - vyd * sn), nvy = f * (vxd * sn + vyd * cs);
      vxd = nvx; vyd = nvy;
      counts[1]++;
    } else if (r < nel + nex + niz) {
      const double ea = e - 15.8;
      const double e1 = u[3] * ea, e2 = ea - e1;
      const double f1 = std::sqrt(e1 / e);
      const double nvx = f1 * (vxd * cs - vyd * sn), nvy = f1 * (vxd * sn + vyd * cs);
      double w[4];
      draw4((uint32_t)id[p], s32, 1u, TAG_MCC, seed, w);
      double sn2, cs2;
      sincos2pi(w[0], &sn2, &cs2);
      const double sp2 = sp * std::sqrt(e2 / e);
      births.push_back({id[p], x[p], y[p], (float)(sp2 * cs2), (float)(sp2 * sn2)});
      vxd = nvx; vyd = nvy;
      counts[2]++;
    } else {
      counts[3]++;
      continue;
    }
    vx[p] = (float)vxd;
    vy[p] = (float)vyd;
  }
  std::sort(births.begin(), births.end(),
            [](const Birth& a, const Birth& b) { return a.parent < b.parent; });
  int64_t m = n;
  for (size_t k = 0; k < births.size() && m < cap; ++k, ++m) {
    x[m] = births[k].x; y[m] = births[k].y; vx[m] = births[k].vx; vy[m] = births[k].vy;
    id[m] = (int32_t)(next_id + (int64_t)k);
  }
  return m;
}

// DSMC slot of thesisCode/dsmcCollisions/cellTrackerAdvanced/src/main.cc:288-350 (partner
// search per cell after the push; GPU twin cellTrackerAdvanced/src/pic.cu:791-864), with
// NTC hard-sphere pair selection (Bird) in rounds of disjoint pairs.  Per cell of n >= 2
// particles: M = floor(mr) + [u < frac(mr)] candidates, mr = 1/2 n (n-1) coeff gmax.
// Candidate k belongs to round k / P (P = n / 2 pairs per round) and is the pair of slots
// (pi(2m), pi(2m+1)), m = k % P, under the round's pseudo-random permutation pi of the
// cell's slots; the pairs of a round are disjoint, so a round's candidates are
// independent.  Candidate k is accepted when u gm < g with gm = max(gmax at the round's
// start, g); the pair's centre-of-mass velocity is kept and its relative velocity turned
// to a uniform direction.  After the round gmax = max(gmax, all g of the round).
// Streams: counter {cell, step, 0xffffffff, DSMC} for M, {cell, step, k, DSMC} for
// candidate k, {cell, step, 0x80000000 | round, DSMP} for the permutation keys.
int64_t or_dsmc_step_f32(int ncells, const int32_t* offsets, float* vx, float* vy,
                         float* gmax, double coeff, uint64_t seed, int64_t step) {
  const uint32_t s32 = (uint32_t)step;
  const uint32_t skey[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  int64_t total = 0;
  for (int c = 0; c < ncells; ++c) {
    const int b = offsets[c], n = offsets[c + 1] - b;
    if (n < 2) continue;
    double gm = (double)gmax[c];
    double u[4];
    draw4((uint32_t)c, s32, 0xffffffffu, TAG_DSMC, seed, u);
    const double mr = 0.5 * (double)n * (double)(n - 1) * coeff * gm;
    const double fl = std::floor(mr);
    int64_t M = (int64_t)fl + ((u[0] < mr - fl) ? 1 : 0);
    if (M > 0x7fffffffll) M = 0x7fffffffll;
    const int64_t P = n / 2;
    const int half = pair_half_bits((uint32_t)n);
    for (int64_t r0 = 0; r0 < M; r0 += P) {
      const uint32_t ctr[4] = {(uint32_t)c, s32, 0x80000000u | (uint32_t)(r0 / P), TAG_DSMP};
      uint32_t key[4];
      philox(ctr, skey, key);
      const int64_t npr = std::min<int64_t>(P, M - r0);
      double gr = gm;
      for (int64_t m = 0; m < npr; ++m) {
        const int64_t pi = b + pair_perm((uint32_t)(2 * m), (uint32_t)n, half, key);
        const int64_t pj = b + pair_perm((uint32_t)(2 * m + 1), (uint32_t)n, half, key);
        draw4((uint32_t)c, s32, (uint32_t)(r0 + m), TAG_DSMC, seed, u);
        const double vix = vx[pi], viy = vy[pi], vjx = vx[pj], vjy = vy[pj];
        const double gx = vix - vjx, gy = viy - vjy;
        const double g = std::sqrt(gx * gx + gy * gy);
        const double ge = g > gm ? g : gm;
        if (g > gr) gr = g;
        if (u[0] * ge < g) {
          double sn, cs;
          sincos2pi(u[1], &sn, &cs);
          const double cx = 0.5 * (vix + vjx), cy = 0.5 * (viy + vjy);
          const double hx = 0.5 * g * cs, hy = 0.5 * g * sn;
          vx[pi] = (float)(cx + hx);
          vy[pi] = (float)(cy + hy);
          vx[pj] = (float)(cx - hx);
          vy[pj] = (float)(cy - hy);
          ++total;
        }
      }
      gm = gr;
    }
    gmax[c] = (float)gm;
  }
  return total;
}

// The pseudo-random slot permutation of one DSMC round (for tests): out[t] = pi(t).
void or_dsmc_pair_perm(uint32_t n, int64_t cell, int64_t step, int64_t round, uint64_t seed,
                       uint32_t* out) {
  const uint32_t skey[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  const uint32_t ctr[4] = {(uint32_t)cell, (uint32_t)step, 0x80000000u | (uint32_t)round, TAG_DSMP};
  uint32_t key[4];
  philox(ctr, skey, key);
  const int half = pair_half_bits(n);
  for (uint32_t t = 0; t < n; ++t) out[t] = pair_perm(t, n, half, key);
}

// ------------------------------------------------- the reference's own collide arithmetic
// Uniform stream of the reference's host collisions: std::uniform_real_distribution<double>
// {0, 1} over std::mt19937 gen{seed} (thesisCode/mccCollisions/src/main.cc:136,
// particle.cc:117; thesisCode/dsmcCollisions/cellTrackerAdvanced/src/main.cc:327-328).
void or_std_uniform_stream(uint32_t seed, int64_t count, double* out) {
  std::mt19937 gen{seed};
  std::uniform_real_distribution<double> dist{0.0, 1.0};
  for (int64_t k = 0; k < count; ++k) out[k] = dist(gen);
}

// MCC collide of thesisCode/mccCollisions/src/particle.cc:112-127 called for particles
// 0..n-1 in order (main.cc:181-183) on the uniform stream u: r = u[k]; when prob > r the
// velocity becomes (2(u1 + u2 + u3 - 1.5), 2(u4 + u5 + u6 - 1.5)) with the next six
// uniforms.  Returns the uniforms consumed, or -1 when the stream ends first.
int64_t or_mcc_reference(int64_t n, double* vx, double* vy, double prob, const double* u,
                         int64_t nu) {
  int64_t k = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (k >= nu) return -1;
    const double r = u[k];
    if (prob > r) {
      if (k + 7 > nu) return -1;
      vx[i] = 2.0 * (u[k + 1] + u[k + 2] + u[k + 3] - 1.5);
      vy[i] = 2.0 * (u[k + 4] + u[k + 5] + u[k + 6] - 1.5);
      k += 7;
    } else {
      k += 1;
    }
  }
  return k;
}

// DSMC collide of thesisCode/dsmcCollisions/cellTrackerAdvanced/src/main.cc:288-350 on
// cell lists built as cellTracking does (:269-278, ascending particle index; the cell key
// here is cx*ny + cy, the reference GPU's, which partitions the particles like the host's
// cx + cy*ny when nx = ny): for each particle i in order, the nearest other particle t of
// its cell (Euclidean distance, strictly smaller wins, first found on ties, initial bound
// 1e8); kept only when t > i; then r = u[k] and, when prob > r,
// v_i = (m1, m2), v_t = (-m1, -m2) with m = 2(u1 + u2 + u3 - 1.5) from the next six
// uniforms.  cx = (int)(x/dx) as the reference Particle keeps it (particle.cc:23-24).
// Returns the uniforms consumed (-1 when the stream ends first); *ncoll = collisions.
int64_t or_dsmc_reference(int64_t n, const double* x, const double* y, double* vx,
                          double* vy, int nx, int ny, double lx, double ly, double prob,
                          const double* u, int64_t nu, int64_t* ncoll) {
  const double dx = lx / nx, dy = ly / ny;
  std::vector<std::vector<int64_t>> cells((size_t)nx * ny);
  for (int64_t i = 0; i < n; ++i) {
    const int cx = (int)(x[i] / dx), cy = (int)(y[i] / dy);
    cells[(size_t)cx * ny + cy].push_back(i);
  }
  int64_t k = 0, nc = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int cx = (int)(x[i] / dx), cy = (int)(y[i] / dy);
    double nearest = 1e8;
    int64_t target = -1;
    for (int64_t t : cells[(size_t)cx * ny + cy]) {
      if (t == i) continue;
      const double d = std::sqrt((x[i] - x[t]) * (x[i] - x[t]) + (y[i] - y[t]) * (y[i] - y[t]));
      if (d < nearest) {
        nearest = d;
        target = t;
      }
    }
    if (target < i) continue;
    if (k >= nu) return -1;
    const double r = u[k];
    if (prob > r) {
      if (k + 7 > nu) return -1;
      const double m1 = 2.0 * (u[k + 1] + u[k + 2] + u[k + 3] - 1.5);
      const double m2 = 2.0 * (u[k + 4] + u[k + 5] + u[k + 6] - 1.5);
      vx[i] = m1;
      vy[i] = m2;
      vx[target] = -m1;
      vy[target] = -m2;
      k += 7;
      ++nc;
    } else {
      k += 1;
    }
  }
  if (ncoll) *ncoll = nc;
  return k;
}

// ---------------------------------------------------------------- adjoint
// Reverse mode of the continuous-weight FP64 run, with the whole trajectory kept in memory
// (small cases).  E = M rho with M = -Dc P (P: spectral Green operator, symmetric; Dc:
// periodic central difference, antisymmetric), so M^T (lx, ly) = P (Dcx lx + Dcy ly).
void or_adjoint_mt(int nx, int ny, double lx, double ly, const double* lex, const double* ley,
                   double* lrho) {
  const size_t G = (size_t)nx * ny;
  std::vector<double> ax(G), ay(G), bx(G), by(G), g(G);
  efield_impl(nx, ny, lx, ly, lex, ax.data(), ay.data());  // ax = -Dcx lex
  efield_impl(nx, ny, lx, ly, ley, bx.data(), by.data());  // by = -Dcy ley
  for (size_t i = 0; i < G; ++i) g[i] = -(ax[i] + by[i]);
  poisson_impl(nx, ny, lx, ly, g.data(), lrho, nullptr);
}

namespace {
struct Stencil {
  size_t nd[4];
  double w[4], wx[4], wy[4];
};
Stencil stencil(double x, double y, int nx, int ny, double dx, double dy) {
  int cx, cy;
  double fx, fy;
  locate(x, dx, nx, &cx, &fx);
  locate(y, dy, ny, &cy, &fy);
  const int cxp = wrapi(cx + 1, nx), cyp = wrapi(cy + 1, ny);
  Stencil s;
  s.nd[0] = (size_t)cx * ny + cy;
  s.nd[1] = (size_t)cxp * ny + cy;
  s.nd[2] = (size_t)cx * ny + cyp;
  s.nd[3] = (size_t)cxp * ny + cyp;
  s.w[0] = (1 - fx) * (1 - fy); s.w[1] = fx * (1 - fy); s.w[2] = (1 - fx) * fy; s.w[3] = fx * fy;
  s.wx[0] = -(1 - fy) / dx; s.wx[1] = (1 - fy) / dx; s.wx[2] = -fy / dx; s.wx[3] = fy / dx;
  s.wy[0] = -(1 - fx) / dy; s.wy[1] = -fx / dy; s.wy[2] = (1 - fx) / dy; s.wy[3] = fx / dy;
  return s;
}
}  // namespace

double or_adjoint_f64(int64_t n, const double* xl, const double* yl, const double* vx0,
                      const double* vy0, int nmodes, const int32_t* mx, const int32_t* my,
                      const double* amp, int nx, int ny, double lx, double ly, double dt,
                      double qm, double qp, int nsteps, double* grad_a, double* grad_vx,
                      double* grad_vy) {
  const size_t G = (size_t)nx * ny, N = (size_t)n;
  const double dx = lx / nx, dy = ly / ny, dA = dx * dy;
  std::vector<double> kx(nmodes), ky(nmodes);
  for (int m = 0; m < nmodes; ++m) {
    kx[m] = 2.0 * kPi * mx[m] / lx;
    ky[m] = 2.0 * kPi * my[m] / ly;
  }
  // forward with the trajectory: X[k], Y[k] for k = 0..nsteps, EX/EY[k] for k = 0..nsteps
  std::vector<double> X((nsteps + 1) * N), Y((nsteps + 1) * N);
  std::vector<double> EX((nsteps + 1) * G), EY((nsteps + 1) * G);
  std::vector<double> vx(vx0, vx0 + N), vy(vy0, vy0 + N), rho(G), phi(G);
  for (size_t p = 0; p < N; ++p) {
    double d = 0.0;
    for (int m = 0; m < nmodes; ++m) d += amp[m] * std::cos(kx[m] * xl[p] + ky[m] * yl[p]);
    X[p] = wrap_pos(xl[p] + d, lx);
    Y[p] = yl[p];
  }
  for (int k = 0; k <= nsteps; ++k) {
    double* x = &X[k * N];
    double* y = &Y[k * N];
    double* ex = &EX[k * G];
    double* ey = &EY[k * G];
    or_deposit_f64(n, x, y, nx, ny, lx, ly, qp, rho.data());
    poisson_impl(nx, ny, lx, ly, rho.data(), phi.data(), nullptr);
    efield_impl(nx, ny, lx, ly, phi.data(), ex, ey);
    if (k == nsteps) break;
    std::memcpy(&X[(k + 1) * N], x, sizeof(double) * N);
    std::memcpy(&Y[(k + 1) * N], y, sizeof(double) * N);
    push_cont(n, &X[(k + 1) * N], &Y[(k + 1) * N], vx.data(), vy.data(), ex, ey, nx, ny, lx, ly,
              dt, qm, k == 0 ? 0.5 : 1.0, 1, nullptr, nullptr, nullptr);
  }
  const double* exN = &EX[(size_t)nsteps * G];
  const double* eyN = &EY[(size_t)nsteps * G];
  double loss = 0.0;
  for (size_t i = 0; i < G; ++i) loss += exN[i] * exN[i] + eyN[i] * eyN[i];
  loss *= dA;
  if (!grad_a && !grad_vx && !grad_vy) return loss;
  // reverse sweep
  std::vector<double> ax(N, 0.0), ay(N, 0.0), bx(N, 0.0), by(N, 0.0);
  std::vector<double> lex(G), ley(G), lrho(G);
  for (size_t i = 0; i < G; ++i) { lex[i] = 2.0 * dA * exN[i]; ley[i] = 2.0 * dA * eyN[i]; }
  or_adjoint_mt(nx, ny, lx, ly, lex.data(), ley.data(), lrho.data());
  const double cq = qp / dA;
  for (int k = nsteps;; --k) {
    // deposit Jacobian at x_k
    const double* x = &X[(size_t)k * N];
    const double* y = &Y[(size_t)k * N];
    for (size_t p = 0; p < N; ++p) {
      const Stencil s = stencil(x[p], y[p], nx, ny, dx, dy);
      for (int q = 0; q < 4; ++q) {
        ax[p] += cq * s.wx[q] * lrho[s.nd[q]];
        ay[p] += cq * s.wy[q] * lrho[s.nd[q]];
      }
    }
    if (k == 0) break;
    // step k-1: drift, then kick at x_{k-1} with E_{k-1}
    const int j = k - 1;
    const double ck = (j == 0 ? 0.5 : 1.0) * dt * qm;
    const double* xp = &X[(size_t)j * N];
    const double* yp = &Y[(size_t)j * N];
    const double* ex = &EX[(size_t)j * G];
    const double* ey = &EY[(size_t)j * G];
    std::fill(lex.begin(), lex.end(), 0.0);
    std::fill(ley.begin(), ley.end(), 0.0);
    for (size_t p = 0; p < N; ++p) {
      bx[p] += dt * ax[p];
      by[p] += dt * ay[p];
      const Stencil s = stencil(xp[p], yp[p], nx, ny, dx, dy);
      double gxx = 0, gxy = 0, gyx = 0, gyy = 0;
      for (int q = 0; q < 4; ++q) {
        lex[s.nd[q]] += ck * s.w[q] * bx[p];
        ley[s.nd[q]] += ck * s.w[q] * by[p];
        gxx += s.wx[q] * ex[s.nd[q]];
        gxy += s.wy[q] * ex[s.nd[q]];
        gyx += s.wx[q] * ey[s.nd[q]];
        gyy += s.wy[q] * ey[s.nd[q]];
      }
      ax[p] += ck * (bx[p] * gxx + by[p] * gyx);
      ay[p] += ck * (bx[p] * gxy + by[p] * gyy);
    }
    or_adjoint_mt(nx, ny, lx, ly, lex.data(), ley.data(), lrho.data());
  }
  if (grad_a)
    for (int m = 0; m < nmodes; ++m) {
      double s = 0.0;
      for (size_t p = 0; p < N; ++p) s += ax[p] * std::cos(kx[m] * xl[p] + ky[m] * yl[p]);
      grad_a[m] = s;
    }
  if (grad_vx) std::memcpy(grad_vx, bx.data(), sizeof(double) * N);
  if (grad_vy) std::memcpy(grad_vy, by.data(), sizeof(double) * N);
  return loss;
}

void or_sincos2pi(double u, double* s, double* c) { sincos2pi(u, s, c); }
double or_log(double x) { return log_own(x); }

}  // extern "C"
