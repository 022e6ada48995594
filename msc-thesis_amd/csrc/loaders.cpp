// loaders.cpp -- host-side initial particle loaders of the C ABI.
//
// pic_load_reference follows thesisCode/collisionless/double/src/main.cc:201-244 (lattice,
// std::default_random_engine(1234) + std::normal_distribution<double>, mean removed, drift
// added).  The quiet-start Landau and two-stream loaders have no reference counterpart;
// their velocities come from Philox4x32-10 so that any particle can be regenerated alone.
#include <cmath>
#include <cstdint>
#include <random>
#include <thread>
#include <vector>

#include "engine.h"

using namespace pic;

namespace {

void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[0] = n0; c[1] = (uint32_t)p1; c[2] = n2; c[3] = (uint32_t)p0;
  }
}

void box_muller(uint64_t seed, uint32_t p, double* n1, double* n2) {
  uint32_t c[4] = {p, 0u, 0u, PIC_TAG_LOAD};
  philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double u0 = ((double)c[0] + 0.5) * 2.3283064365386963e-10;
  const double u1 = ((double)c[1] + 0.5) * 2.3283064365386963e-10;
  const double r = std::sqrt(-2.0 * std::log(u0));
  const double th = 2.0 * 3.14159265358979323846 * u1;
  *n1 = r * std::cos(th);
  *n2 = r * std::sin(th);
}

void lattice(int nx, int ny, int ppc, int* npx, int* npy) {
  int a = 1;
  for (int d = 1; d <= ppc; ++d)
    if (ppc % d == 0 && (double)d >= std::sqrt((double)ppc) - 1e-9) { a = d; break; }
  *npx = nx * a;
  *npy = ny * (ppc / a);
}

double wrap(double x, double L) {
  if (x >= L) x -= L;
  else if (x < 0) x += L;
  if (!(x >= 0 && x < L)) {
    x -= L * std::floor(x / L);
    if (!(x >= 0 && x < L)) x = 0;
  }
  return x;
}

template <typename F>
void parallel_rows(int n, F f) {
  unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  if (n < 64) nt = 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([=] {
      for (int i = (int)t; i < n; i += (int)nt) f(i);
    });
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

int64_t pic_lattice_count(int nx, int ny, int ppc) {
  if (nx < 1 || ny < 1 || ppc < 1) return 0;
  int a, b;
  lattice(nx, ny, ppc, &a, &b);
  return (int64_t)a * b;
}

int pic_load_landau(int nx, int ny, double lx, double ly, int ppc, double vth, double alpha,
                    int mode, uint64_t seed, double* x, double* y, double* vx, double* vy) {
  if (!x || !y || !vx || !vy || nx < 1 || ny < 1 || ppc < 1 || !(lx > 0) || !(ly > 0))
    return fail(PIC_ERR_ARG, "bad loader arguments");
  int npx, npy;
  lattice(nx, ny, ppc, &npx, &npy);
  if ((int64_t)npx * npy > 0x7fff0000) return fail(PIC_ERR_ARG, "too many particles");
  const double k = 2.0 * 3.14159265358979323846 * mode / lx;
  parallel_rows(npx, [&](int i) {
    const double target = (i + 0.5) * lx / npx;
    double xs = target;
    if (mode != 0 && alpha != 0.0) {
      for (int it = 0; it < 60; ++it) {
        const double f = xs + (alpha / k) * std::sin(k * xs) - target;
        const double step = f / (1.0 + alpha * std::cos(k * xs));
        xs -= step;
        if (std::fabs(step) < 1e-15 * lx) break;
      }
    }
    xs = wrap(xs, lx);
    for (int j = 0; j < npy; ++j) {
      const size_t p = (size_t)i * npy + j;
      x[p] = xs;
      y[p] = (j + 0.5) * ly / npy;
      double n1, n2;
      box_muller(seed, (uint32_t)p, &n1, &n2);
      vx[p] = vth * n1;
      vy[p] = vth * n2;
    }
  });
  return PIC_OK;
}

int pic_load_two_stream(int nx, int ny, double lx, double ly, int ppc, double v0, double vth,
                        double amp, int mode, uint64_t seed, double* x, double* y, double* vx,
                        double* vy) {
  if (!x || !y || !vx || !vy || nx < 1 || ny < 1 || ppc < 1 || !(lx > 0) || !(ly > 0))
    return fail(PIC_ERR_ARG, "bad loader arguments");
  int npx, npy;
  lattice(nx, ny, ppc, &npx, &npy);
  if ((int64_t)npx * npy > 0x7fff0000) return fail(PIC_ERR_ARG, "too many particles");
  const double k = 2.0 * 3.14159265358979323846 * mode / lx;
  parallel_rows(npx, [&](int i) {
    const double xl = (i + 0.5) * lx / npx;
    const double xp = wrap(xl + amp * std::cos(k * xl), lx);
    for (int j = 0; j < npy; ++j) {
      const size_t p = (size_t)i * npy + j;
      x[p] = xp;
      y[p] = (j + 0.5) * ly / npy;
      double n1, n2;
      box_muller(seed, (uint32_t)p, &n1, &n2);
      vx[p] = ((p & 1) ? -v0 : v0) + vth * n1;
      vy[p] = vth * n2;
    }
  });
  return PIC_OK;
}

int pic_load_reference(int npx, int npy, int xdim, int ydim, int nx, int ny, double vtx,
                       double vty, double vdx, double vdy, double* x, double* y, double* vx,
                       double* vy) {
  if (!x || !y || !vx || !vy || npx < 1 || npy < 1 || nx < 1 || ny < 1 || xdim < 0 || ydim < 0)
    return fail(PIC_ERR_ARG, "bad loader arguments");
  std::default_random_engine gen(1234);
  std::normal_distribution<double> nd{0.0, 1.0};
  const double xoff = 0.5 * (double)(xdim / nx), yoff = 0.5 * (double)(ydim / ny);
  for (int i = 0; i < npy; ++i) {
    const double yc = ((double)ydim / npy) * (double)i + yoff;
    for (int j = 0; j < npx; ++j) {
      const size_t p = (size_t)i * npx + j;
      x[p] = ((double)xdim / npx) * (double)j + xoff;
      y[p] = yc;
      vx[p] = vtx * nd(gen);
      vy[p] = vty * nd(gen);
    }
  }
  const size_t n = (size_t)npx * npy;
  double sx = 0, sy = 0;
  for (size_t p = 0; p < n; ++p) { sx += vx[p]; sy += vy[p]; }
  const double inv = 1.0 / (double)n;
  sx = inv * sx - vdx;
  sy = inv * sy - vdy;
  for (size_t p = 0; p < n; ++p) { vx[p] -= sx; vy[p] -= sy; }
  return PIC_OK;
}

}  // extern "C"
