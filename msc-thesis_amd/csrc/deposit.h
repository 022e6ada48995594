// deposit.h -- CIC charge weights in int64 fixed point and their global deposit, shared by
// the step, the stage operators and the adjoint's forward pass.
#pragma once
#include "pic_device.h"

template <typename R>
struct Geo {
  int nx, ny;
  R dx, dy, Lx, Ly;
  R idx, idy;  // 1/dx, 1/dy rounded once from FP64
};

template <typename R>
__device__ __forceinline__ Geo<R> make_geo(int nx, int ny, double lx, double ly) {
  Geo<R> g;
  g.nx = nx; g.ny = ny;
  const double dx = lx / nx, dy = ly / ny;
  g.dx = (R)dx; g.dy = (R)dy; g.Lx = (R)lx; g.Ly = (R)ly;
  g.idx = (R)(1.0 / dx); g.idy = (R)(1.0 / dy);
  return g;
}

template <typename R>
__device__ __forceinline__ void deposit_weights(R xr, R yr, const Geo<R>& g, int& cx, int& cy,
                                                long long w[4]) {
  R fx, fy;
  locate_w(xr, g.idx, g.nx, cx, fx);
  locate_w(yr, g.idy, g.ny, cy, fy);
  // per-axis weights in [0, 2^QB]: 32-bit, one 32x32->64 product per node
  const unsigned one = 1u << Fix<R>::QB;
  unsigned ix1 = (unsigned)qweight(fx), iy1 = (unsigned)qweight(fy);
  unsigned ix0 = one - ix1, iy0 = one - iy1;
  if (Fix<R>::QB <= 23) {
    // the weights fit in 24 bits (they are <= 2^QB): with the high byte known clear the
    // compiler uses the full-rate 24-bit multiplies for the 48-bit products; the masks
    // change no value
    ix0 &= 0xFFFFFFu; iy0 &= 0xFFFFFFu;
    ix1 &= 0xFFFFFFu; iy1 &= 0xFFFFFFu;
  }
  w[0] = (long long)((unsigned long long)ix0 * iy0);
  w[1] = (long long)((unsigned long long)ix1 * iy0);
  w[2] = (long long)((unsigned long long)ix0 * iy1);
  w[3] = (long long)((unsigned long long)ix1 * iy1);
}

template <typename R>
__device__ __forceinline__ void deposit_global(unsigned long long* rho, unsigned long long* rlo,
                                               int cx, int cy, const long long w[4],
                                               const Geo<R>& g) {
  const int cxp = (cx + 1 == g.nx) ? 0 : cx + 1, cyp = (cy + 1 == g.ny) ? 0 : cy + 1;
  const long long nd[4] = {(long long)cx * g.ny + cy, (long long)cxp * g.ny + cy,
                           (long long)cx * g.ny + cyp, (long long)cxp * g.ny + cyp};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (sizeof(R) == 8) {
      atomicAdd(&rho[nd[k]], (unsigned long long)(w[k] >> 30));
      atomicAdd(&rlo[nd[k]], (unsigned long long)(w[k] & ((1ll << 30) - 1)));
    } else {
      atomicAdd(&rho[nd[k]], (unsigned long long)w[k]);
    }
  }
}
