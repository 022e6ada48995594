// k_adjoint.hip -- reverse-mode (adjoint) kernels of the collisionless step.
//
// Forward step k (no reference counterpart; it is the reference step of
// thesisCode/collisionless/double/src/main.cc:137-172 written as a map):
//   rho_k = D(x_k)                         CIC deposit, charge qp per particle
//   E_k   = M rho_k                        spectral Poisson + central-difference symbol
//   v_{k+1/2} = v_{k-1/2} + c_k dt qm G(x_k) E_k      (c_0 = 1/2 from synchronous v_0)
//   x_{k+1}   = x_k + dt v_{k+1/2}
// Loss L = sum |E_N|^2 dx dy.  Reverse step k (lambda_x, lambda_v = dL/dx, dL/dv):
//   lv' = lv + dt lx                                   (drift)
//   lE  = c_k dt qm sum_p W(x_p) lv'_p                 (gather is the transpose of deposit)
//   lrho = M^T lE                                      (conjugate Fourier symbols)
//   lx += c_k dt qm (grad W . E_k) lv' + (qp/dA) grad W . lrho
// The CIC weights are piecewise bilinear; their derivatives are taken on the continuous
// fractions (the fixed-point quantisation of the forward weights is below 2^-18 of a cell).
#include <type_traits>
#include "fft_lds.h"
#include "deposit.h"
#include "launch.h"
#include "adjoint_launch.h"

namespace {

template <typename R>
using R2T = typename std::conditional<sizeof(R) == 4, float2, double2>::type;

template <typename R>
__device__ __forceinline__ Geo<R> geo_of(const AdjGrid& g) {
  return make_geo<R>(g.nx, g.ny, g.lx, g.ly);
}

// x0 = xl + sum_m a_m cos(kx_m xl + ky_m yl) (wrapped), y0 = yl; deposits x0's charge
// unless hi is NULL
template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_adj_init(AdjGrid g, const double* __restrict__ xl,
                                                     const double* __restrict__ yl, int64_t n,
                                                     const double* __restrict__ amps,
                                                     const double* __restrict__ mkx,
                                                     const double* __restrict__ mky, int nm,
                                                     R* __restrict__ x, R* __restrict__ y,
                                                     unsigned long long* hi,
                                                     unsigned long long* lo) {
  const Geo<R> q = geo_of<R>(g);
  for (int64_t p = blockIdx.x * (int64_t)PIC_WG + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * PIC_WG) {
    const double a = xl[p], b = yl[p];
    double d = 0.0;
    for (int m = 0; m < nm; ++m) d += amps[m] * cos(mkx[m] * a + mky[m] * b);
    const R xr = wrap_pos((R)wrap_pos(a + d, g.lx), (R)g.lx);
    const R yr = (R)b;
    x[p] = xr;
    y[p] = yr;
    if (hi) {
      int cx, cy;
      long long w[4];
      deposit_weights<R>(xr, yr, q, cx, cy, w);
      deposit_global<R>(hi, lo, cx, cy, w, q);
    }
  }
}

// One forward step of particle-indexed arrays: gather with the fixed-point weights, kick
// (factor c), drift from (xi, yi) to (xo, yo) (may alias), then deposit the new position's
// charge into hi/lo (skipped when hi is NULL).  ck != NULL receives the state before the
// step (checkpoint layout x | y | vx | vy, n each).
template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_adj_push(AdjGrid g, const R* xi, const R* yi,
                                                     R* xo, R* yo, R* __restrict__ vx,
                                                     R* __restrict__ vy,
                                                     const R* __restrict__ Epad, int64_t n,
                                                     double c, R* __restrict__ ck,
                                                     unsigned long long* hi,
                                                     unsigned long long* lo) {
  const Geo<R> q = geo_of<R>(g);
  const R Lx = (R)g.lx, Ly = (R)g.ly;
  const R hk = (R)(c * g.dt * g.qm), dtr = (R)g.dt;
  const R qs = (R)(1.0 / (double)(1 << Fix<R>::QB));
  const R2T<R>* E = reinterpret_cast<const R2T<R>*>(Epad);
  const int ny1 = g.ny + 1;
  for (int64_t p = blockIdx.x * (int64_t)PIC_WG + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * PIC_WG) {
    R xr = xi[p], yr = yi[p], vxr = vx[p], vyr = vy[p];
    if (ck) {
      ck[p] = xr;
      ck[n + p] = yr;
      ck[2 * n + p] = vxr;
      ck[3 * n + p] = vyr;
    }
    int cx, cy;
    R fx, fy;
    locate_w(xr, q.idx, g.nx, cx, fx);
    locate_w(yr, q.idy, g.ny, cy, fy);
    const R gx1 = (R)qweight(fx) * qs, gy1 = (R)qweight(fy) * qs;
    const R gx0 = (R)1 - gx1, gy0 = (R)1 - gy1;
    const long long i00 = (long long)cx * ny1 + cy;
    const R2T<R> e00 = E[i00], e01 = E[i00 + 1], e10 = E[i00 + ny1], e11 = E[i00 + ny1 + 1];
    R epx = (gx0 * gy0) * e00.x;
    epx = epx + (gx1 * gy0) * e10.x;
    epx = epx + (gx0 * gy1) * e01.x;
    epx = epx + (gx1 * gy1) * e11.x;
    R epy = (gx0 * gy0) * e00.y;
    epy = epy + (gx1 * gy0) * e10.y;
    epy = epy + (gx0 * gy1) * e01.y;
    epy = epy + (gx1 * gy1) * e11.y;
    vxr = vxr + hk * epx;
    vyr = vyr + hk * epy;
    xr = wrap_pos((R)(xr + dtr * vxr), Lx);
    yr = wrap_pos((R)(yr + dtr * vyr), Ly);
    xo[p] = xr;
    yo[p] = yr;
    vx[p] = vxr;
    vy[p] = vyr;
    if (hi) {
      long long w[4];
      deposit_weights<R>(xr, yr, q, cx, cy, w);
      deposit_global<R>(hi, lo, cx, cy, w, q);
    }
  }
}

// restore a checkpoint (x | y | vx | vy) into (x, y) and (vx, vy); deposit x's charge into
// hi/lo unless hi is NULL
template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_adj_restore(AdjGrid g, const R* __restrict__ ck,
                                                        int64_t n, R* __restrict__ x,
                                                        R* __restrict__ y, R* __restrict__ vx,
                                                        R* __restrict__ vy,
                                                        unsigned long long* hi,
                                                        unsigned long long* lo) {
  const Geo<R> q = geo_of<R>(g);
  for (int64_t p = blockIdx.x * (int64_t)PIC_WG + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * PIC_WG) {
    const R xr = ck[p], yr = ck[n + p];
    x[p] = xr;
    y[p] = yr;
    vx[p] = ck[2 * n + p];
    vy[p] = ck[3 * n + p];
    if (hi) {
      int cx, cy;
      long long w[4];
      deposit_weights<R>(xr, yr, q, cx, cy, w);
      deposit_global<R>(hi, lo, cx, cy, w, q);
    }
  }
}

__global__ __launch_bounds__(PIC_WG) void k_adj_zero(double* __restrict__ a, long long n) {
  for (long long q = blockIdx.x * (long long)PIC_WG + threadIdx.x; q < n;
       q += (long long)gridDim.x * PIC_WG)
    a[q] = 0.0;
}

// continuous CIC weights and their gradients at x (double arithmetic)
struct CicD {
  int cx, cy;                        // cell
  int c00, c10, c01, c11;            // node indices (periodic)
  double w[4], wx[4], wy[4];         // W, dW/dx, dW/dy in node order 00,10,01,11
};

template <typename R>
__device__ __forceinline__ CicD cic_d(R xr, R yr, const AdjGrid& g) {
  int cx, cy;
  R f;
  locate(xr, (R)g.dx, g.nx, cx, f);
  locate(yr, (R)g.dy, g.ny, cy, f);
  double fx = (double)xr / g.dx - cx, fy = (double)yr / g.dy - cy;
  fx = fx < 0 ? 0 : (fx > 1 ? 1 : fx);
  fy = fy < 0 ? 0 : (fy > 1 ? 1 : fy);
  const int cxp = (cx + 1 == g.nx) ? 0 : cx + 1, cyp = (cy + 1 == g.ny) ? 0 : cy + 1;
  CicD r;
  r.cx = cx; r.cy = cy;
  r.c00 = cx * g.ny + cy; r.c10 = cxp * g.ny + cy; r.c01 = cx * g.ny + cyp; r.c11 = cxp * g.ny + cyp;
  const double ix = 1.0 / g.dx, iy = 1.0 / g.dy;
  r.w[0] = (1 - fx) * (1 - fy); r.w[1] = fx * (1 - fy); r.w[2] = (1 - fx) * fy; r.w[3] = fx * fy;
  r.wx[0] = -(1 - fy) * ix; r.wx[1] = (1 - fy) * ix; r.wx[2] = -fy * ix; r.wx[3] = fy * ix;
  r.wy[0] = -(1 - fx) * iy; r.wy[1] = -fx * iy; r.wy[2] = (1 - fx) * iy; r.wy[3] = fx * iy;
  return r;
}

// Fused reverse pass: finish step k (deposit Jacobian at x_k with lrho_k), then, when
// xp != NULL, start step k-1: drift, deposit of lv' into lE (scaled by ck1), kick Jacobian
// at x_{k-1} with the field E_{k-1}.
template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_adj_pass(AdjGrid g, const R* __restrict__ xk,
                                                     const R* __restrict__ yk,
                                                     const double* __restrict__ lrho,
                                                     double cq, const R* __restrict__ xp,
                                                     const R* __restrict__ yp,
                                                     const R* __restrict__ Eprev, double ck1,
                                                     double* __restrict__ lx,
                                                     double* __restrict__ ly,
                                                     double* __restrict__ lvx,
                                                     double* __restrict__ lvy,
                                                     double* __restrict__ lE, int64_t n) {
  const int G = g.nx * g.ny;
  const R2T<R>* E = reinterpret_cast<const R2T<R>*>(Eprev);
  const int ny1 = g.ny + 1;
  for (int64_t p = blockIdx.x * (int64_t)PIC_WG + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * PIC_WG) {
    double ax = lx[p], ay = ly[p];
    {
      const CicD c = cic_d<R>(xk[p], yk[p], g);
      const int nd[4] = {c.c00, c.c10, c.c01, c.c11};
      double sx = 0.0, sy = 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double l = lrho[nd[q]];
        sx += c.wx[q] * l;
        sy += c.wy[q] * l;
      }
      ax += cq * sx;
      ay += cq * sy;
    }
    if (xp) {
      double bx = lvx[p] + g.dt * ax, by = lvy[p] + g.dt * ay;
      const CicD c = cic_d<R>(xp[p], yp[p], g);
      const int nd[4] = {c.c00, c.c10, c.c01, c.c11};
      // the field at the nodes (padded array: node (i,j) at i*(ny+1)+j)
      double exx = 0, exy = 0, eyx = 0, eyy = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = nd[q] / g.ny, j = nd[q] % g.ny;
        const R2T<R> e = E[(long long)i * ny1 + j];
        exx += c.wx[q] * (double)e.x; exy += c.wy[q] * (double)e.x;
        eyx += c.wx[q] * (double)e.y; eyy += c.wy[q] * (double)e.y;
        atomicAdd(&lE[nd[q]], ck1 * c.w[q] * bx);
        atomicAdd(&lE[G + nd[q]], ck1 * c.w[q] * by);
      }
      ax += ck1 * (bx * exx + by * eyx);
      ay += ck1 * (bx * exy + by * eyy);
      lvx[p] = bx;
      lvy[p] = by;
    }
    lx[p] = ax;
    ly[p] = ay;
  }
}

// ---- windowed forms.  The particles are counting-sorted by tile at every checkpoint
// (stable, with their original indices), so within a segment the particles of a push item
// stay near its tile: the field is read from an LDS window of the tile + halo, and the
// deposits (int64 charge forward, FP64 E sensitivity in reverse) are summed in an LDS
// window and flushed with one global atomic per touched node.  Particles that drifted out
// of the window since the checkpoint use the global arrays directly.  Per particle the
// arithmetic is that of k_adj_push / k_adj_pass.

// item {tile x, tile y, split} -> particle range [f0, f1) and window origin
__device__ __forceinline__ void adj_item(const AdjGrid& g, const int4 it,
                                         const int* __restrict__ toff, int& f0, int& f1,
                                         int& x0, int& y0) {
  const int tb = it.x * g.tiles_y + it.y;
  const long long s0 = toff[tb], F = toff[tb + 1] - s0;
  f0 = (int)(s0 + F * it.z / g.splits);
  f1 = (int)(s0 + F * (it.z + 1) / g.splits);
  x0 = it.x * g.tile - g.halo;
  y0 = it.y * g.tile - g.halo;
}

// PUSH = false: deposit the charge of the positions as they are (no field, no motion)
template <typename R, bool PUSH>
__global__ __launch_bounds__(PIC_WG) void k_adj_push_win(AdjGrid g,
                                                         const int4* __restrict__ items,
                                                         const int* __restrict__ toff,
                                                         R* __restrict__ x, R* __restrict__ y,
                                                         R* __restrict__ vx,
                                                         R* __restrict__ vy,
                                                         const R* __restrict__ Epad, double c,
                                                         unsigned long long* hi,
                                                         unsigned long long* lo) {
  extern __shared__ unsigned long long wsh[];
  const int W = g.win, nw = W * W;
  constexpr int NQ = (sizeof(R) == 8) ? 2 : 1;  // int64 words per node (hi, lo)
  unsigned long long* whi = wsh;
  unsigned long long* wlo = wsh + nw;            // FP64 particles only
  R2T<R>* ew = reinterpret_cast<R2T<R>*>(wsh + NQ * nw);
  int f0, f1, x0, y0;
  adj_item(g, items[blockIdx.x], toff, f0, f1, x0, y0);
  const int ny1 = g.ny + 1;
  const R2T<R>* E = reinterpret_cast<const R2T<R>*>(Epad);
  for (int i = threadIdx.x; i < NQ * nw; i += PIC_WG) wsh[i] = 0ull;
  if (PUSH)
    for (int i = threadIdx.x; i < nw; i += PIC_WG) {
      const int r = i / W, cc = i - r * W;
      ew[i] = E[(long long)pwrap(x0 + r, g.nx) * ny1 + pwrap(y0 + cc, g.ny)];
    }
  __syncthreads();
  const Geo<R> q = geo_of<R>(g);
  const R Lx = (R)g.lx, Ly = (R)g.ly;
  const R hk = (R)(c * g.dt * g.qm), dtr = (R)g.dt;
  const R qs = (R)(1.0 / (double)(1 << Fix<R>::QB));
  for (int p = f0 + threadIdx.x; p < f1; p += PIC_WG) {
    R xr = x[p], yr = y[p];
    int cx, cy, rx, ry;
    if (PUSH) {
      R vxr = vx[p], vyr = vy[p];
      R fx, fy;
      locate_w(xr, q.idx, g.nx, cx, fx);
      locate_w(yr, q.idy, g.ny, cy, fy);
      const R gx1 = (R)qweight(fx) * qs, gy1 = (R)qweight(fy) * qs;
      const R gx0 = (R)1 - gx1, gy0 = (R)1 - gy1;
      R2T<R> e00, e01, e10, e11;
      rx = pwrap(cx - x0, g.nx);
      ry = pwrap(cy - y0, g.ny);
      if (rx < W - 1 && ry < W - 1) {
        const int i0 = rx * W + ry;
        e00 = ew[i0]; e01 = ew[i0 + 1]; e10 = ew[i0 + W]; e11 = ew[i0 + W + 1];
      } else {
        const long long i00 = (long long)cx * ny1 + cy;
        e00 = E[i00]; e01 = E[i00 + 1]; e10 = E[i00 + ny1]; e11 = E[i00 + ny1 + 1];
      }
      R epx = (gx0 * gy0) * e00.x;
      epx = epx + (gx1 * gy0) * e10.x;
      epx = epx + (gx0 * gy1) * e01.x;
      epx = epx + (gx1 * gy1) * e11.x;
      R epy = (gx0 * gy0) * e00.y;
      epy = epy + (gx1 * gy0) * e10.y;
      epy = epy + (gx0 * gy1) * e01.y;
      epy = epy + (gx1 * gy1) * e11.y;
      vxr = vxr + hk * epx;
      vyr = vyr + hk * epy;
      xr = wrap_pos((R)(xr + dtr * vxr), Lx);
      yr = wrap_pos((R)(yr + dtr * vyr), Ly);
      x[p] = xr;
      y[p] = yr;
      vx[p] = vxr;
      vy[p] = vyr;
    }
    long long w[4];
    deposit_weights<R>(xr, yr, q, cx, cy, w);
    rx = pwrap(cx - x0, g.nx);
    ry = pwrap(cy - y0, g.ny);
    if (rx < W - 1 && ry < W - 1) {
      const int i0 = rx * W + ry;
      const int off[4] = {0, W, 1, W + 1};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (sizeof(R) == 8) {
          atomicAdd(&whi[i0 + off[k]], (unsigned long long)(w[k] >> 30));
          atomicAdd(&wlo[i0 + off[k]], (unsigned long long)(w[k] & ((1ll << 30) - 1)));
        } else {
          atomicAdd(&whi[i0 + off[k]], (unsigned long long)w[k]);
        }
      }
    } else {
      deposit_global<R>(hi, lo, cx, cy, w, q);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nw; i += PIC_WG) {
    const unsigned long long v = whi[i];
    const unsigned long long vl = (sizeof(R) == 8) ? wlo[i] : 0ull;
    if (v | vl) {
      const int r = i / W, cc = i - r * W;
      const long long nd = (long long)pwrap(x0 + r, g.nx) * g.ny + pwrap(y0 + cc, g.ny);
      if (v) atomicAdd(&hi[nd], v);
      if (sizeof(R) == 8 && vl) atomicAdd(&lo[nd], vl);
    }
  }
}

template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_adj_pass_win(AdjGrid g,
                                                         const int4* __restrict__ items,
                                                         const int* __restrict__ toff,
                                                         const R* __restrict__ xk,
                                                         const R* __restrict__ yk,
                                                         const double* __restrict__ lrho,
                                                         double cq, const R* __restrict__ xp,
                                                         const R* __restrict__ yp,
                                                         const R* __restrict__ Eprev,
                                                         double ck1, double* __restrict__ lx,
                                                         double* __restrict__ ly,
                                                         double* __restrict__ lvx,
                                                         double* __restrict__ lvy,
                                                         double* __restrict__ lE) {
  extern __shared__ double wl[];  // E sensitivity window: x components, then y
  const int W = g.win, nw = W * W;
  const int G = g.nx * g.ny;
  int f0, f1, x0, y0;
  adj_item(g, items[blockIdx.x], toff, f0, f1, x0, y0);
  for (int i = threadIdx.x; i < 2 * nw; i += PIC_WG) wl[i] = 0.0;
  __syncthreads();
  const R2T<R>* E = reinterpret_cast<const R2T<R>*>(Eprev);
  const int ny1 = g.ny + 1;
  for (int p = f0 + threadIdx.x; p < f1; p += PIC_WG) {
    double ax = lx[p], ay = ly[p];
    {
      const CicD c = cic_d<R>(xk[p], yk[p], g);
      const int nd[4] = {c.c00, c.c10, c.c01, c.c11};
      double sx = 0.0, sy = 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double l = lrho[nd[q]];
        sx += c.wx[q] * l;
        sy += c.wy[q] * l;
      }
      ax += cq * sx;
      ay += cq * sy;
    }
    const double bx = lvx[p] + g.dt * ax, by = lvy[p] + g.dt * ay;
    const CicD c = cic_d<R>(xp[p], yp[p], g);
    const int nd[4] = {c.c00, c.c10, c.c01, c.c11};
    const int rx = pwrap(c.cx - x0, g.nx), ry = pwrap(c.cy - y0, g.ny);
    const bool inw = rx < W - 1 && ry < W - 1;
    const int i0 = rx * W + ry;
    const int off[4] = {0, W, 1, W + 1};
    double exx = 0, exy = 0, eyx = 0, eyy = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = nd[q] / g.ny, j = nd[q] % g.ny;
      const R2T<R> e = E[(long long)i * ny1 + j];
      exx += c.wx[q] * (double)e.x; exy += c.wy[q] * (double)e.x;
      eyx += c.wx[q] * (double)e.y; eyy += c.wy[q] * (double)e.y;
      if (inw) {
        atomicAdd(&wl[i0 + off[q]], ck1 * c.w[q] * bx);
        atomicAdd(&wl[nw + i0 + off[q]], ck1 * c.w[q] * by);
      } else {
        atomicAdd(&lE[nd[q]], ck1 * c.w[q] * bx);
        atomicAdd(&lE[G + nd[q]], ck1 * c.w[q] * by);
      }
    }
    ax += ck1 * (bx * exx + by * eyx);
    ay += ck1 * (bx * exy + by * eyy);
    lvx[p] = bx;
    lvy[p] = by;
    lx[p] = ax;
    ly[p] = ay;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nw; i += PIC_WG) {
    const double sx = wl[i], sy = wl[nw + i];
    if (sx != 0.0 || sy != 0.0) {
      const int r = i / W, cc = i - r * W;
      const int nd = pwrap(x0 + r, g.nx) * g.ny + pwrap(y0 + cc, g.ny);
      if (sx != 0.0) atomicAdd(&lE[nd], sx);
      if (sy != 0.0) atomicAdd(&lE[G + nd], sy);
    }
  }
}

// four particle arrays into the order before a sort: b[k][from[d]] = a[k][d]
__global__ __launch_bounds__(PIC_WG) void k_adj_reorder(const int* __restrict__ from,
                                                        const double* __restrict__ a,
                                                        double* __restrict__ b, int64_t n) {
  for (int64_t d = blockIdx.x * (int64_t)PIC_WG + threadIdx.x; d < n;
       d += (int64_t)gridDim.x * PIC_WG) {
    const int64_t o = from[d];
#pragma unroll
    for (int k = 0; k < 4; ++k) b[k * n + o] = a[k * n + d];
  }
}

__global__ __launch_bounds__(PIC_WG) void k_adj_iota(int* __restrict__ id, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)PIC_WG + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * PIC_WG)
    id[i] = (int)i;
}

// lE = s * E (unpadded FP64 two-component array from the padded field)
template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_adj_seed(AdjGrid g, const R* __restrict__ Epad,
                                                     double s, double* __restrict__ lE) {
  const int G = g.nx * g.ny;
  const R2T<R>* E = reinterpret_cast<const R2T<R>*>(Epad);
  for (int q = blockIdx.x * PIC_WG + threadIdx.x; q < G; q += gridDim.x * PIC_WG) {
    const int i = q / g.ny, j = q % g.ny;
    const R2T<R> e = E[(long long)i * (g.ny + 1) + j];
    lE[q] = s * (double)e.x;
    lE[G + q] = s * (double)e.y;
  }
}

// sum over nodes of Ex^2 + Ey^2 (block partials; the host adds them in order)
template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_adj_energy(AdjGrid g, const R* __restrict__ Epad,
                                                       double* __restrict__ part) {
  __shared__ double red[4];
  const int G = g.nx * g.ny;
  const R2T<R>* E = reinterpret_cast<const R2T<R>*>(Epad);
  double s = 0.0;
  for (int q = blockIdx.x * PIC_WG + threadIdx.x; q < G; q += gridDim.x * PIC_WG) {
    const int i = q / g.ny, j = q % g.ny;
    const R2T<R> e = E[(long long)i * (g.ny + 1) + j];
    s += (double)e.x * (double)e.x + (double)e.y * (double)e.y;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// ---- M^T by a complex 2D FFT of the packed pair z = lEx + i lEy
// rows: z[i][j] -> A[k][i] = FFT_y(z[i])[k]  (transposed, all k)
// (lE is zeroed as it is read, ready for the next reverse deposit)
__global__ __launch_bounds__(PIC_WG) void k_adj_rows_fwd(AdjGrid g, double* __restrict__ lE,
                                                         double2* __restrict__ A,
                                                         const double2* __restrict__ tw) {
  extern __shared__ double2 s[];
  const int i = xcd_item(blockIdx.x, gridDim.x), ny = g.ny, G = g.nx * g.ny;
  for (int j = threadIdx.x; j < ny; j += PIC_WG) {
    s[bitrev(j, g.logny)] = make_double2(lE[i * ny + j], lE[G + i * ny + j]);
    lE[i * ny + j] = 0.0;
    lE[G + i * ny + j] = 0.0;
  }
  __syncthreads();
  fft_lds(s, ny, g.logny, -1, tw);
  for (int k = threadIdx.x; k < ny; k += PIC_WG) A[(long long)k * g.nx + i] = s[k];
}

// columns k and k' = -k: FFT_x, unpack lEx_hat / lEy_hat, apply the conjugate symbols,
// inverse FFT_x; writes R_hat columns back into A.
__global__ __launch_bounds__(PIC_WG) void k_adj_cols(AdjGrid g, double2* __restrict__ A,
                                                     const double2* __restrict__ tw) {
  extern __shared__ double2 s[];
  const int nx = g.nx, ny = g.ny;
  const int k = blockIdx.x;            // 0 .. ny/2
  const int k2 = (ny - k) & (ny - 1);  // partner column
  double2* a = s;                      // column k
  double2* b = s + nx;                 // column k2
  double2* ra = s + 2 * nx;
  double2* rb = s + 3 * nx;
  for (int i = threadIdx.x; i < nx; i += PIC_WG) {
    a[bitrev(i, g.lognx)] = A[(long long)k * nx + i];
    b[bitrev(i, g.lognx)] = A[(long long)k2 * nx + i];
  }
  __syncthreads();
  fft_lds_rows(a, 2, nx, g.lognx, -1, tw);  // a, b consecutive
  const double twopi = 6.283185307179586476925;
  for (int m = threadIdx.x; m < nx; m += PIC_WG) {
    const int m2 = (nx - m) & (nx - 1);
    const int mm = (2 * m < nx) ? m : m - nx;
    const double kxv = twopi * mm / g.lx;
    const double sx = sin2pi_frac(m, nx, tw) / g.dx;
    for (int side = 0; side < 2; ++side) {
      const int kk = side ? k2 : k;
      const double2 zk = side ? b[m] : a[m];      // Z(m, kk)
      const double2 zc = side ? a[m2] : b[m2];    // Z(-m, -kk)
      const double2 hx = make_double2(0.5 * (zk.x + zc.x), 0.5 * (zk.y - zc.y));
      const double2 hy = make_double2(0.5 * (zk.y + zc.y), -0.5 * (zk.x - zc.x));
      const int kw = (2 * kk < ny) ? kk : kk - ny;
      const double kyv = twopi * kw / g.ly;
      const double sy = sin2pi_frac(kk, ny, tw) / g.dy;
      const double k2v = kxv * kxv + kyv * kyv;
      const double gr = (m == 0 && kk == 0) ? 0.0 : 1.0 / k2v;
      // R = i gr (sx hx + sy hy)
      const double re = sx * hx.x + sy * hy.x, im = sx * hx.y + sy * hy.y;
      const double2 r = make_double2(-gr * im, gr * re);
      (side ? rb : ra)[bitrev(m, g.lognx)] = r;
    }
  }
  __syncthreads();
  fft_lds_rows(ra, 2, nx, g.lognx, +1, tw);  // ra, rb consecutive
  for (int i = threadIdx.x; i < nx; i += PIC_WG) {
    A[(long long)k * nx + i] = ra[i];
    if (k2 != k) A[(long long)k2 * nx + i] = rb[i];
  }
}

// rows: inverse FFT_y of R_hat row i -> lrho[i][j] = Re / (nx ny)
__global__ __launch_bounds__(PIC_WG) void k_adj_rows_inv(AdjGrid g, const double2* __restrict__ A,
                                                         double* __restrict__ lrho,
                                                         const double2* __restrict__ tw) {
  extern __shared__ double2 s[];
  const int i = xcd_item(blockIdx.x, gridDim.x), ny = g.ny;
  for (int k = threadIdx.x; k < ny; k += PIC_WG) s[bitrev(k, g.logny)] = A[(long long)k * g.nx + i];
  __syncthreads();
  fft_lds(s, ny, g.logny, +1, tw);
  const double norm = 1.0 / ((double)g.nx * ny);
  for (int j = threadIdx.x; j < ny; j += PIC_WG) lrho[i * ny + j] = s[j].x * norm;
}

// dL/da_m = sum_p lx_p cos(kx_m xl_p + ky_m yl_p): block partials [block][m]
__global__ __launch_bounds__(PIC_WG) void k_adj_amp_grad(const double* __restrict__ lx,
                                                         const double* __restrict__ xl,
                                                         const double* __restrict__ yl,
                                                         int64_t n, const double* __restrict__ mkx,
                                                         const double* __restrict__ mky, int nm,
                                                         double* __restrict__ part) {
  __shared__ double red[4];
  for (int m = 0; m < nm; ++m) {
    double s = 0.0;
    for (int64_t p = blockIdx.x * (int64_t)PIC_WG + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * PIC_WG)
      s += lx[p] * cos(mkx[m] * xl[p] + mky[m] * yl[p]);
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[(long long)blockIdx.x * nm + m] = s;
  }
}

__global__ __launch_bounds__(PIC_WG) void k_adj_amp_reduce(const double* __restrict__ part,
                                                           int nblocks, int nm, double scale,
                                                           double* __restrict__ out) {
  if (threadIdx.x < nm) {
    // in block order; the loads of eight blocks are issued before their adds
    double s = 0.0;
    int b = 0;
    for (; b + 8 <= nblocks; b += 8) {
      double v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = part[(long long)(b + k) * nm + threadIdx.x];
#pragma unroll
      for (int k = 0; k < 8; ++k) s += v[k];
    }
    for (; b < nblocks; ++b) s += part[(long long)b * nm + threadIdx.x];
    out[threadIdx.x] = scale * s;
  }
}

template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_adj_to_r(const double* __restrict__ a, double s,
                                                     R* __restrict__ out, int64_t n) {
  for (int64_t p = blockIdx.x * (int64_t)PIC_WG + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * PIC_WG)
    out[p] = (R)(s * a[p]);
}

inline int gridn(long long n) {
  long long b = (n + PIC_WG - 1) / PIC_WG;
  return (int)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

}  // namespace

template <typename R>
void adj_init(hipStream_t s, const AdjGrid& g, const double* xl, const double* yl, int64_t n,
              const double* amps, const double* mkx, const double* mky, int nm, R* x, R* y,
              unsigned long long* hi, unsigned long long* lo) {
  k_adj_init<R><<<gridn(n), PIC_WG, 0, s>>>(g, xl, yl, n, amps, mkx, mky, nm, x, y, hi, lo);
}
template <typename R>
void adj_push(hipStream_t s, const AdjGrid& g, const R* xi, const R* yi, R* xo, R* yo, R* vx,
              R* vy, const R* E, int64_t n, double c, R* ck, unsigned long long* hi,
              unsigned long long* lo) {
  k_adj_push<R><<<gridn(n), PIC_WG, 0, s>>>(g, xi, yi, xo, yo, vx, vy, E, n, c, ck, hi, lo);
}
template <typename R>
void adj_restore(hipStream_t s, const AdjGrid& g, const R* ck, int64_t n, R* x, R* y, R* vx,
                 R* vy, unsigned long long* hi, unsigned long long* lo) {
  k_adj_restore<R><<<gridn(n), PIC_WG, 0, s>>>(g, ck, n, x, y, vx, vy, hi, lo);
}
void adj_zero(hipStream_t s, double* a, long long n) {
  k_adj_zero<<<gridn(n), PIC_WG, 0, s>>>(a, n);
}
template <typename R>
void adj_pass(hipStream_t s, const AdjGrid& g, const R* xk, const R* yk, const double* lrho,
              double cq, const R* xp, const R* yp, const R* Eprev, double ck1, double* lx,
              double* ly, double* lvx, double* lvy, double* lE, int64_t n) {
  k_adj_pass<R><<<gridn(n), PIC_WG, 0, s>>>(g, xk, yk, lrho, cq, xp, yp, Eprev, ck1, lx, ly, lvx,
                                            lvy, lE, n);
}
template <typename R>
void adj_push_win(hipStream_t s, const AdjGrid& g, const int4* items, const int* toff, R* x,
                  R* y, R* vx, R* vy, const R* E, double c, unsigned long long* hi,
                  unsigned long long* lo) {
  const size_t nw = (size_t)g.win * g.win;
  const size_t lds = nw * ((sizeof(R) == 8 ? 2 : 1) * sizeof(unsigned long long) +
                           2 * sizeof(R));
  if (E)
    k_adj_push_win<R, true><<<g.nitems, PIC_WG, lds, s>>>(g, items, toff, x, y, vx, vy, E, c,
                                                          hi, lo);
  else
    k_adj_push_win<R, false><<<g.nitems, PIC_WG, lds, s>>>(g, items, toff, x, y, vx, vy, E, c,
                                                           hi, lo);
}
template <typename R>
void adj_pass_win(hipStream_t s, const AdjGrid& g, const int4* items, const int* toff,
                  const R* xk, const R* yk, const double* lrho, double cq, const R* xp,
                  const R* yp, const R* Eprev, double ck1, double* lx, double* ly,
                  double* lvx, double* lvy, double* lE) {
  const size_t lds = 2 * sizeof(double) * (size_t)g.win * g.win;
  k_adj_pass_win<R><<<g.nitems, PIC_WG, lds, s>>>(g, items, toff, xk, yk, lrho, cq, xp, yp,
                                                  Eprev, ck1, lx, ly, lvx, lvy, lE);
}
void adj_reorder(hipStream_t s, const int* from, const double* a, double* b, int64_t n) {
  k_adj_reorder<<<gridn(n), PIC_WG, 0, s>>>(from, a, b, n);
}
void adj_iota(hipStream_t s, int* id, int64_t n) { k_adj_iota<<<gridn(n), PIC_WG, 0, s>>>(id, n); }
template <typename R>
void adj_seed(hipStream_t s, const AdjGrid& g, const R* E, double sc, double* lE) {
  k_adj_seed<R><<<gridn((long long)g.nx * g.ny), PIC_WG, 0, s>>>(g, E, sc, lE);
}
template <typename R>
void adj_energy(hipStream_t s, const AdjGrid& g, const R* E, double* part, int nblocks) {
  k_adj_energy<R><<<nblocks, PIC_WG, 0, s>>>(g, E, part);
}
void adj_mt(hipStream_t s, const AdjGrid& g, double* lE, double2* A, double* lrho,
            const double2* tw) {
  k_adj_rows_fwd<<<g.nx, PIC_WG, sizeof(double2) * g.ny, s>>>(g, lE, A, tw);
  k_adj_cols<<<g.ny / 2 + 1, PIC_WG, sizeof(double2) * 4 * g.nx, s>>>(g, A, tw);
  k_adj_rows_inv<<<g.nx, PIC_WG, sizeof(double2) * g.ny, s>>>(g, A, lrho, tw);
}
void adj_amp_grad(hipStream_t s, const double* lx, const double* xl, const double* yl, int64_t n,
                  const double* mkx, const double* mky, int nm, double* part, int nblocks,
                  double scale, double* out) {
  if (nm <= 0) return;
  k_adj_amp_grad<<<nblocks, PIC_WG, 0, s>>>(lx, xl, yl, n, mkx, mky, nm, part);
  k_adj_amp_reduce<<<1, PIC_WG, 0, s>>>(part, nblocks, nm, scale, out);
}
template <typename R>
void adj_to_r(hipStream_t s, const double* a, double sc, R* out, int64_t n) {
  k_adj_to_r<R><<<gridn(n), PIC_WG, 0, s>>>(a, sc, out, n);
}

#define ADJ_INST(R)                                                                           \
  template void adj_init<R>(hipStream_t, const AdjGrid&, const double*, const double*,       \
                            int64_t, const double*, const double*, const double*, int, R*,   \
                            R*, unsigned long long*, unsigned long long*);                   \
  template void adj_push<R>(hipStream_t, const AdjGrid&, const R*, const R*, R*, R*, R*, R*,  \
                            const R*, int64_t, double, R*, unsigned long long*,              \
                            unsigned long long*);                                            \
  template void adj_restore<R>(hipStream_t, const AdjGrid&, const R*, int64_t, R*, R*, R*,    \
                               R*, unsigned long long*, unsigned long long*);                \
  template void adj_pass<R>(hipStream_t, const AdjGrid&, const R*, const R*, const double*,   \
                            double, const R*, const R*, const R*, double, double*, double*,  \
                            double*, double*, double*, int64_t);                             \
  template void adj_push_win<R>(hipStream_t, const AdjGrid&, const int4*, const int*, R*, R*,  \
                                R*, R*, const R*, double, unsigned long long*,               \
                                unsigned long long*);                                        \
  template void adj_pass_win<R>(hipStream_t, const AdjGrid&, const int4*, const int*,         \
                                const R*, const R*, const double*, double, const R*,         \
                                const R*, const R*, double, double*, double*, double*,       \
                                double*, double*);                                           \
  template void adj_seed<R>(hipStream_t, const AdjGrid&, const R*, double, double*);          \
  template void adj_energy<R>(hipStream_t, const AdjGrid&, const R*, double*, int);           \
  template void adj_to_r<R>(hipStream_t, const double*, double, R*, int64_t);
ADJ_INST(float)
ADJ_INST(double)
