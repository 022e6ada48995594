// adjoint_launch.h -- launchers of the adjoint kernels (internal to libpicamd.so).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct AdjGrid {
  int nx, ny, lognx, logny;
  double lx, ly, dx, dy, dt, qm;
  // tiling of the windowed kernels (particles tile-sorted at every checkpoint)
  int tile, halo, win, tiles_y, splits, nitems;
};

template <typename R>
void adj_init(hipStream_t s, const AdjGrid& g, const double* xl, const double* yl, int64_t n,
              const double* amps, const double* mkx, const double* mky, int nm, R* x, R* y,
              unsigned long long* hi, unsigned long long* lo);
template <typename R>
void adj_push(hipStream_t s, const AdjGrid& g, const R* xi, const R* yi, R* xo, R* yo, R* vx,
              R* vy, const R* E, int64_t n, double c, R* ck, unsigned long long* hi,
              unsigned long long* lo);
template <typename R>
void adj_restore(hipStream_t s, const AdjGrid& g, const R* ck, int64_t n, R* x, R* y, R* vx,
                 R* vy, unsigned long long* hi, unsigned long long* lo);
void adj_zero(hipStream_t s, double* a, long long n);
template <typename R>
void adj_pass(hipStream_t s, const AdjGrid& g, const R* xk, const R* yk, const double* lrho,
              double cq, const R* xp, const R* yp, const R* Eprev, double ck1, double* lx,
              double* ly, double* lvx, double* lvy, double* lE, int64_t n);
// forward step of tile-sorted particles (in place) with LDS windows for the field and
// the new charge; items {tile x, tile y, split} over the tile offsets `toff`.  E == NULL:
// deposit the charge of the positions as they are, no push
template <typename R>
void adj_push_win(hipStream_t s, const AdjGrid& g, const int4* items, const int* toff, R* x,
                  R* y, R* vx, R* vy, const R* E, double c, unsigned long long* hi,
                  unsigned long long* lo);
// reverse step of tile-sorted particles; the E sensitivity is summed in an LDS window
template <typename R>
void adj_pass_win(hipStream_t s, const AdjGrid& g, const int4* items, const int* toff,
                  const R* xk, const R* yk, const double* lrho, double cq, const R* xp,
                  const R* yp, const R* Eprev, double ck1, double* lx, double* ly,
                  double* lvx, double* lvy, double* lE);
// four particle arrays (a: 4 x n) into the order before a sort: b[k][from[d]] = a[k][d]
void adj_reorder(hipStream_t s, const int* from, const double* a, double* b, int64_t n);
void adj_iota(hipStream_t s, int* id, int64_t n);
template <typename R>
void adj_seed(hipStream_t s, const AdjGrid& g, const R* E, double sc, double* lE);
template <typename R>
void adj_energy(hipStream_t s, const AdjGrid& g, const R* E, double* part, int nblocks);
void adj_mt(hipStream_t s, const AdjGrid& g, double* lE, double2* A, double* lrho,
            const double2* tw);
void adj_amp_grad(hipStream_t s, const double* lx, const double* xl, const double* yl, int64_t n,
                  const double* mkx, const double* mky, int nm, double* part, int nblocks,
                  double scale, double* out);
template <typename R>
void adj_to_r(hipStream_t s, const double* a, double sc, R* out, int64_t n);
