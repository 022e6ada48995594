// k_field.hip -- FP64 spectral Poisson solve and electric field, hand-written LDS FFT.
//
// Reference stages replaced (thesisCode/collisionless/double/src/pic.cu):
//   cufftExecD2Z :551, fftPoissonSolver :147-194, cufftExecZ2D :557,
//   normaliseTransform :218-232, copyD2D :203-210, computeElectricField :242-283.
// Three kernels per step:
//   rows    : fixed-point rho (two rows packed as one complex row, groups of row pairs
//             per workgroup) -> row spectra, written transposed as S[ky][x]; clears the
//             other rho parity buffer
//   columns : column FFT, phi_hat = rho_hat / |k|^2 (k = 2 pi m / L, mean removed),
//             Ex_hat = -i sin(kx dx)/dx phi_hat, Ey_hat likewise (the Fourier symbol of the
//             periodic central difference), inverse column FFTs
//   field   : Ex + i Ey from one inverse complex row FFT per row (groups of rows per
//             workgroup); writes the padded (Ex, Ey) node array and the field energy
//             (fixed-order reduction)
// Inside a captured sequence of steps the three kernels also run the previous step's
// end-of-step work in extra workgroups after their own items (step_tail.h): its
// diagnostics row (rows) and step counter (columns), and on the step after a count step
// the fused sort's count-table scan (rows, columns, field).
#include <type_traits>
#include "fft_lds.h"
#include "launch.h"
#include "step_tail.h"

namespace {

__device__ __forceinline__ double fixed_to_density(unsigned long long hi, unsigned long long lo,
                                                   int fp64, double scale) {
  if (fp64)
    return ((double)(long long)hi * 9.313225746154785e-10 +
            (double)(long long)lo * 8.673617379884035e-19) * scale;
  return (double)(long long)hi * scale;
}

__device__ __forceinline__ int ilog2(int v) { return 31 - __clz(v); }

// Group of PP row pairs (rows i0 .. i0 + 2PP - 1): pair q is one complex FFT of
// row i0+2q + i row i0+2q+1; the half spectra are written to S[k][i0 .. i0+2PP) with
// consecutive threads on consecutive rows.
__global__ __launch_bounds__(PIC_WG) void k_fft_rows(FieldArgs a) {
  extern __shared__ double2 lds[];
  if (blockIdx.x >= a.n_row_pair_items) {
    // extra workgroups: the previous step's diagnostics rows, then the count-table sums
    int g = blockIdx.x - a.n_row_pair_items;
    if (a.fold_finish) {
      if (g < a.nruns) {
        step_finish_run(a.runs, a.st, a.ppush, a.partials, a.diag, g, false);
        return;
      }
      g -= a.nruns;
    }
    fscan_part(a.runs, a.ftable, a.fchunk_sums, a.fstride, a.fscan_items, g);
    return;
  }
  const int2 it = a.row_pair_items[xcd_item(blockIdx.x, a.n_row_pair_items)];
  const RunDesc rd = a.runs[it.x];
  const int nx = rd.nx, ny = rd.ny, logny = rd.logny;
  const int PP = rd.fft_pp;
  const int i0 = it.y;
  const long long g0 = rd.gbase + (long long)i0 * ny;
  double2* twl = lds + a.row_lds;
  fft_stage_twiddles(twl, ny, a.tw);
  if (a.rho_in) {
    const double* r = a.rho_in + (g0 - rd.gbase);
    for (int t = threadIdx.x; t < PP * ny; t += PIC_WG) {
      const int q = t >> logny, j = t & (ny - 1);
      lds[q * ny + bitrev(j, logny)] =
          make_double2(r[(long long)(2 * q) * ny + j], r[(long long)(2 * q + 1) * ny + j]);
    }
  } else {
    const int par = (int)((a.st[it.x].step + a.fold_finish) & 1);
    const unsigned long long* rh = a.rho + par * a.gsum;
    const int zpar = a.clear_read ? par : 1 - par;
    unsigned long long* zw = a.rho_w + zpar * a.gsum;
    const unsigned long long* rl = a.fp64 ? a.rho_lo + par * a.gsum : nullptr;
    unsigned long long* zl = a.fp64 ? a.rho_lo_w + zpar * a.gsum : nullptr;
    for (int t = threadIdx.x; t < PP * ny; t += PIC_WG) {
      const int q = t >> logny, j = t & (ny - 1);
      const long long e0 = g0 + (long long)(2 * q) * ny + j, e1 = e0 + ny;
      const double d0 = fixed_to_density(rh[e0], a.fp64 ? rl[e0] : 0ull, a.fp64, rd.rho_scale);
      const double d1 = fixed_to_density(rh[e1], a.fp64 ? rl[e1] : 0ull, a.fp64, rd.rho_scale);
      lds[q * ny + bitrev(j, logny)] = make_double2(d0, d1);
      zw[e0] = 0ull;
      zw[e1] = 0ull;
      if (a.fp64) { zl[e0] = 0ull; zl[e1] = 0ull; }
    }
  }
  __syncthreads();
  fft_lds_rows(lds, PP, ny, logny, -1, twl, logny);
  double2* S = a.S + rd.sbase;
  const int lg = ilog2(2 * PP);
  for (int t = threadIdx.x; t < (ny / 2 + 1) * 2 * PP; t += PIC_WG) {
    const int rr = t & (2 * PP - 1), k = t >> lg, q = rr >> 1;
    const double2 zk = lds[q * ny + k], zm = lds[q * ny + ((ny - k) & (ny - 1))];
    const double2 v = (rr & 1) ? make_double2(0.5 * (zk.y + zm.y), -0.5 * (zk.x - zm.x))
                               : make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
    S[(long long)k * nx + i0 + rr] = v;
  }
}

// One column: forward FFT, Green 1/k^2, derivative symbols, inverse FFTs.  The LDS rows are
// padded when PAD (fft_slot): a 16-byte load of 64 lanes 128 B apart, the first pass's
// pattern, hits the same banks.
template <int PAD>
__device__ __forceinline__ void fft_col_item(const FieldArgs& a, double2* lds) {
  const int2 it = a.col_items[blockIdx.x];
  const RunDesc rd = a.runs[it.x];
  const int nx = rd.nx, ny = rd.ny, lognx = rd.lognx;
  const int k = it.y;
  // nxp slots per row; A, B, C back to back (row r of a buffer starts at slot r nxp)
  const int nxp = fft_slot<PAD>(nx), mxp = a.max_nx + (a.max_nx >> 3);
  double2* col = lds;
  double2* A = lds + nxp;
  double2* B = lds + 2 * nxp;
  double2* C = lds + 3 * nxp;
  double2* twl = lds + (a.want_phi ? 4 : 3) * mxp;
  double2* S = a.S + rd.sbase + (long long)k * nx;
  fft_stage_twiddles(twl, nx, a.tw);
  for (int i = threadIdx.x; i < nx; i += PIC_WG) col[fft_slot<PAD>(bitrev(i, lognx))] = S[i];
  __syncthreads();
  fft_lds_rows<PAD>(col, 1, nx, lognx, -1, twl, lognx);
  if (k == 0 && threadIdx.x == 0 && !a.rho_in) {
    // the one reader and writer of the run's step counter in this kernel
    const long long stp = a.st[it.x].step;
    if (a.diag) {
      const double2 r1 = col[1];
      a.diag[(rd.dbase + ((stp + a.fold_finish) % PIC_DIAG_CAP)) * 8 + 4] =
          sqrt(r1.x * r1.x + r1.y * r1.y) / ((double)nx * ny);
    }
    if (a.fold_finish) {  // the previous step's end (its row came from the rows pass)
      a.st[it.x].step = stp + 1;
      a.st[it.x].vsync = 0;
    }
  }
  const double twopi = 6.283185307179586476925;
  const int kk = (2 * k < ny) ? k : k - ny;
  const double ky = twopi * kk / rd.ly;
  const double sy = sin2pi_frac(k, ny, a.tw) / rd.dy;  // one value per workgroup
  for (int m = threadIdx.x; m < nx; m += PIC_WG) {
    const int mm = (2 * m < nx) ? m : m - nx;
    const double kx = twopi * mm / rd.lx;
    const double k2 = kx * kx + ky * ky;
    const double2 r = col[fft_slot<PAD>(m)];
    double2 ph = make_double2(0.0, 0.0);
    if (m != 0 || k != 0) ph = make_double2(r.x / k2, r.y / k2);
    const double sx = sin2pi_frac(m, nx, twl, lognx) / rd.dx;
    const int br = fft_slot<PAD>(bitrev(m, lognx));
    A[br] = make_double2(sx * ph.y, -sx * ph.x);
    B[br] = make_double2(sy * ph.y, -sy * ph.x);
    if (a.want_phi) C[br] = ph;
  }
  __syncthreads();
  fft_lds_rows<PAD>(A, a.want_phi ? 3 : 2, nx, lognx, +1, twl, lognx);  // A, B, C consecutive
  double2* Sx = a.Sx + rd.sbase + (long long)k * nx;
  double2* Sy = a.Sy + rd.sbase + (long long)k * nx;
  for (int i = threadIdx.x; i < nx; i += PIC_WG) {
    const int q = fft_slot<PAD>(i);
    Sx[i] = A[q];
    Sy[i] = B[q];
    if (a.want_phi) S[i] = C[q];
  }
}

__global__ __launch_bounds__(PIC_WG) void k_fft_cols(FieldArgs a) {
  extern __shared__ double2 lds[];
  if (blockIdx.x >= a.n_col_items) {  // extra workgroups: the count-table scan, one per run
    fscan_base(a.runs, a.st, a.fchunk_sums, a.fstride, a.ftoff, blockIdx.x - a.n_col_items);
    return;
  }
  // padded rows need whole groups of 8 values (every grid is a power of two >= 4)
  if (a.runs[a.col_items[blockIdx.x].x].nx >= 8) fft_col_item<1>(a, lds);
  else fft_col_item<0>(a, lds);
}

// MODE 0: row i of (Ex + i Ey) -> padded E + field energy.  MODE 1: rows (i, i+1) of phi.
// MODE 0: group of RF rows of (Ex + i Ey), one inverse complex FFT per row -> padded E and
// the field energy.  MODE 1: group of PP row pairs of phi (two real rows per FFT).  The
// spectra are read along x (consecutive threads on consecutive rows).
template <typename R, int MODE>
__global__ __launch_bounds__(PIC_WG) void k_fft_field(FieldArgs a) {
  extern __shared__ double2 lds[];
  __shared__ double red[4];
  const int nitems = (MODE == 0) ? a.n_row_items : a.n_row_pair_items;
  if (blockIdx.x >= nitems) {  // extra workgroups (MODE 0): the count-table slots
    fscan_rows(a.runs, a.ftable, a.fchunk_sums, a.fstride, a.fscan_items, blockIdx.x - nitems);
    return;
  }
  const int item = xcd_item(blockIdx.x, nitems);
  const int2 it = (MODE == 0) ? a.row_items[item] : a.row_pair_items[item];
  const RunDesc rd = a.runs[it.x];
  const int nx = rd.nx, ny = rd.ny, logny = rd.logny;
  const int i0 = it.y;
  // rows held in LDS (MODE 0: RF rows; MODE 1: PP packed pairs) and rows read per row slot
  const int nr = (MODE == 0) ? rd.fft_rf : rd.fft_pp;
  const int lg = ilog2(nr);
  const double2* SA = (MODE == 0) ? a.Sx + rd.sbase + i0 : a.S + rd.sbase + i0;
  const double2* SB = (MODE == 0) ? a.Sy + rd.sbase + i0 : a.S + rd.sbase + i0 + 1;
  double2* twl = lds + a.row_lds;
  fft_stage_twiddles(twl, ny, a.tw);
  for (int t = threadIdx.x; t < nr * ny; t += PIC_WG) {
    const int rr = t & (nr - 1), k = t >> lg;
    const int col = (MODE == 0) ? rr : 2 * rr;  // offset of row slot rr from i0
    double2 A, B;
    if (2 * k <= ny) {
      A = SA[(long long)k * nx + col];
      B = SB[(long long)k * nx + col];
      if (k == 0 || 2 * k == ny) { A.y = 0.0; B.y = 0.0; }
    } else {
      const int kk = ny - k;
      A = SA[(long long)kk * nx + col];
      B = SB[(long long)kk * nx + col];
      A.y = -A.y;
      B.y = -B.y;
    }
    lds[rr * ny + bitrev(k, logny)] = make_double2(A.x - B.y, A.y + B.x);
  }
  __syncthreads();
  fft_lds_rows(lds, nr, ny, logny, +1, twl, logny);
  const double norm = 1.0 / ((double)nx * ny);
  if (MODE == 1) {
    double* ph = a.phi + rd.gbase;
    for (int t = threadIdx.x; t < nr * ny; t += PIC_WG) {
      const int q = t >> logny, j = t & (ny - 1);
      ph[(long long)(i0 + 2 * q) * ny + j] = lds[q * ny + j].x * norm;
      ph[(long long)(i0 + 2 * q + 1) * ny + j] = lds[q * ny + j].y * norm;
    }
    return;
  }
  typedef typename std::conditional<sizeof(R) == 4, float2, double2>::type R2;
  R2* E = reinterpret_cast<R2*>(a.E) + rd.ebase;
  double en = 0.0;
  for (int t = threadIdx.x; t < nr * ny; t += PIC_WG) {
    const int rr = t >> logny, j = t & (ny - 1), i = i0 + rr;
    const double ex = lds[rr * ny + j].x * norm, ey = lds[rr * ny + j].y * norm;
    en += ex * ex + ey * ey;
    R2 v;
    v.x = (R)ex;
    v.y = (R)ey;
    E[(long long)i * (ny + 1) + j] = v;
    if (j == 0) E[(long long)i * (ny + 1) + ny] = v;
    if (i == 0) {
      E[(long long)nx * (ny + 1) + j] = v;
      if (j == 0) E[(long long)nx * (ny + 1) + ny] = v;
    }
  }
  // field-energy partial of this row group (k_step_finish sums them per run in row order)
  if (MODE != 0 || !a.partials) return;
  const double s = block_sum(en, red);
  if (threadIdx.x == 0) a.partials[item] = s;
}

__global__ __launch_bounds__(PIC_WG) void k_op_efield(const double* __restrict__ phi,
                                                      double* __restrict__ ex,
                                                      double* __restrict__ ey, int nx, int ny,
                                                      double dx, double dy) {
  const long long G = (long long)nx * ny;
  for (long long q = blockIdx.x * (long long)PIC_WG + threadIdx.x; q < G;
       q += (long long)gridDim.x * PIC_WG) {
    const int i = (int)(q / ny), j = (int)(q % ny);
    const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
    const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
    ex[q] = (phi[(long long)im * ny + j] - phi[(long long)ip * ny + j]) / (2.0 * dx);
    ey[q] = (phi[(long long)i * ny + jm] - phi[(long long)i * ny + jp]) / (2.0 * dy);
  }
}

template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_unpad_E(const R* __restrict__ E, int nx, int ny,
                                                    double* __restrict__ ex,
                                                    double* __restrict__ ey) {
  const long long G = (long long)nx * ny;
  for (long long q = blockIdx.x * (long long)PIC_WG + threadIdx.x; q < G;
       q += (long long)gridDim.x * PIC_WG) {
    const long long i = q / ny, j = q % ny;
    const long long e = i * (ny + 1) + j;
    ex[q] = (double)E[2 * e];
    ey[q] = (double)E[2 * e + 1];
  }
}

__global__ __launch_bounds__(PIC_WG) void k_rho_to_double(const RunDesc* __restrict__ runs,
                                                          const RunState* __restrict__ st,
                                                          int run,
                                                          const unsigned long long* rho,
                                                          const unsigned long long* rlo,
                                                          long long gsum, int fp64,
                                                          double background,
                                                          double* __restrict__ out) {
  const RunDesc rd = runs[run];
  const int par = (int)(st[run].step & 1);
  const long long G = (long long)rd.nx * rd.ny;
  const unsigned long long* h = rho + par * gsum + rd.gbase;
  const unsigned long long* l = fp64 ? rlo + par * gsum + rd.gbase : nullptr;
  for (long long q = blockIdx.x * (long long)PIC_WG + threadIdx.x; q < G;
       q += (long long)gridDim.x * PIC_WG)
    out[q] = fixed_to_density(h[q], fp64 ? l[q] : 0ull, fp64, rd.rho_scale) + background;
}

inline int grid_for(long long n) {
  long long b = (n + PIC_WG - 1) / PIC_WG;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

}  // namespace

void launch_field(hipStream_t s, const FieldArgs& a, hipEvent_t* ev) {
  // data, then the n/2 twiddles of the transform length
  const size_t lrow = ((size_t)a.row_lds + a.max_ny / 2) * sizeof(double2);
  const size_t lcol = (size_t)(a.max_nx + a.max_nx / 8) * sizeof(double2) *
                          (a.want_phi ? 4 : 3) +
                      (size_t)(a.max_nx / 2) * sizeof(double2);
  // extra workgroups of the folded end-of-step work
  const int xrows = (a.fold_finish ? a.nruns : 0) + (a.fold_scan ? a.nfscan_items : 0);
  const int xcols = a.fold_scan ? a.nruns : 0, xfield = a.fold_scan ? a.nfscan_items : 0;
  if (ev) (void)hipEventRecord(ev[0], s);
  k_fft_rows<<<a.n_row_pair_items + xrows, PIC_WG, lrow, s>>>(a);
  if (ev) (void)hipEventRecord(ev[1], s);
  k_fft_cols<<<a.n_col_items + xcols, PIC_WG, lcol, s>>>(a);
  if (ev) (void)hipEventRecord(ev[2], s);
  if (a.fp64)
    k_fft_field<double, 0><<<a.n_row_items + xfield, PIC_WG, lrow, s>>>(a);
  else
    k_fft_field<float, 0><<<a.n_row_items + xfield, PIC_WG, lrow, s>>>(a);
  if (a.want_phi && a.phi) k_fft_field<float, 1><<<a.n_row_pair_items, PIC_WG, lrow, s>>>(a);
  if (ev) (void)hipEventRecord(ev[3], s);
}

void launch_op_efield(hipStream_t s, const double* phi, double* ex, double* ey, int nx, int ny,
                      double lx, double ly) {
  k_op_efield<<<grid_for((long long)nx * ny), PIC_WG, 0, s>>>(phi, ex, ey, nx, ny, lx / nx,
                                                             ly / ny);
}

void launch_unpad_E(hipStream_t s, const void* E, int fp64, int nx, int ny, double* ex,
                    double* ey) {
  if (fp64)
    k_unpad_E<double><<<grid_for((long long)nx * ny), PIC_WG, 0, s>>>((const double*)E, nx, ny,
                                                                      ex, ey);
  else
    k_unpad_E<float><<<grid_for((long long)nx * ny), PIC_WG, 0, s>>>((const float*)E, nx, ny, ex,
                                                                     ey);
}

void launch_rho_to_double(hipStream_t s, const RunDesc* runs, RunState* st, int run,
                          const unsigned long long* rho, const unsigned long long* rho_lo,
                          long long gsum, int fp64, double background, double* out) {
  k_rho_to_double<<<512, PIC_WG, 0, s>>>(runs, st, run, rho, rho_lo, gsum, fp64, background,
                                         out);
}
