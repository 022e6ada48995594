// fft_lds.h -- radix-2 FP64 FFT on LDS-resident data, shared by the field and adjoint kernels.
#pragma once
#include "pic_device.h"

__device__ __forceinline__ int bitrev(int v, int logn) {
  return (int)(__brev((unsigned)v) >> (32 - logn));
}

__device__ __forceinline__ void fft_bfly(double2& u, double2& v, double2 w) {
  const double tr = fma(v.x, w.x, -v.y * w.y);
  const double ti = fma(v.x, w.y, v.y * w.x);
  const double2 a = u;
  u = make_double2(a.x + tr, a.y + ti);
  v = make_double2(a.x - tr, a.y - ti);
}

// LDS slot of element i of a padded buffer (PAD = 1: one spare slot after every 8 values,
// so that the 8 values a thread of the first pass takes are not 128 B apart from the next
// thread's: a 16-byte load of 64 lanes at that stride hits the same banks).  Rows of n >= 8
// values stay back to back: row r starts at slot r (n + n/8).
template <int PAD>
__device__ __forceinline__ int fft_slot(int i) {
  return PAD ? i + (i >> 3) : i;
}

// Stages s .. s+K-1 of the radix-2 transform below, in registers: a thread loads the 2^K
// elements i0 + m h (h = 2^(s-1)) that those stages combine, applies the same butterflies
// with the same twiddles in the same order as stage-by-stage radix-2 (so the result is
// bit-identical to it), and stores them back.
template <int K, int PAD>
__device__ __forceinline__ void fft_lds_pass(double2* a, int rows, int n, int logn, int s,
                                             double ws, const double2* __restrict__ tw,
                                             int logtw) {
  const int h = 1 << (s - 1);
  const int gmask = (n >> K) - 1;
  const int total = rows * (n >> K);
  for (int b = threadIdx.x; b < total; b += PIC_WG) {
    const int r = b >> (logn - K), g = b & gmask;
    const int j = g & (h - 1);
    const int e0 = r * n + ((g >> (s - 1)) << (s - 1 + K)) + j;
    double2 v[1 << K];
#pragma unroll
    for (int m = 0; m < (1 << K); ++m) v[m] = a[fft_slot<PAD>(e0 + m * h)];
#pragma unroll
    for (int t = 0; t < K; ++t) {
      const int tsh = logtw - s - t;
#pragma unroll
      for (int m = 0; m < (1 << K); ++m) {
        if (m & (1 << t)) continue;
        double2 w = tw[(j + (m & ((1 << t) - 1)) * h) << tsh];
        w.y *= ws;
        fft_bfly(v[m], v[m + (1 << t)], w);
      }
    }
#pragma unroll
    for (int m = 0; m < (1 << K); ++m) a[fft_slot<PAD>(e0 + m * h)] = v[m];
  }
}

// In-place decimation-in-time FFTs of `rows` rows of n = 2^logn complex values each (row r
// at a + r*n), every row stored in bit-reversed order; sign -1 forward, +1 inverse
// (unnormalised).  tw[j] = exp(-2 pi i j / 2^logtw), j < 2^logtw / 2, logtw >= logn: the
// global table (logtw = PIC_FFT_LOG) or a copy of the n-point twiddles in LDS
// (fft_stage_twiddles, logtw = logn).  Radix-2 butterflies, three stages per pass in
// registers (one barrier per pass), the groups of all rows spread over the workgroup.  All
// threads of the workgroup (PIC_WG) must call it.  PAD: the buffer's layout (fft_slot).
template <int PAD = 0>
__device__ __forceinline__ void fft_lds_rows(double2* a, int rows, int n, int logn, int sign,
                                             const double2* __restrict__ tw,
                                             int logtw = PIC_FFT_LOG) {
  const double ws = (sign < 0) ? 1.0 : -1.0;
  int s = 1;
  for (; s + 2 <= logn; s += 3) {
    fft_lds_pass<3, PAD>(a, rows, n, logn, s, ws, tw, logtw);
    __syncthreads();
  }
  if (s + 1 == logn) {
    fft_lds_pass<2, PAD>(a, rows, n, logn, s, ws, tw, logtw);
    __syncthreads();
  } else if (s == logn) {
    fft_lds_pass<1, PAD>(a, rows, n, logn, s, ws, tw, logtw);
    __syncthreads();
  }
}

// the n/2 twiddles of an n-point transform, from the global table into LDS (twl[j] =
// exp(-2 pi i j / n)); the caller synchronises before the transform
__device__ __forceinline__ void fft_stage_twiddles(double2* twl, int n,
                                                   const double2* __restrict__ tw) {
  const int st = PIC_FFT_MAX / n;
  for (int j = threadIdx.x; j < n / 2; j += PIC_WG) twl[j] = tw[j * st];
}

// one row
__device__ __forceinline__ void fft_lds(double2* a, int n, int logn, int sign,
                                        const double2* __restrict__ tw) {
  fft_lds_rows(a, 1, n, logn, sign, tw);
}

// sin(2 pi m / n) from a twiddle table of 2^logtw points (the global one by default, or the
// n-point copy in LDS with logtw = log2 n)
__device__ __forceinline__ double sin2pi_frac(int m, int n, const double2* __restrict__ tw,
                                              int logtw = PIC_FFT_LOG) {
  const int st = (1 << logtw) / n;
  if (m == 0 || 2 * m == n) return 0.0;
  if (2 * m < n) return -tw[m * st].y;
  return tw[(n - m) * st].y;
}
