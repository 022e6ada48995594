// fft_lds.h -- radix-2 FP64 FFT on LDS-resident data, shared by the field and adjoint kernels.
#pragma once
#include "pic_device.h"

__device__ __forceinline__ int bitrev(int v, int logn) {
  return (int)(__brev((unsigned)v) >> (32 - logn));
}

// In-place decimation-in-time FFTs of `rows` rows of n = 2^logn complex values each (row r
// at a + r*n), every row stored in bit-reversed order; sign -1 forward, +1 inverse
// (unnormalised).  tw[j] = exp(-2 pi i j / PIC_FFT_MAX), j < PIC_FFT_MAX/2.  The butterflies
// of all rows of a stage are spread over the workgroup, one barrier per stage.  All threads
// of the workgroup (PIC_WG) must call it.
__device__ __forceinline__ void fft_lds_rows(double2* a, int rows, int n, int logn, int sign,
                                             const double2* __restrict__ tw) {
  const double ws = (sign < 0) ? 1.0 : -1.0;
  const int half = n >> 1;
  const int total = rows * half;
  for (int s = 1; s <= logn; ++s) {
    const int h = 1 << (s - 1);
    const int ts = PIC_FFT_MAX >> s;
    for (int b = threadIdx.x; b < total; b += PIC_WG) {
      const int r = b >> (logn - 1), bb = b & (half - 1);
      const int j = bb & (h - 1);
      double2* row = a + r * n;
      const int i0 = ((bb >> (s - 1)) << s) + j;
      const int i1 = i0 + h;
      double2 w = tw[j * ts];
      w.y *= ws;
      const double2 u = row[i0], v = row[i1];
      const double tr = fma(v.x, w.x, -v.y * w.y);
      const double ti = fma(v.x, w.y, v.y * w.x);
      row[i0] = make_double2(u.x + tr, u.y + ti);
      row[i1] = make_double2(u.x - tr, u.y - ti);
    }
    __syncthreads();
  }
}

// one row
__device__ __forceinline__ void fft_lds(double2* a, int n, int logn, int sign,
                                        const double2* __restrict__ tw) {
  fft_lds_rows(a, 1, n, logn, sign, tw);
}

// sin(2 pi m / n) from the twiddle table
__device__ __forceinline__ double sin2pi_frac(int m, int n, const double2* __restrict__ tw) {
  const int st = PIC_FFT_MAX / n;
  if (m == 0 || 2 * m == n) return 0.0;
  if (2 * m < n) return -tw[m * st].y;
  return tw[(n - m) * st].y;
}
