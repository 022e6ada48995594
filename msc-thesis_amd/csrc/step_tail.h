// step_tail.h -- the end-of-step work: the diagnostics row of a step (the fixed-order
// reduction of the per-workgroup partial sums) and the scan of the fused sort's count
// table.  Each part is a workgroup body: its own kernels in k_particles.hip run it, and
// inside a captured sequence of steps the field kernels of the next step run it in extra
// workgroups instead (a launch less per part; the step's field solve does not read what
// these parts write, and kernel order makes their results visible to the next push).
#pragma once
#include "launch.h"

// In-place exclusive scan of a[0..n) by one 256-thread workgroup (tiles of 256 values
// with a carried prefix); returns the total in every thread.
__device__ __forceinline__ int block_exclusive_scan(int* a, int n, int* wsum) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int carry = 0;
  for (int base = 0; base < n; base += PIC_WG) {
    const int i = base + threadIdx.x;
    const int v = (i < n) ? a[i] : 0;
    int inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(inc, o, 64);
      if (lane >= o) inc += t;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int wpre = 0;
    for (int k = 0; k < w; ++k) wpre += wsum[k];
    const int tile_total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    if (i < n) a[i] = carry + wpre + inc - v;
    carry += tile_total;
    __syncthreads();
  }
  return carry;
}

// The diagnostics row of run `run`'s step st.step from the per-workgroup partial sums of
// the field pass (field energy, in row-group order) and of the push (kinetic energy and
// momentum, in item order).  advance: also move the run to the next step (step counter,
// synchronous-velocity flag); without it the caller's kernel order does that (the field
// kernels of the next step read the counter, so their own finishing workgroups leave it
// to the column pass).  One PIC_WG workgroup.
__device__ __forceinline__ void step_finish_run(const RunDesc* __restrict__ runs,
                                                RunState* __restrict__ st,
                                                const double* __restrict__ ppush,
                                                const double* __restrict__ prows,
                                                double* __restrict__ diag, int run,
                                                bool advance) {
  __shared__ double red[4][4];
  // the step counter and particle count do not depend on the sums: loaded first, so their
  // latency overlaps the partials'
  long long stp = 0;
  int np = 0;
  if (threadIdx.x == 0) {
    stp = st[run].step;
    np = st[run].np;
  }
  const RunDesc rd = runs[run];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  // each thread adds its items i, i + PIC_WG, ... in that order; FINISH_U of them are loaded
  // together (past the end: the last item, added as zero), so a run with many push items
  // (the collision runs' tail chunks) waits for one load latency per FINISH_U items
  constexpr int FINISH_U = 8;
  const int pe = rd.push_item_end;
  for (int i0 = rd.push_item_begin + threadIdx.x; i0 < pe; i0 += FINISH_U * PIC_WG) {
    double a0[FINISH_U], a1[FINISH_U], a2[FINISH_U];
#pragma unroll
    for (int u = 0; u < FINISH_U; ++u) {
      const int i = i0 + u * PIC_WG;
      const long long ic = 3 * (long long)(i < pe ? i : pe - 1);
      a0[u] = ppush[ic + 0];
      a1[u] = ppush[ic + 1];
      a2[u] = ppush[ic + 2];
    }
#pragma unroll
    for (int u = 0; u < FINISH_U; ++u) {
      const bool ok = i0 + u * PIC_WG < pe;
      v[0] += ok ? a0[u] : 0.0;
      v[1] += ok ? a1[u] : 0.0;
      v[2] += ok ? a2[u] : 0.0;
    }
  }
  for (int q = rd.row_item_begin + threadIdx.x; q < rd.row_item_end; q += PIC_WG) v[3] += prows[q];
  // the four workgroup sums in one pass (each: wave sums, then the four waves in order)
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
    for (int k = 0; k < 4; ++k) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    const double s0 = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
    const double s1 = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
    const double s2 = ((red[2][0] + red[2][1]) + red[2][2]) + red[2][3];
    const double s3 = ((red[3][0] + red[3][1]) + red[3][2]) + red[3][3];
    const double mp = rd.qp / rd.qm;
    double* row = diag + (rd.dbase + (stp % PIC_DIAG_CAP)) * 8;
    row[0] = 0.5 * s3 * rd.dx * rd.dy;
    row[1] = 0.5 * mp * s0;
    row[2] = mp * s1;
    row[3] = mp * s2;
    row[5] = (double)np;
    // the collision count of the next step starts at zero (a collision stage adds to its
    // step's row, before or after that step's end runs)
    diag[(rd.dbase + ((stp + 1) % PIC_DIAG_CAP)) * 8 + 6] = 0.0;
    if (advance) {
      st[run].step = stp + 1;
      st[run].vsync = 0;
    }
  }
}

// ---- scan of the fused sort's count table.  Row-major: row c (one item of the push grid)
// holds that item's count for every tile, so each push workgroup reads and writes one
// contiguous row.  The destination slot of (tile b, row c) is
//   base[b] + sum_{c' < c} count[c'][b],   base[b] = sum_{b' < b} sum_c count[c][b'],
// i.e. tile-major order with rows in item order.  Rows are scanned in groups of
// FSCAN_ROWS: fscan_part sums each group's columns, fscan_base scans the group sums down
// each column and the column totals across tiles (one workgroup per run; it also writes
// the tile offsets of the new order), fscan_rows writes the slots.

// FSCAN_UNROLL independent loads per thread are issued before any of them is used
#define FSCAN_UNROLL 32

// row group g (an index into items)
__device__ __forceinline__ void fscan_part(const RunDesc* __restrict__ runs,
                                           const int* __restrict__ table, int* __restrict__ part,
                                           int stride, const int2* __restrict__ items, int g) {
  const int2 it = items[g];
  const RunDesc rd = runs[it.x];
  const int nb = rd.nbuckets;
  const int r0 = it.y * FSCAN_ROWS, r1 = min(r0 + FSCAN_ROWS, rd.fchunks);
  for (int b = threadIdx.x; b < nb; b += PIC_WG) {
    const int* col = table + rd.ftbase + b;
    int sum = 0;
    for (int c0 = r0; c0 < r1; c0 += FSCAN_UNROLL) {
      int v[FSCAN_UNROLL];
#pragma unroll
      for (int k = 0; k < FSCAN_UNROLL; ++k)
        v[k] = (c0 + k < r1) ? col[(long long)(c0 + k) * nb] : 0;
#pragma unroll
      for (int k = 0; k < FSCAN_UNROLL; ++k) sum += v[k];
    }
    part[(long long)g * stride + b] = sum;
  }
}

__device__ __forceinline__ void fscan_base(const RunDesc* __restrict__ runs,
                                           RunState* __restrict__ st, int* __restrict__ part,
                                           int stride, int* __restrict__ toff, int run) {
  __shared__ int tot[512];  // one per tile (the fused sort allows at most 512)
  __shared__ int wsum[4];
  const RunDesc rd = runs[run];
  if (rd.fchunks <= 0) return;  // uniform: this run sorts with the separate kernels
  const int nb = rd.nbuckets, g0 = rd.fscan_item_begin, g1 = rd.fscan_item_end;
  // column totals over the row groups
  for (int b = threadIdx.x; b < nb; b += PIC_WG) {
    int sum = 0;
    for (int gi = g0; gi < g1; gi += FSCAN_UNROLL) {
      int v[FSCAN_UNROLL];
#pragma unroll
      for (int k = 0; k < FSCAN_UNROLL; ++k)
        v[k] = (gi + k < g1) ? part[(long long)(gi + k) * stride + b] : 0;
#pragma unroll
      for (int k = 0; k < FSCAN_UNROLL; ++k) sum += v[k];
    }
    tot[b] = sum;
  }
  __syncthreads();
  block_exclusive_scan(tot, nb, wsum);
  // each group's first slot per tile: the tile base plus the groups before it
  for (int b = threadIdx.x; b < nb; b += PIC_WG) {
    int run_slot = tot[b];
    toff[rd.toffbase + b] = run_slot;
    for (int gi = g0; gi < g1; gi += FSCAN_UNROLL) {
      int v[FSCAN_UNROLL];
#pragma unroll
      for (int k = 0; k < FSCAN_UNROLL; ++k)
        v[k] = (gi + k < g1) ? part[(long long)(gi + k) * stride + b] : 0;
#pragma unroll
      for (int k = 0; k < FSCAN_UNROLL; ++k) {
        if (gi + k < g1) part[(long long)(gi + k) * stride + b] = run_slot;
        run_slot += v[k];
      }
    }
  }
  if (threadIdx.x == 0) {
    const int np = st[run].np;
    toff[rd.toffbase + nb] = np;
    st[run].nsorted = np;
  }
}

// row group g (an index into items)
__device__ __forceinline__ void fscan_rows(const RunDesc* __restrict__ runs,
                                           int* __restrict__ table,
                                           const int* __restrict__ part, int stride,
                                           const int2* __restrict__ items, int g) {
  const int2 it = items[g];
  const RunDesc rd = runs[it.x];
  const int nb = rd.nbuckets;
  const int r0 = it.y * FSCAN_ROWS, r1 = min(r0 + FSCAN_ROWS, rd.fchunks);
  for (int b = threadIdx.x; b < nb; b += PIC_WG) {
    int* col = table + rd.ftbase + b;
    int run_slot = part[(long long)g * stride + b];
    for (int c0 = r0; c0 < r1; c0 += FSCAN_UNROLL) {
      int v[FSCAN_UNROLL];
#pragma unroll
      for (int k = 0; k < FSCAN_UNROLL; ++k)
        v[k] = (c0 + k < r1) ? col[(long long)(c0 + k) * nb] : 0;
#pragma unroll
      for (int k = 0; k < FSCAN_UNROLL; ++k) {
        if (c0 + k < r1) col[(long long)(c0 + k) * nb] = run_slot;
        run_slot += v[k];
      }
    }
  }
}
