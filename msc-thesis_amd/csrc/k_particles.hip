// k_particles.hip -- particle-side kernels: cell sort, fused gather+push+deposit, outputs.
//
// Reference stages replaced (thesisCode/collisionless/double/src/pic.cu):
//   currentCell        :69-87    -> cell key computed inside the sort and the push
//   chargeAssignment   :99-138   -> deterministic fixed-point CIC deposit in LDS windows
//   forceCalculation   :298-355  -> bilinear gather fused into the push
//   positionUpdate     :372-420  -> leapfrog push, periodic wrap into [0, L)
//   cellTrackerSetup   thesisCode/dsmcCollisions/cellTrackerAdvanced/src/pic.cu:419-445
//                                -> stable counting sort by key cx*ny+cy (ties by id)
#include <climits>
#include <type_traits>
#include "launch.h"
#include "deposit.h"
#include "step_tail.h"

namespace {

// ========================================================================== sort
#define HWIN 8192  // LDS histogram window (keys) of the rank kernel

// Cell numbering of the engine's own sorts: tile-major (tile t = (cx/T)*tiles_y + cy/T,
// then cx%T, cy%T), so that every tile and every cell is one contiguous range.  The
// operator pic_op_sort sorts whatever keys it is given (cx*ny+cy from pic_op_cell_keys).
// Tile sizes are powers of two and cells are clamped to >= 0: shifts and masks.
__device__ __forceinline__ int engine_cell_key(int cx, int cy, const RunDesc& rd) {
  const int sh = __builtin_ctz((unsigned)rd.tile), m = rd.tile - 1;
  const int t = (cx >> sh) * rd.tiles_y + (cy >> sh);
  return (((t << sh) + (cx & m)) << sh) + (cy & m);
}

template <typename R, bool FROM_POS>
__global__ __launch_bounds__(PIC_WG) void k_sort_rank(const RunDesc* __restrict__ runs,
                                                      const RunState* __restrict__ st,
                                                      const R* __restrict__ x,
                                                      const R* __restrict__ y,
                                                      const int* __restrict__ keys_in,
                                                      int* __restrict__ keys,
                                                      int* __restrict__ rank,
                                                      int* __restrict__ counts,
                                                      const int2* __restrict__ items) {
  __shared__ int hist[HWIN];
  __shared__ int red[8];
  // slice blockIdx % PIC_SORT_SLICES of chunk item blockIdx / PIC_SORT_SLICES
  const int2 it = items[blockIdx.x / PIC_SORT_SLICES];
  const RunDesc rd = runs[it.x];
  const int np = st[it.x].np;
  const int p0 = it.y + (blockIdx.x % PIC_SORT_SLICES) * (PIC_CHUNK / PIC_SORT_SLICES);
  const int p1 = min(p0 + PIC_CHUNK / PIC_SORT_SLICES, np);
  if (p0 >= p1) return;  // uniform across the workgroup
  const long long pb = rd.pbase;
  const int nx = rd.nx, ny = rd.ny;
  const R dxr = (R)rd.dx, dyr = (R)rd.dy;
  int* cnt = counts + rd.cbase;
  // a slice is RANK_U particles per thread: their loads go out together (lanes past p1 load
  // particle p1 - 1, no branches) and their keys stay in registers for the later passes
  constexpr int RANK_U = (PIC_CHUNK / PIC_SORT_SLICES) / PIC_WG;
  static_assert(RANK_U * PIC_WG == PIC_CHUNK / PIC_SORT_SLICES, "slice = whole rounds");
  int key[RANK_U];
  bool act[RANK_U];
  int kmin = INT_MAX, kmax = INT_MIN;
  if (FROM_POS) {
    R xv[RANK_U], yv[RANK_U];
#pragma unroll
    for (int u = 0; u < RANK_U; ++u) {
      const int p = p0 + threadIdx.x + u * PIC_WG;
      act[u] = p < p1;
      const int pl = act[u] ? p : p1 - 1;
      xv[u] = x[pb + pl];
      yv[u] = y[pb + pl];
    }
#pragma unroll
    for (int u = 0; u < RANK_U; ++u) {
      int cx, cy;
      R f;
      locate(xv[u], dxr, nx, cx, f);
      locate(yv[u], dyr, ny, cy, f);
      key[u] = engine_cell_key(cx, cy, rd);
      if (act[u]) keys[pb + p0 + threadIdx.x + u * PIC_WG] = key[u];
    }
  } else {
#pragma unroll
    for (int u = 0; u < RANK_U; ++u) {
      const int p = p0 + threadIdx.x + u * PIC_WG;
      act[u] = p < p1;
      key[u] = keys_in[pb + (act[u] ? p : p1 - 1)];
    }
  }
#pragma unroll
  for (int u = 0; u < RANK_U; ++u)
    if (act[u]) {
      kmin = min(kmin, key[u]);
      kmax = max(kmax, key[u]);
    }
  kmin = wave_min_i(kmin);
  kmax = wave_max_i(kmax);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { red[w] = kmin; red[4 + w] = kmax; }
  __syncthreads();
  kmin = min(min(red[0], red[1]), min(red[2], red[3]));
  kmax = max(max(red[4], red[5]), max(red[6], red[7]));
  const long long span = (long long)kmax - kmin + 1;
  if (span <= HWIN) {
    for (int k = threadIdx.x; k < span; k += PIC_WG) hist[k] = 0;
    __syncthreads();
    int rk[RANK_U];
#pragma unroll
    for (int u = 0; u < RANK_U; ++u)
      if (act[u]) rk[u] = atomicAdd(&hist[key[u] - kmin], 1);
    __syncthreads();
    for (int k = threadIdx.x; k < span; k += PIC_WG) {
      int c = hist[k];
      if (c) hist[k] = atomicAdd(&cnt[kmin + k], c);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < RANK_U; ++u)
      if (act[u]) rank[pb + p0 + threadIdx.x + u * PIC_WG] = rk[u] + hist[key[u] - kmin];
  } else {
#pragma unroll
    for (int u = 0; u < RANK_U; ++u)
      if (act[u]) rank[pb + p0 + threadIdx.x + u * PIC_WG] = atomicAdd(&cnt[key[u]], 1);
  }
}

// exclusive scan of the counts of one chunk of PIC_CELL_CHUNK cells; with `clear` the counts
// are zeroed once read (the engine's sorts keep them zero between sorts, so no fill launch
// precedes the rank kernel)
__global__ __launch_bounds__(PIC_WG) void k_scan_local(const RunDesc* __restrict__ runs,
                                                       int* __restrict__ counts,
                                                       int* __restrict__ offsets,
                                                       int* __restrict__ chunk_sums,
                                                       const int2* __restrict__ items,
                                                       int clear) {
  __shared__ int wsum[4];
  const int2 it = items[blockIdx.x];
  const RunDesc rd = runs[it.x];
  const int G = rd.nx * rd.ny;
  const int c0 = it.y + threadIdx.x * 8;
  int v[8], s = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    int c = c0 + k;
    v[k] = (c < G) ? counts[rd.cbase + c] : 0;
    s += v[k];
  }
  if (clear) {
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (c0 + k < G) counts[rd.cbase + c0 + k] = 0;
  }
  // inclusive wave scan of per-thread totals
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int inc = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int wpre = 0;
  for (int k = 0; k < w; ++k) wpre += wsum[k];
  int ex = wpre + inc - s;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    int c = c0 + k;
    if (c < G) offsets[rd.cbase + c] = ex;
    ex += v[k];
  }
  if (threadIdx.x == PIC_WG - 1) chunk_sums[blockIdx.x] = wpre + inc;
}

// The chunk's base added to its offsets: the base is the sum of the run's chunk totals
// before it (integer adds, so any order gives the same base), added up by the workgroup
// itself, which saves a launch that scanned the chunk totals.  With toff (a tile-major cell
// numbering: tile t owns cells [t T^2, (t + 1) T^2)) also the tile offsets toff[t] =
// offsets[t T^2].  The run's last chunk writes the total: offsets[G], nsorted and the
// closing tile offset; the first workgroup clears the count of cells too large for the
// within-cell sort's small kernel (which appends to it next).
__global__ __launch_bounds__(PIC_WG) void k_scan_add(const RunDesc* __restrict__ runs,
                                                     RunState* __restrict__ st,
                                                     int* __restrict__ offsets,
                                                     const int* __restrict__ chunk_sums,
                                                     const int2* __restrict__ items,
                                                     int* __restrict__ toff,
                                                     int* __restrict__ big_count) {
  __shared__ int wsum[PIC_WG / 64];
  const int2 it = items[blockIdx.x];
  const RunDesc rd = runs[it.x];
  const int G = rd.nx * rd.ny;
  const int TT = rd.tile * rd.tile;
  int part = 0;
  for (int i = rd.cell_item_begin + (int)threadIdx.x; i < (int)blockIdx.x; i += PIC_WG)
    part += chunk_sums[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = part;
  __syncthreads();
  int add = 0;
#pragma unroll
  for (int w = 0; w < PIC_WG / 64; ++w) add += wsum[w];
  for (int k = threadIdx.x; k < PIC_CELL_CHUNK; k += PIC_WG) {
    int c = it.y + k;
    if (c < G) {
      const int o = offsets[rd.cbase + c] + add;
      offsets[rd.cbase + c] = o;
      if (toff && c % TT == 0) toff[rd.toffbase + c / TT] = o;
    }
  }
  if (threadIdx.x == 0) {
    if ((int)blockIdx.x == rd.cell_item_end - 1) {
      const int total = add + chunk_sums[blockIdx.x];
      offsets[rd.cbase + G] = total;
      st[it.x].nsorted = total;
      if (toff) toff[rd.toffbase + rd.nbuckets] = total;
    }
    if (blockIdx.x == 0) *big_count = 0;
  }
}

__global__ __launch_bounds__(PIC_WG) void k_sort_scatter(const RunDesc* __restrict__ runs,
                                                         const RunState* __restrict__ st,
                                                         const int* __restrict__ keys,
                                                         const int* __restrict__ rank,
                                                         const int* __restrict__ offsets,
                                                         int* __restrict__ perm,
                                                         const int2* __restrict__ items) {
  const int2 it = items[blockIdx.x / PIC_SORT_SLICES];
  const RunDesc rd = runs[it.x];
  const int np = st[it.x].np;
  const int p0 = it.y + (blockIdx.x % PIC_SORT_SLICES) * (PIC_CHUNK / PIC_SORT_SLICES);
  const int p1 = min(p0 + PIC_CHUNK / PIC_SORT_SLICES, np);
  for (int p = p0 + threadIdx.x; p < p1; p += PIC_WG) {
    int d = offsets[rd.cbase + keys[rd.pbase + p]] + rank[rd.pbase + p];
    perm[rd.pbase + d] = p;
  }
}

__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int m) {
  unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
  lo = __shfl_xor(lo, m, 64);
  hi = __shfl_xor(hi, m, 64);
  return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ unsigned long long shfl_u64(unsigned long long v, int src) {
  unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
  lo = __shfl(lo, src, 64);
  hi = __shfl(hi, src, 64);
  return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ unsigned long long seg_key(const int* id, long long pb, int slot) {
  unsigned k = id ? (unsigned)id[pb + slot] : (unsigned)slot;
  return ((unsigned long long)k << 32) | (unsigned)slot;
}

// Canonical order inside each cell: ascending key (id, or the source index when id is
// NULL, which makes the sort stable).  One wavefront per cell: <= 64 entries by a bitonic
// network across lanes, <= 256 entries by counting ranks (4 entries per lane, keys are
// unique); larger cells are queued for k_segsort_big.  Items cover PIC_SEG_CELLS cells
// (one per wave).
// 65-256 entries of one cell, sorted by one wave: a bitonic network over 64 * NR keys held
// NR per lane (key e at lane e % 64 of register e / 64; padding ~0 sorts last).  Distances of
// 64 and more pair registers of one lane, shorter ones lanes of one register.  (Ranking each
// key by counting the smaller ones took n steps per wave.)
template <int NR>
__device__ __forceinline__ void bitonic_regs(const int* __restrict__ id, int* __restrict__ perm,
                                             long long pb, int b, int n, int lane) {
  unsigned long long k4[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int j = lane + 64 * k;
    k4[k] = (j < n) ? seg_key(id, pb, perm[pb + b + j]) : ~0ull;
  }
#pragma unroll
  for (int size = 2; size <= 64 * NR; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride >= 64) {
        const int sr = stride >> 6;
#pragma unroll
        for (int k = 0; k < NR; ++k) {
          if (k & sr) continue;  // k is the lower element of the pair (k, k | sr)
          const bool up = ((64 * k) & size) == 0;
          const unsigned long long a0 = k4[k], a1 = k4[k | sr];
          const unsigned long long mn = a0 < a1 ? a0 : a1, mx = a0 < a1 ? a1 : a0;
          k4[k] = up ? mn : mx;
          k4[k | sr] = up ? mx : mn;
        }
      } else {
        const bool lower = (lane & stride) == 0;
#pragma unroll
        for (int k = 0; k < NR; ++k) {
          const unsigned long long o = shfl_xor_u64(k4[k], stride);
          const bool up = ((lane + 64 * k) & size) == 0;
          const unsigned long long mn = k4[k] < o ? k4[k] : o, mx = k4[k] < o ? o : k4[k];
          k4[k] = (lower == up) ? mn : mx;
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NR; ++k)
    if (lane + 64 * k < n) perm[pb + b + lane + 64 * k] = (int)(unsigned)(k4[k] & 0xffffffffull);
}

__global__ __launch_bounds__(PIC_WG) void k_segsort_small(const RunDesc* __restrict__ runs,
                                                          const int* __restrict__ offsets,
                                                          int* __restrict__ perm,
                                                          const int* __restrict__ id,
                                                          int* __restrict__ big_list,
                                                          int* __restrict__ big_count,
                                                          const int2* __restrict__ items) {
  const int2 it = items[blockIdx.x];
  const RunDesc rd = runs[it.x];
  const int G = rd.nx * rd.ny;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int cend = min(it.y + PIC_SEG_CELLS, G);
  const long long pb = rd.pbase;
  for (int c = it.y + w; c < cend; c += 4) {
    const int b = offsets[rd.cbase + c], e = offsets[rd.cbase + c + 1];
    const int n = e - b;
    if (n <= 1) continue;
    if (n > 256) {
      if (lane == 0) {
        int slot = atomicAdd(big_count, 1);
        big_list[2 * slot] = it.x;
        big_list[2 * slot + 1] = c;
      }
      continue;
    }
    if (n > 128) {
      bitonic_regs<4>(id, perm, pb, b, n, lane);
      continue;
    }
    if (n > 64) {
      bitonic_regs<2>(id, perm, pb, b, n, lane);
      continue;
    }
    unsigned long long v = (lane < n) ? seg_key(id, pb, perm[pb + b + lane]) : ~0ull;
    // the network over the first m = 2^ceil(log2 n) lanes only (the others hold padding)
    const int m = 1 << (32 - __clz(n - 1));
    for (int k = 2; k <= m; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        unsigned long long o = shfl_xor_u64(v, j);
        bool up = (lane & k) == 0;
        bool lower = (lane & j) == 0;
        unsigned long long mn = v < o ? v : o, mx = v < o ? o : v;
        v = (lower == up) ? mn : mx;
      }
    }
    if (lane < n) perm[pb + b + lane] = (int)(unsigned)(v & 0xffffffffull);
  }
}

// Large cells: bitonic sort by one workgroup, in LDS up to 4096 entries, else in a
// global scratch region [2*(pbase+b), 2*(pbase+b) + m) with m = next pow2 >= n (<= 2n).
__global__ __launch_bounds__(PIC_WG) void k_segsort_big(const RunDesc* __restrict__ runs,
                                                        const int* __restrict__ offsets,
                                                        int* __restrict__ perm,
                                                        const int* __restrict__ id,
                                                        const int* __restrict__ big_list,
                                                        const int* __restrict__ big_count,
                                                        unsigned long long* __restrict__ scratch) {
  __shared__ unsigned long long s[4096];
  const int nbig = *big_count;
  for (int q = blockIdx.x; q < nbig; q += gridDim.x) {
    const int run = big_list[2 * q], c = big_list[2 * q + 1];
    const RunDesc rd = runs[run];
    const long long pb = rd.pbase;
    const int b = offsets[rd.cbase + c], e = offsets[rd.cbase + c + 1];
    const int n = e - b;
    int m = 1;
    while (m < n) m <<= 1;
    unsigned long long* a = (m <= 4096) ? s : scratch + 2 * (pb + b);
    for (int i = threadIdx.x; i < m; i += PIC_WG)
      a[i] = (i < n) ? seg_key(id, pb, perm[pb + b + i]) : ~0ull;
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < m; i += PIC_WG) {
          int l = i ^ j;
          if (l > i) {
            unsigned long long u = a[i], v = a[l];
            bool up = (i & k) == 0;
            if ((u > v) == up) { a[i] = v; a[l] = u; }
          }
        }
        __syncthreads();
      }
    }
    for (int i = threadIdx.x; i < n; i += PIC_WG)
      perm[pb + b + i] = (int)(unsigned)(a[i] & 0xffffffffull);
    __syncthreads();
  }
}

template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_sort_apply(const RunDesc* __restrict__ runs,
                                                       const RunState* __restrict__ st,
                                                       const int* __restrict__ perm,
                                                       PartBufs<R> src, PartBufs<R> dst,
                                                       const int2* __restrict__ items) {
  const int2 it = items[blockIdx.x / PIC_SORT_SLICES];
  const RunDesc rd = runs[it.x];
  const int np = st[it.x].np;
  const int d0 = it.y + (blockIdx.x % PIC_SORT_SLICES) * (PIC_CHUNK / PIC_SORT_SLICES);
  const int p1 = min(d0 + PIC_CHUNK / PIC_SORT_SLICES, np);
  const long long pb = rd.pbase;
  // APPLY_U slots per thread per round, every load of a round issued before any store (one
  // slot per thread and round waited for its permutation entry, then for its five reads);
  // lanes past p1 read slot p1 - 1 (loads without branches) and store nothing
  constexpr int APPLY_U = 4;
  for (int d = d0 + threadIdx.x; d < p1; d += APPLY_U * PIC_WG) {
    int sl[APPLY_U];
#pragma unroll
    for (int u = 0; u < APPLY_U; ++u) sl[u] = perm[pb + min(d + u * PIC_WG, p1 - 1)];
    R x[APPLY_U], y[APPLY_U], vx[APPLY_U], vy[APPLY_U];
    int idv[APPLY_U];
#pragma unroll
    for (int u = 0; u < APPLY_U; ++u) {
      const long long s = pb + sl[u];
      x[u] = src.x[s]; y[u] = src.y[s]; vx[u] = src.vx[s]; vy[u] = src.vy[s];
      idv[u] = src.id[s];
    }
#pragma unroll
    for (int u = 0; u < APPLY_U; ++u) {
      const int dd = d + u * PIC_WG;
      if (dd < p1) {
        dst.x[pb + dd] = x[u]; dst.y[pb + dd] = y[u]; dst.vx[pb + dd] = vx[u];
        dst.vy[pb + dd] = vy[u]; dst.id[pb + dd] = idv[u];
      }
    }
  }
}

// ------------------------------------------------------------ tile sort of the step
// Stable counting sort of the particles by tile (sort bucket):
//   1. k_tile_hist    per chunk of schunk particles: count of each bucket -> table[b][chunk]
//   2. k_tscan_*      exclusive scan of the table in bucket-major order (per run)
//   3. k_tile_scatter one wavefront per chunk walks its particles in index order; lanes
//                     with equal buckets get consecutive slots (ballot ranks), so the
//                     order inside a bucket is the previous order: deterministic and
//                     stable without any per-cell sort.
// The bucket of a position is its tile.  Every pass of one sort evaluates this same
// function, so any deterministic cell choice is consistent: the cell comes from the
// reciprocal multiply (locate_w), the tile from shifts (tile sizes are powers of two).  No
// per-particle division: the count and scatter passes run it once per particle.
template <typename R>
__device__ __forceinline__ int bucket_of(R x, R y, const Geo<R>& g, const RunDesc& rd) {
  int cx, cy;
  R f;
  locate_w(x, g.idx, g.nx, cx, f);
  locate_w(y, g.idy, g.ny, cy, f);
  const int sh = __builtin_ctz((unsigned)rd.tile);
  return (cx >> sh) * rd.tiles_y + (cy >> sh);
}

// Key of the tile sorts: the tile of the position extrapolated by rd.sort_lead (about half
// the span of the steps until the next sort; 0 in runs with collisions, whose cell lists
// need the tiles of the positions themselves).  The particles of a tile range then stay
// within about half a cycle's drift of their tile, instead of drifting a whole cycle away
// from it, so far fewer of them leave the tile's LDS windows (gathers and deposits through
// HBM).  Any key gives the same trajectories; both passes of one sort evaluate it on the
// same stored values.
template <typename R>
__device__ __forceinline__ int bucket_lead(R x, R y, R vx, R vy, const Geo<R>& g,
                                           const RunDesc& rd) {
  if (rd.sort_lead == 0.0) return bucket_of<R>(x, y, g, rd);
  const R t = (R)rd.sort_lead;
  return bucket_of<R>(wrap_pos(fmar(t, vx, x), g.Lx), wrap_pos(fmar(t, vy, y), g.Ly), g, rd);
}

// The same key with its place relative to tile (tx, ty), for the scatter step's ranking:
// bits 0-15 the key, bits 16 and up the neighbour code (dx + 1) * 3 + (dy + 1) of a key
// tile within one tile of (tx, ty) along each axis (periodically), else 9.  With fewer than
// 3 tiles along an axis the neighbours are not distinct tiles: every key gets 9.
#define PUSH_NCODES 9
template <typename R>
__device__ __forceinline__ int bucket_lead_code(R x, R y, R vx, R vy, const Geo<R>& g,
                                                const RunDesc& rd, int tx, int ty) {
  if (rd.sort_lead != 0.0) {
    const R t = (R)rd.sort_lead;
    x = wrap_pos(fmar(t, vx, x), g.Lx);
    y = wrap_pos(fmar(t, vy, y), g.Ly);
  }
  int cx, cy;
  R f;
  locate_w(x, g.idx, g.nx, cx, f);
  locate_w(y, g.idy, g.ny, cy, f);
  const int sh = __builtin_ctz((unsigned)rd.tile);
  const int kx = cx >> sh, ky = cy >> sh;
  int dx = kx - tx, dy = ky - ty;
  if (dx > 1) dx -= rd.tiles_x;
  else if (dx < -1) dx += rd.tiles_x;
  if (dy > 1) dy -= rd.tiles_y;
  else if (dy < -1) dy += rd.tiles_y;
  const bool near = rd.tiles_x >= 3 && rd.tiles_y >= 3 && dx >= -1 && dx <= 1 && dy >= -1 &&
                    dy <= 1;
  const int code = near ? (dx + 1) * 3 + (dy + 1) : PUSH_NCODES;
  return (kx * rd.tiles_y + ky) | (code << 16);
}

// the tile of neighbour code c of tile (tx, ty) (bucket_lead_code)
__device__ __forceinline__ int code_tile(int c, int tx, int ty, const RunDesc& rd) {
  int kx = tx + c / 3 - 1, ky = ty + c % 3 - 1;
  kx = kx < 0 ? kx + rd.tiles_x : (kx >= rd.tiles_x ? kx - rd.tiles_x : kx);
  ky = ky < 0 ? ky + rd.tiles_y : (ky >= rd.tiles_y ? ky - rd.tiles_y : ky);
  return kx * rd.tiles_y + ky;
}

// lanes below this one with their bit set in m
__device__ __forceinline__ int lanes_below(unsigned long long m) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_tile_hist(const RunDesc* __restrict__ runs,
                                                      const RunState* __restrict__ st,
                                                      const R* __restrict__ x,
                                                      const R* __restrict__ y,
                                                      const R* __restrict__ vx,
                                                      const R* __restrict__ vy,
                                                      int* __restrict__ table,
                                                      const int2* __restrict__ items) {
  extern __shared__ int h[];
  const int2 it = items[blockIdx.x];
  const RunDesc rd = runs[it.x];
  const Geo<R> g = make_geo<R>(rd.nx, rd.ny, rd.lx, rd.ly);
  const int np = st[it.x].np;
  const int nb = rd.nbuckets;
  const int chunk = it.y / rd.schunk;
  for (int b = threadIdx.x; b < nb; b += PIC_WG) h[b] = 0;
  __syncthreads();
  const int p1 = min(it.y + rd.schunk, np);
  // four particles per thread per pass, loads first
  for (int q = it.y + threadIdx.x; q < p1; q += 4 * PIC_WG) {
    R xs[4], ys[4], vxs[4], vys[4];
    const bool lead = rd.sort_lead != 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int p = q + u * PIC_WG;
      xs[u] = p < p1 ? x[rd.pbase + p] : (R)0;
      ys[u] = p < p1 ? y[rd.pbase + p] : (R)0;
      vxs[u] = (lead && p < p1) ? vx[rd.pbase + p] : (R)0;
      vys[u] = (lead && p < p1) ? vy[rd.pbase + p] : (R)0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (q + u * PIC_WG < p1)
        atomicAdd(&h[bucket_lead<R>(xs[u], ys[u], vxs[u], vys[u], g, rd)], 1);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += PIC_WG)
    table[rd.tbase + (long long)b * rd.nchunks + chunk] = h[b];
}

// exclusive scan of one PIC_CELL_CHUNK slice of a run's tile table, in place
__global__ __launch_bounds__(PIC_WG) void k_tscan_local(const RunDesc* __restrict__ runs,
                                                        int* __restrict__ table,
                                                        int* __restrict__ chunk_sums,
                                                        const int2* __restrict__ items) {
  __shared__ int wsum[4];
  const int2 it = items[blockIdx.x];
  const RunDesc rd = runs[it.x];
  const long long len = (long long)rd.nbuckets * rd.nchunks, tb = rd.tbase;
  const long long c0 = it.y + threadIdx.x * 8;
  int v[8], s = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    v[k] = (c0 + k < len) ? table[tb + c0 + k] : 0;
    s += v[k];
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int inc = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int wpre = 0;
  for (int k = 0; k < w; ++k) wpre += wsum[k];
  int ex = wpre + inc - s;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (c0 + k < len) table[tb + c0 + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == PIC_WG - 1) chunk_sums[blockIdx.x] = wpre + inc;
}

__global__ __launch_bounds__(PIC_WG) void k_tscan_chunks(const RunDesc* __restrict__ runs,
                                                         int* __restrict__ chunk_sums) {
  __shared__ int wsum[4];
  const RunDesc rd = runs[blockIdx.x];
  block_exclusive_scan(chunk_sums + rd.tscan_item_begin,
                       rd.tscan_item_end - rd.tscan_item_begin, wsum);
}

__global__ __launch_bounds__(PIC_WG) void k_tscan_add(const RunDesc* __restrict__ runs,
                                                      int* __restrict__ table,
                                                      const int* __restrict__ chunk_sums,
                                                      const int2* __restrict__ items) {
  const int2 it = items[blockIdx.x];
  const RunDesc rd = runs[it.x];
  const long long len = (long long)rd.nbuckets * rd.nchunks, tb = rd.tbase;
  const int add = chunk_sums[blockIdx.x];
  for (int k = threadIdx.x; k < PIC_CELL_CHUNK; k += PIC_WG) {
    const long long c = (long long)it.y + k;
    if (c < len) table[tb + c] += add;
  }
}

// ---- scan of the fused sort's count table (the workgroup bodies are in step_tail.h)
__global__ __launch_bounds__(PIC_WG) void k_fscan_part(const RunDesc* __restrict__ runs,
                                                       const int* __restrict__ table,
                                                       int* __restrict__ part, int stride,
                                                       const int2* __restrict__ items) {
  fscan_part(runs, table, part, stride, items, blockIdx.x);
}

__global__ __launch_bounds__(PIC_WG) void k_fscan_base(const RunDesc* __restrict__ runs,
                                                       RunState* __restrict__ st,
                                                       int* __restrict__ part, int stride,
                                                       int* __restrict__ toff) {
  fscan_base(runs, st, part, stride, toff, blockIdx.x);
}

__global__ __launch_bounds__(PIC_WG) void k_fscan_rows(const RunDesc* __restrict__ runs,
                                                       int* __restrict__ table,
                                                       const int* __restrict__ part,
                                                       int stride,
                                                       const int2* __restrict__ items) {
  fscan_rows(runs, table, part, stride, items, blockIdx.x);
}

template <typename R>
__global__ __launch_bounds__(64) void k_tile_scatter(const RunDesc* __restrict__ runs,
                                                     RunState* __restrict__ st,
                                                     PartBufs<R> src, PartBufs<R> dst,
                                                     const int* __restrict__ table,
                                                     int* __restrict__ toff,
                                                     const int2* __restrict__ items) {
  extern __shared__ int base[];
  const int2 it = items[blockIdx.x];
  const RunDesc rd = runs[it.x];
  const Geo<R> g = make_geo<R>(rd.nx, rd.ny, rd.lx, rd.ly);
  const int np = st[it.x].np;
  const int nb = rd.nbuckets;
  const int chunk = it.y / rd.schunk;
  const int lane = threadIdx.x;
  for (int b = lane; b < nb; b += 64) {
    const int v = table[rd.tbase + (long long)b * rd.nchunks + chunk];
    base[b] = v;
    if (chunk == 0) toff[rd.toffbase + b] = v;
  }
  if (chunk == 0 && lane == 0) {
    toff[rd.toffbase + nb] = np;
    st[it.x].nsorted = np;
  }
  __syncthreads();
  const long long pb = rd.pbase;
  const int p1 = min(it.y + rd.schunk, np);
  // SCAT_U groups of 64 particles per pass: all their loads are issued first (the wave's
  // memory latency is paid once per pass), then the groups are ranked in index order
  constexpr int SCAT_U = 4;
  for (int q0 = it.y; q0 < p1; q0 += 64 * SCAT_U) {
    R vx[SCAT_U], vy[SCAT_U], vvx[SCAT_U], vvy[SCAT_U];
    int vid[SCAT_U], key[SCAT_U];
#pragma unroll
    for (int u = 0; u < SCAT_U; ++u) {
      // lanes past p1 load particle p1 - 1 (p1 > q0 here): loads without branches
      const int p = q0 + 64 * u + lane;
      const int pl = p < p1 ? p : p1 - 1;
      vx[u] = src.x[pb + pl];
      vy[u] = src.y[pb + pl];
      vvx[u] = src.vx[pb + pl];
      vvy[u] = src.vy[pb + pl];
      vid[u] = src.id[pb + pl];
    }
#pragma unroll
    for (int u = 0; u < SCAT_U; ++u) {
      const int p = q0 + 64 * u + lane;
      key[u] = (p < p1) ? bucket_lead<R>(vx[u], vy[u], vvx[u], vvy[u], g, rd) : -1;
    }
#pragma unroll
    for (int u = 0; u < SCAT_U; ++u) {
      const int p = q0 + 64 * u + lane;
      const bool act = p < p1;
      unsigned long long rem = __ballot(act);
      int d = 0;
      while (rem) {
        const int leader = __ffsll((long long)rem) - 1;
        const int lk = __builtin_amdgcn_readlane(key[u], leader);
        const unsigned long long m = __ballot(act && key[u] == lk);
        const int b0 = base[lk];
        if (act && key[u] == lk) d = b0 + __popcll(m & ((1ull << lane) - 1ull));
        __syncthreads();
        if (lane == leader) base[lk] = b0 + __popcll(m);
        __syncthreads();
        rem &= ~m;
      }
      if (act) {
        dst.x[pb + d] = vx[u]; dst.y[pb + d] = vy[u]; dst.vx[pb + d] = vvx[u];
        dst.vy[pb + d] = vvy[u];
        dst.id[pb + d] = vid[u];
      }
    }
  }
}

__global__ void k_fill_i32(int* p, long long n, int v) {
  for (long long i = blockIdx.x * (long long)PIC_WG + threadIdx.x; i < n;
       i += (long long)gridDim.x * PIC_WG)
    p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* p, long long n) {
  for (long long i = blockIdx.x * (long long)PIC_WG + threadIdx.x; i < n;
       i += (long long)gridDim.x * PIC_WG)
    p[i] = 0ull;
}

// ===================================================================== push
// One workgroup per (tile, split): the particles of the tile are the T column ranges
// [off[cx*ny + ty*T], off[cx*ny + ty*T + T]) of the cell-sorted arrays.  The new charge
// is accumulated in an LDS window of (T+2H+1)^2 nodes around the tile with int64 atomics
// and flushed once per workgroup; particles that drifted outside the window since the
// last sort deposit straight to HBM.  Gather reads the padded E array (L2-resident).
// gather + kick + drift of one particle; returns synchronous-velocity sums
// E window in LDS: ew[r*W + c] = E at node (x0 + r, y0 + c) (periodic), or NULL
template <typename R>
struct EWin {
  typedef typename std::conditional<sizeof(R) == 4, float2, double2>::type R2;
  const R2* ew;
  int W, x0, y0;
};

template <typename R>
__device__ __forceinline__ void push_one(R& xr, R& yr, R& vxr, R& vyr, const R* __restrict__ E,
                                         const Geo<R>& g, R hk, R dtr, double skc, double& ke,
                                         double& px, double& py, const EWin<R>& w) {
  int cx, cy;
  R fx, fy;
  locate_w(xr, g.idx, g.nx, cx, fx);
  locate_w(yr, g.idy, g.ny, cy, fy);
  const R qs = (R)(1.0 / (double)(1 << Fix<R>::QB));
  const R gx1 = (R)qweight(fx) * qs, gy1 = (R)qweight(fy) * qs;
  const R gx0 = (R)1 - gx1, gy0 = (R)1 - gy1;
  const R g00 = gx0 * gy0, g10 = gx1 * gy0, g01 = gx0 * gy1, g11 = gx1 * gy1;
  typedef typename EWin<R>::R2 R2;
  R2 e00, e01, e10, e11;
  int rx = cx - w.x0, ry = cy - w.y0;
  rx = pwrap(rx, g.nx);
  ry = pwrap(ry, g.ny);
  if (w.ew && rx < w.W - 1 && ry < w.W - 1) {
    const int i0 = rx * w.W + ry;
    e00 = w.ew[i0]; e01 = w.ew[i0 + 1]; e10 = w.ew[i0 + w.W]; e11 = w.ew[i0 + w.W + 1];
  } else {
    const R2* e = reinterpret_cast<const R2*>(E);
    const long long i00 = (long long)cx * (g.ny + 1) + cy;
    e00 = e[i00]; e01 = e[i00 + 1]; e10 = e[i00 + g.ny + 1]; e11 = e[i00 + g.ny + 2];
  }
  R epx = g00 * e00.x;
  epx = fmar(g10, e10.x, epx);
  epx = fmar(g01, e01.x, epx);
  epx = fmar(g11, e11.x, epx);
  R epy = g00 * e00.y;
  epy = fmar(g10, e10.y, epy);
  epy = fmar(g01, e01.y, epy);
  epy = fmar(g11, e11.y, epy);
  const double vsx = (double)vxr + skc * (double)epx, vsy = (double)vyr + skc * (double)epy;
  ke += vsx * vsx + vsy * vsy;
  px += vsx;
  py += vsy;
  vxr = fmar(hk, epx, vxr);
  vyr = fmar(hk, epy, vyr);
  xr = wrap_pos(fmar(dtr, vxr, xr), g.Lx);
  yr = wrap_pos(fmar(dtr, vyr, yr), g.Ly);
}

#include "push_kernel.inc"

// synchronous velocities for output: v + 1/2 dt qm E(x) when a half kick is pending
template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_half_kick_out(const RunDesc* __restrict__ runs,
                                                          const RunState* __restrict__ st,
                                                          int run, PartBufs<R> p,
                                                          const R* __restrict__ Eall,
                                                          R* __restrict__ ovx,
                                                          R* __restrict__ ovy,
                                                          const int2* __restrict__ items) {
  const int2 it = items[blockIdx.x];
  if (it.x != run) return;
  const RunDesc rd = runs[run];
  const RunState rs = st[run];
  const Geo<R> g = make_geo<R>(rd.nx, rd.ny, rd.lx, rd.ly);
  const R* E = Eall + 2 * rd.ebase;
  const R hk = rs.vsync ? (R)0 : (R)(0.5 * rd.dt * rd.qm);
  const int p1 = min(it.y + PIC_CHUNK, rs.np);
  for (int q = it.y + threadIdx.x; q < p1; q += PIC_WG) {
    const long long i = rd.pbase + q;
    R xr = p.x[i], yr = p.y[i], vxr = p.vx[i], vyr = p.vy[i];
    double ke = 0, px = 0, py = 0;
    R x2 = xr, y2 = yr;
    EWin<R> none;
    none.ew = nullptr; none.W = 0; none.x0 = 0; none.y0 = 0;
    push_one<R>(x2, y2, vxr, vyr, E, g, hk, (R)0, 0.0, ke, px, py, none);
    ovx[q] = vxr;
    ovy[q] = vyr;
  }
}

// scatter slot q of run `run` to output index id[q] (id order)
template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_unpermute(const RunDesc* __restrict__ runs,
                                                      const RunState* __restrict__ st, int run,
                                                      PartBufs<R> p, const R* __restrict__ vxs,
                                                      const R* __restrict__ vys,
                                                      R* __restrict__ ox, R* __restrict__ oy,
                                                      R* __restrict__ ovx, R* __restrict__ ovy,
                                                      const int2* __restrict__ items) {
  const int2 it = items[blockIdx.x];
  if (it.x != run) return;
  const RunDesc rd = runs[run];
  const int np = st[run].np;
  const int p1 = min(it.y + PIC_CHUNK, np);
  for (int q = it.y + threadIdx.x; q < p1; q += PIC_WG) {
    const long long i = rd.pbase + q;
    const int k = p.id[i];
    if (k < 0 || k >= np) continue;
    ox[k] = p.x[i];
    oy[k] = p.y[i];
    ovx[k] = vxs ? vxs[q] : p.vx[i];
    ovy[k] = vys ? vys[q] : p.vy[i];
  }
}

template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_check_finite(const RunDesc* __restrict__ runs,
                                                         const RunState* __restrict__ st,
                                                         PartBufs<R> p,
                                                         const int2* __restrict__ items,
                                                         int* flag) {
  const int2 it = items[blockIdx.x];
  const RunDesc rd = runs[it.x];
  const int p1 = min(it.y + PIC_CHUNK, st[it.x].np);
  bool bad = false;
  for (int q = it.y + threadIdx.x; q < p1; q += PIC_WG) {
    const long long i = rd.pbase + q;
    R s = p.x[i] + p.y[i] + p.vx[i] + p.vy[i];
    if (!isfinite(s)) bad = true;
  }
  if (bad) atomicOr(flag, 1);
}

// ===================================================================== ops
template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_op_cell_keys(const R* __restrict__ x,
                                                         const R* __restrict__ y, int64_t n,
                                                         int nx, int ny, double dx, double dy,
                                                         int* __restrict__ keys) {
  for (int64_t p = blockIdx.x * (int64_t)PIC_WG + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * PIC_WG) {
    int cx, cy;
    R f;
    locate(x[p], (R)dx, nx, cx, f);
    locate(y[p], (R)dy, ny, cy, f);
    keys[p] = cx * ny + cy;
  }
}

template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_op_gather_push(R* x, R* y, R* vx, R* vy, int64_t n,
                                                           const double* __restrict__ ex,
                                                           const double* __restrict__ ey,
                                                           int nx, int ny, double lx, double ly,
                                                           double dt, double qm, double c,
                                                           double* partials) {
  __shared__ double red[4];
  const Geo<R> g = make_geo<R>(nx, ny, lx, ly);
  const R hk = (R)(c * dt * qm), dtr = (R)dt;
  const double skc = (c - 0.5) * dt * qm;
  double ke = 0, px = 0, py = 0;
  for (int64_t p = blockIdx.x * (int64_t)PIC_WG + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * PIC_WG) {
    R xr = x[p], yr = y[p], vxr = vx[p], vyr = vy[p];
    int cx, cy;
    R fx, fy;
    locate_w(xr, g.idx, nx, cx, fx);
    locate_w(yr, g.idy, ny, cy, fy);
    const R qs = (R)(1.0 / (double)(1 << Fix<R>::QB));
    const R gx1 = (R)qweight(fx) * qs, gy1 = (R)qweight(fy) * qs;
    const R gx0 = (R)1 - gx1, gy0 = (R)1 - gy1;
    const R g00 = gx0 * gy0, g10 = gx1 * gy0, g01 = gx0 * gy1, g11 = gx1 * gy1;
    const int cxp = (cx + 1 == nx) ? 0 : cx + 1, cyp = (cy + 1 == ny) ? 0 : cy + 1;
    const long long n00 = (long long)cx * ny + cy, n10 = (long long)cxp * ny + cy,
                    n01 = (long long)cx * ny + cyp, n11 = (long long)cxp * ny + cyp;
    R epx = g00 * (R)ex[n00];
    epx = fmar(g10, (R)ex[n10], epx);
    epx = fmar(g01, (R)ex[n01], epx);
    epx = fmar(g11, (R)ex[n11], epx);
    R epy = g00 * (R)ey[n00];
    epy = fmar(g10, (R)ey[n10], epy);
    epy = fmar(g01, (R)ey[n01], epy);
    epy = fmar(g11, (R)ey[n11], epy);
    const double vsx = (double)vxr + skc * (double)epx, vsy = (double)vyr + skc * (double)epy;
    ke += vsx * vsx + vsy * vsy;
    vxr = fmar(hk, epx, vxr);
    vyr = fmar(hk, epy, vyr);
    x[p] = wrap_pos(fmar(dtr, vxr, xr), g.Lx);
    y[p] = wrap_pos(fmar(dtr, vyr, yr), g.Ly);
    vx[p] = vxr;
    vy[p] = vyr;
  }
  const double s = block_sum(ke, red);
  if (partials && threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <typename R>
__global__ __launch_bounds__(PIC_WG) void k_op_deposit(const R* __restrict__ x,
                                                       const R* __restrict__ y, int64_t n,
                                                       int nx, int ny, double lx, double ly,
                                                       unsigned long long* hi,
                                                       unsigned long long* lo) {
  const Geo<R> g = make_geo<R>(nx, ny, lx, ly);
  for (int64_t p = blockIdx.x * (int64_t)PIC_WG + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * PIC_WG) {
    int cx, cy;
    long long w[4];
    deposit_weights<R>(x[p], y[p], g, cx, cy, w);
    deposit_global<R>(hi, lo, cx, cy, w, g);
  }
}

__global__ __launch_bounds__(PIC_WG) void k_op_fixed_to_double(const unsigned long long* hi,
                                                               const unsigned long long* lo,
                                                               long long G, int fp64,
                                                               double scale, double* out) {
  for (long long q = blockIdx.x * (long long)PIC_WG + threadIdx.x; q < G;
       q += (long long)gridDim.x * PIC_WG) {
    if (fp64)
      out[q] = ((double)(long long)hi[q] * 9.313225746154785e-10 +
                (double)(long long)lo[q] * 8.673617379884035e-19) * scale;
    else
      out[q] = (double)(long long)hi[q] * scale;
  }
}

__global__ __launch_bounds__(PIC_WG) void k_op_validate_keys(const int* keys, int64_t n,
                                                             int ncells, int* flag) {
  bool bad = false;
  for (int64_t p = blockIdx.x * (int64_t)PIC_WG + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * PIC_WG)
    if (keys[p] < 0 || keys[p] >= ncells) bad = true;
  if (bad) atomicOr(flag, 1);
}

__global__ void k_reduce_sum(const double* parts, int n, double* out) {
  __shared__ double red[4];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += PIC_WG) v += parts[i];
  v = block_sum(v, red);
  if (threadIdx.x == 0) *out = v;
}

inline int grid_for(long long n) {
  long long b = (n + PIC_WG - 1) / PIC_WG;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

}  // namespace

// ====================================================================== launchers
void launch_fill_i32(hipStream_t s, int* p, long long n, int v) {
  if (n > 0) k_fill_i32<<<grid_for(n), PIC_WG, 0, s>>>(p, n, v);
}
void launch_fill_u64(hipStream_t s, unsigned long long* p, long long n) {
  if (n > 0) k_fill_u64<<<grid_for(n), PIC_WG, 0, s>>>(p, n);
}
template <typename R>
void launch_op_deposit(hipStream_t s, const R* x, const R* y, int64_t n, int nx, int ny,
                       double lx, double ly, unsigned long long* hi, unsigned long long* lo) {
  k_op_deposit<R><<<grid_for(n), PIC_WG, 0, s>>>(x, y, n, nx, ny, lx, ly, hi, lo);
}
template void launch_op_deposit<float>(hipStream_t, const float*, const float*, int64_t, int, int,
                                       double, double, unsigned long long*, unsigned long long*);
template void launch_op_deposit<double>(hipStream_t, const double*, const double*, int64_t, int,
                                        int, double, double, unsigned long long*,
                                        unsigned long long*);
void launch_op_fixed_to_double(hipStream_t s, const unsigned long long* hi,
                               const unsigned long long* lo, long long G, int fp64, double scale,
                               double* out) {
  k_op_fixed_to_double<<<grid_for(G), PIC_WG, 0, s>>>(hi, lo, G, fp64, scale, out);
}
void launch_op_validate_keys(hipStream_t s, const int* keys, int64_t n, int ncells, int* flag) {
  k_op_validate_keys<<<grid_for(n), PIC_WG, 0, s>>>(keys, n, ncells, flag);
}
void launch_reduce_sum(hipStream_t s, const double* parts, int n, double* out) {
  k_reduce_sum<<<1, PIC_WG, 0, s>>>(parts, n, out);
}

static void sort_tail(hipStream_t s, const RunDesc* runs, RunState* st, const int* id,
                      SortBufs sb, const int2* citems, int ncitems) {
  k_segsort_small<<<sb.nsegitems, PIC_WG, 0, s>>>(runs, sb.offsets, sb.perm, id, sb.big_list,
                                                  sb.big_count, sb.segitems);
  k_segsort_big<<<256, PIC_WG, 0, s>>>(runs, sb.offsets, sb.perm, id, sb.big_list, sb.big_count,
                                       sb.scratch);
}

// clear: zero the counts once scanned; toff: also write the tile offsets (nullptr: not)
static void sort_scan(hipStream_t s, const RunDesc* runs, RunState* st, int nruns, SortBufs sb,
                      const int2* citems, int ncitems, int clear, int* toff) {
  k_scan_local<<<ncitems, PIC_WG, 0, s>>>(runs, sb.counts, sb.offsets, sb.chunk_sums, citems,
                                          clear);
  k_scan_add<<<ncitems, PIC_WG, 0, s>>>(runs, st, sb.offsets, sb.chunk_sums, citems, toff,
                                        sb.big_count);
}

template <typename R>
void launch_sort(hipStream_t s, const RunDesc* runs, RunState* st, int nruns, PartBufs<R> src,
                 PartBufs<R> dst, SortBufs sb, const int2* pitems, int npitems,
                 const int2* citems, int ncitems, long long ctotal, bool move_data) {
  // sb.counts is zero here: zeroed at allocation and by the scan of every sort since
  (void)ctotal;
  k_sort_rank<R, true><<<npitems * PIC_SORT_SLICES, PIC_WG, 0, s>>>(runs, st, src.x, src.y,
                                                                   nullptr, sb.keys,
                                                 sb.rank, sb.counts, pitems);
  sort_scan(s, runs, st, nruns, sb, citems, ncitems, 1, sb.toff);
  k_sort_scatter<<<npitems * PIC_SORT_SLICES, PIC_WG, 0, s>>>(runs, st, sb.keys, sb.rank,
                                                              sb.offsets, sb.perm, pitems);
  sort_tail(s, runs, st, src.id, sb, citems, ncitems);
  if (move_data)
    k_sort_apply<R><<<npitems * PIC_SORT_SLICES, PIC_WG, 0, s>>>(runs, st, sb.perm, src, dst,
                                                                 pitems);
}

template <typename R>
void launch_tile_sort(hipStream_t s, const RunDesc* runs, RunState* st, int nruns,
                      PartBufs<R> src, PartBufs<R> dst, SortBufs sb, const int2* pitems,
                      int npitems) {
  const size_t lds = sizeof(int) * (size_t)sb.max_buckets;
  k_tile_hist<R><<<npitems, PIC_WG, lds, s>>>(runs, st, src.x, src.y, src.vx, src.vy, sb.table,
                                               pitems);
  k_tscan_local<<<sb.ntscan_items, PIC_WG, 0, s>>>(runs, sb.table, sb.tchunk_sums,
                                                   sb.tscan_items);
  k_tscan_chunks<<<nruns, PIC_WG, 0, s>>>(runs, sb.tchunk_sums);
  k_tscan_add<<<sb.ntscan_items, PIC_WG, 0, s>>>(runs, sb.table, sb.tchunk_sums,
                                                 sb.tscan_items);
  k_tile_scatter<R><<<npitems, 64, lds, s>>>(runs, st, src, dst, sb.table, sb.toff, pitems);
}

void launch_fused_scan(hipStream_t s, const RunDesc* runs, RunState* st, int nruns,
                       SortBufs sb, int* toff_new) {
  k_fscan_part<<<sb.nfscan_items, PIC_WG, 0, s>>>(runs, sb.ftable, sb.fchunk_sums,
                                                  sb.max_buckets, sb.fscan_items);
  k_fscan_base<<<nruns, PIC_WG, 0, s>>>(runs, st, sb.fchunk_sums, sb.max_buckets, toff_new);
  k_fscan_rows<<<sb.nfscan_items, PIC_WG, 0, s>>>(runs, sb.ftable, sb.fchunk_sums,
                                                  sb.max_buckets, sb.fscan_items);
}

void launch_sort_keys_only(hipStream_t s, const RunDesc* runs, RunState* st, const int* keys,
                           SortBufs sb, const int2* pitems, int npitems, const int2* citems,
                           int ncitems, long long ctotal) {
  launch_fill_i32(s, sb.counts, ctotal, 0);
  k_sort_rank<float, false><<<npitems * PIC_SORT_SLICES, PIC_WG, 0, s>>>(runs, st, nullptr,
                                                                        nullptr, keys,
                                                      sb.keys, sb.rank, sb.counts, pitems);
  sort_scan(s, runs, st, 1, sb, citems, ncitems, 0, nullptr);
  k_sort_scatter<<<npitems * PIC_SORT_SLICES, PIC_WG, 0, s>>>(runs, st, keys, sb.rank, sb.offsets,
                                                              sb.perm, pitems);
  sort_tail(s, runs, st, nullptr, sb, citems, ncitems);
}

// End of a step, one workgroup per run: the diagnostics row, then the step counter.
// Kernel order makes the partials visible (no grid-wide fences in the producers).
__global__ __launch_bounds__(PIC_WG) void k_step_finish(const RunDesc* __restrict__ runs,
                                                        RunState* __restrict__ st,
                                                        const double* __restrict__ ppush,
                                                        const double* __restrict__ prows,
                                                        double* __restrict__ diag) {
  step_finish_run(runs, st, ppush, prows, diag, blockIdx.x, true);
}

void launch_step_finish(hipStream_t s, const RunDesc* runs, RunState* st, int nruns,
                        const double* ppush, const double* prows, double* diag) {
  k_step_finish<<<nruns, PIC_WG, 0, s>>>(runs, st, ppush, prows, diag);
}

template <typename R>
void launch_push(hipStream_t s, const PushArgs<R>& a) {
  if (a.nitems <= 0) return;
  if (a.mode != 0)
    k_push<R, 1, 0><<<a.nitems, PUSH_NT, a.lds_bytes, s>>>(a);
  else if (a.sort_mode == 1)
    k_push<R, 0, 1><<<a.nitems, PUSH_NT, a.lds_bytes, s>>>(a);
  else if (a.sort_mode == 2)
    k_push<R, 0, 2><<<a.nitems, PUSH_NT, a.lds_bytes, s>>>(a);
  else
    k_push<R, 0, 0><<<a.nitems, PUSH_NT, a.lds_bytes, s>>>(a);
}

template <typename R>
void launch_half_kick_out(hipStream_t s, const RunDesc* runs, RunState* st, int run,
                          PartBufs<R> p, const R* E, R* out_vx, R* out_vy, const int2* pitems,
                          int npitems) {
  k_half_kick_out<R><<<npitems, PIC_WG, 0, s>>>(runs, st, run, p, E, out_vx, out_vy, pitems);
}

template <typename R>
void launch_unpermute(hipStream_t s, const RunDesc* runs, RunState* st, int run, PartBufs<R> p,
                      const R* vxs, const R* vys, R* ox, R* oy, R* ovx, R* ovy,
                      const int2* pitems, int npitems) {
  k_unpermute<R><<<npitems, PIC_WG, 0, s>>>(runs, st, run, p, vxs, vys, ox, oy, ovx, ovy, pitems);
}

template <typename R>
void launch_check_finite(hipStream_t s, const RunDesc* runs, RunState* st, PartBufs<R> p,
                         const int2* pitems, int npitems, int* flag) {
  k_check_finite<R><<<npitems, PIC_WG, 0, s>>>(runs, st, p, pitems, flag);
}

template <typename R>
void launch_op_cell_keys(hipStream_t s, const R* x, const R* y, int64_t n, int nx, int ny,
                         double lx, double ly, int* keys) {
  k_op_cell_keys<R><<<grid_for(n), PIC_WG, 0, s>>>(x, y, n, nx, ny, lx / nx, ly / ny, keys);
}

template <typename R>
void launch_op_gather_push(hipStream_t s, R* x, R* y, R* vx, R* vy, int64_t n, const double* ex,
                           const double* ey, int nx, int ny, double lx, double ly, double dt,
                           double qm, double c, double* ke_partials, int nblocks) {
  k_op_gather_push<R><<<nblocks, PIC_WG, 0, s>>>(x, y, vx, vy, n, ex, ey, nx, ny, lx, ly, dt, qm,
                                                 c, ke_partials);
}

#define PIC_INST(R)                                                                          \
  template void launch_sort<R>(hipStream_t, const RunDesc*, RunState*, int, PartBufs<R>,    \
                               PartBufs<R>, SortBufs, const int2*, int, const int2*, int,   \
                               long long, bool);                                            \
  template void launch_tile_sort<R>(hipStream_t, const RunDesc*, RunState*, int, PartBufs<R>,\
                                    PartBufs<R>, SortBufs, const int2*, int);                 \
  template void launch_push<R>(hipStream_t, const PushArgs<R>&);                            \
  template void launch_half_kick_out<R>(hipStream_t, const RunDesc*, RunState*, int,        \
                                        PartBufs<R>, const R*, R*, R*, const int2*, int);   \
  template void launch_unpermute<R>(hipStream_t, const RunDesc*, RunState*, int,            \
                                    PartBufs<R>, const R*, const R*, R*, R*, R*, R*,        \
                                    const int2*, int);                                      \
  template void launch_check_finite<R>(hipStream_t, const RunDesc*, RunState*, PartBufs<R>, \
                                       const int2*, int, int*);                             \
  template void launch_op_cell_keys<R>(hipStream_t, const R*, const R*, int64_t, int, int,  \
                                       double, double, int*);                               \
  template void launch_op_gather_push<R>(hipStream_t, R*, R*, R*, R*, int64_t,              \
                                         const double*, const double*, int, int, double,    \
                                         double, double, double, double, double*, int);
PIC_INST(float)
PIC_INST(double)
