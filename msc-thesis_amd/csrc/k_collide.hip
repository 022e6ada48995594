// k_collide.hip -- collision stages on FP32 particles, driven by stateless Philox streams.
//
// MCC occupies the slot of thesisCode/mccCollisions/src/pic.cu:435-463 (collision after the
// push) with electron-neutral null-collision physics: analytic cross sections (elastic
// 1e-19/(1+E/10) m^2, excitation 2e-21(1-11.5/E) above 11.5 eV, ionisation
// 3e-20 ln(E/15.8)/E above 15.8 eV), isotropic scattering in the velocity plane, births
// appended after a deterministic ordering by parent id.
// DSMC occupies the slot of thesisCode/dsmcCollisions/cellTrackerAdvanced/src/pic.cu:791-864
// with NTC hard-sphere like-particle collisions on the cell-sorted arrays, candidates in
// rounds of disjoint pairs.
// Random streams: counter {particle id | cell, step, draw, tag}, key = seed.
#include "launch.h"

namespace {

__device__ __forceinline__ void cross_sections(double e, double& sel, double& sex, double& siz) {
  sel = 1e-19 / (1.0 + e / 10.0);
  sex = (e > 11.5) ? 2e-21 * (1.0 - 11.5 / e) : 0.0;
  siz = (e > 15.8) ? 3e-20 * log_own(e / 15.8) / e : 0.0;
}

// Null-collision test, every particle: word 0 of Philox2x32-10 with counter (id, step) and
// key seed_lo ^ seed_hi ^ 'MCC1'; the particles whose uniform falls below p_coll = 1 -
// exp(-nu_max dt) (about 1 in 100) go to the step's collision list.  (A Philox4x32-10 block
// per particle, of which the test used one word, made the kernel VALU-bound at ~100
// instructions per particle.)  Workgroup b appends its hits to its own region
// [b * per_block, (b + 1) * per_block) of the list through an LDS counter (one LDS atomic per
// wave; one global counter for all of them serialised ~60k atomics per step on a single
// address) and stores its count.  The collisions run on the list in a second launch, so the
// waves do not carry the ~1 % of lanes through the long FP64 collision branch while the
// other 63 wait.
#define MCC_NULL_U 8
__global__ __launch_bounds__(PIC_WG) void k_mcc_null(MccArgs a, int n_upper, int per_block) {
  __shared__ int s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  const int np = min(a.st->np, n_upper);
  const long long step = a.from_state_step ? a.st->step - (a.finish_after ? 0 : 1) : a.step;
  const uint32_t s32 = (uint32_t)step;
  const uint32_t nk = (uint32_t)a.seed ^ (uint32_t)(a.seed >> 32) ^ PIC_TAG_MCC;
  const int lane = threadIdx.x & 63;
  int* list = a.coll_list + (long long)blockIdx.x * per_block;
  // MCC_NULL_U ids per thread loaded together (the last id for lanes past the end: always in
  // bounds), then tested in grid-stride order: one load per test exposed the load latency
  // at every step of the loop.
  const int stride = gridDim.x * PIC_WG;
  const int* __restrict__ ids = a.id + a.pbase;
  for (int q0 = blockIdx.x * PIC_WG + threadIdx.x; q0 < np; q0 += MCC_NULL_U * stride) {
    uint32_t idv[MCC_NULL_U];
#pragma unroll
    for (int j = 0; j < MCC_NULL_U; ++j) idv[j] = (uint32_t)ids[min(q0 + j * stride, np - 1)];
#pragma unroll
    for (int j = 0; j < MCC_NULL_U; ++j) {
      const int q = q0 + j * stride;
      uint32_t ctr[2] = {idv[j], s32};
      philox2x32_10(ctr, nk);
      const bool hit = q < np && u01(ctr[0]) < a.p_coll;
      const unsigned long long mask = __ballot(hit);
      if (mask) {
        const int leader = __builtin_ctzll(mask);
        int base = 0;
        if (lane == leader) base = atomicAdd(&s_cnt, (int)__popcll(mask));
        base = __shfl(base, leader, 64);
        if (hit) list[base + (int)__popcll(mask & ((1ull << lane) - 1ull))] = q;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) a.coll_count[blockIdx.x] = s_cnt;
}

// The collisions of the step's list (workgroup b takes region b of the null test): a
// Philox4x32-10 block keyed by (id, step), the process chosen by u1 nu_max dt against
// the cumulative nu_k(E) dt, isotropic scattering in the velocity plane; an ionisation
// appends a birth (ordered by parent id afterwards).
__global__ __launch_bounds__(PIC_WG) void k_mcc(MccArgs a, int per_block) {
  __shared__ int cnt[4];
  if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int ncand = a.coll_count[blockIdx.x];
  const int* list = a.coll_list + (long long)blockIdx.x * per_block;
  const long long step = a.from_state_step ? a.st->step - (a.finish_after ? 0 : 1) : a.step;
  const uint32_t s32 = (uint32_t)step;
  for (int k = threadIdx.x; k < ncand; k += PIC_WG) {
    const long long i = a.pbase + list[k];
    const uint32_t pid = (uint32_t)a.id[i];
    double u[4];
    draw4(pid, s32, 0u, PIC_TAG_MCC, a.seed, u);
    double vx = (double)a.vx[i], vy = (double)a.vy[i];
    const double v2 = vx * vx + vy * vy;
    const double sp = sqrt(v2);
    const double e = a.v_unit2_half_me_over_e * v2;
    double sel, sex, siz;
    cross_sections(e, sel, sex, siz);
    const double nel = a.ng_vunit_dtphys * sel * sp;
    const double nex = a.ng_vunit_dtphys * sex * sp;
    const double niz = a.ng_vunit_dtphys * siz * sp;
    const double r = u[1] * a.nu_max_dt;
    double cs, sn;
    sincos2pi_own(u[2], sn, cs);
    if (r < nel) {
      const double f = sqrt(1.0 - a.mass_ratio2 * (1.0 - cs));
      const double nvx = f * (vx * cs - vy * sn), nvy = f * (vx * sn + vy * cs);
      vx = nvx; vy = nvy;
      atomicAdd(&cnt[0], 1);
    } else if (r < nel + nex) {
      const double f = sqrt((e - 11.5) / e);
      const double nvx = f * (vx * cs - vy * sn), nvy = f * (vx * sn + vy * cs);
      vx = nvx; vy = nvy;
      atomicAdd(&cnt[1], 1);
    } else if (r < nel + nex + niz) {
      const double ea = e - 15.8;
      const double e1 = u[3] * ea, e2 = ea - e1;
      const double f1 = sqrt(e1 / e);
      const double nvx = f1 * (vx * cs - vy * sn), nvy = f1 * (vx * sn + vy * cs);
      double w[4];
      draw4(pid, s32, 1u, PIC_TAG_MCC, a.seed, w);
      double cs2, sn2;
      sincos2pi_own(w[0], sn2, cs2);
      const double sp2 = sp * sqrt(e2 / e);
      const int slot = atomicAdd(a.birth_count, 1);
      if (slot < a.cap) {
        a.birth_parent[slot] = (int)pid;
        a.birth_buf[4 * (long long)slot + 0] = a.x[i];
        a.birth_buf[4 * (long long)slot + 1] = a.y[i];
        a.birth_buf[4 * (long long)slot + 2] = (float)(sp2 * cs2);
        a.birth_buf[4 * (long long)slot + 3] = (float)(sp2 * sn2);
      }
      vx = nvx; vy = nvy;
      atomicAdd(&cnt[2], 1);
    } else {
      atomicAdd(&cnt[3], 1);
      continue;
    }
    a.vx[i] = (float)vx;
    a.vy[i] = (float)vy;
  }
  __syncthreads();
  // the workgroup's counts; k_mcc_births adds them up (thousands of workgroups adding to the
  // same four counters would serialise on those addresses)
  if (threadIdx.x < 4) a.coll_partials[4 * blockIdx.x + threadIdx.x] = cnt[threadIdx.x];
}

// Births of the step: order by parent id (bitonic, one workgroup), assign ids
// next_id + rank, append after np, deposit their charge into the current parity.  First the
// step's collision counts: the per-workgroup counts of k_mcc added up (into the cumulative
// counters and the diagnostics row of the step).
__global__ __launch_bounds__(PIC_WG) void k_mcc_births(MccArgs a, unsigned long long* scratch,
                                                       unsigned long long* rho,
                                                       unsigned long long* rho_lo_unused,
                                                       long long gsum, const RunDesc* runs,
                                                       int run, int nblocks) {
  __shared__ unsigned long long s[2048];
  __shared__ unsigned long long red[4][4];
  {
    // the k_mcc workgroups' counts, eight per thread loaded together (integer sums: any
    // order gives the same totals)
    unsigned long long c[4] = {0, 0, 0, 0};
    constexpr int BU = 8;
    for (int b0 = threadIdx.x; b0 < nblocks; b0 += BU * PIC_WG) {
      int4 v[BU];
#pragma unroll
      for (int u = 0; u < BU; ++u) {
        const int b = b0 + u * PIC_WG;
        v[u] = reinterpret_cast<const int4*>(a.coll_partials)[b < nblocks ? b : nblocks - 1];
      }
#pragma unroll
      for (int u = 0; u < BU; ++u)
        if (b0 + u * PIC_WG < nblocks) {
          c[0] += (unsigned long long)v[u].x;
          c[1] += (unsigned long long)v[u].y;
          c[2] += (unsigned long long)v[u].z;
          c[3] += (unsigned long long)v[u].w;
        }
    }
    for (int k = 0; k < 4; ++k)
      for (int o = 32; o > 0; o >>= 1) c[k] += __shfl_xor(c[k], o, 64);
    if ((threadIdx.x & 63) == 0)
      for (int k = 0; k < 4; ++k) red[k][threadIdx.x >> 6] = c[k];
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t[4];
      for (int k = 0; k < 4; ++k) t[k] = red[k][0] + red[k][1] + red[k][2] + red[k][3];
      if (a.counts)
        for (int k = 0; k < 4; ++k) a.counts[k] += (long long)t[k];
      if (a.diag) {
        const long long step = a.from_state_step ? a.st->step - (a.finish_after ? 0 : 1) : a.step;
        a.diag[(step % PIC_DIAG_CAP) * 8 + 6] += (double)(t[0] + t[1] + t[2]);
      }
    }
  }
  const int nb = min(*a.birth_count, a.cap);
  if (nb == 0) return;
  int m = 1;
  while (m < nb) m <<= 1;
  unsigned long long* k = (m <= 2048) ? s : scratch;
  for (int t = threadIdx.x; t < m; t += PIC_WG)
    k[t] = (t < nb) ? (((unsigned long long)(unsigned)a.birth_parent[t] << 32) | (unsigned)t)
                    : ~0ull;
  __syncthreads();
  for (int kk = 2; kk <= m; kk <<= 1)
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < m; t += PIC_WG) {
        const int l = t ^ j;
        if (l > t) {
          const unsigned long long u = k[t], v = k[l];
          const bool up = (t & kk) == 0;
          if ((u > v) == up) { k[t] = v; k[l] = u; }
        }
      }
      __syncthreads();
    }
  const RunDesc rd = runs[run];
  const int np = a.st->np;
  const long long nid = a.st->next_id;
  const int room = max(0, a.cap - np);
  const int take = min(nb, room);
  const int par = (int)((a.st->step + (a.finish_after ? 1 : 0)) & 1);  // the next step's charge
  unsigned long long* rh = rho + par * gsum + rd.gbase;
  const float idx = (float)(1.0 / rd.dx), idy = (float)(1.0 / rd.dy);
  for (int t = threadIdx.x; t < take; t += PIC_WG) {
    const int src = (int)(k[t] & 0xffffffffull);
    const long long d = a.pbase + np + t;
    const float bx = a.birth_buf[4 * (long long)src + 0], by = a.birth_buf[4 * (long long)src + 1];
    a.x[d] = bx;
    a.y[d] = by;
    a.vx[d] = a.birth_buf[4 * (long long)src + 2];
    a.vy[d] = a.birth_buf[4 * (long long)src + 3];
    a.id[d] = (int)(nid + t);
    int cx, cy;
    float fx, fy;
    locate_w(bx, idx, rd.nx, cx, fx);
    locate_w(by, idy, rd.ny, cy, fy);
    const long long one = 1ll << 18;
    const long long ix1 = qweight(fx), iy1 = qweight(fy), ix0 = one - ix1, iy0 = one - iy1;
    const int cxp = (cx + 1 == rd.nx) ? 0 : cx + 1, cyp = (cy + 1 == rd.ny) ? 0 : cy + 1;
    atomicAdd(&rh[(long long)cx * rd.ny + cy], (unsigned long long)(ix0 * iy0));
    atomicAdd(&rh[(long long)cxp * rd.ny + cy], (unsigned long long)(ix1 * iy0));
    atomicAdd(&rh[(long long)cx * rd.ny + cyp], (unsigned long long)(ix0 * iy1));
    atomicAdd(&rh[(long long)cxp * rd.ny + cyp], (unsigned long long)(ix1 * iy1));
    if (a.count_table) {
      // this step counted the tiles for the sort that the next step fuses into its push:
      // the birth (index np + t, pushed by tail item (np + t) / PIC_TAIL_CHUNK, which follows
      // the run's tile items) is counted there, by the tile of its position (collision runs
      // key on positions)
      const int sh = __builtin_ctz((unsigned)rd.tile);
      const int key = (cx >> sh) * rd.tiles_y + (cy >> sh);
      const int ntail = (rd.pcap + PIC_TAIL_CHUNK - 1) / PIC_TAIL_CHUNK;
      const long long row =
          rd.ftbase + (long long)(rd.fchunks - ntail + (np + t) / PIC_TAIL_CHUNK) * rd.nbuckets;
      atomicAdd(&a.count_table[row + key], 1);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a.st->np = np + take;
    a.st->next_id = nid + take;
    if (take < nb) a.st->err |= 1;
    *a.birth_count = 0;
  }
}


// Pair selection of a DSMC round: slot t of [0, n) through a 12-round balanced Feistel
// network on 2*half bits (the smallest even width with 2^(2 half) >= n; round i keyed by
// key[i % 4] ^ i*0x632BE5AB), repeated until the value is below n (cycle walking: the walk
// stays on the cycle through t, which holds t itself, so it ends).  The pairs
// (pi(2m), pi(2m+1)) of a round are disjoint.
__device__ __forceinline__ uint32_t feistel_f(uint32_t r, uint32_t k, int half) {
  uint32_t h = (r ^ k) * 0x9E3779B1u;
  h ^= h >> 15;
  h *= 0x85EBCA77u;
  h ^= h >> 13;
  return h >> (32 - half);
}

__device__ __forceinline__ uint32_t pair_perm(uint32_t t, uint32_t n, int half,
                                              const uint32_t key[4]) {
  const uint32_t mask = (1u << half) - 1u;
  uint32_t x = t;
  do {
    uint32_t l = x >> half, r = x & mask;
#pragma unroll
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

__device__ __forceinline__ int pair_half_bits(uint32_t n) {
  int b = 32 - __builtin_clz(n - 1u);  // n >= 2
  b += b & 1;
  return b >> 1;
}

__device__ __forceinline__ double wave_max_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Row-major cell cx*ny + cy of cell c of the offsets' numbering: the engine's sorts number
// cells tile-major (tile t = (cx >> sh)*tiles_y + (cy >> sh), then cx, cy inside the tile);
// the operator's offsets are row-major already (tile_shift < 0).  The per-cell maximum
// relative speed and the random streams are keyed by the row-major cell, so a run gives
// the same collisions under any tiling.
__device__ __forceinline__ int dsmc_cell_rm(int c, const DsmcArgs& a) {
  if (a.tile_shift < 0) return c;
  const int sh = a.tile_shift, sh2 = 2 * sh;
  const int t = c >> sh2, w = c & ((1 << sh2) - 1);
  const int tx = t / a.tiles_y, ty = t - tx * a.tiles_y;
  const int cx = (tx << sh) + (w >> sh), cy = (ty << sh) + (w & ((1 << sh) - 1));
  return cx * a.ny + cy;
}

// One candidate: the pair of slots (pi(2m), pi(2m+1)) of the cell's round permutation,
// accepted when u gm < g with gm = max(gmax at the round's start, g); an accepted pair keeps
// its centre-of-mass velocity and turns its relative velocity to a uniform direction.
// Returns g (for the round's new gmax); counts acceptances in acc.
__device__ __forceinline__ double dsmc_candidate(const DsmcArgs& a, int b, uint32_t n,
                                                 int half, const uint32_t key[4], int cr,
                                                 uint32_t s32, long long k, uint32_t m,
                                                 double gm, unsigned int& acc) {
  const long long pi = b + (long long)pair_perm(2u * m, n, half, key);
  const long long pj = b + (long long)pair_perm(2u * m + 1u, n, half, key);
  double w[4];
  draw4((uint32_t)cr, s32, (uint32_t)k, PIC_TAG_DSMC, a.seed, w);
  const double vix = a.vx[pi], viy = a.vy[pi], vjx = a.vx[pj], vjy = a.vy[pj];
  const double gx = vix - vjx, gy = viy - vjy;
  const double g = sqrt(gx * gx + gy * gy);
  const double ge = g > gm ? g : gm;
  if (w[0] * ge < g) {
    double sn, cs;
    sincos2pi_own(w[1], sn, cs);
    const double cx = 0.5 * (vix + vjx), cy = 0.5 * (viy + vjy);
    const double hx = 0.5 * g * cs, hy = 0.5 * g * sn;
    a.vx[pi] = (float)(cx + hx);
    a.vy[pi] = (float)(cy + hy);
    a.vx[pj] = (float)(cx - hx);
    a.vy[pj] = (float)(cy - hy);
    ++acc;
  }
  return g;
}

// NTC hard-sphere collisions in rounds of disjoint pairs.  A cell of n particles has
// M = floor(mr) + [u < frac(mr)] candidates, mr = 1/2 n (n-1) coeff gmax, dealt to rounds of
// P = n/2 pairs; the pairs of a round are disjoint, so its candidates are independent.
// After a round gmax = max(gmax, every g of the round) (a maximum: exact, order-free).
//
// One workgroup takes DSMC_CELLS consecutive cells: the per-cell set-up (offsets, gmax,
// the draw of M, the first round's permutation key) runs a thread per cell, then the
// first-round candidates of all its cells are dealt to the workgroup's threads together
// (most cells need one round of a few candidates), and the few cells with more rounds go
// on one at a time with the threads over the round's pairs.  Velocities are read and
// written in HBM directly (only the candidates' particles are touched); a round that
// follows another first waits for its stores (__syncthreads: workgroup-scope fence).
#define DSMC_CELLS 64
#define DSMC_THREADS 256
__global__ __launch_bounds__(DSMC_THREADS) void k_dsmc(DsmcArgs a) {
  __shared__ int s_b[DSMC_CELLS], s_n[DSMC_CELLS], s_cr[DSMC_CELLS], s_half[DSMC_CELLS];
  __shared__ int s_first[DSMC_CELLS + 1];
  __shared__ long long s_M[DSMC_CELLS];
  __shared__ uint32_t s_key[DSMC_CELLS][4];
  __shared__ double s_gm[DSMC_CELLS];
  __shared__ unsigned long long s_gr[DSMC_CELLS];  // bits of the largest g (g >= 0: ordered)
  __shared__ unsigned long long s_more;
  const int tid = threadIdx.x;
  const long long step = a.st ? a.st->step - (a.finish_after ? 0 : 1) : a.step;
  const uint32_t s32 = (uint32_t)step;
  const uint32_t sk0 = (uint32_t)a.seed, sk1 = (uint32_t)(a.seed >> 32);
  unsigned int acc = 0;
  for (int c0 = blockIdx.x * DSMC_CELLS; c0 < a.ncells; c0 += gridDim.x * DSMC_CELLS) {
    // ---- per-cell set-up, the first wave, a lane per cell
    if (tid < DSMC_CELLS) {
      const int c = c0 + tid;
      int b = 0, n = 0, cr = 0;
      long long M = 0;
      double gm = 0.0;
      if (c < a.ncells) {
        b = a.offsets[c];
        n = a.offsets[c + 1] - b;
        if (n >= 2) {
          cr = dsmc_cell_rm(c, a);
          gm = (double)a.gmax[cr];
          double u[4];
          draw4((uint32_t)cr, s32, 0xffffffffu, PIC_TAG_DSMC, a.seed, u);
          const double mr = 0.5 * (double)n * (double)(n - 1) * a.coeff * gm;
          const double fl = floor(mr);
          M = (long long)fl + ((u[0] < mr - fl) ? 1 : 0);
          if (M > 0x7fffffffll) M = 0x7fffffffll;
        }
      }
      const int P = n >> 1;
      const int r1 = M > 0 ? (int)(M < P ? M : P) : 0;  // candidates of the first round
      int incl = r1;                                    // inclusive scan over the lanes
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (tid >= o) incl += t;
      }
      s_b[tid] = b;
      s_n[tid] = n;
      s_cr[tid] = cr;
      s_M[tid] = M;
      s_gm[tid] = gm;
      s_gr[tid] = (unsigned long long)__double_as_longlong(gm);
      s_first[tid + 1] = incl;
      if (tid == 0) s_first[0] = 0;
      const unsigned long long more = __ballot(M > (long long)P && r1 > 0);
      if (tid == 0) s_more = more;
      if (r1 > 0) {
        uint32_t key[4] = {(uint32_t)cr, s32, 0x80000000u, PIC_TAG_DSMP};
        philox4x32_10(key, sk0, sk1);
        s_key[tid][0] = key[0]; s_key[tid][1] = key[1];
        s_key[tid][2] = key[2]; s_key[tid][3] = key[3];
        s_half[tid] = pair_half_bits((uint32_t)n);
      }
    }
    __syncthreads();
    // ---- first round of every cell, candidates dealt to the threads
    const int total = s_first[DSMC_CELLS];
    for (int t = tid; t < total; t += DSMC_THREADS) {
      int q = 0;  // cell of candidate t: s_first[q] <= t < s_first[q + 1]
#pragma unroll
      for (int w = DSMC_CELLS / 2; w > 0; w >>= 1)
        if (s_first[q + w] <= t) q += w;
      const uint32_t m = (uint32_t)(t - s_first[q]);
      const uint32_t key[4] = {s_key[q][0], s_key[q][1], s_key[q][2], s_key[q][3]};
      const double g = dsmc_candidate(a, s_b[q], (uint32_t)s_n[q], s_half[q], key, s_cr[q], s32,
                                      (long long)m, m, s_gm[q], acc);
      atomicMax(&s_gr[q], (unsigned long long)__double_as_longlong(g));
    }
    __syncthreads();
    // ---- cells with further rounds, one at a time
    unsigned long long more = s_more;
    while (more) {
      const int q = __builtin_ctzll(more);
      more &= more - 1;
      const long long Mq = s_M[q];
      const int nq = s_n[q], Pq = nq >> 1, bq = s_b[q], crq = s_cr[q], hq = s_half[q];
      for (long long r0 = Pq; r0 < Mq; r0 += Pq) {
        // the previous round's stores and gmax are in (__syncthreads above / below)
        const double gq = __longlong_as_double((long long)s_gr[q]);
        uint32_t key[4] = {(uint32_t)crq, s32, 0x80000000u | (uint32_t)(r0 / Pq), PIC_TAG_DSMP};
        philox4x32_10(key, sk0, sk1);
        const int npr = (int)((Mq - r0 < Pq) ? (Mq - r0) : Pq);
        __syncthreads();  // every thread has read s_gr[q] before it moves
        for (int m = tid; m < npr; m += DSMC_THREADS) {
          const double g = dsmc_candidate(a, bq, (uint32_t)nq, hq, key, crq, s32, r0 + m,
                                          (uint32_t)m, gq, acc);
          atomicMax(&s_gr[q], (unsigned long long)__double_as_longlong(g));
        }
        __syncthreads();
      }
    }
    if (tid < DSMC_CELLS && c0 + tid < a.ncells && s_first[tid + 1] > s_first[tid])
      a.gmax[s_cr[tid]] = (float)__longlong_as_double((long long)s_gr[tid]);
    __syncthreads();  // LDS reused by the next group of cells
  }
  // accepted pairs: integer sum, order-free
  unsigned int t = acc;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  if ((tid & 63) == 0 && t) {
    if (a.ncoll) atomicAdd((unsigned long long*)a.ncoll, (unsigned long long)t);
    if (a.diag) atomicAdd(&a.diag[(step % PIC_DIAG_CAP) * 8 + 6], (double)t);
  }
}

}  // namespace

void mcc_list_layout(long long n_upper, int* blocks, int* per_block) {
  long long b = (n_upper + PIC_WG - 1) / PIC_WG;
  b = b < 1 ? 1 : (b > PIC_MCC_BLOCKS ? PIC_MCC_BLOCKS : b);
  const long long per = (n_upper + b * PIC_WG - 1) / (b * PIC_WG) * PIC_WG;  // grid-stride share
  *blocks = (int)b;
  *per_block = (int)(per < PIC_WG ? PIC_WG : per);
}

void launch_mcc(hipStream_t s, const MccArgs& a, int n_upper) {
  int blocks, per_block;
  mcc_list_layout(n_upper, &blocks, &per_block);
  k_mcc_null<<<blocks, PIC_WG, 0, s>>>(a, n_upper, per_block);
  k_mcc<<<blocks, PIC_WG, 0, s>>>(a, per_block);
}

void launch_mcc_births(hipStream_t s, const MccArgs& a, unsigned long long* scratch,
                       unsigned long long* rho, long long gsum, const RunDesc* runs, int run,
                       int n_upper) {
  int blocks, per_block;
  mcc_list_layout(n_upper, &blocks, &per_block);
  k_mcc_births<<<1, PIC_WG, 0, s>>>(a, scratch, rho, nullptr, gsum, runs, run, blocks);
}

void launch_dsmc(hipStream_t s, const DsmcArgs& a) {
  int blocks = (a.ncells + DSMC_CELLS - 1) / DSMC_CELLS;
  blocks = blocks < 1 ? 1 : (blocks > 65536 ? 65536 : blocks);
  k_dsmc<<<blocks, DSMC_THREADS, 0, s>>>(a);
}
