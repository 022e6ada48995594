// k_refcollide.hip -- the reference's own collide arithmetic on an injected uniform stream.
//
// The reference's host collisions draw from one std::mt19937 in particle order:
//   MCC  thesisCode/mccCollisions/src/particle.cc:112-127: r; if prob > r six more draws
//        give v = (2(u1 + u2 + u3 - 1.5), 2(u4 + u5 + u6 - 1.5));
//   DSMC thesisCode/dsmcCollisions/cellTrackerAdvanced/src/main.cc:288-350: the same draws
//        for every particle whose nearest cell partner has a larger index, and the partner
//        gets the opposite velocity.
// Consumer j (a particle that draws) starts at stream position s_j: s_0 = 0 and
// s_{j+1} = s_j + (prob > u[s_j] ? 7 : 1), a sequential chain.  It is resolved in parallel
// by chunks of the stream: a consumer's start lies at most 6 positions past the end of the
// previous chunk, so each chunk's walk is tabulated for the 7 possible entry offsets
// (k_chain_maps), one thread composes the tables in chunk order (k_chain_scan), and each
// chunk then walks again from its known entry and writes its consumers' starts
// (k_chain_starts).
#include "launch.h"

namespace {

__global__ __launch_bounds__(PIC_WG) void k_chain_maps(const double* __restrict__ u, long long nu,
                                                       double prob, long long nchunks,
                                                       int2* __restrict__ maps) {
  const long long t = blockIdx.x * (long long)PIC_WG + threadIdx.x;
  if (t >= nchunks * 7) return;
  const long long j = t / 7;
  const int e = (int)(t - 7 * j);
  const long long next = (j + 1) * REF_CHAIN_CHUNK;
  const long long end = next < nu ? next : nu;
  long long pos = j * REF_CHAIN_CHUNK + e;
  int cnt = 0;
  while (pos < end) {
    ++cnt;
    pos += (prob > u[pos]) ? 7 : 1;
  }
  maps[t] = make_int2((int)(pos - next), cnt);  // exit offset into chunk j + 1 (0..6)
}

// one thread: entry offset and first consumer of every chunk, and the consumers the
// stream starts in total (the loads of chunk j + 1 do not depend on the entry chosen in
// chunk j, so they are issued one chunk ahead)
__global__ void k_chain_scan(const int2* __restrict__ maps, long long nchunks,
                             int* __restrict__ entry, long long* __restrict__ base,
                             long long* __restrict__ total) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int e = 0;
  long long b = 0;
  int2 cur[7], nxt[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) cur[k] = nchunks > 0 ? maps[k] : make_int2(0, 0);
  for (long long j = 0; j < nchunks; ++j) {
#pragma unroll
    for (int k = 0; k < 7; ++k) nxt[k] = (j + 1 < nchunks) ? maps[(j + 1) * 7 + k] : make_int2(0, 0);
    entry[j] = e;
    base[j] = b;
    int2 m = cur[0];
#pragma unroll
    for (int k = 1; k < 7; ++k)
      if (e == k) m = cur[k];
    b += m.y;
    e = m.x;
#pragma unroll
    for (int k = 0; k < 7; ++k) cur[k] = nxt[k];
  }
  *total = b;
}

__global__ __launch_bounds__(PIC_WG) void k_chain_starts(const double* __restrict__ u,
                                                         long long nu, double prob,
                                                         long long nchunks,
                                                         const int* __restrict__ entry,
                                                         const long long* __restrict__ base,
                                                         long long m,
                                                         long long* __restrict__ start) {
  const long long j = blockIdx.x * (long long)PIC_WG + threadIdx.x;
  if (j >= nchunks) return;
  const long long next = (j + 1) * REF_CHAIN_CHUNK;
  const long long end = next < nu ? next : nu;
  long long pos = j * REF_CHAIN_CHUNK + entry[j], idx = base[j];
  while (pos < end && idx < m) {
    start[idx++] = pos;
    pos += (prob > u[pos]) ? 7 : 1;
  }
}

// uniforms taken by the first m consumers (last start + its draws), or -1 when the stream
// cannot serve them all
__global__ void k_chain_consumed(const double* __restrict__ u, long long nu, double prob,
                                 const long long* __restrict__ start,
                                 const long long* __restrict__ total, long long m,
                                 long long* __restrict__ consumed) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (m == 0) { *consumed = 0; return; }
  if (*total < m) { *consumed = -1; return; }
  const long long s = start[m - 1];
  const long long e = s + ((prob > u[s]) ? 7 : 1);
  *consumed = e <= nu ? e : -1;
}

// Bird's approximate unit normal of the reference, 2*(U1 + U2 + U3 - 1.5), left to right
__device__ __forceinline__ double bird_normal(const double* u) {
  return 2.0 * (u[0] + u[1] + u[2] - 1.5);
}

__global__ __launch_bounds__(PIC_WG) void k_mcc_reference(double* __restrict__ vx,
                                                          double* __restrict__ vy, long long n,
                                                          double prob,
                                                          const double* __restrict__ u,
                                                          const long long* __restrict__ start) {
  for (long long i = blockIdx.x * (long long)PIC_WG + threadIdx.x; i < n;
       i += (long long)gridDim.x * PIC_WG) {
    const long long s = start[i];
    if (prob > u[s]) {
      vx[i] = bird_normal(u + s + 1);
      vy[i] = bird_normal(u + s + 4);
    }
  }
}

__global__ __launch_bounds__(PIC_WG) void k_ref_cell_keys(const double* __restrict__ x,
                                                          const double* __restrict__ y,
                                                          long long n, int nx, int ny,
                                                          double dx, double dy,
                                                          int* __restrict__ keys,
                                                          int* __restrict__ bad) {
  for (long long i = blockIdx.x * (long long)PIC_WG + threadIdx.x; i < n;
       i += (long long)gridDim.x * PIC_WG) {
    const double tx = x[i] / dx, ty = y[i] / dy;
    // (int)(x/dx) as the reference's Particle keeps it; outside the grid is an input error
    if (!(tx >= 0.0 && tx < (double)nx && ty >= 0.0 && ty < (double)ny)) {
      atomicOr(bad, 1);
      keys[i] = 0;
      continue;
    }
    keys[i] = (int)tx * ny + (int)ty;
  }
}

// Nearest other particle of the cell (cell lists in ascending particle index), strictly
// smaller distance wins (the first found on ties), initial bound 1e8; kept only when its
// index is larger, else -1.  valid[i] = 0 for a drawing particle, 1 otherwise (sort key).
__global__ __launch_bounds__(PIC_WG) void k_ref_nearest(const double* __restrict__ x,
                                                        const double* __restrict__ y,
                                                        long long n,
                                                        const int* __restrict__ keys,
                                                        const int* __restrict__ perm,
                                                        const int* __restrict__ offsets,
                                                        int* __restrict__ target,
                                                        int* __restrict__ drawkey) {
  for (long long i = blockIdx.x * (long long)PIC_WG + threadIdx.x; i < n;
       i += (long long)gridDim.x * PIC_WG) {
    const int c = keys[i];
    const double xi = x[i], yi = y[i];
    double nearest = 1e8;
    int t = -1;
    for (int s = offsets[c]; s < offsets[c + 1]; ++s) {
      const int q = perm[s];
      if (q == (int)i) continue;
      const double ddx = xi - x[q], ddy = yi - y[q];
      const double d = sqrt(ddx * ddx + ddy * ddy);
      if (d < nearest) {
        nearest = d;
        t = q;
      }
    }
    if (t < (int)i) t = -1;
    target[i] = t;
    drawkey[i] = t >= 0 ? 0 : 1;
  }
}

// consumer j = the j-th drawing particle: draw, and record the pair's reset velocity
__global__ __launch_bounds__(PIC_WG) void k_dsmc_reference_draw(
    long long m, double prob, const double* __restrict__ u, const long long* __restrict__ start,
    const int* __restrict__ consumers, const int* __restrict__ target, double* __restrict__ m1,
    double* __restrict__ m2, int* __restrict__ self, int* __restrict__ lastw,
    unsigned long long* __restrict__ ncoll) {
  unsigned int acc = 0;
  for (long long j = blockIdx.x * (long long)PIC_WG + threadIdx.x; j < m;
       j += (long long)gridDim.x * PIC_WG) {
    const long long s = start[j];
    if (prob > u[s]) {
      const int i = consumers[j];
      m1[j] = bird_normal(u + s + 1);
      m2[j] = bird_normal(u + s + 4);
      self[i] = (int)j;
      atomicMax(&lastw[target[i]], (int)j);
      ++acc;
    }
  }
  if (acc) atomicAdd(ncoll, (unsigned long long)acc);
}

// The reference writes v_i and then v_target for i = 0, 1, ...: a particle's last write is
// its own draw when it has one (its partners all have smaller indices), else the draw of
// the largest index that chose it as partner (opposite velocity).
__global__ __launch_bounds__(PIC_WG) void k_dsmc_reference_apply(
    double* __restrict__ vx, double* __restrict__ vy, long long n, const double* __restrict__ m1,
    const double* __restrict__ m2, const int* __restrict__ self, const int* __restrict__ lastw) {
  for (long long p = blockIdx.x * (long long)PIC_WG + threadIdx.x; p < n;
       p += (long long)gridDim.x * PIC_WG) {
    const int j = self[p];
    if (j >= 0) {
      vx[p] = m1[j];
      vy[p] = m2[j];
    } else if (lastw[p] >= 0) {
      vx[p] = -m1[lastw[p]];
      vy[p] = -m2[lastw[p]];
    }
  }
}

inline int grid_of(long long n) {
  long long b = (n + PIC_WG - 1) / PIC_WG;
  return (int)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

}  // namespace

void launch_ref_chain(hipStream_t s, const double* u, long long nu, double prob, long long m,
                      RefChainBufs b) {
  const long long nchunks = (nu + REF_CHAIN_CHUNK - 1) / REF_CHAIN_CHUNK;
  if (nchunks > 0) {
    k_chain_maps<<<(int)((nchunks * 7 + PIC_WG - 1) / PIC_WG), PIC_WG, 0, s>>>(u, nu, prob,
                                                                               nchunks, b.maps);
  }
  k_chain_scan<<<1, 64, 0, s>>>(b.maps, nchunks, b.entry, b.base, b.total);
  if (nchunks > 0) {
    k_chain_starts<<<(int)((nchunks + PIC_WG - 1) / PIC_WG), PIC_WG, 0, s>>>(
        u, nu, prob, nchunks, b.entry, b.base, m, b.start);
  }
  k_chain_consumed<<<1, 64, 0, s>>>(u, nu, prob, b.start, b.total, m, b.consumed);
}

void launch_mcc_reference(hipStream_t s, double* vx, double* vy, long long n, double prob,
                          const double* u, const long long* start) {
  if (n > 0) k_mcc_reference<<<grid_of(n), PIC_WG, 0, s>>>(vx, vy, n, prob, u, start);
}

void launch_ref_cell_keys(hipStream_t s, const double* x, const double* y, long long n, int nx,
                          int ny, double dx, double dy, int* keys, int* bad) {
  if (n > 0) k_ref_cell_keys<<<grid_of(n), PIC_WG, 0, s>>>(x, y, n, nx, ny, dx, dy, keys, bad);
}

void launch_ref_nearest(hipStream_t s, const double* x, const double* y, long long n,
                        const int* keys, const int* perm, const int* offsets, int* target,
                        int* drawkey) {
  if (n > 0)
    k_ref_nearest<<<grid_of(n), PIC_WG, 0, s>>>(x, y, n, keys, perm, offsets, target, drawkey);
}

void launch_dsmc_reference(hipStream_t s, double* vx, double* vy, long long n, long long m,
                           double prob, const double* u, const long long* start,
                           const int* consumers, const int* target, double* m1, double* m2,
                           int* self, int* lastw, unsigned long long* ncoll) {
  if (m > 0)
    k_dsmc_reference_draw<<<grid_of(m), PIC_WG, 0, s>>>(m, prob, u, start, consumers, target, m1,
                                                        m2, self, lastw, ncoll);
  if (n > 0) k_dsmc_reference_apply<<<grid_of(n), PIC_WG, 0, s>>>(vx, vy, n, m1, m2, self, lastw);
}
