// k_collide.hip -- collision stages on FP32 particles, driven by stateless Philox streams.
//
// MCC occupies the slot of thesisCode/mccCollisions/src/pic.cu:435-463 (collision after the
// push) with electron-neutral null-collision physics: analytic cross sections (elastic
// 1e-19/(1+E/10) m^2, excitation 2e-21(1-11.5/E) above 11.5 eV, ionisation
// 3e-20 ln(E/15.8)/E above 15.8 eV), isotropic scattering in the velocity plane, births
// appended after a deterministic ordering by parent id.
// DSMC occupies the slot of thesisCode/dsmcCollisions/cellTrackerAdvanced/src/pic.cu:791-864
// with NTC hard-sphere like-particle collisions on the cell-sorted arrays.
// Random streams: counter {particle id | cell, step, draw, tag}, key = seed.
#include "launch.h"

namespace {

__device__ __forceinline__ void cross_sections(double e, double& sel, double& sex, double& siz) {
  sel = 1e-19 / (1.0 + e / 10.0);
  sex = (e > 11.5) ? 2e-21 * (1.0 - 11.5 / e) : 0.0;
  siz = (e > 15.8) ? 3e-20 * log_own(e / 15.8) / e : 0.0;
}

__global__ __launch_bounds__(PIC_WG) void k_mcc(MccArgs a, int n_upper) {
  __shared__ int cnt[4];
  if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int np = min(a.st->np, n_upper);
  const long long step = a.from_state_step ? a.st->step - 1 : a.step;
  const uint32_t s32 = (uint32_t)step;
  for (int q = blockIdx.x * PIC_WG + threadIdx.x; q < np; q += gridDim.x * PIC_WG) {
    const long long i = a.pbase + q;
    const uint32_t pid = (uint32_t)a.id[i];
    double u[4];
    draw4(pid, s32, 0u, PIC_TAG_MCC, a.seed, u);
    if (u[0] >= a.p_coll) continue;
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
  if (threadIdx.x < 4 && cnt[threadIdx.x] && a.counts)
    atomicAdd((unsigned long long*)&a.counts[threadIdx.x], (unsigned long long)cnt[threadIdx.x]);
}

// Births of the step: order by parent id (bitonic, one workgroup), assign ids
// next_id + rank, append after np, deposit their charge into the current parity.
__global__ __launch_bounds__(PIC_WG) void k_mcc_births(MccArgs a, unsigned long long* scratch,
                                                       unsigned long long* rho,
                                                       unsigned long long* rho_lo_unused,
                                                       long long gsum, const RunDesc* runs,
                                                       int run) {
  __shared__ unsigned long long s[2048];
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
  const int par = (int)(a.st->step & 1);
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
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a.st->np = np + take;
    a.st->next_id = nid + take;
    if (take < nb) a.st->err |= 1;
    *a.birth_count = 0;
  }
}

#define DSMC_WAVES 4        // waves per workgroup; each takes one cell at a time
#define DSMC_LDS_MAX 1024    // particles of a cell held in LDS (larger cells stay in HBM)

__device__ __forceinline__ double readlane_f64(double x, int t) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffll), t);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), t);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// candidate k of cell c: the pair (i, j), i != j, the acceptance uniform and the direction
// of the post-collision relative velocity
__device__ __forceinline__ void dsmc_candidate(const DsmcArgs& a, int c, uint32_t s32,
                                               long long k, int n, int& i, int& j, double& u2,
                                               double& sn, double& cs) {
  double u[4];
  draw4((uint32_t)c, s32, (uint32_t)k, PIC_TAG_DSMC, a.seed, u);
  i = (int)(u[0] * n);
  i = i > n - 1 ? n - 1 : i;
  j = (int)(u[1] * (n - 1));
  j = j > n - 2 ? n - 2 : j;
  if (j >= i) ++j;
  u2 = u[2];
  sincos2pi_own(u[3], sn, cs);
}

// NTC pass of one cell: candidates in order k = 0 .. M-1; a candidate with relative speed
// g raises the running bound gm to g, and is accepted when u2 gm < g: the pair's centre-
// of-mass velocity is kept and its relative velocity turned to the drawn direction
__device__ __forceinline__ void dsmc_pair(double vix, double viy, double vjx, double vjy,
                                          double& gm, double u2, double sn, double cs,
                                          bool& hit, float2& ni, float2& nj) {
  const double gx = vix - vjx, gy = viy - vjy;
  const double g = sqrt(gx * gx + gy * gy);
  if (g > gm) gm = g;
  hit = u2 * gm < g;
  const double cx = 0.5 * (vix + vjx), cy = 0.5 * (viy + vjy);
  const double hx = 0.5 * g * cs, hy = 0.5 * g * sn;
  ni = make_float2((float)(cx + hx), (float)(cy + hy));
  nj = make_float2((float)(cx - hx), (float)(cy - hy));
}

// One wave per cell.  The cell's velocities are staged in LDS; the wave draws the random
// numbers of 64 candidates at once (one per lane), then steps through them in order with
// wave-uniform values (the pass is sequential: a particle may take part in several pairs),
// so the serial chain per candidate is LDS reads and arithmetic instead of dependent HBM
// round trips.  The candidates, their order and every operation are those of the one-cell
// sequential pass, so the result is bit-identical to it.
__global__ __launch_bounds__(64 * DSMC_WAVES) void k_dsmc(DsmcArgs a) {
  __shared__ float2 vs[DSMC_WAVES][DSMC_LDS_MAX];
  __shared__ unsigned int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long step = a.st ? a.st->step - 1 : a.step;
  const uint32_t s32 = (uint32_t)step;
  float2* sv = vs[wv];
  unsigned int acc = 0;  // the same in every lane
  for (int c = blockIdx.x * DSMC_WAVES + wv; c < a.ncells; c += gridDim.x * DSMC_WAVES) {
    const int b = a.offsets[c], n = a.offsets[c + 1] - b;
    if (n < 2) continue;
    double gm = (double)a.gmax[c];
    double u[4];
    draw4((uint32_t)c, s32, 0xffffffffu, PIC_TAG_DSMC, a.seed, u);
    const double mr = 0.5 * (double)n * (double)(n - 1) * a.coeff * gm;
    const double fl = floor(mr);
    const long long M = (long long)fl + ((u[0] < mr - fl) ? 1 : 0);
    if (n > DSMC_LDS_MAX) {
      // too many particles for the LDS slot: the same pass on the velocities in HBM
      if (lane == 0) {
        for (long long k = 0; k < M; ++k) {
          int i, j;
          double u2, sn, cs;
          dsmc_candidate(a, c, s32, k, n, i, j, u2, sn, cs);
          const long long pi = b + i, pj = b + j;
          bool hit;
          float2 ni, nj;
          dsmc_pair(a.vx[pi], a.vy[pi], a.vx[pj], a.vy[pj], gm, u2, sn, cs, hit, ni, nj);
          if (hit) {
            a.vx[pi] = ni.x; a.vy[pi] = ni.y;
            a.vx[pj] = nj.x; a.vy[pj] = nj.y;
            ++acc;
          }
        }
        a.gmax[c] = (float)gm;
      }
      continue;
    }
    for (int t = lane; t < n; t += 64) sv[t] = make_float2(a.vx[b + t], a.vy[b + t]);
    for (long long k0 = 0; k0 < M; k0 += 64) {
      int ci = 0, cj = 1;
      double cu2 = 0.0, csn = 0.0, ccs = 1.0;
      if (k0 + lane < M) dsmc_candidate(a, c, s32, k0 + lane, n, ci, cj, cu2, csn, ccs);
      const int tn = (M - k0 < 64) ? (int)(M - k0) : 64;
      for (int t = 0; t < tn; ++t) {
        const int i = __builtin_amdgcn_readlane(ci, t), j = __builtin_amdgcn_readlane(cj, t);
        const float2 fi = sv[i], fj = sv[j];
        bool hit;
        float2 ni, nj;
        dsmc_pair(fi.x, fi.y, fj.x, fj.y, gm, readlane_f64(cu2, t), readlane_f64(csn, t),
                  readlane_f64(ccs, t), hit, ni, nj);
        if (hit) {
          if (lane == 0) {
            sv[i] = ni;
            sv[j] = nj;
          }
          ++acc;
        }
      }
    }
    for (int t = lane; t < n; t += 64) {
      const float2 v = sv[t];
      a.vx[b + t] = v.x;
      a.vy[b + t] = v.y;
    }
    if (lane == 0) a.gmax[c] = (float)gm;
  }
  if (lane == 0 && acc) atomicAdd(&cnt, acc);
  __syncthreads();
  if (threadIdx.x == 0 && cnt && a.ncoll)
    atomicAdd((unsigned long long*)a.ncoll, (unsigned long long)cnt);
}

}  // namespace

void launch_mcc(hipStream_t s, const MccArgs& a, int n_upper) {
  int blocks = (n_upper + PIC_WG - 1) / PIC_WG;
  blocks = blocks < 1 ? 1 : (blocks > 4096 ? 4096 : blocks);
  k_mcc<<<blocks, PIC_WG, 0, s>>>(a, n_upper);
}

void launch_mcc_births(hipStream_t s, const MccArgs& a, unsigned long long* scratch,
                       unsigned long long* rho, long long gsum, const RunDesc* runs, int run) {
  k_mcc_births<<<1, PIC_WG, 0, s>>>(a, scratch, rho, nullptr, gsum, runs, run);
}

void launch_dsmc(hipStream_t s, const DsmcArgs& a) {
  int blocks = (a.ncells + DSMC_WAVES - 1) / DSMC_WAVES;
  blocks = blocks < 1 ? 1 : (blocks > 16384 ? 16384 : blocks);
  k_dsmc<<<blocks, 64 * DSMC_WAVES, 0, s>>>(a);
}
