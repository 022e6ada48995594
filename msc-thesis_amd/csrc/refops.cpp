// refops.cpp -- the reference's own collide arithmetic as stage operators (C ABI).
//
// pic_op_mcc_reference  thesisCode/mccCollisions/src/particle.cc:112-127, called for every
//                       particle in order (mccCollisions/src/main.cc:181-183);
// pic_op_dsmc_reference thesisCode/dsmcCollisions/cellTrackerAdvanced/src/main.cc:269-350
//                       (cellTracking + collisions).
// The caller injects the uniform stream the reference would draw (its std::mt19937 +
// uniform_real_distribution<double>, or any other); the operators resolve which uniforms
// each particle takes on the device (k_refcollide.hip) and return how many they used, so
// the next call continues the stream.  FP64 particles, as the reference's.
#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.h"

using namespace pic;

namespace {

struct Scratch {
  std::vector<void*> ptrs;
  ~Scratch() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  hipError_t get(T** p, size_t count) {
    hipError_t e = dalloc(p, count);
    if (e == hipSuccess) ptrs.push_back((void*)*p);
    return e;
  }
};

// consumer starts of m consumers on u[nu]; *used = uniforms they take
int resolve_chain(hipStream_t s, Scratch& sc, const double* u, long long nu, double prob,
                  long long m, long long** start, long long* used) {
  const long long nchunks = (nu + REF_CHAIN_CHUNK - 1) / REF_CHAIN_CHUNK;
  RefChainBufs b;
  PIC_HIP(sc.get(&b.maps, 7 * std::max(nchunks, 1LL)));
  PIC_HIP(sc.get(&b.entry, std::max(nchunks, 1LL)));
  PIC_HIP(sc.get(&b.base, std::max(nchunks, 1LL)));
  PIC_HIP(sc.get(&b.total, 1));
  PIC_HIP(sc.get(&b.start, std::max(m, 1LL)));
  PIC_HIP(sc.get(&b.consumed, 1));
  launch_ref_chain(s, u, nu, prob, m, b);
  PIC_HIP(hipGetLastError());
  long long c = -1;
  PIC_HIP(hipMemcpyAsync(&c, b.consumed, sizeof(c), hipMemcpyDeviceToHost, s));
  PIC_HIP(hipStreamSynchronize(s));
  if (c < 0) return fail(PIC_ERR_ARG, "uniform stream too short for the particles that draw");
  *start = b.start;
  *used = c;
  return PIC_OK;
}

bool finite_prob(double p) { return std::isfinite(p); }

}  // namespace

extern "C" {

int pic_op_mcc_reference(double* vx, double* vy, int64_t n, double prob, const double* u,
                         int64_t nu, int64_t* consumed, void* stream) {
  if (n < 0 || nu < 0 || !consumed || !finite_prob(prob)) return fail(PIC_ERR_ARG, "bad argument");
  if (n > 0 && (!vx || !vy || !u)) return fail(PIC_ERR_ARG, "NULL array");
  hipStream_t s = (hipStream_t)stream;
  Scratch sc;
  long long* start = nullptr;
  long long used = 0;
  int rc = resolve_chain(s, sc, u, nu, prob, n, &start, &used);
  if (rc) return rc;
  launch_mcc_reference(s, vx, vy, n, prob, u, start);
  PIC_HIP(hipGetLastError());
  PIC_HIP(hipStreamSynchronize(s));
  *consumed = used;
  return PIC_OK;
}

int pic_op_dsmc_reference(const double* x, const double* y, double* vx, double* vy, int64_t n,
                          const pic_grid* g, double prob, const double* u, int64_t nu,
                          int64_t* consumed, int64_t* ncoll, void* stream) {
  if (!g || n < 0 || n > 0x7fff0000 || nu < 0 || !consumed || !finite_prob(prob))
    return fail(PIC_ERR_ARG, "bad argument");
  if (n > 0 && (!x || !y || !vx || !vy || !u)) return fail(PIC_ERR_ARG, "NULL array");
  pic_config c;
  pic_config_default(&c);
  c.nx = g->nx; c.ny = g->ny; c.lx = g->lx; c.ly = g->ly;
  int rc = validate_config(c);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int G = g->nx * g->ny;
  Scratch sc;
  int *keys, *bad, *perm, *counts, *offsets, *target, *drawkey, *cons, *ccount, *coff;
  PIC_HIP(sc.get(&keys, std::max<int64_t>(n, 1)));
  PIC_HIP(sc.get(&bad, 1));
  PIC_HIP(sc.get(&perm, std::max<int64_t>(n, 1)));
  PIC_HIP(sc.get(&counts, G));
  PIC_HIP(sc.get(&offsets, (size_t)G + 1));
  PIC_HIP(sc.get(&target, std::max<int64_t>(n, 1)));
  PIC_HIP(sc.get(&drawkey, std::max<int64_t>(n, 1)));
  PIC_HIP(sc.get(&cons, std::max<int64_t>(n, 1)));
  PIC_HIP(sc.get(&ccount, 2));
  PIC_HIP(sc.get(&coff, 3));
  // the reference's cells, (int)(x/dx) and (int)(y/dy), keyed cx*ny + cy
  PIC_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
  launch_ref_cell_keys(s, x, y, n, g->nx, g->ny, g->lx / g->nx, g->ly / g->ny, keys, bad);
  int hbad = 0;
  PIC_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  PIC_HIP(hipStreamSynchronize(s));
  if (hbad) return fail(PIC_ERR_ARG, "particle outside [0, lx) x [0, ly)");
  // cell lists in ascending particle index: the stable counting sort by cell
  if ((rc = pic_op_sort(keys, n, G, perm, counts, offsets, stream))) return rc;
  launch_ref_nearest(s, x, y, n, keys, perm, offsets, target, drawkey);
  // the particles that draw, in ascending index: stable sort by (draws ? 0 : 1)
  if ((rc = pic_op_sort(drawkey, n, 2, cons, ccount, coff, stream))) return rc;
  int hoff[2] = {0, 0};
  PIC_HIP(hipMemcpyAsync(hoff, coff, sizeof(hoff), hipMemcpyDeviceToHost, s));
  PIC_HIP(hipStreamSynchronize(s));
  const long long m = hoff[1];
  long long* start = nullptr;
  long long used = 0;
  if ((rc = resolve_chain(s, sc, u, nu, prob, m, &start, &used))) return rc;
  double *m1, *m2;
  int *self, *lastw;
  unsigned long long* nc;
  PIC_HIP(sc.get(&m1, std::max(m, 1LL)));
  PIC_HIP(sc.get(&m2, std::max(m, 1LL)));
  PIC_HIP(sc.get(&self, std::max<int64_t>(n, 1)));
  PIC_HIP(sc.get(&lastw, std::max<int64_t>(n, 1)));
  PIC_HIP(sc.get(&nc, 1));
  launch_fill_i32(s, self, n, -1);
  launch_fill_i32(s, lastw, n, -1);
  PIC_HIP(hipMemsetAsync(nc, 0, sizeof(unsigned long long), s));
  launch_dsmc_reference(s, vx, vy, n, m, prob, u, start, cons, target, m1, m2, self, lastw, nc);
  PIC_HIP(hipGetLastError());
  unsigned long long hn = 0;
  PIC_HIP(hipMemcpyAsync(&hn, nc, sizeof(hn), hipMemcpyDeviceToHost, s));
  PIC_HIP(hipStreamSynchronize(s));
  *consumed = used;
  if (ncoll) *ncoll = (int64_t)hn;
  return PIC_OK;
}

}  // extern "C"
