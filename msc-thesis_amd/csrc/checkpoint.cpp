// checkpoint.cpp -- versioned binary snapshot of one run, and restart from it.
// No reference counterpart (the reference has no restart; its -f flag only truncates a file,
// thesisCode/collisionless/double/src/main.cc:101-108).
//
// File layout (little-endian, all fields naturally aligned):
//   CkptHeader                      magic "PICAMDCK", format version, the run's pic_config,
//                                   particle count, completed steps, velocity phase, next
//                                   id, macro-particle charge, collision counters, size of
//                                   the DSMC per-cell maxima
//   x[n], y[n], vx[n], vy[n]        particle state in id order, in the run's precision;
//                                   velocities as stored (staggered unless vsync = 1)
//   gmax[ngmax]                     float, DSMC only: per-cell maximum relative speed of
//                                   cell cx*ny + cy (row-major, so a run may restart under
//                                   another tiling), ngmax = nx*ny + 1
//   uint64 checksum                 FNV-1a over everything before it
// A restart continues bit for bit: the charge is redeposited from the positions with the
// same integer arithmetic, and the random streams are keyed by particle id and step.
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "engine.h"

using namespace pic;

namespace {

constexpr char kMagic[8] = {'P', 'I', 'C', 'A', 'M', 'D', 'C', 'K'};
constexpr uint32_t kVersion = 2;  // 2: DSMC per-cell maxima in row-major cell order

struct CkptHeader {
  char magic[8];
  uint32_t version;
  uint32_t precision;
  pic_config cfg;
  int64_t n;
  int64_t step;
  int32_t vsync;
  int32_t reserved;
  int64_t next_id;
  double qp;
  int64_t coll[4];
  int64_t ngmax;
};

struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void* p, size_t bytes) {
    const unsigned char* c = (const unsigned char*)p;
    for (size_t i = 0; i < bytes; ++i) {
      h ^= c[i];
      h *= 1099511628211ull;
    }
  }
};

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

bool same_run(const pic_config& a, const pic_config& b) {
  return a.nx == b.nx && a.ny == b.ny && a.lx == b.lx && a.ly == b.ly && a.dt == b.dt &&
         a.qm == b.qm && a.precision == b.precision && a.collide == b.collide &&
         a.seed == b.seed;
}

int read_header(const char* path, File& fh, CkptHeader& h) {
  fh.f = std::fopen(path, "rb");
  if (!fh.f) return fail(PIC_ERR_ARG, std::string("cannot open ") + path);
  if (std::fread(&h, sizeof(h), 1, fh.f) != 1) return fail(PIC_ERR_ARG, "checkpoint too short");
  if (std::memcmp(h.magic, kMagic, 8) != 0) return fail(PIC_ERR_ARG, "not a checkpoint file");
  if (h.version != kVersion) return fail(PIC_ERR_UNSUPPORTED, "checkpoint format version");
  if (h.n < 1 || h.ngmax < 0 || (h.precision != PIC_FP32 && h.precision != PIC_FP64))
    return fail(PIC_ERR_ARG, "corrupt checkpoint header");
  return PIC_OK;
}

}  // namespace

extern "C" {

int pic_save_checkpoint(pic_ctx* ctx, int run, const char* path) {
  if (!ctx || !path) return fail(PIC_ERR_ARG, "NULL argument");
  Ctx& x = ctx->impl;
  if (run < 0 || run >= x.nruns || !x.has_particles[run])
    return fail(PIC_ERR_STATE, "run has no particles");
  PIC_HIP(hipSetDevice(x.device));
  PIC_HIP(hipStreamSynchronize(x.stream));
  RunState s;
  PIC_HIP(hipMemcpy(&s, x.d_st + run, sizeof(s), hipMemcpyDeviceToHost));
  if (s.err) return fail(PIC_ERR_OVERFLOW, "run is in an error state");
  const int64_t n = s.np;
  std::vector<double> hx(n), hy(n), hvx(n), hvy(n);
  int rc = pic_get_particles(ctx, run, hx.data(), hy.data(), hvx.data(), hvy.data(), n, 0);
  if (rc) return rc;
  CkptHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, kMagic, 8);
  h.version = kVersion;
  h.precision = (uint32_t)x.cfg[run].precision;
  h.cfg = x.cfg[run];
  h.n = n;
  h.step = s.step;
  h.vsync = s.vsync;
  h.next_id = s.next_id;
  h.qp = x.rd[run].qp;
  PIC_HIP(hipMemcpy(h.coll, x.d_coll_counts + 4 * run, sizeof(h.coll), hipMemcpyDeviceToHost));
  std::vector<float> gm;
  if (x.cfg[run].collide == PIC_COLLIDE_DSMC) {
    const RunDesc& d = x.rd[run];
    gm.resize((size_t)d.nx * d.ny + 1);
    PIC_HIP(hipMemcpy(gm.data(), x.gmax + d.cbase, sizeof(float) * gm.size(), hipMemcpyDeviceToHost));
  }
  h.ngmax = (int64_t)gm.size();
  File fh;
  fh.f = std::fopen(path, "wb");
  if (!fh.f) return fail(PIC_ERR_ARG, std::string("cannot write ") + path);
  Fnv sum;
  auto put = [&](const void* p, size_t bytes) {
    sum.add(p, bytes);
    return std::fwrite(p, 1, bytes, fh.f) == bytes;
  };
  bool ok = put(&h, sizeof(h));
  for (const std::vector<double>* a : {&hx, &hy, &hvx, &hvy}) {
    if (x.fp64) {
      ok = ok && put(a->data(), sizeof(double) * n);
    } else {
      std::vector<float> f(a->begin(), a->end());  // the stored values are floats: exact
      ok = ok && put(f.data(), sizeof(float) * n);
    }
  }
  if (!gm.empty()) ok = ok && put(gm.data(), sizeof(float) * gm.size());
  ok = ok && std::fwrite(&sum.h, sizeof(sum.h), 1, fh.f) == 1;
  ok = (std::fflush(fh.f) == 0) && ok;
  if (!ok) return fail(PIC_ERR_ARG, std::string("write failed: ") + path);
  return PIC_OK;
}

int pic_checkpoint_info(const char* path, pic_config* cfg, int64_t* n, int64_t* step) {
  if (!path) return fail(PIC_ERR_ARG, "NULL path");
  File fh;
  CkptHeader h;
  int rc = read_header(path, fh, h);
  if (rc) return rc;
  if (cfg) *cfg = h.cfg;
  if (n) *n = h.n;
  if (step) *step = h.step;
  return PIC_OK;
}

int pic_load_checkpoint(pic_ctx* ctx, int run, const char* path) {
  if (!ctx || !path) return fail(PIC_ERR_ARG, "NULL argument");
  Ctx& x = ctx->impl;
  if (run < 0 || run >= x.nruns) return fail(PIC_ERR_ARG, "run index out of range");
  File fh;
  CkptHeader h;
  int rc = read_header(path, fh, h);
  if (rc) return rc;
  if (!same_run(h.cfg, x.cfg[run]))
    return fail(PIC_ERR_INVALID_CONFIG,
                "checkpoint grid, domain, dt, qm, precision, collisions or seed differ from the run");
  // sizes from the header are checked before anything is allocated or the run touched
  const int64_t n = h.n;
  const int64_t cap = x.allocated ? (int64_t)x.rd[run].pcap : (int64_t)0x7fff0000;
  if (n > cap)
    return fail(PIC_ERR_OVERFLOW, "checkpoint holds " + std::to_string(n) +
                                      " particles, more than the run capacity");
  const int64_t want_gm =
      x.cfg[run].collide == PIC_COLLIDE_DSMC ? (int64_t)x.cfg[run].nx * x.cfg[run].ny + 1 : 0;
  if (h.ngmax != want_gm) return fail(PIC_ERR_ARG, "corrupt checkpoint (per-cell table size)");
  const size_t rs = x.fp64 ? 8 : 4;
  std::vector<unsigned char> raw;
  std::vector<float> gm;
  try {
    raw.resize(rs * 4 * (size_t)n);
    gm.resize((size_t)h.ngmax);
  } catch (const std::bad_alloc&) {
    return fail(PIC_ERR_OOM, "host memory for the checkpoint");
  }
  Fnv sum;
  sum.add(&h, sizeof(h));
  if (std::fread(raw.data(), 1, raw.size(), fh.f) != raw.size())
    return fail(PIC_ERR_ARG, "checkpoint truncated");
  sum.add(raw.data(), raw.size());
  if (h.ngmax) {
    if (std::fread(gm.data(), sizeof(float), gm.size(), fh.f) != gm.size())
      return fail(PIC_ERR_ARG, "checkpoint truncated");
    sum.add(gm.data(), sizeof(float) * gm.size());
  }
  uint64_t stored = 0;
  if (std::fread(&stored, sizeof(stored), 1, fh.f) != 1 || stored != sum.h)
    return fail(PIC_ERR_ARG, "checkpoint checksum mismatch");
  std::vector<double> a;
  try {
    a.resize(4 * (size_t)n);
  } catch (const std::bad_alloc&) {
    return fail(PIC_ERR_OOM, "host memory for the checkpoint");
  }
  for (size_t i = 0; i < a.size(); ++i) {
    if (x.fp64) {
      double v;
      std::memcpy(&v, raw.data() + 8 * i, 8);
      a[i] = v;
    } else {
      float v;
      std::memcpy(&v, raw.data() + 4 * i, 4);
      a[i] = v;
    }
  }
  Restore r;
  r.step = h.step;
  r.vsync = h.vsync;
  r.next_id = h.next_id;
  r.qp = h.qp;
  rc = set_particles_impl(ctx, run, a.data(), a.data() + n, a.data() + 2 * n, a.data() + 3 * n, n,
                          false, &r);
  if (rc) return rc;
  PIC_HIP(hipMemcpy(x.d_coll_counts + 4 * run, h.coll, sizeof(h.coll), hipMemcpyHostToDevice));
  if (h.ngmax) {
    // per-cell maxima in row-major cell order: independent of the tiling in use
    const RunDesc& d = x.rd[run];
    PIC_HIP(hipMemcpy(x.gmax + d.cbase, gm.data(), sizeof(float) * gm.size(), hipMemcpyHostToDevice));
  }
  return PIC_OK;
}

}  // extern "C"
