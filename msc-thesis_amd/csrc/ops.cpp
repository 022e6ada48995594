// ops.cpp -- per-stage operators of the C ABI on caller-owned device arrays.
//
// The reference keeps each stage as a file-static kernel configured through
// __constant__ symbols (thesisCode/collisionless/double/src/pic.cu:13-26, 69-420); these
// entry points expose the same stages with explicit arguments so each can be dropped into
// another driver or tested alone.  Temporaries are allocated per call, except in
// pic_op_deposit_noalloc (a static device array).
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "engine.h"

using namespace pic;

namespace {

struct DevMem {
  std::vector<void*> ptrs;
  ~DevMem() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  hipError_t get(T** p, size_t count) {
    hipError_t e = dalloc(p, count);
    if (e == hipSuccess) ptrs.push_back((void*)*p);
    return e;
  }
};

int check_grid(const pic_grid* g) {
  if (!g) return fail(PIC_ERR_ARG, "NULL grid");
  pic_config c;
  pic_config_default(&c);
  c.nx = g->nx; c.ny = g->ny; c.lx = g->lx; c.ly = g->ly;
  return validate_config(c);
}

// single-run descriptor for the operator kernels (all bases zero)
RunDesc op_desc(int nx, int ny, double lx, double ly, long long pcap) {
  RunDesc d;
  std::memset(&d, 0, sizeof(d));
  d.schunk = PIC_CHUNK;
  d.nx = nx; d.ny = ny;
  int l = 0;
  while ((1 << l) < nx) ++l;
  d.lognx = l;
  l = 0;
  while ((1 << l) < ny) ++l;
  d.logny = l;
  d.lx = lx; d.ly = ly; d.dx = lx / nx; d.dy = ly / ny;
  d.pcap = (int)pcap;
  fft_groups(nx, nx, ny, &d.fft_pp, &d.fft_rf);
  return d;
}

}  // namespace

extern "C" {

int pic_op_cell_keys(const void* x, const void* y, int64_t n, int precision, const pic_grid* g,
                     int32_t* keys, void* stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!x || !y || !keys))) return fail(PIC_ERR_ARG, "bad arrays");
  if (n == 0) return PIC_OK;
  hipStream_t s = (hipStream_t)stream;
  if (precision == PIC_FP64)
    launch_op_cell_keys<double>(s, (const double*)x, (const double*)y, n, g->nx, g->ny, g->lx,
                                g->ly, keys);
  else if (precision == PIC_FP32)
    launch_op_cell_keys<float>(s, (const float*)x, (const float*)y, n, g->nx, g->ny, g->lx, g->ly,
                               keys);
  else
    return fail(PIC_ERR_ARG, "precision must be PIC_FP32 or PIC_FP64");
  PIC_HIP(hipGetLastError());
  return PIC_OK;
}

int pic_op_sort(const int32_t* keys, int64_t n, int ncells, int32_t* perm, int32_t* counts,
                int32_t* offsets, void* stream) {
  if (n < 0 || ncells < 1 || n > 0x7fff0000) return fail(PIC_ERR_ARG, "bad sizes");
  if (!perm || !counts || !offsets || (n > 0 && !keys)) return fail(PIC_ERR_ARG, "NULL array");
  hipStream_t s = (hipStream_t)stream;
  DevMem m;
  int* flag;
  PIC_HIP(m.get(&flag, 1));
  PIC_HIP(hipMemsetAsync(flag, 0, sizeof(int), s));
  if (n > 0) launch_op_validate_keys(s, keys, n, ncells, flag);
  int h = 0;
  PIC_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, s));
  PIC_HIP(hipStreamSynchronize(s));
  if (h) return fail(PIC_ERR_ARG, "key outside [0, ncells)");
  // one run with nx = ncells, ny = 1 (G = ncells); counts/offsets are the caller's arrays
  RunDesc d = op_desc(ncells, 1, 1.0, 1.0, n);
  RunDesc* dd;
  RunState* st;
  PIC_HIP(m.get(&dd, 1));
  PIC_HIP(m.get(&st, 1));
  RunState hs;
  std::memset(&hs, 0, sizeof(hs));
  hs.np = (int)n;
  std::vector<int2> pit, cit, sit;
  for (long long q = 0; q < std::max<long long>(n, 1); q += PIC_CHUNK) pit.push_back(make_int2(0, (int)q));
  for (int c = 0; c < ncells; c += PIC_CELL_CHUNK) cit.push_back(make_int2(0, c));
  for (int c = 0; c < ncells; c += PIC_SEG_CELLS) sit.push_back(make_int2(0, c));
  d.cell_item_begin = 0;
  d.cell_item_end = (int)cit.size();
  int2 *dp, *dc, *ds;
  PIC_HIP(m.get(&dp, pit.size()));
  PIC_HIP(m.get(&dc, cit.size()));
  PIC_HIP(m.get(&ds, sit.size()));
  PIC_HIP(hipMemcpyAsync(ds, sit.data(), sizeof(int2) * sit.size(), hipMemcpyHostToDevice, s));
  SortBufs sb;
  std::memset(&sb, 0, sizeof(sb));
  sb.segitems = ds;
  sb.nsegitems = (int)sit.size();
  PIC_HIP(m.get(&sb.rank, std::max<int64_t>(n, 1)));
  PIC_HIP(m.get(&sb.chunk_sums, cit.size()));
  PIC_HIP(m.get(&sb.big_list, 2 * (size_t)ncells));
  PIC_HIP(m.get(&sb.big_count, 1));
  PIC_HIP(m.get(&sb.scratch, 2 * (size_t)std::max<int64_t>(n, 1)));
  sb.keys = nullptr;
  sb.perm = perm;
  sb.counts = counts;
  // offsets need ncells + 1 entries: the caller's array
  sb.offsets = offsets;
  PIC_HIP(hipMemcpyAsync(dd, &d, sizeof(d), hipMemcpyHostToDevice, s));
  PIC_HIP(hipMemcpyAsync(st, &hs, sizeof(hs), hipMemcpyHostToDevice, s));
  PIC_HIP(hipMemcpyAsync(dp, pit.data(), sizeof(int2) * pit.size(), hipMemcpyHostToDevice, s));
  PIC_HIP(hipMemcpyAsync(dc, cit.data(), sizeof(int2) * cit.size(), hipMemcpyHostToDevice, s));
  launch_sort_keys_only(s, dd, st, keys, sb, dp, (int)pit.size(), dc, (int)cit.size(), ncells);
  PIC_HIP(hipGetLastError());
  PIC_HIP(hipStreamSynchronize(s));
  return PIC_OK;
}

int pic_op_deposit(const void* x, const void* y, int64_t n, int precision, const pic_grid* g,
                   double qp, double* rho, void* stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  if (n < 0 || !rho || (n > 0 && (!x || !y))) return fail(PIC_ERR_ARG, "bad arrays");
  if (precision != PIC_FP32 && precision != PIC_FP64) return fail(PIC_ERR_ARG, "bad precision");
  hipStream_t s = (hipStream_t)stream;
  const long long G = (long long)g->nx * g->ny;
  DevMem m;
  unsigned long long *hi, *lo = nullptr;
  PIC_HIP(m.get(&hi, G));
  PIC_HIP(hipMemsetAsync(hi, 0, sizeof(unsigned long long) * G, s));
  if (precision == PIC_FP64) {
    PIC_HIP(m.get(&lo, G));
    PIC_HIP(hipMemsetAsync(lo, 0, sizeof(unsigned long long) * G, s));
  }
  if (n > 0) {
    if (precision == PIC_FP64)
      launch_op_deposit<double>(s, (const double*)x, (const double*)y, n, g->nx, g->ny, g->lx,
                                g->ly, hi, lo);
    else
      launch_op_deposit<float>(s, (const float*)x, (const float*)y, n, g->nx, g->ny, g->lx, g->ly,
                               hi, lo);
  }
  const double dx = g->lx / g->nx, dy = g->ly / g->ny;
  const double scale = qp / (dx * dy) * (precision == PIC_FP64 ? 1.0 : std::ldexp(1.0, -36));
  launch_op_fixed_to_double(s, hi, lo, G, precision == PIC_FP64, scale, rho);
  PIC_HIP(hipGetLastError());
  PIC_HIP(hipStreamSynchronize(s));
  return PIC_OK;
}

}  // extern "C"

// ---- allocation-free deposit: the fixed-point accumulators live in a file-scope device
// array instead of per-call hipMalloc.  Every device that runs the library holds its own
// copy (16 MiB) until the process ends.  In host code the array's name is only a
// stand-in symbol: its device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned long long scratch[2 * PIC_DEPOSIT_NOALLOC_MAX_NODES];

namespace {
std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[64];  // device address of `scratch` per device ordinal

// Device address of `scratch` on the current device, looked up on the first call there.
int noalloc_scratch(unsigned long long** out) {
  int dev = 0;
  PIC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(PIC_ERR_UNSUPPORTED, "device ordinal above 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(PIC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (unsigned long long*)p;
  }
  *out = g_scratch_ptr[dev];
  return PIC_OK;
}
}  // namespace

extern "C" {

int pic_op_deposit_noalloc(const void* x, const void* y, int64_t n, int precision,
                           const pic_grid* g, double qp, double* rho, void* stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  if (n < 0 || !rho || (n > 0 && (!x || !y))) return fail(PIC_ERR_ARG, "bad arrays");
  if (precision != PIC_FP32 && precision != PIC_FP64) return fail(PIC_ERR_ARG, "bad precision");
  const long long G = (long long)g->nx * g->ny;
  if (G > PIC_DEPOSIT_NOALLOC_MAX_NODES)
    return fail(PIC_ERR_UNSUPPORTED, "pic_op_deposit_noalloc: nx*ny = " + std::to_string(G) +
                                         " exceeds its static scratch of " +
                                         std::to_string(PIC_DEPOSIT_NOALLOC_MAX_NODES) + " nodes");
  unsigned long long* hi = nullptr;
  if ((rc = noalloc_scratch(&hi))) return rc;
  unsigned long long* lo = hi + PIC_DEPOSIT_NOALLOC_MAX_NODES;
  hipStream_t s = (hipStream_t)stream;
  // cleared by a fill kernel (the previous call's sums are still there)
  launch_fill_u64(s, hi, G);
  if (precision == PIC_FP64) launch_fill_u64(s, lo, G);
  if (n > 0) {
    if (precision == PIC_FP64)
      launch_op_deposit<double>(s, (const double*)x, (const double*)y, n, g->nx, g->ny, g->lx,
                                g->ly, hi, lo);
    else
      launch_op_deposit<float>(s, (const float*)x, (const float*)y, n, g->nx, g->ny, g->lx, g->ly,
                               hi, nullptr);
  }
  const double dx = g->lx / g->nx, dy = g->ly / g->ny;
  const double scale = qp / (dx * dy) * (precision == PIC_FP64 ? 1.0 : std::ldexp(1.0, -36));
  launch_op_fixed_to_double(s, hi, precision == PIC_FP64 ? lo : nullptr, G,
                            precision == PIC_FP64, scale, rho);
  PIC_HIP(hipGetLastError());
  PIC_HIP(hipStreamSynchronize(s));
  return PIC_OK;
}

int pic_op_poisson(const double* rho, double* phi, const pic_grid* g, void* stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  if (!rho || !phi) return fail(PIC_ERR_ARG, "NULL array");
  hipStream_t s = (hipStream_t)stream;
  const int nx = g->nx, ny = g->ny;
  DevMem m;
  RunDesc d = op_desc(nx, ny, g->lx, g->ly, 0);
  d.row_item_begin = 0;
  d.row_item_end = nx / d.fft_rf;
  RunDesc* dd;
  RunState* st;
  PIC_HIP(m.get(&dd, 1));
  PIC_HIP(m.get(&st, 1));
  PIC_HIP(hipMemsetAsync(st, 0, sizeof(RunState), s));
  PIC_HIP(hipMemcpyAsync(dd, &d, sizeof(d), hipMemcpyHostToDevice, s));
  const long long ss = (long long)(ny / 2 + 1) * nx;
  double2 *S, *Sx, *Sy, *tw;
  float* E;
  PIC_HIP(m.get(&S, ss));
  PIC_HIP(m.get(&Sx, ss));
  PIC_HIP(m.get(&Sy, ss));
  PIC_HIP(m.get(&E, 2 * (size_t)(nx + 1) * (ny + 1)));
  PIC_HIP(m.get(&tw, PIC_FFT_MAX / 2));
  std::vector<double2> h(PIC_FFT_MAX / 2);
  for (int j = 0; j < PIC_FFT_MAX / 2; ++j) {
    const long double a = 2.0L * 3.141592653589793238462643383279L * j / PIC_FFT_MAX;
    h[j] = make_double2((double)cosl(a), (double)-sinl(a));
  }
  PIC_HIP(hipMemcpyAsync(tw, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, s));
  std::vector<int2> rp, cl, rw;
  for (int i = 0; i < nx; i += 2 * d.fft_pp) rp.push_back(make_int2(0, i));
  for (int k = 0; k <= ny / 2; ++k) cl.push_back(make_int2(0, k));
  for (int i = 0; i < nx; i += d.fft_rf) rw.push_back(make_int2(0, i));
  int2 *drp, *dcl, *drw;
  PIC_HIP(m.get(&drp, rp.size()));
  PIC_HIP(m.get(&dcl, cl.size()));
  PIC_HIP(m.get(&drw, rw.size()));
  PIC_HIP(hipMemcpyAsync(drp, rp.data(), sizeof(int2) * rp.size(), hipMemcpyHostToDevice, s));
  PIC_HIP(hipMemcpyAsync(dcl, cl.data(), sizeof(int2) * cl.size(), hipMemcpyHostToDevice, s));
  PIC_HIP(hipMemcpyAsync(drw, rw.data(), sizeof(int2) * rw.size(), hipMemcpyHostToDevice, s));
  FieldArgs a;
  std::memset(&a, 0, sizeof(a));
  a.runs = dd; a.st = st;
  a.S = S; a.Sx = Sx; a.Sy = Sy; a.E = E; a.phi = phi; a.tw = tw;
  a.row_pair_items = drp; a.n_row_pair_items = (int)rp.size();
  a.col_items = dcl; a.n_col_items = (int)cl.size();
  a.row_items = drw; a.n_row_items = (int)rw.size();
  a.nruns = 1; a.fp64 = 0; a.want_phi = 1; a.max_nx = nx; a.max_ny = ny;
  a.row_lds = std::max(d.fft_pp, d.fft_rf) * ny;
  a.rho_in = rho;
  launch_field(s, a, nullptr);
  PIC_HIP(hipGetLastError());
  PIC_HIP(hipStreamSynchronize(s));
  return PIC_OK;
}

int pic_op_efield(const double* phi, double* ex, double* ey, const pic_grid* g, void* stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  if (!phi || !ex || !ey) return fail(PIC_ERR_ARG, "NULL array");
  launch_op_efield((hipStream_t)stream, phi, ex, ey, g->nx, g->ny, g->lx, g->ly);
  PIC_HIP(hipGetLastError());
  return PIC_OK;
}

int pic_op_gather_push(void* x, void* y, void* vx, void* vy, int64_t n, int precision,
                       const double* ex, const double* ey, const pic_grid* g, double dt,
                       double qm, double c, double* ke, void* stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  if (n < 0 || !ex || !ey || (n > 0 && (!x || !y || !vx || !vy))) return fail(PIC_ERR_ARG, "bad arrays");
  if (precision != PIC_FP32 && precision != PIC_FP64) return fail(PIC_ERR_ARG, "bad precision");
  hipStream_t s = (hipStream_t)stream;
  long long b = (n + PIC_WG - 1) / PIC_WG;
  const int nb = (int)std::max(1LL, std::min(b, 4096LL));
  DevMem m;
  double* parts = nullptr;
  if (ke) PIC_HIP(m.get(&parts, nb));
  if (precision == PIC_FP64)
    launch_op_gather_push<double>(s, (double*)x, (double*)y, (double*)vx, (double*)vy, n, ex, ey,
                                  g->nx, g->ny, g->lx, g->ly, dt, qm, c, parts, nb);
  else
    launch_op_gather_push<float>(s, (float*)x, (float*)y, (float*)vx, (float*)vy, n, ex, ey, g->nx,
                                 g->ny, g->lx, g->ly, dt, qm, c, parts, nb);
  if (ke) launch_reduce_sum(s, parts, nb, ke);
  PIC_HIP(hipGetLastError());
  PIC_HIP(hipStreamSynchronize(s));
  return PIC_OK;
}

int pic_op_mcc(float* x, float* y, float* vx, float* vy, int32_t* id, int64_t* n_io, int64_t cap,
               const pic_config* cfg, int64_t step, int64_t* counts, void* stream) {
  if (!x || !y || !vx || !vy || !id || !n_io || !cfg) return fail(PIC_ERR_ARG, "NULL argument");
  if (*n_io < 0 || cap < *n_io || cap > 0x7fff0000) return fail(PIC_ERR_ARG, "bad sizes");
  pic_config c = *cfg;
  c.collide = PIC_COLLIDE_MCC;
  c.precision = PIC_FP32;
  int rc = validate_config(c);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  DevMem m;
  RunState* st;
  RunDesc* dd;
  int *bc, *bp, *cl, *cc, *cp;
  float* bb;
  unsigned long long *scratch, *rho;
  PIC_HIP(m.get(&st, 1));
  PIC_HIP(m.get(&dd, 1));
  PIC_HIP(m.get(&bc, 1));
  PIC_HIP(m.get(&bp, std::max<int64_t>(cap, 1)));
  PIC_HIP(m.get(&bb, 4 * std::max<int64_t>(cap, 1)));
  int lb, lp;
  mcc_list_layout(std::max<int64_t>(*n_io, 1), &lb, &lp);
  PIC_HIP(m.get(&cl, (size_t)lb * lp));
  PIC_HIP(m.get(&cc, PIC_MCC_BLOCKS));
  PIC_HIP(m.get(&cp, 4 * PIC_MCC_BLOCKS));
  PIC_HIP(m.get(&scratch, 2 * std::max<int64_t>(cap, 1)));
  PIC_HIP(m.get(&rho, 2 * (size_t)c.nx * c.ny));
  RunState hs;
  std::memset(&hs, 0, sizeof(hs));
  hs.np = (int)*n_io;
  // next_id = largest id + 1 is the caller's contract: ids are 0..n-1 here
  hs.next_id = *n_io;
  hs.step = step;
  RunDesc d = op_desc(c.nx, c.ny, c.lx, c.ly, cap);
  PIC_HIP(hipMemcpyAsync(st, &hs, sizeof(hs), hipMemcpyHostToDevice, s));
  PIC_HIP(hipMemcpyAsync(dd, &d, sizeof(d), hipMemcpyHostToDevice, s));
  PIC_HIP(hipMemsetAsync(bc, 0, sizeof(int), s));
  PIC_HIP(hipMemsetAsync(rho, 0, sizeof(unsigned long long) * 2 * c.nx * c.ny, s));
  MccArgs a;
  std::memset(&a, 0, sizeof(a));
  a.x = x; a.y = y; a.vx = vx; a.vy = vy; a.id = id;
  a.st = st; a.pbase = 0; a.cap = (int)cap;
  const double kE = 1.602176634e-19, kMe = 9.1093837015e-31;
  a.v_unit2_half_me_over_e = 0.5 * kMe * c.mcc_v_unit * c.mcc_v_unit / kE;
  a.ng_vunit_dtphys = c.mcc_ng * c.mcc_v_unit * c.mcc_dt_phys * c.dt;
  const double numax_dt = mcc_nu_max(c) * c.mcc_dt_phys * c.dt;
  a.nu_max_dt = numax_dt;
  a.p_coll = 1.0 - std::exp(-numax_dt);
  a.mass_ratio2 = 2.0 * kMe / c.mcc_neutral_mass;
  a.seed = c.seed;
  a.counts = (long long*)counts;
  a.birth_count = bc; a.birth_parent = bp; a.birth_buf = bb;
  a.coll_list = cl; a.coll_count = cc; a.coll_partials = cp;
  a.from_state_step = 0;
  a.step = step;
  launch_mcc(s, a, (int)*n_io);
  launch_mcc_births(s, a, scratch, rho, (long long)c.nx * c.ny, dd, 0, (int)*n_io);
  PIC_HIP(hipGetLastError());
  PIC_HIP(hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s));
  PIC_HIP(hipStreamSynchronize(s));
  *n_io = hs.np;
  if (hs.err & 1) return fail(PIC_ERR_OVERFLOW, "births exceeded capacity");
  return PIC_OK;
}

int pic_op_dsmc(float* vx, float* vy, const int32_t* id, const int32_t* offsets, int ncells,
                float* gmax, double coeff, uint64_t seed, int64_t step, int64_t* ncoll,
                void* stream) {
  if (!vx || !vy || !id || !offsets || !gmax || ncells < 1 || !(coeff >= 0))
    return fail(PIC_ERR_ARG, "bad argument");
  DsmcArgs a;
  std::memset(&a, 0, sizeof(a));
  (void)id;  // cells hold their particles in the caller's (canonical) order
  a.vx = vx; a.vy = vy; a.offsets = offsets; a.gmax = gmax;
  a.ncells = ncells; a.tile_shift = -1; a.tiles_y = 1; a.ny = 1; a.coeff = coeff; a.seed = seed; a.step = step; a.st = nullptr;
  a.ncoll = (long long*)ncoll;
  launch_dsmc((hipStream_t)stream, a);
  PIC_HIP(hipGetLastError());
  return PIC_OK;
}

}  // extern "C"
