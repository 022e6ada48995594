// adjoint.cpp -- differentiable collisionless run (C ABI half): forward with particle
// checkpoints every K steps, reverse sweep by segments with recomputation of the K steps in
// between.  The forward also keeps the field of every step (nsteps x (nx+1)(ny+1) x 2 reals,
// far smaller than the particle state), so the recomputation is particle pushes only: no
// deposit and no field solve.  At every checkpoint the particles are counting-sorted by
// tile (the checkpoint holds the sorted state and the original index of every slot), so the
// deposits of the segment's forward pushes and reverse passes sum in LDS windows; the
// per-particle adjoint arrays follow the order of the segment being reversed.
// No reference counterpart; the forward step is the reference step order of
// thesisCode/collisionless/double/src/main.cc:137-172 (deposit, Poisson, E, gather, push).
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "adjoint_launch.h"
#include "engine.h"

using namespace pic;

struct pic_adjoint {
  pic_config cfg;
  int fp64 = 0;
  size_t rs = 4;
  int64_t n = 0;
  int nm = 0, nsteps = 0, K = 1, nseg = 0;
  int device = 0;
  hipStream_t s = nullptr;
  AdjGrid g;
  double qp = 0;
  // device buffers
  double *xl = nullptr, *yl = nullptr, *mkx = nullptr, *mky = nullptr;
  void *x = nullptr, *y = nullptr, *vx = nullptr, *vy = nullptr;
  void* ckpt = nullptr;                       // nseg x 4 x n
  void *trx = nullptr, *try_ = nullptr;       // (K+1) x n
  void* trE = nullptr;                        // nsteps x padded field (E_0 .. E_{N-1})
  double *lE = nullptr, *lrho = nullptr;
  unsigned long long *hi = nullptr, *lo = nullptr;
  void* E = nullptr;
  double2 *S = nullptr, *Sx = nullptr, *Sy = nullptr, *A = nullptr, *tw = nullptr;
  RunDesc* dd = nullptr;
  RunState* st = nullptr;
  int2 *rp = nullptr, *cl = nullptr, *rw = nullptr;
  int nrp = 0, ncl = 0, nrw = 0;
  // tile sort at the checkpoints and windowed kernels
  int nb = 0;                                 // tiles
  int* ids = nullptr;                         // nseg x n: per slot of segment s, its slot in
                                              // the order before the sort (segment s - 1, or
                                              // the original order for s = 0)
  int* iota = nullptr;                        // n: 0, 1, ..., n - 1
  int* toffs = nullptr;                       // nseg x (nb + 1) tile offsets
  int4* items = nullptr;                      // {tile x, tile y, split} push / pass items
  int2* pitems = nullptr;                     // sort chunk items
  int npitems = 0;
  SortBufs sb;
  double* lam[2] = {nullptr, nullptr};        // 4 x n each: (lx, ly, lvx, lvy) of a segment
  double* res = nullptr;                      // the one that ends in the original order
  int row_lds = 0;  // complex values per row-pass workgroup
  double* part = nullptr;
  bool forward_done = false;
  hipGraphExec_t gfwd = nullptr, gbwd = nullptr;  // captured step loops (made on first use)
  std::vector<void*> all;

  ~pic_adjoint() {
    (void)hipSetDevice(device);
    if (s) (void)hipStreamSynchronize(s);
    if (gfwd) (void)hipGraphExecDestroy(gfwd);
    if (gbwd) (void)hipGraphExecDestroy(gbwd);
    for (void* p : all) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
  }
  template <typename T>
  hipError_t get(T** p, size_t count) {
    hipError_t e = dalloc(p, count);
    if (e == hipSuccess) all.push_back((void*)*p);
    return e;
  }
  hipError_t getb(void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e == hipSuccess) all.push_back(*p);
    return e;
  }
  size_t epad() const { return (size_t)2 * (g.nx + 1) * (g.ny + 1); }  // reals
  char* at(void* base, size_t elems) const { return (char*)base + elems * rs; }
};

namespace {

constexpr int kEnergyBlocks = 256;

// charge of the fixed-point sums (the row pass reads and clears them) -> E at `E` (padded,
// particle precision)
void field(pic_adjoint& a, void* E) {
  FieldArgs f;
  std::memset(&f, 0, sizeof(f));
  f.runs = a.dd; f.st = a.st;
  f.S = a.S; f.Sx = a.Sx; f.Sy = a.Sy; f.E = E; f.phi = nullptr; f.tw = a.tw;
  f.row_pair_items = a.rp; f.n_row_pair_items = a.nrp;
  f.col_items = a.cl; f.n_col_items = a.ncl;
  f.row_items = a.rw; f.n_row_items = a.nrw;
  f.nruns = 1; f.fp64 = a.fp64; f.want_phi = 0; f.max_nx = a.g.nx; f.max_ny = a.g.ny;
  f.row_lds = a.row_lds;
  f.rho = a.hi; f.rho_w = a.hi; f.rho_lo = a.lo; f.rho_lo_w = a.lo;
  f.gsum = (long long)a.g.nx * a.g.ny;
  f.clear_read = 1;
  launch_field(a.s, f, nullptr);
}

inline double kick(int k) { return k == 0 ? 0.5 : 1.0; }

// Forward steps 0..N-1 from the state (x, y, vx, vy) whose charge is already in hi/lo,
// saving (x, y, vx, vy) every K steps; then E_N and its energy partials.
template <typename R>
void record_forward(pic_adjoint& a) {
  R *x = (R*)a.x, *y = (R*)a.y, *vx = (R*)a.vx, *vy = (R*)a.vy;
  const size_t n = (size_t)a.n;
  for (int k = 0; k < a.nsteps; ++k) {
    R* Ek = (R*)a.at(a.trE, (size_t)k * a.epad());
    const int seg = k / a.K;
    const int* toff = a.toffs + (size_t)seg * (a.nb + 1);
    if (k > 0) field(a, Ek);  // the charge of x_k (deposited by the previous push)
    if (k % a.K == 0) {
      // checkpoint = the state sorted by tile; ids = each slot's slot before the sort
      R* ck = (R*)a.at(a.ckpt, (size_t)seg * 4 * n);
      PartBufs<R> src{x, y, vx, vy, a.iota};
      PartBufs<R> dst{ck, ck + n, ck + 2 * n, ck + 3 * n, a.ids + (size_t)seg * n};
      SortBufs sb = a.sb;
      sb.toff = a.toffs + (size_t)seg * (a.nb + 1);
      launch_tile_sort<R>(a.s, a.dd, a.st, 1, src, dst, sb, a.pitems, a.npitems);
      R* cur[4] = {x, y, vx, vy};
      for (int q = 0; q < 4; ++q)
        (void)hipMemcpyAsync(cur[q], ck + q * n, a.rs * n, hipMemcpyDeviceToDevice, a.s);
    }
    if (k == 0) {  // the initial charge, deposited once the particles are sorted
      adj_push_win<R>(a.s, a.g, a.items, toff, x, y, vx, vy, nullptr, 0.0, a.hi, a.lo);
      field(a, Ek);
    }
    adj_push_win<R>(a.s, a.g, a.items, toff, x, y, vx, vy, Ek, kick(k), a.hi, a.lo);
  }
  field(a, a.E);
  adj_energy<R>(a.s, a.g, (const R*)a.E, a.part, kEnergyBlocks);
}

// Reverse sweep for dL = 1 (the caller scales the outputs): seed from E_N, then for each
// segment restore its checkpoint, recompute its steps into the trajectory buffers (pushes
// in the stored fields of the forward) and run the reverse steps over them.
template <typename R>
void record_backward(pic_adjoint& a) {
  const size_t n = (size_t)a.n;
  adj_seed<R>(a.s, a.g, (const R*)a.E, 2.0 * a.g.dx * a.g.dy, a.lE);
  adj_mt(a.s, a.g, a.lE, a.A, a.lrho, a.tw);
  int cur = 0;  // adjoint arrays of the segment being reversed: lam[cur]
  adj_zero(a.s, a.lam[cur], 4 * a.n);
  auto L = [&](int k) { return a.lam[cur] + (size_t)k * n; };
  const double cq = a.qp / (a.g.dx * a.g.dy);
  auto X = [&](int slot) { return (R*)a.at(a.trx, (size_t)slot * n); };
  auto Y = [&](int slot) { return (R*)a.at(a.try_, (size_t)slot * n); };
  auto Ek = [&](int k) { return (R*)a.at(a.trE, (size_t)k * a.epad()); };
  for (int seg = a.nseg - 1; seg >= 0; --seg) {
    const int k0 = seg * a.K, k1 = std::min(k0 + a.K, a.nsteps);
    const int* toff = a.toffs + (size_t)seg * (a.nb + 1);
    if (seg < a.nseg - 1) {
      // adjoint arrays from the order of segment seg + 1 to the order of segment seg: slot d
      // of seg + 1 came from slot ids[seg + 1][d] of seg (mostly nearby: the same tile)
      adj_reorder(a.s, a.ids + (size_t)(seg + 1) * n, a.lam[cur], a.lam[1 - cur], a.n);
      cur = 1 - cur;
    }
    adj_restore<R>(a.s, a.g, (const R*)a.at(a.ckpt, (size_t)seg * 4 * n), a.n, X(0), Y(0),
                   (R*)a.vx, (R*)a.vy, nullptr, nullptr);
    for (int k = k0; k < k1; ++k) {
      const int sl = k - k0;
      adj_push<R>(a.s, a.g, X(sl), Y(sl), X(sl + 1), Y(sl + 1), (R*)a.vx, (R*)a.vy, Ek(k), a.n,
                  kick(k), nullptr, nullptr, nullptr);
    }
    for (int k = k1; k > k0; --k) {
      const int sl = k - k0;
      adj_pass_win<R>(a.s, a.g, a.items, toff, X(sl), Y(sl), a.lrho, cq, X(sl - 1), Y(sl - 1),
                      Ek(k - 1), kick(k - 1) * a.g.dt * a.g.qm, L(0), L(1), L(2), L(3), a.lE);
      adj_mt(a.s, a.g, a.lE, a.A, a.lrho, a.tw);
    }
  }
  // deposit Jacobian of step 0 at x_0 (slot 0 of segment 0)
  adj_pass<R>(a.s, a.g, X(0), Y(0), a.lrho, cq, nullptr, nullptr, nullptr, 0.0, L(0), L(1),
              L(2), L(3), a.lE, a.n);
  // results in the original particle order (segment 0 was sorted from it)
  adj_reorder(a.s, a.ids, a.lam[cur], a.lam[1 - cur], a.n);
  a.res = a.lam[1 - cur];
}

template <typename F>
int capture(pic_adjoint& a, hipGraphExec_t* out, F&& body) {
  hipGraph_t g = nullptr;
  PIC_HIP(hipStreamBeginCapture(a.s, hipStreamCaptureModeThreadLocal));
  body();
  hipError_t e = hipStreamEndCapture(a.s, &g);
  if (e != hipSuccess)
    return fail(PIC_ERR_HIP, std::string("adjoint stream capture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess)
    return fail(PIC_ERR_HIP, std::string("adjoint graph instantiate: ") + hipGetErrorString(e));
  return PIC_OK;
}

template <typename R>
int forward_t(pic_adjoint& a, const double* amps, const void* vx0, const void* vy0,
              double* loss) {
  if (!a.gfwd) {
    int rc = capture(a, &a.gfwd, [&] { record_forward<R>(a); });
    if (rc) return rc;
  }
  adj_init<R>(a.s, a.g, a.xl, a.yl, a.n, amps, a.mkx, a.mky, a.nm, (R*)a.x, (R*)a.y, nullptr,
              nullptr);
  PIC_HIP(hipMemcpyAsync(a.vx, vx0, a.rs * a.n, hipMemcpyDeviceToDevice, a.s));
  PIC_HIP(hipMemcpyAsync(a.vy, vy0, a.rs * a.n, hipMemcpyDeviceToDevice, a.s));
  PIC_HIP(hipGraphLaunch(a.gfwd, a.s));
  std::vector<double> h(kEnergyBlocks);
  PIC_HIP(hipMemcpyAsync(h.data(), a.part, sizeof(double) * kEnergyBlocks,
                         hipMemcpyDeviceToHost, a.s));
  PIC_HIP(hipStreamSynchronize(a.s));
  double sum = 0.0;
  for (double v : h) sum += v;
  if (loss) *loss = sum * a.g.dx * a.g.dy;
  a.forward_done = true;
  return PIC_OK;
}

template <typename R>
int backward_t(pic_adjoint& a, double dloss, double* grad_amps, void* gvx, void* gvy) {
  if (!a.gbwd) {
    int rc = capture(a, &a.gbwd, [&] { record_backward<R>(a); });
    if (rc) return rc;
  }
  PIC_HIP(hipGraphLaunch(a.gbwd, a.s));
  const double* lx0 = a.res;                 // original order
  const double* lvx0 = a.res + 2 * (size_t)a.n;
  const double* lvy0 = a.res + 3 * (size_t)a.n;
  if (grad_amps)
    adj_amp_grad(a.s, lx0, a.xl, a.yl, a.n, a.mkx, a.mky, a.nm, a.part, kEnergyBlocks, dloss,
                 grad_amps);
  if (gvx) adj_to_r<R>(a.s, lvx0, dloss, (R*)gvx, a.n);
  if (gvy) adj_to_r<R>(a.s, lvy0, dloss, (R*)gvy, a.n);
  PIC_HIP(hipGetLastError());
  PIC_HIP(hipStreamSynchronize(a.s));
  return PIC_OK;
}

}  // namespace

extern "C" {

int pic_adjoint_create(const pic_config* cfg, const double* xl, const double* yl, int64_t n,
                       int nmodes, const int32_t* mx, const int32_t* my, int nsteps,
                       int checkpoint_every, pic_adjoint** out) {
  if (!cfg || !xl || !yl || !out || n < 1 || n > 0x7fffffff || nmodes < 0 || nmodes > 64 ||
      nsteps < 1 ||
      checkpoint_every < 1 || (nmodes > 0 && (!mx || !my)))
    return fail(PIC_ERR_ARG, "bad arguments to pic_adjoint_create");
  int rc = validate_config(*cfg);
  if (rc) return rc;
  if (cfg->collide != PIC_COLLIDE_NONE)
    return fail(PIC_ERR_UNSUPPORTED, "the adjoint covers the collisionless step only");
  *out = nullptr;
  std::unique_ptr<pic_adjoint> a(new pic_adjoint());
  a->cfg = *cfg;
  a->device = cfg->device;
  PIC_HIP(hipSetDevice(a->device));
  a->fp64 = cfg->precision == PIC_FP64;
  a->rs = a->fp64 ? 8 : 4;
  a->n = n;
  a->nm = nmodes;
  a->nsteps = nsteps;
  a->K = checkpoint_every;
  a->nseg = (nsteps + a->K - 1) / a->K;
  int l = 0;
  AdjGrid& g = a->g;
  g.nx = cfg->nx; g.ny = cfg->ny;
  while ((1 << l) < g.nx) ++l;
  g.lognx = l;
  l = 0;
  while ((1 << l) < g.ny) ++l;
  g.logny = l;
  g.lx = cfg->lx; g.ly = cfg->ly; g.dx = g.lx / g.nx; g.dy = g.ly / g.ny;
  g.dt = cfg->dt; g.qm = cfg->qm;
  a->qp = cfg->qp != 0.0 ? cfg->qp : -(cfg->lx * cfg->ly) / (double)n;
  PIC_HIP(hipStreamCreateWithFlags(&a->s, hipStreamNonBlocking));
  const size_t G = (size_t)g.nx * g.ny;
  pic_adjoint& A = *a;
  PIC_HIP(A.get(&A.xl, n));
  PIC_HIP(A.get(&A.yl, n));
  PIC_HIP(A.get(&A.mkx, std::max(nmodes, 1)));
  PIC_HIP(A.get(&A.mky, std::max(nmodes, 1)));
  PIC_HIP(A.getb(&A.x, A.rs * n));
  PIC_HIP(A.getb(&A.y, A.rs * n));
  PIC_HIP(A.getb(&A.vx, A.rs * n));
  PIC_HIP(A.getb(&A.vy, A.rs * n));
  PIC_HIP(A.getb(&A.ckpt, A.rs * 4 * n * A.nseg));
  PIC_HIP(A.getb(&A.trx, A.rs * n * (A.K + 1)));
  PIC_HIP(A.getb(&A.try_, A.rs * n * (A.K + 1)));
  PIC_HIP(A.getb(&A.trE, A.rs * A.epad() * (size_t)nsteps));
  PIC_HIP(A.get(&A.lam[0], 4 * (size_t)n));
  PIC_HIP(A.get(&A.lam[1], 4 * (size_t)n));
  PIC_HIP(A.get(&A.lE, 2 * G));
  PIC_HIP(A.get(&A.lrho, G));
  PIC_HIP(A.get(&A.hi, G));
  PIC_HIP(A.get(&A.lo, G));
  PIC_HIP(A.getb(&A.E, A.rs * A.epad()));
  const size_t ss = (size_t)(g.ny / 2 + 1) * g.nx;
  PIC_HIP(A.get(&A.S, ss));
  PIC_HIP(A.get(&A.Sx, ss));
  PIC_HIP(A.get(&A.Sy, ss));
  PIC_HIP(A.get(&A.A, G));
  PIC_HIP(A.get(&A.tw, PIC_FFT_MAX / 2));
  PIC_HIP(A.get(&A.dd, 1));
  PIC_HIP(A.get(&A.st, 1));
  PIC_HIP(A.get(&A.part, kEnergyBlocks * std::max(nmodes, 1)));
  PIC_HIP(hipMemcpy(A.xl, xl, sizeof(double) * n, hipMemcpyHostToDevice));
  PIC_HIP(hipMemcpy(A.yl, yl, sizeof(double) * n, hipMemcpyHostToDevice));
  std::vector<double> kx(std::max(nmodes, 1), 0.0), ky(std::max(nmodes, 1), 0.0);
  const double twopi = 6.283185307179586476925;
  for (int m = 0; m < nmodes; ++m) {
    kx[m] = twopi * mx[m] / cfg->lx;
    ky[m] = twopi * my[m] / cfg->ly;
  }
  PIC_HIP(hipMemcpy(A.mkx, kx.data(), sizeof(double) * kx.size(), hipMemcpyHostToDevice));
  PIC_HIP(hipMemcpy(A.mky, ky.data(), sizeof(double) * ky.size(), hipMemcpyHostToDevice));
  std::vector<double2> h(PIC_FFT_MAX / 2);
  for (int j = 0; j < PIC_FFT_MAX / 2; ++j) {
    const long double ang = 2.0L * 3.141592653589793238462643383279L * j / PIC_FFT_MAX;
    h[j] = make_double2((double)cosl(ang), (double)-sinl(ang));
  }
  PIC_HIP(hipMemcpy(A.tw, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice));
  RunDesc d;
  std::memset(&d, 0, sizeof(d));
  d.nx = g.nx; d.ny = g.ny; d.lognx = g.lognx; d.logny = g.logny;
  d.lx = g.lx; d.ly = g.ly; d.dx = g.dx; d.dy = g.dy;
  d.rho_scale = a->qp / (g.dx * g.dy) * (a->fp64 ? 1.0 : std::ldexp(1.0, -36));
  fft_groups(g.nx, g.nx, g.ny, &d.fft_pp, &d.fft_rf);
  A.row_lds = std::max(d.fft_pp, d.fft_rf) * g.ny;
  d.row_item_begin = 0; d.row_item_end = g.nx / d.fft_rf;
  // Tiles of the windowed kernels: about an eighth of the grid side (4 .. 32 cells), a halo
  // of up to 16 cells (the drift of a segment), a window that never wraps onto itself, and
  // about 2048 particles per item.
  {
    const int m = std::min(g.nx, g.ny);
    int T = 4;
    while (T * 2 <= std::min(32, m / 8)) T *= 2;
    T = std::max(1, std::min(T, m / 2));
    g.tile = T;
    g.halo = std::max(0, std::min(16, (m - T - 1) / 2));
    g.win = T + 2 * g.halo + 1;
    g.tiles_y = g.ny / T;
    A.nb = (g.nx / T) * g.tiles_y;
    g.splits = (int)std::max<int64_t>(1, std::min<int64_t>(64, (n + 2048LL * A.nb - 1) /
                                                                  (2048LL * A.nb)));
    g.nitems = A.nb * g.splits;
    d.tile = T; d.halo = g.halo; d.tiles_x = g.nx / T; d.tiles_y = g.tiles_y;
    d.win = g.win; d.splits = g.splits; d.nbuckets = A.nb;
    // sort chunks of 1024 particles: the adjoint's runs are small (config 5: 512 chunks)
    d.schunk = 1024;
    d.nchunks = (int)((n + d.schunk - 1) / d.schunk);
    d.pcap = (int)n; d.tbase = 0; d.toffbase = 0; d.pbase = 0;
    std::vector<int2> ts, pi;
    for (long long s = 0; s < (long long)A.nb * d.nchunks; s += PIC_CELL_CHUNK)
      ts.push_back(make_int2(0, (int)s));
    d.tscan_item_begin = 0; d.tscan_item_end = (int)ts.size();
    for (int64_t s = 0; s < n; s += d.schunk) pi.push_back(make_int2(0, (int)s));
    std::vector<int4> it;
    for (int tx = 0; tx < g.nx / T; ++tx)
      for (int ty = 0; ty < g.tiles_y; ++ty)
        for (int sp = 0; sp < g.splits; ++sp) it.push_back(make_int4(tx, ty, sp, 0));
    int2* dts = nullptr;
    PIC_HIP(A.get(&dts, ts.size()));
    PIC_HIP(hipMemcpy(dts, ts.data(), sizeof(int2) * ts.size(), hipMemcpyHostToDevice));
    PIC_HIP(A.get(&A.pitems, pi.size()));
    PIC_HIP(hipMemcpy(A.pitems, pi.data(), sizeof(int2) * pi.size(), hipMemcpyHostToDevice));
    A.npitems = (int)pi.size();
    PIC_HIP(A.get(&A.items, it.size()));
    PIC_HIP(hipMemcpy(A.items, it.data(), sizeof(int4) * it.size(), hipMemcpyHostToDevice));
    std::memset(&A.sb, 0, sizeof(A.sb));
    PIC_HIP(A.get(&A.sb.table, (size_t)A.nb * d.nchunks));
    PIC_HIP(A.get(&A.sb.tchunk_sums, ts.size()));
    A.sb.tscan_items = dts;
    A.sb.ntscan_items = (int)ts.size();
    A.sb.max_buckets = A.nb;
    PIC_HIP(A.get(&A.ids, (size_t)A.nseg * n));
    PIC_HIP(A.get(&A.iota, n));
    PIC_HIP(A.get(&A.toffs, (size_t)A.nseg * (A.nb + 1)));
    adj_iota(A.s, A.iota, n);
    PIC_HIP(hipStreamSynchronize(A.s));
  }
  PIC_HIP(hipMemcpy(A.dd, &d, sizeof(d), hipMemcpyHostToDevice));
  RunState st0;
  std::memset(&st0, 0, sizeof(st0));
  st0.np = (int)n;
  PIC_HIP(hipMemcpy(A.st, &st0, sizeof(RunState), hipMemcpyHostToDevice));
  PIC_HIP(hipMemset(A.hi, 0, sizeof(unsigned long long) * G));
  PIC_HIP(hipMemset(A.lo, 0, sizeof(unsigned long long) * G));
  std::vector<int2> rp, cl, rw;
  for (int i = 0; i < g.nx; i += 2 * d.fft_pp) rp.push_back(make_int2(0, i));
  for (int k = 0; k <= g.ny / 2; ++k) cl.push_back(make_int2(0, k));
  for (int i = 0; i < g.nx; i += d.fft_rf) rw.push_back(make_int2(0, i));
  PIC_HIP(A.get(&A.rp, rp.size()));
  PIC_HIP(A.get(&A.cl, cl.size()));
  PIC_HIP(A.get(&A.rw, rw.size()));
  PIC_HIP(hipMemcpy(A.rp, rp.data(), sizeof(int2) * rp.size(), hipMemcpyHostToDevice));
  PIC_HIP(hipMemcpy(A.cl, cl.data(), sizeof(int2) * cl.size(), hipMemcpyHostToDevice));
  PIC_HIP(hipMemcpy(A.rw, rw.data(), sizeof(int2) * rw.size(), hipMemcpyHostToDevice));
  A.nrp = (int)rp.size(); A.ncl = (int)cl.size(); A.nrw = (int)rw.size();
  *out = a.release();
  return PIC_OK;
}

int pic_adjoint_destroy(pic_adjoint* adj) {
  delete adj;
  return PIC_OK;
}

int pic_adjoint_forward(pic_adjoint* a, const double* amps, const void* vx0, const void* vy0,
                        double* loss) {
  if (!a || !vx0 || !vy0 || (a->nm > 0 && !amps)) return fail(PIC_ERR_ARG, "bad arguments");
  PIC_HIP(hipSetDevice(a->device));
  return a->fp64 ? forward_t<double>(*a, amps, vx0, vy0, loss)
                 : forward_t<float>(*a, amps, vx0, vy0, loss);
}

int pic_adjoint_backward(pic_adjoint* a, double dloss, double* grad_amps, void* grad_vx,
                         void* grad_vy) {
  if (!a) return fail(PIC_ERR_ARG, "NULL adjoint");
  if (!a->forward_done) return fail(PIC_ERR_STATE, "pic_adjoint_backward before forward");
  PIC_HIP(hipSetDevice(a->device));
  return a->fp64 ? backward_t<double>(*a, dloss, grad_amps, grad_vx, grad_vy)
                 : backward_t<float>(*a, dloss, grad_amps, grad_vx, grad_vy);
}

int pic_adjoint_stream(pic_adjoint* a, void** stream) {
  if (!a || !stream) return fail(PIC_ERR_ARG, "NULL argument");
  *stream = (void*)a->s;
  return PIC_OK;
}

}  // extern "C"
