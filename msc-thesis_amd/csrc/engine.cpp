// engine.cpp -- host driver of the PIC step: context, buffers, step sequencing, graphs,
// and the context half of the C ABI (include/pic_capi.h).
//
// The reference host function picFloatGpu (thesisCode/collisionless/double/src/pic.cu:
// 429-611) allocates, plans, copies, loops and frees in one call; here those phases are
// separate so the step loop runs on resident data: pic_create allocates once,
// pic_set_particles uploads and sorts, pic_step enqueues captured macro-steps.
#include "engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>

namespace pic {

thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
int fail(int code, const std::string& m) {
  g_err = m;
  return code;
}

static bool is_pow2(long long v) { return v > 0 && (v & (v - 1)) == 0; }
static int ilog2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}
static long long round_up(long long v, long long m) { return (v + m - 1) / m * m; }

int validate_config(const pic_config& c) {
  if (!is_pow2(c.nx) || !is_pow2(c.ny) || c.nx < 4 || c.ny < 4 || c.nx > PIC_FFT_MAX ||
      c.ny > PIC_FFT_MAX)
    return fail(PIC_ERR_INVALID_CONFIG, "nx and ny must be powers of two in [4, 2048]");
  if (!(c.lx > 0) || !(c.ly > 0) || !std::isfinite(c.lx) || !std::isfinite(c.ly))
    return fail(PIC_ERR_INVALID_CONFIG, "lx and ly must be positive and finite");
  if (!(c.dt > 0) || !std::isfinite(c.dt))
    return fail(PIC_ERR_INVALID_CONFIG, "dt must be positive and finite");
  if (c.qm == 0 || !std::isfinite(c.qm) || !std::isfinite(c.qp))
    return fail(PIC_ERR_INVALID_CONFIG, "qm must be non-zero and finite");
  if (c.precision != PIC_FP32 && c.precision != PIC_FP64)
    return fail(PIC_ERR_INVALID_CONFIG, "precision must be PIC_FP32 or PIC_FP64");
  if (c.capacity < 0 || c.capacity > (int64_t)0x7fff0000)
    return fail(PIC_ERR_INVALID_CONFIG, "capacity must be in [0, 2^31 - 2^16]");
  if (c.collide < PIC_COLLIDE_NONE || c.collide > PIC_COLLIDE_DSMC)
    return fail(PIC_ERR_INVALID_CONFIG, "collide must be a PIC_COLLIDE_* value");
  if (c.collide != PIC_COLLIDE_NONE && c.precision != PIC_FP32)
    return fail(PIC_ERR_UNSUPPORTED, "collision stages run on FP32 particles");
  if (c.collide == PIC_COLLIDE_MCC &&
      (!(c.mcc_ng > 0) || !(c.mcc_v_unit > 0) || !(c.mcc_dt_phys > 0) || !(c.mcc_neutral_mass > 0)))
    return fail(PIC_ERR_INVALID_CONFIG, "MCC needs mcc_ng, mcc_v_unit, mcc_dt_phys, mcc_neutral_mass > 0");
  if (c.collide == PIC_COLLIDE_DSMC && (!(c.dsmc_coeff >= 0) || !(c.dsmc_gmax0 > 0)))
    return fail(PIC_ERR_INVALID_CONFIG, "DSMC needs dsmc_coeff >= 0 and dsmc_gmax0 > 0");
  if (c.tile != 0 && (!is_pow2(c.tile) || c.tile > std::min(c.nx, c.ny) / 2 || c.tile > 64))
    return fail(PIC_ERR_INVALID_CONFIG, "tile must be a power of two <= min(nx, ny)/2 and <= 64");
  if (c.halo < 0 || c.splits < 0 || c.splits > 64 || c.sort_every < 0)
    return fail(PIC_ERR_INVALID_CONFIG, "halo, splits (<= 64) and sort_every must be >= 0");
  return PIC_OK;
}

static const double kE = 1.602176634e-19, kMe = 9.1093837015e-31, kEps0 = 8.8541878128e-12;

double mcc_nu_max(const pic_config& c) {
  // max over E of ng * sigma_T(E) * v(E), E in (0, 1000] eV, v = sqrt(2 E e / m_e)
  double best = 0.0;
  for (int i = 1; i <= 200000; ++i) {
    const double e = 1000.0 * i / 200000.0;
    const double sel = 1e-19 / (1.0 + e / 10.0);
    const double sex = (e > 11.5) ? 2e-21 * (1.0 - 11.5 / e) : 0.0;
    const double siz = (e > 15.8) ? 3e-20 * std::log(e / 15.8) / e : 0.0;
    const double v = std::sqrt(2.0 * e * kE / kMe);
    best = std::max(best, c.mcc_ng * (sel + sex + siz) * v);
  }
  return best * 1.0001;  // margin so that no particle's total rate exceeds nu_max
}

static Tuning default_tuning(const pic_config& c, int tile, int halo, int splits) {
  Tuning t;
  const int m = std::min(c.nx, c.ny);
  t.tile = std::min(tile ? tile : 16, m / 2);
  // at most 4096 tiles (the sort keeps one counter per tile in LDS)
  while ((c.nx / t.tile) * (c.ny / t.tile) > 4096 && t.tile < m / 2) t.tile *= 2;
  // the window (tile + 2 halo + 1 nodes) must not wrap onto itself: win <= min(nx, ny)
  const int hmax = std::max(0, (m - t.tile - 1) / 2);
  t.halo = std::min(halo ? halo : 16, hmax);
  const int ntiles = (c.nx / t.tile) * (c.ny / t.tile);
  // about 8192 particles per 512-thread workgroup (1 to 64 workgroups per tile)
  const long long per_tile = std::max<long long>(c.capacity, 1) / ntiles;
  t.splits = splits ? splits : (int)std::max(1LL, std::min(64LL, (per_tile + 8191) / 8192));
  t.per_tile = splits == 0 && c.collide == PIC_COLLIDE_DSMC;
  return t;
}

// Items (run, tile x, tile y, split | splits << 16) of run r: t.splits per tile, or with
// per-tile splits (and the tile offsets of the current order on the host) one split per
// PUSH_SPLIT_TARGET particles of the tile, 1 to 64.
void Ctx::push_items_for(int r, const int* toff_h) {
  RunDesc& d = rd[r];
  const Tuning& t = tune[r];
  constexpr long long PUSH_SPLIT_TARGET = 4096;
  d.push_item_begin = (int)push_items.size();
  for (int tx = 0; tx < d.tiles_x; ++tx)
    for (int ty = 0; ty < d.tiles_y; ++ty) {
      int S = t.splits;
      if (t.per_tile && toff_h) {
        const long long tb = d.toffbase + (long long)tx * d.tiles_y + ty;
        const long long n = toff_h[tb + 1] - toff_h[tb];
        S = (int)std::max(1LL, std::min(64LL, (n + PUSH_SPLIT_TARGET - 1) / PUSH_SPLIT_TARGET));
      }
      for (int s = 0; s < S; ++s) push_items.push_back(make_int4(r, tx, ty, s | (S << 16)));
    }
  if (cfg[r].collide == PIC_COLLIDE_MCC) {
    // tail items: the particles appended since the last sort (births), PIC_TAIL_CHUNK
    // indices each
    const long long preg = round_up(std::max<long long>(d.pcap, 1), PIC_TAIL_CHUNK);
    for (long long s = 0; s < preg; s += PIC_TAIL_CHUNK)
      push_items.push_back(make_int4(r, -1, (int)s, 0));
  }
  d.push_item_end = (int)push_items.size();
}

// Per-tile splits follow the tile counts of the current order: rebuilt from them (after a
// sort) when a run asks for them.  Only the push items, their partials and the captured
// sequences change.
int Ctx::rebalance_push_items() {
  bool any = false;
  for (int r = 0; r < nruns; ++r) any = any || (tune[r].per_tile && has_particles[r]);
  if (!any) return PIC_OK;
  PIC_HIP(hipStreamSynchronize(stream));
  std::vector<int> toff_h((size_t)toffsum);
  PIC_HIP(hipMemcpy(toff_h.data(), toffs[cur], sizeof(int) * toffsum, hipMemcpyDeviceToHost));
  push_items.clear();
  for (int r = 0; r < nruns; ++r) push_items_for(r, has_particles[r] ? toff_h.data() : nullptr);
  if (d_push) (void)hipFree(d_push);
  PIC_HIP(dalloc(&d_push, push_items.size()));
  PIC_HIP(hipMemcpy(d_push, push_items.data(), sizeof(int4) * push_items.size(),
                    hipMemcpyHostToDevice));
  if (partials_push) (void)hipFree(partials_push);
  PIC_HIP(dalloc(&partials_push, push_items.size() * 3));
  free_graphs();
  return upload_runs();
}

Ctx::~Ctx() {
  if (device >= 0) (void)hipSetDevice(device);
  free_graphs();
  auto F = [](void* p) { if (p) (void)hipFree(p); };
  F(d_runs); F(d_st);
  for (int i = 0; i < 2; ++i) { F(px[i]); F(py[i]); F(pvx[i]); F(pvy[i]); F(pid[i]); }
  F(rho); F(rho_lo); F(E); F(S); F(Sx); F(Sy); F(phi); F(tw);
  F(sb.keys); F(sb.rank); F(sb.perm); F(sb.counts); F(sb.offsets); F(sb.chunk_sums);
  F(sb.big_list); F(sb.big_count); F(sb.scratch); F(sb.table); F(sb.tchunk_sums);
  F(toffs[0]); F(toffs[1]); F(sb.ftable); F(sb.fchunk_sums); F(d_fscan_items);
  F(d_pitems); F(d_citems); F(d_rowpair); F(d_cols); F(d_rows); F(d_push);
  F(d_segitems); F(d_tscan_items);
  F(partials_push); F(partials_rows); F(diag);
  F(d_coll_counts); F(birth_count); F(birth_parent); F(birth_buf); F(gmax); F(flag);
  F(coll_list); F(coll_count); F(coll_partials);
  for (auto& e : ev) if (e) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

void Ctx::free_graphs() {
  for (int i = 0; i < 2; ++i) {
    if (macro[i]) (void)hipGraphExecDestroy(macro[i]);
    macro[i] = nullptr;
    macro_k[i] = 0;
  }
}

int Ctx::upload_runs() {
  PIC_HIP(hipMemcpy(d_runs, rd.data(), sizeof(RunDesc) * nruns, hipMemcpyHostToDevice));
  return PIC_OK;
}

int Ctx::build_items() {
  pitems.clear(); citems.clear(); rowpair.clear(); cols.clear(); rows.clear(); push_items.clear();
  segitems.clear(); tscan_items.clear();
  lds_push = 0;
  long long tsum = 0;
  toffsum = 0;
  int maxb = 1;
  long long rows_total = 0;  // rows of all runs: sets the FFT row grouping
  for (int r = 0; r < nruns; ++r) rows_total += cfg[r].nx;
  for (int r = 0; r < nruns; ++r) {
    RunDesc& d = rd[r];
    const Tuning& t = tune[r];
    d.tile = t.tile; d.halo = t.halo; d.splits = t.splits;
    d.tiles_x = d.nx / t.tile; d.tiles_y = d.ny / t.tile;
    d.win = t.tile + 2 * t.halo + 1;
    // LDS: charge window (1 or 2 int64 words per node, padded to 16 B) + field window
    const int nwin = d.win * d.win * (fp64 ? 2 : 1);
    lds_push = std::max(lds_push, ((nwin + 1) & ~1) * 8 + d.win * d.win * (fp64 ? 16 : 8));
    const long long preg = round_up(std::max<long long>(d.pcap, 1), PIC_CHUNK);
    d.nbuckets = d.tiles_x * d.tiles_y;
    d.nchunks = (int)(preg / PIC_CHUNK);
    d.schunk = PIC_CHUNK;
    if (d.nbuckets > 4096) return fail(PIC_ERR_INVALID_CONFIG, "more than 4096 tiles: use a larger tile");
    if ((long long)d.nbuckets * d.nchunks >= (1ll << 31))
      return fail(PIC_ERR_INVALID_CONFIG, "tile table too large: use a larger tile");
    maxb = std::max(maxb, d.nbuckets);
    d.tbase = tsum; tsum += (long long)d.nbuckets * d.nchunks;
    d.toffbase = toffsum; toffsum += d.nbuckets + 1;
    d.key_tile_major = 1;
    d.tscan_item_begin = (int)tscan_items.size();
    for (long long s = 0; s < (long long)d.nbuckets * d.nchunks; s += PIC_CELL_CHUNK)
      tscan_items.push_back(make_int2(r, (int)s));
    d.tscan_item_end = (int)tscan_items.size();
    for (long long s = 0; s < preg; s += PIC_CHUNK) pitems.push_back(make_int2(r, (int)s));
    d.cell_item_begin = (int)citems.size();
    for (int c = 0; c < d.nx * d.ny; c += PIC_CELL_CHUNK) citems.push_back(make_int2(r, c));
    d.cell_item_end = (int)citems.size();
    for (int c = 0; c < d.nx * d.ny; c += PIC_SEG_CELLS) segitems.push_back(make_int2(r, c));
    push_items_for(r, nullptr);
    d.fchunks = 0;
    fft_groups(rows_total, d.nx, d.ny, &d.fft_pp, &d.fft_rf);
    for (int i = 0; i < d.nx; i += 2 * d.fft_pp) rowpair.push_back(make_int2(r, i));
    for (int k = 0; k <= d.ny / 2; ++k) cols.push_back(make_int2(r, k));
    d.row_item_begin = (int)rows.size();
    for (int i = 0; i < d.nx; i += d.fft_rf) rows.push_back(make_int2(r, i));
    d.row_item_end = (int)rows.size();
  }
  if (lds_push > 160 * 1024) return fail(PIC_ERR_INVALID_CONFIG, "deposit window exceeds LDS");
  // Sort fused into the push: needs no cell lists (no DSMC), a cadence of at least two
  // steps (the count is taken one step ahead), and room in LDS for one counter per tile
  // (int) and two per tile and wave (16-bit pass counts).  MCC births join it through the
  // tail items and the births kernel (the count step's births are counted as they are
  // appended).
  const int fused_lds = (PUSH_NW + 1) * maxb * (int)sizeof(int);
  fused = !any_dsmc && sort_every >= 2 && maxb <= 512 &&
          lds_push + fused_lds <= 160 * 1024;
  fscan_items.clear();
  long long fsum = 0;
  for (int r = 0; r < nruns && fused; ++r) {
    RunDesc& d = rd[r];
    d.fchunks = d.push_item_end - d.push_item_begin;
    d.ftbase = fsum;
    fsum += (long long)d.nbuckets * d.fchunks;
    d.fscan_item_begin = (int)fscan_items.size();
    for (int g = 0; (long long)g * FSCAN_ROWS < d.fchunks; ++g)
      fscan_items.push_back(make_int2(r, g));  // groups of FSCAN_ROWS table rows
    d.fscan_item_end = (int)fscan_items.size();
  }
  if (fsum >= (1ll << 31)) fused = false;
  if (fused) lds_push += fused_lds;
  // Tile sorts key on positions extrapolated to the middle of the next cycle.  Fused: the
  // pushes of one sorted layout gather at 1 .. sort_every and deposit at 2 .. sort_every + 1
  // steps after the state the keys are taken from; separate (at a step's start): 0 ..
  // sort_every - 1 and 1 .. sort_every.  Not in runs with collisions (their cell lists and
  // births use the tiles of the positions themselves).
  for (int r = 0; r < nruns; ++r)
    rd[r].sort_lead = cfg[r].collide != PIC_COLLIDE_NONE ? 0.0
                      : fused ? 0.5 * (sort_every + 2) * rd[r].dt
                              : 0.5 * sort_every * rd[r].dt;
  auto up = [&](auto& vec, auto** dptr) -> int {
    using T = typename std::remove_reference<decltype(vec)>::type::value_type;
    if (*dptr) (void)hipFree(*dptr);
    PIC_HIP(dalloc(dptr, vec.size()));
    if (!vec.empty())
      PIC_HIP(hipMemcpy(*dptr, vec.data(), sizeof(T) * vec.size(), hipMemcpyHostToDevice));
    return PIC_OK;
  };
  int rc;
  if ((rc = up(pitems, &d_pitems))) return rc;
  if ((rc = up(citems, &d_citems))) return rc;
  if ((rc = up(rowpair, &d_rowpair))) return rc;
  if ((rc = up(cols, &d_cols))) return rc;
  if ((rc = up(rows, &d_rows))) return rc;
  if ((rc = up(push_items, &d_push))) return rc;
  if ((rc = up(segitems, &d_segitems))) return rc;
  if ((rc = up(tscan_items, &d_tscan_items))) return rc;
  if ((rc = up(fscan_items, &d_fscan_items))) return rc;
  sb.fscan_items = d_fscan_items;
  sb.nfscan_items = (int)fscan_items.size();
  if (sb.ftable) (void)hipFree(sb.ftable);
  PIC_HIP(dalloc(&sb.ftable, fused ? fsum : 1));
  if (sb.fchunk_sums) (void)hipFree(sb.fchunk_sums);
  PIC_HIP(dalloc(&sb.fchunk_sums, std::max<size_t>(1, fscan_items.size() * (size_t)maxb)));
  sb.segitems = d_segitems;
  sb.nsegitems = (int)segitems.size();
  sb.tscan_items = d_tscan_items;
  sb.ntscan_items = (int)tscan_items.size();
  sb.max_buckets = maxb;
  if (sb.table) (void)hipFree(sb.table);
  PIC_HIP(dalloc(&sb.table, tsum));
  if (sb.tchunk_sums) (void)hipFree(sb.tchunk_sums);
  PIC_HIP(dalloc(&sb.tchunk_sums, tscan_items.size()));
  for (int k = 0; k < 2; ++k) {
    if (toffs[k]) (void)hipFree(toffs[k]);
    PIC_HIP(dalloc(&toffs[k], toffsum));
    PIC_HIP(hipMemset(toffs[k], 0, sizeof(int) * toffsum));
  }
  sb.toff = toffs[1 - cur];
  if (partials_push) (void)hipFree(partials_push);
  PIC_HIP(dalloc(&partials_push, push_items.size() * 3));
  if (partials_rows) (void)hipFree(partials_rows);
  PIC_HIP(dalloc(&partials_rows, rows.size()));
  if (sb.chunk_sums) (void)hipFree(sb.chunk_sums);
  PIC_HIP(dalloc(&sb.chunk_sums, citems.size()));
  free_graphs();
  return upload_runs();
}

int Ctx::alloc() {
  PIC_HIP(hipSetDevice(device));
  ptotal = gsum = esum = ssum = ctotal = 0;
  max_nx = max_ny = 0;
  any_mcc = any_dsmc = false;
  for (int r = 0; r < nruns; ++r) {
    const pic_config& c = cfg[r];
    RunDesc& d = rd[r];
    std::memset(&d, 0, sizeof(d));
    d.nx = c.nx; d.ny = c.ny; d.lognx = ilog2(c.nx); d.logny = ilog2(c.ny);
    d.lx = c.lx; d.ly = c.ly; d.dx = c.lx / c.nx; d.dy = c.ly / c.ny;
    d.dt = c.dt; d.qm = c.qm; d.qp = c.qp;
    d.rho_scale = d.qp / (d.dx * d.dy) * (fp64 ? 1.0 : std::ldexp(1.0, -36));
    d.pcap = (int)c.capacity;
    d.pbase = ptotal; ptotal += round_up(std::max<long long>(c.capacity, 1), PIC_CHUNK);
    d.gbase = gsum; gsum += (long long)c.nx * c.ny;
    d.ebase = esum; esum += (long long)(c.nx + 1) * (c.ny + 1);
    d.sbase = ssum; ssum += (long long)(c.ny / 2 + 1) * c.nx;
    d.cbase = ctotal; ctotal += (long long)c.nx * c.ny + 1;
    d.dbase = (long long)r * PIC_DIAG_CAP;
    max_nx = std::max(max_nx, c.nx); max_ny = std::max(max_ny, c.ny);
    any_mcc |= c.collide == PIC_COLLIDE_MCC;
    any_dsmc |= c.collide == PIC_COLLIDE_DSMC;
  }
  // device memory the buffers below take (particle double buffers, sort keys / ranks /
  // permutation / segment scratch, MCC birth buffers, grids): refuse before allocating
  // anything when the device cannot hold it
  {
    long long maxcap = 0;
    for (auto& c : cfg) maxcap = std::max<long long>(maxcap, c.capacity);
    const double need = (double)ptotal * (8.0 * rsize + 8 + 28) +
                        (any_mcc ? 24.0 * maxcap : 0.0) + (double)gsum * (fp64 ? 40 : 24) +
                        (double)esum * 2 * rsize + 48.0 * ssum + 12.0 * ctotal;
    size_t free_b = 0, total_b = 0;
    PIC_HIP(hipMemGetInfo(&free_b, &total_b));
    if (need > (double)free_b)
      return fail(PIC_ERR_OOM, "the runs need " + std::to_string((long long)(need / 1e9)) +
                                   " GB of device memory, " +
                                   std::to_string((long long)(free_b / 1e9)) + " GB are free");
  }
  PIC_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  for (auto& e : ev) PIC_HIP(hipEventCreate(&e));
  PIC_HIP(dalloc(&d_runs, nruns));
  PIC_HIP(dalloc(&d_st, nruns));
  const size_t rs = rsize;
  for (int i = 0; i < 2; ++i) {
    PIC_HIP(hipMalloc(&px[i], rs * ptotal));
    PIC_HIP(hipMalloc(&py[i], rs * ptotal));
    PIC_HIP(hipMalloc(&pvx[i], rs * ptotal));
    PIC_HIP(hipMalloc(&pvy[i], rs * ptotal));
    PIC_HIP(dalloc(&pid[i], ptotal));
  }
  PIC_HIP(dalloc(&rho, 2 * gsum));
  if (fp64) PIC_HIP(dalloc(&rho_lo, 2 * gsum));
  PIC_HIP(hipMalloc(&E, rs * 2 * esum));
  PIC_HIP(dalloc(&S, ssum));
  PIC_HIP(dalloc(&Sx, ssum));
  PIC_HIP(dalloc(&Sy, ssum));
  PIC_HIP(dalloc(&phi, gsum));
  PIC_HIP(dalloc(&tw, PIC_FFT_MAX / 2));
  PIC_HIP(dalloc(&sb.keys, ptotal));
  PIC_HIP(dalloc(&sb.rank, ptotal));
  PIC_HIP(dalloc(&sb.perm, ptotal));
  PIC_HIP(dalloc(&sb.counts, ctotal));
  // the cell sorts keep the counts zero between sorts (their scan clears what it read)
  PIC_HIP(hipMemset(sb.counts, 0, sizeof(int) * (size_t)ctotal));
  PIC_HIP(dalloc(&sb.offsets, ctotal));
  PIC_HIP(dalloc(&sb.big_list, 2 * ctotal));
  PIC_HIP(dalloc(&sb.big_count, 1));
  PIC_HIP(dalloc(&sb.scratch, 2 * ptotal));
  PIC_HIP(dalloc(&diag, (size_t)nruns * PIC_DIAG_CAP * 8));
  PIC_HIP(dalloc(&d_coll_counts, 4 * nruns));
  PIC_HIP(dalloc(&flag, 1));
  if (any_mcc) {
    long long maxcap = 0;
    for (auto& c : cfg) maxcap = std::max<long long>(maxcap, c.capacity);
    PIC_HIP(dalloc(&birth_count, 1));
    PIC_HIP(dalloc(&birth_parent, maxcap));
    PIC_HIP(dalloc(&birth_buf, 4 * maxcap));
    PIC_HIP(hipMemset(birth_count, 0, sizeof(int)));
    int lb, lp;
    mcc_list_layout(maxcap, &lb, &lp);
    PIC_HIP(dalloc(&coll_list, (size_t)lb * lp));
    PIC_HIP(dalloc(&coll_count, PIC_MCC_BLOCKS));
    PIC_HIP(dalloc(&coll_partials, 4 * PIC_MCC_BLOCKS));
  }
  if (any_dsmc) PIC_HIP(dalloc(&gmax, ctotal));
  // twiddles exp(-2 pi i j / PIC_FFT_MAX)
  std::vector<double2> h(PIC_FFT_MAX / 2);
  for (int j = 0; j < PIC_FFT_MAX / 2; ++j) {
    const long double a = 2.0L * 3.141592653589793238462643383279L * j / PIC_FFT_MAX;
    h[j] = make_double2((double)cosl(a), (double)-sinl(a));
  }
  PIC_HIP(hipMemcpy(tw, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice));
  PIC_HIP(hipMemset(d_st, 0, sizeof(RunState) * nruns));
  PIC_HIP(hipMemset(rho, 0, sizeof(unsigned long long) * 2 * gsum));
  if (fp64) PIC_HIP(hipMemset(rho_lo, 0, sizeof(unsigned long long) * 2 * gsum));
  PIC_HIP(hipMemset(diag, 0, sizeof(double) * (size_t)nruns * PIC_DIAG_CAP * 8));
  PIC_HIP(hipMemset(d_coll_counts, 0, sizeof(long long) * 4 * nruns));
  PIC_HIP(hipMemset(E, 0, rs * 2 * esum));
  tune.resize(nruns);
  for (int r = 0; r < nruns; ++r)
    tune[r] = default_tuning(cfg[r], cfg[r].tile, cfg[r].halo, cfg[r].splits);
  int rc = build_items();
  if (rc) return rc;
  allocated = true;
  return PIC_OK;
}

FieldArgs Ctx::field_args(int want_phi) const {
  FieldArgs a;
  std::memset(&a, 0, sizeof(a));
  a.runs = d_runs; a.st = d_st;
  a.rho = rho; a.rho_w = rho; a.rho_lo = rho_lo; a.rho_lo_w = rho_lo; a.gsum = gsum;
  a.S = S; a.Sx = Sx; a.Sy = Sy; a.E = E; a.phi = phi; a.tw = tw;
  a.row_pair_items = d_rowpair; a.n_row_pair_items = (int)rowpair.size();
  a.col_items = d_cols; a.n_col_items = (int)cols.size();
  a.row_items = d_rows; a.n_row_items = (int)rows.size();
  a.partials = partials_rows; a.diag = diag;
  a.nruns = nruns; a.fp64 = fp64; a.want_phi = want_phi;
  a.max_nx = max_nx; a.max_ny = max_ny;
  a.row_lds = 0;
  for (const RunDesc& d : rd) a.row_lds = std::max(a.row_lds, std::max(d.fft_pp, d.fft_rf) * d.ny);
  a.rho_in = nullptr;
  return a;
}

template <typename R>
PushArgs<R> Ctx::push_args(int mode, int sort_mode) const {
  PushArgs<R> a;
  std::memset(&a, 0, sizeof(a));
  a.runs = d_runs; a.st = d_st;
  PartBufs<R> b = bufs<R>(cur);
  a.x = b.x; a.y = b.y; a.vx = b.vx; a.vy = b.vy;
  a.E = (const R*)E;
  a.rho = rho; a.rho_lo = rho_lo; a.gsum = gsum;
  a.off = toffs[cur];
  a.items = d_push; a.nitems = (int)push_items.size();
  a.nruns = nruns;
  a.partials = partials_push;
  a.mode = mode; a.lds_bytes = lds_push;
  a.sort_mode = fused ? sort_mode : 0;
  a.ftable = sb.ftable;
  a.dst = bufs<R>(1 - cur);
  a.id = pid[cur];
  return a;
}

int Ctx::launch_field_stage(int want_phi, hipEvent_t* ev4) {
  FieldArgs a = field_args(want_phi);
  launch_field(stream, a, ev4);
  return PIC_OK;
}

int Ctx::launch_push_stage(int mode, int sort_mode) {
  if (fp64) launch_push<double>(stream, push_args<double>(mode, sort_mode));
  else launch_push<float>(stream, push_args<float>(mode, sort_mode));
  return PIC_OK;
}

int Ctx::launch_sort_stage() {
  sb.toff = toffs[1 - cur];  // offsets of the order being written
  if (!any_dsmc) {
    // tile sort: the step only needs tiles to be contiguous
    if (fp64)
      launch_tile_sort<double>(stream, d_runs, d_st, nruns, bufs<double>(cur),
                               bufs<double>(1 - cur), sb, d_pitems, (int)pitems.size());
    else
      launch_tile_sort<float>(stream, d_runs, d_st, nruns, bufs<float>(cur), bufs<float>(1 - cur),
                              sb, d_pitems, (int)pitems.size());
    cur ^= 1;
    return PIC_OK;
  }
  // DSMC needs per-cell lists: full cell sort with tile-major cell numbering
  if (fp64)
    launch_sort<double>(stream, d_runs, d_st, nruns, bufs<double>(cur), bufs<double>(1 - cur), sb,
                        d_pitems, (int)pitems.size(), d_citems, (int)citems.size(), ctotal, true);
  else
    launch_sort<float>(stream, d_runs, d_st, nruns, bufs<float>(cur), bufs<float>(1 - cur), sb,
                       d_pitems, (int)pitems.size(), d_citems, (int)citems.size(), ctotal, true);
  cur ^= 1;
  return PIC_OK;
}

double Ctx::numax_dt_of(int r) {
  if ((int)numax_dt.size() != nruns) numax_dt.assign(nruns, -1.0);
  if (numax_dt[r] < 0) numax_dt[r] = mcc_nu_max(cfg[r]) * cfg[r].mcc_dt_phys * cfg[r].dt;
  return numax_dt[r];
}

int Ctx::launch_collide_stage() {
  for (int r = 0; r < nruns; ++r) {
    const pic_config& c = cfg[r];
    const RunDesc& d = rd[r];
    if (c.collide == PIC_COLLIDE_MCC) {
      MccArgs a;
      std::memset(&a, 0, sizeof(a));
      PartBufs<float> b = bufs<float>(cur);
      a.x = b.x; a.y = b.y; a.vx = b.vx; a.vy = b.vy; a.id = b.id;
      a.st = d_st + r; a.pbase = d.pbase; a.cap = d.pcap;
      a.v_unit2_half_me_over_e = 0.5 * kMe * c.mcc_v_unit * c.mcc_v_unit / kE;
      a.ng_vunit_dtphys = c.mcc_ng * c.mcc_v_unit * c.mcc_dt_phys * c.dt;
      const double numax_dt = numax_dt_of(r);
      a.nu_max_dt = numax_dt;
      a.p_coll = 1.0 - std::exp(-numax_dt);
      a.mass_ratio2 = 2.0 * kMe / c.mcc_neutral_mass;
      a.seed = c.seed;
      a.counts = d_coll_counts + 4 * r;
      a.birth_count = birth_count; a.birth_parent = birth_parent; a.birth_buf = birth_buf;
      a.coll_list = coll_list; a.coll_count = coll_count; a.coll_partials = coll_partials;
      a.count_table = collide_count_table;
      a.from_state_step = 1;
      a.finish_after = finish_late() ? 1 : 0;
      a.diag = diag + d.dbase * 8;
      launch_mcc(stream, a, d.pcap);
      launch_mcc_births(stream, a, sb.scratch, rho, gsum, d_runs, r, d.pcap);
    }
  }
  if (any_dsmc) {
    launch_sort_stage();
    for (int r = 0; r < nruns; ++r) {
      const pic_config& c = cfg[r];
      if (c.collide != PIC_COLLIDE_DSMC) continue;
      const RunDesc& d = rd[r];
      PartBufs<float> b = bufs<float>(cur);
      DsmcArgs a;
      std::memset(&a, 0, sizeof(a));
      a.vx = b.vx + d.pbase; a.vy = b.vy + d.pbase;
      a.offsets = sb.offsets + d.cbase; a.gmax = gmax + d.cbase;
      a.ncells = d.nx * d.ny; a.coeff = c.dsmc_coeff; a.seed = c.seed;
      // the cell sort numbers cells tile-major; gmax and the streams are row-major
      a.tile_shift = __builtin_ctz((unsigned)d.tile); a.tiles_y = d.tiles_y; a.ny = d.ny;
      a.st = d_st + r; a.ncoll = d_coll_counts + 4 * r;
      a.finish_after = finish_late() ? 1 : 0;
      a.diag = diag + d.dbase * 8;
      launch_dsmc(stream, a);
    }
  }
  return PIC_OK;
}

int Ctx::launch_one_step(bool allow_sort, hipEvent_t* ev, bool defer_tail) {
  {
    FieldArgs fa = field_args(0);
    if (fin_pending) {
      fa.fold_finish = 1;
      fa.ppush = partials_push;
    }
    if (scan_pending) {
      fa.fold_scan = 1;
      fa.ftable = sb.ftable; fa.fchunk_sums = sb.fchunk_sums; fa.fstride = sb.max_buckets;
      fa.fscan_items = sb.fscan_items; fa.nfscan_items = sb.nfscan_items;
      fa.ftoff = scan_toff;
    }
    launch_field(stream, fa, ev);
    fin_pending = scan_pending = false;
  }
  // fused sort: the push one step before the sort step counts the tiles of the positions
  // it produces (then the counts are scanned), the push of the sort step scatters
  int smode = 0;
  if (fused && allow_sort) {
    if (steps_since_sort == sort_every - 2) smode = 1;
    else if (steps_since_sort == sort_every - 1) smode = 2;
  }
  launch_push_stage(0, smode);
  if (ev) (void)hipEventRecord(ev[7], stream);  // the push stage is the k_push launch alone
  // the step's end (diagnostics row, step counter) runs after the collision stage: MCC
  // births then enter the row's particle count, and inside a captured sequence the end
  // folds into the next step's field kernels as in a collisionless run
  auto finish = [&]() {
    if (defer_tail) fin_pending = true;
    else launch_step_finish(stream, d_runs, d_st, nruns, partials_push, partials_rows, diag);
  };
  if (!finish_late()) finish();
  // with MCC the count table is scanned after the collision stage, whose births add their
  // tiles to it
  const bool scan_after_collide = any_mcc;
  if (smode == 1) {
    sb.toff = toffs[1 - cur];
    if (defer_tail) {
      scan_pending = true;
      scan_toff = toffs[1 - cur];
    } else if (!scan_after_collide) {
      launch_fused_scan(stream, d_runs, d_st, nruns, sb, toffs[1 - cur]);
    }
  }
  // the sort step's push wrote the new order into the other buffer: the collision stage
  // works on that one
  if (smode == 2) cur ^= 1;
  if (ev) (void)hipEventRecord(ev[4], stream);
  collide_count_table = (smode == 1) ? sb.ftable : nullptr;
  launch_collide_stage();
  collide_count_table = nullptr;
  if (finish_late()) finish();
  if (smode == 1 && !defer_tail && scan_after_collide)
    launch_fused_scan(stream, d_runs, d_st, nruns, sb, toffs[1 - cur]);
  if (ev) (void)hipEventRecord(ev[5], stream);
  if (any_dsmc) {
    steps_since_sort = 0;
  } else if (allow_sort) {
    ++steps_since_sort;
    if (smode == 2) {
      steps_since_sort = 0;
    } else if (!fused && steps_since_sort >= sort_every) {
      launch_sort_stage();
      steps_since_sort = 0;
    }
  }
  return PIC_OK;
}

int Ctx::capture_macro(int which) {
  const int k = macro_len();
  const int saved_cur = cur, saved_since = steps_since_sort;
  hipGraph_t g = nullptr;
  PIC_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  // inside the sequence each step's end-of-step work runs in the next step's field kernels;
  // the last step ends with its own kernels, so a replay starts and ends like eager steps
  const bool fold = true;
  for (int i = 0; i < k; ++i) launch_one_step(true, nullptr, fold && i + 1 < k);
  hipError_t e = hipStreamEndCapture(stream, &g);
  const int flip = cur ^ saved_cur;  // the sorts inside switch buffers (one per sort)
  cur = saved_cur;
  steps_since_sort = saved_since;
  fin_pending = scan_pending = false;
  if (e != hipSuccess) return fail(PIC_ERR_HIP, std::string("stream capture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&macro[which], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(PIC_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
  macro_k[which] = k;
  macro_flip[which] = flip;
  return PIC_OK;
}

int Ctx::step(int n) {
  const int k = macro_len();
  while (n > 0) {
    if (use_graphs && steps_since_sort == 0 && n >= k) {
      const int w = cur;
      if (!macro[w]) {
        int rc = capture_macro(w);
        if (rc) return rc;
      }
      PIC_HIP(hipGraphLaunch(macro[w], stream));
      cur ^= macro_flip[w];
      n -= k;
    } else {
      launch_one_step(true);
      --n;
    }
  }
  PIC_HIP(hipGetLastError());
  return PIC_OK;
}

}  // namespace pic

using namespace pic;

// ============================================================================ C ABI
#define CTX_CHECK(c)                                                  \
  do {                                                                \
    if (!(c)) return fail(PIC_ERR_ARG, "NULL context");               \
    PIC_HIP(hipSetDevice((c)->impl.device));                          \
  } while (0)

extern "C" {

const char* pic_last_error(void) { return g_err.c_str(); }
int pic_version(void) { return 1; }

void pic_config_default(pic_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->nx = 32; c->ny = 32;
  c->lx = 4.0 * 3.14159265358979323846; c->ly = c->lx;
  c->dt = 0.05; c->qm = -1.0; c->qp = 0.0;
  c->precision = PIC_FP32;
  c->seed = 1234;
  c->mcc_neutral_mass = 39.948 * 1.66053906660e-27;
  c->dsmc_gmax0 = 4.0;
}

int pic_mcc_preset(pic_config* c, double ng, double te_ev, double ne) {
  if (!c || !(ng > 0) || !(te_ev > 0) || !(ne > 0)) return fail(PIC_ERR_ARG, "bad MCC preset");
  const double wp = std::sqrt(ne * kE * kE / (kEps0 * kMe));
  const double vu = std::sqrt(te_ev * kE / kMe);  // lambda_D * omega_p
  c->collide = PIC_COLLIDE_MCC;
  c->mcc_ng = ng;
  c->mcc_v_unit = vu;
  c->mcc_dt_phys = 1.0 / wp;
  if (!(c->mcc_neutral_mass > 0)) c->mcc_neutral_mass = 39.948 * 1.66053906660e-27;
  const double numax = mcc_nu_max(*c);
  c->dt = 0.01 * wp / numax;  // nu_max * dt_phys = 0.01
  return PIC_OK;
}

double pic_mcc_nu_max(const pic_config* c) {
  if (!c) return 0.0;
  return mcc_nu_max(*c) * c->mcc_dt_phys * c->dt;
}

static int create_impl(const pic_config* cfgs, int nruns, pic_ctx** out) {
  if (!cfgs || !out || nruns < 1) return fail(PIC_ERR_ARG, "bad arguments to create");
  *out = nullptr;
  for (int r = 0; r < nruns; ++r) {
    int rc = validate_config(cfgs[r]);
    if (rc) return rc;
    if (cfgs[r].precision != cfgs[0].precision || cfgs[r].device != cfgs[0].device)
      return fail(PIC_ERR_INVALID_CONFIG, "all runs of a batch share precision and device");
    if (nruns > 1 && cfgs[r].capacity <= 0)
      return fail(PIC_ERR_INVALID_CONFIG, "batch runs need capacity > 0");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(PIC_ERR_HIP, "no HIP device available");
  if (cfgs[0].device < 0 || cfgs[0].device >= ndev)
    return fail(PIC_ERR_INVALID_CONFIG, "device ordinal out of range");
  std::unique_ptr<pic_ctx> c(new pic_ctx());
  Ctx& x = c->impl;
  x.device = cfgs[0].device;
  x.nruns = nruns;
  x.fp64 = cfgs[0].precision == PIC_FP64;
  x.rsize = x.fp64 ? 8 : 4;
  x.cfg.assign(cfgs, cfgs + nruns);
  x.rd.resize(nruns);
  x.has_particles.assign(nruns, 0);
  int se = 0;
  for (auto& cc : x.cfg) se = std::max(se, cc.sort_every);
  x.sort_every = se ? se : 4;
  if (nruns > 1 || cfgs[0].capacity > 0) {
    int rc = x.alloc();
    if (rc) return rc;
  }
  *out = c.release();
  return PIC_OK;
}

int pic_create(const pic_config* cfg, pic_ctx** out) { return create_impl(cfg, 1, out); }
int pic_batch_create(const pic_config* cfgs, int nruns, pic_ctx** out) {
  return create_impl(cfgs, nruns, out);
}

int pic_destroy(pic_ctx* ctx) {
  if (!ctx) return PIC_OK;
  if (ctx->impl.stream) (void)hipStreamSynchronize(ctx->impl.stream);
  delete ctx;
  return PIC_OK;
}

int pic_num_runs(pic_ctx* ctx, int* n) {
  if (!ctx || !n) return fail(PIC_ERR_ARG, "NULL argument");
  *n = ctx->impl.nruns;
  return PIC_OK;
}

}  // extern "C"

template <typename R>
static int upload_particles(Ctx& x, int run, const void* px_, const void* py_, const void* pvx_,
                            const void* pvy_, int64_t n, bool device_src) {
  const RunDesc& d = x.rd[run];
  PartBufs<R> b = x.bufs<R>(x.cur);
  if (device_src) {
    PIC_HIP(hipMemcpyAsync(b.x + d.pbase, px_, sizeof(R) * n, hipMemcpyDeviceToDevice, x.stream));
    PIC_HIP(hipMemcpyAsync(b.y + d.pbase, py_, sizeof(R) * n, hipMemcpyDeviceToDevice, x.stream));
    PIC_HIP(hipMemcpyAsync(b.vx + d.pbase, pvx_, sizeof(R) * n, hipMemcpyDeviceToDevice, x.stream));
    PIC_HIP(hipMemcpyAsync(b.vy + d.pbase, pvy_, sizeof(R) * n, hipMemcpyDeviceToDevice, x.stream));
  } else {
    std::vector<R> h(n);
    const double* srcs[4] = {(const double*)px_, (const double*)py_, (const double*)pvx_,
                             (const double*)pvy_};
    R* dsts[4] = {b.x + d.pbase, b.y + d.pbase, b.vx + d.pbase, b.vy + d.pbase};
    for (int k = 0; k < 4; ++k) {
      for (int64_t i = 0; i < n; ++i) h[i] = (R)srcs[k][i];
      PIC_HIP(hipMemcpyAsync(dsts[k], h.data(), sizeof(R) * n, hipMemcpyHostToDevice, x.stream));
      PIC_HIP(hipStreamSynchronize(x.stream));
    }
  }
  std::vector<int> ids(n);
  for (int64_t i = 0; i < n; ++i) ids[i] = (int)i;
  PIC_HIP(hipMemcpyAsync(b.id + d.pbase, ids.data(), sizeof(int) * n, hipMemcpyHostToDevice, x.stream));
  PIC_HIP(hipStreamSynchronize(x.stream));
  return PIC_OK;
}

namespace pic {
int set_particles_impl(pic_ctx* ctx, int run, const void* px_, const void* py_,
                       const void* pvx_, const void* pvy_, int64_t n, bool device_src,
                       const Restore* restore) {
  CTX_CHECK(ctx);
  Ctx& x = ctx->impl;
  if (run < 0 || run >= x.nruns) return fail(PIC_ERR_ARG, "run index out of range");
  if (!px_ || !py_ || !pvx_ || !pvy_) return fail(PIC_ERR_ARG, "NULL particle array");
  if (n < 1) return fail(PIC_ERR_INVALID_CONFIG, "need at least one particle");
  if (!x.allocated) {
    x.cfg[0].capacity = std::max<int64_t>(x.cfg[0].capacity, n);
    int rc = x.alloc();
    if (rc) return rc;
  }
  RunDesc& d = x.rd[run];
  if (n > d.pcap) return fail(PIC_ERR_OVERFLOW, "more particles than the run capacity");
  const pic_config& c = x.cfg[run];
  d.qp = restore ? restore->qp : ((c.qp != 0.0) ? c.qp : -(c.lx * c.ly) / (double)n);
  d.rho_scale = d.qp / (d.dx * d.dy) * (x.fp64 ? 1.0 : std::ldexp(1.0, -36));
  int rc = x.upload_runs();
  if (rc) return rc;
  rc = x.fp64 ? upload_particles<double>(x, run, px_, py_, pvx_, pvy_, n, device_src)
              : upload_particles<float>(x, run, px_, py_, pvx_, pvy_, n, device_src);
  if (rc) return rc;
  RunState s;
  std::memset(&s, 0, sizeof(s));
  s.np = (int)n; s.nsorted = 0; s.vsync = 1; s.step = 0; s.next_id = n;
  if (restore) {
    s.vsync = restore->vsync;
    s.step = restore->step;
    s.next_id = restore->next_id;
  }
  PIC_HIP(hipMemcpy(x.d_st + run, &s, sizeof(s), hipMemcpyHostToDevice));
  const long long G = (long long)d.nx * d.ny;
  for (int p = 0; p < 2; ++p) {
    launch_fill_u64(x.stream, x.rho + p * x.gsum + d.gbase, G);
    if (x.fp64) launch_fill_u64(x.stream, x.rho_lo + p * x.gsum + d.gbase, G);
  }
  if (c.collide == PIC_COLLIDE_DSMC) {
    std::vector<float> g(G + 1, (float)c.dsmc_gmax0);
    PIC_HIP(hipMemcpy(x.gmax + d.cbase, g.data(), sizeof(float) * (G + 1), hipMemcpyHostToDevice));
  }
  PIC_HIP(hipMemsetAsync(x.d_coll_counts + 4 * run, 0, sizeof(long long) * 4, x.stream));
  {
    // the first step's collision count starts at zero (each step's end clears the next's)
    const double zero = 0.0;
    PIC_HIP(hipMemcpyAsync(x.diag + (d.dbase + s.step % PIC_DIAG_CAP) * 8 + PIC_DIAG_NCOLL,
                           &zero, sizeof(double), hipMemcpyHostToDevice, x.stream));
    PIC_HIP(hipStreamSynchronize(x.stream));
  }
  x.launch_sort_stage();
  x.steps_since_sort = 0;
  x.has_particles[run] = 1;
  if ((rc = x.rebalance_push_items())) return rc;
  // deposit the initial charge of this run only
  if (x.fp64) {
    PushArgs<double> a = x.push_args<double>(1);
    a.items = x.d_push + d.push_item_begin;
    a.nitems = d.push_item_end - d.push_item_begin;
    launch_push<double>(x.stream, a);
  } else {
    PushArgs<float> a = x.push_args<float>(1);
    a.items = x.d_push + d.push_item_begin;
    a.nitems = d.push_item_end - d.push_item_begin;
    launch_push<float>(x.stream, a);
  }
  PIC_HIP(hipGetLastError());
  PIC_HIP(hipStreamSynchronize(x.stream));
  x.has_particles[run] = 1;
  return PIC_OK;
}
}  // namespace pic

extern "C" {

int pic_set_particles(pic_ctx* ctx, int run, const double* x, const double* y, const double* vx,
                      const double* vy, int64_t n) {
  return set_particles_impl(ctx, run, x, y, vx, vy, n, false, nullptr);
}
int pic_set_particles_device(pic_ctx* ctx, int run, const void* x, const void* y, const void* vx,
                             const void* vy, int64_t n) {
  return set_particles_impl(ctx, run, x, y, vx, vy, n, true, nullptr);
}

}  // extern "C"

static int read_state(Ctx& x, int run, RunState* s) {
  PIC_HIP(hipStreamSynchronize(x.stream));
  PIC_HIP(hipMemcpy(s, x.d_st + run, sizeof(RunState), hipMemcpyDeviceToHost));
  return PIC_OK;
}

template <typename R>
static int get_particles_t(Ctx& x, int run, void* ox, void* oy, void* ovx, void* ovy, int64_t n,
                           int sync, bool device_dst) {
  RunState s;
  int rc = read_state(x, run, &s);
  if (rc) return rc;
  if (n < s.np) return fail(PIC_ERR_ARG, "output arrays smaller than the particle count");
  const int np = s.np;
  R *t = nullptr;
  PIC_HIP(dalloc(&t, 6 * (size_t)std::max(np, 1)));
  R *tx = t, *ty = t + np, *tvx = t + 2 * (size_t)np, *tvy = t + 3 * (size_t)np;
  R *svx = nullptr, *svy = nullptr;
  if (sync && !s.vsync) {
    svx = t + 4 * (size_t)np;
    svy = t + 5 * (size_t)np;
    x.launch_field_stage(0);
    launch_half_kick_out<R>(x.stream, x.d_runs, x.d_st, run, x.bufs<R>(x.cur), (const R*)x.E, svx,
                            svy, x.d_pitems, (int)x.pitems.size());
  }
  launch_unpermute<R>(x.stream, x.d_runs, x.d_st, run, x.bufs<R>(x.cur), svx, svy, tx, ty, tvx, tvy,
                      x.d_pitems, (int)x.pitems.size());
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(x.stream);
  if (e != hipSuccess) {
    (void)hipFree(t);
    return fail(PIC_ERR_HIP, std::string("get_particles: ") + hipGetErrorString(e));
  }
  void* outs[4] = {ox, oy, ovx, ovy};
  R* srcs[4] = {tx, ty, tvx, tvy};
  for (int k = 0; k < 4; ++k) {
    if (!outs[k]) continue;
    if (device_dst) {
      e = hipMemcpy(outs[k], srcs[k], sizeof(R) * np, hipMemcpyDeviceToDevice);
    } else {
      std::vector<R> h(np);
      e = hipMemcpy(h.data(), srcs[k], sizeof(R) * np, hipMemcpyDeviceToHost);
      for (int i = 0; i < np; ++i) ((double*)outs[k])[i] = (double)h[i];
    }
    if (e != hipSuccess) break;
  }
  (void)hipFree(t);
  if (e != hipSuccess) return fail(PIC_ERR_HIP, std::string("get_particles copy: ") + hipGetErrorString(e));
  return PIC_OK;
}

extern "C" {

int pic_get_particles(pic_ctx* ctx, int run, double* x, double* y, double* vx, double* vy,
                      int64_t n, int sync) {
  CTX_CHECK(ctx);
  Ctx& c = ctx->impl;
  if (run < 0 || run >= c.nruns || !c.has_particles[run]) return fail(PIC_ERR_STATE, "run has no particles");
  return c.fp64 ? get_particles_t<double>(c, run, x, y, vx, vy, n, sync, false)
                : get_particles_t<float>(c, run, x, y, vx, vy, n, sync, false);
}

int pic_get_particles_device(pic_ctx* ctx, int run, void* x, void* y, void* vx, void* vy,
                             int64_t n, int sync) {
  CTX_CHECK(ctx);
  Ctx& c = ctx->impl;
  if (run < 0 || run >= c.nruns || !c.has_particles[run]) return fail(PIC_ERR_STATE, "run has no particles");
  return c.fp64 ? get_particles_t<double>(c, run, x, y, vx, vy, n, sync, true)
                : get_particles_t<float>(c, run, x, y, vx, vy, n, sync, true);
}

int pic_num_particles(pic_ctx* ctx, int run, int64_t* n) {
  CTX_CHECK(ctx);
  if (!n || run < 0 || run >= ctx->impl.nruns) return fail(PIC_ERR_ARG, "bad argument");
  if (!ctx->impl.allocated) { *n = 0; return PIC_OK; }
  RunState s;
  int rc = read_state(ctx->impl, run, &s);
  if (rc) return rc;
  *n = s.np;
  return PIC_OK;
}

int pic_step(pic_ctx* ctx, int nsteps) {
  CTX_CHECK(ctx);
  Ctx& x = ctx->impl;
  if (nsteps < 0) return fail(PIC_ERR_ARG, "nsteps must be >= 0");
  for (int r = 0; r < x.nruns; ++r)
    if (!x.has_particles[r]) return fail(PIC_ERR_STATE, "pic_step before pic_set_particles");
  return x.step(nsteps);
}

int pic_sync(pic_ctx* ctx) {
  CTX_CHECK(ctx);
  Ctx& x = ctx->impl;
  if (!x.allocated) return PIC_OK;
  PIC_HIP(hipStreamSynchronize(x.stream));
  std::vector<RunState> s(x.nruns);
  PIC_HIP(hipMemcpy(s.data(), x.d_st, sizeof(RunState) * x.nruns, hipMemcpyDeviceToHost));
  for (int r = 0; r < x.nruns; ++r) {
    if (s[r].err & 1) return fail(PIC_ERR_OVERFLOW, "particle capacity exceeded (births dropped)");
    if (s[r].err & 2) return fail(PIC_ERR_STATE, "sort slot outside the run (tile counts out of date)");
  }
  return PIC_OK;
}

int pic_check_finite(pic_ctx* ctx) {
  CTX_CHECK(ctx);
  Ctx& x = ctx->impl;
  if (!x.allocated) return PIC_OK;
  PIC_HIP(hipMemsetAsync(x.flag, 0, sizeof(int), x.stream));
  if (x.fp64)
    launch_check_finite<double>(x.stream, x.d_runs, x.d_st, x.bufs<double>(x.cur), x.d_pitems,
                                (int)x.pitems.size(), x.flag);
  else
    launch_check_finite<float>(x.stream, x.d_runs, x.d_st, x.bufs<float>(x.cur), x.d_pitems,
                               (int)x.pitems.size(), x.flag);
  int h = 0;
  PIC_HIP(hipStreamSynchronize(x.stream));
  PIC_HIP(hipMemcpy(&h, x.flag, sizeof(int), hipMemcpyDeviceToHost));
  if (h) return fail(PIC_ERR_NONFINITE, "NaN or Inf in the particle state");
  return PIC_OK;
}

int pic_get_grid(pic_ctx* ctx, int run, int which, double* out) {
  CTX_CHECK(ctx);
  Ctx& x = ctx->impl;
  if (!out || run < 0 || run >= x.nruns) return fail(PIC_ERR_ARG, "bad argument");
  if (!x.has_particles[run]) return fail(PIC_ERR_STATE, "run has no particles");
  const RunDesc& d = x.rd[run];
  const long long G = (long long)d.nx * d.ny;
  double* t = nullptr;
  PIC_HIP(dalloc(&t, 2 * G));
  RunState s;
  int rc = read_state(x, run, &s);
  if (rc) { (void)hipFree(t); return rc; }
  x.launch_field_stage(which == PIC_GRID_PHI);
  const double* src = t;
  if (which == PIC_GRID_RHO) {
    const double bg = -d.qp * s.np / (d.lx * d.ly);
    launch_rho_to_double(x.stream, x.d_runs, x.d_st, run, x.rho, x.rho_lo, x.gsum, x.fp64, bg, t);
  } else if (which == PIC_GRID_PHI) {
    src = x.phi + d.gbase;
  } else if (which == PIC_GRID_EX || which == PIC_GRID_EY) {
    launch_unpad_E(x.stream, (const char*)x.E + x.rsize * 2 * d.ebase, x.fp64, d.nx, d.ny, t, t + G);
    if (which == PIC_GRID_EY) src = t + G;
  } else {
    (void)hipFree(t);
    return fail(PIC_ERR_ARG, "unknown grid selector");
  }
  hipError_t e = hipStreamSynchronize(x.stream);
  if (e == hipSuccess) e = hipMemcpy(out, src, sizeof(double) * G, hipMemcpyDeviceToHost);
  (void)hipFree(t);
  if (e != hipSuccess) return fail(PIC_ERR_HIP, std::string("get_grid: ") + hipGetErrorString(e));
  return PIC_OK;
}

int pic_get_diagnostics(pic_ctx* ctx, int run, int64_t first, int count, double* out) {
  CTX_CHECK(ctx);
  Ctx& x = ctx->impl;
  if (!out || run < 0 || run >= x.nruns || count < 0 || first < 0)
    return fail(PIC_ERR_ARG, "bad argument");
  if (count > PIC_DIAG_CAP) return fail(PIC_ERR_ARG, "count exceeds the diagnostics ring");
  PIC_HIP(hipStreamSynchronize(x.stream));
  const RunDesc& d = x.rd[run];
  for (int i = 0; i < count; ++i) {
    const long long row = (first + i) % PIC_DIAG_CAP;
    PIC_HIP(hipMemcpy(out + (size_t)i * PIC_DIAG_COLS, x.diag + (d.dbase + row) * 8,
                      sizeof(double) * 8, hipMemcpyDeviceToHost));
  }
  return PIC_OK;
}

int pic_current_step(pic_ctx* ctx, int run, int64_t* step) {
  CTX_CHECK(ctx);
  if (!step || run < 0 || run >= ctx->impl.nruns) return fail(PIC_ERR_ARG, "bad argument");
  RunState s;
  int rc = read_state(ctx->impl, run, &s);
  if (rc) return rc;
  *step = s.step;
  return PIC_OK;
}

int pic_collision_counts(pic_ctx* ctx, int run, int64_t* out4) {
  CTX_CHECK(ctx);
  Ctx& x = ctx->impl;
  if (!out4 || run < 0 || run >= x.nruns) return fail(PIC_ERR_ARG, "bad argument");
  PIC_HIP(hipStreamSynchronize(x.stream));
  PIC_HIP(hipMemcpy(out4, x.d_coll_counts + 4 * run, sizeof(long long) * 4, hipMemcpyDeviceToHost));
  return PIC_OK;
}

int pic_stream(pic_ctx* ctx, void** stream) {
  if (!ctx || !stream) return fail(PIC_ERR_ARG, "NULL argument");
  *stream = (void*)ctx->impl.stream;
  return PIC_OK;
}

int pic_use_graphs(pic_ctx* ctx, int enable) {
  if (!ctx) return fail(PIC_ERR_ARG, "NULL context");
  ctx->impl.use_graphs = enable != 0;
  return PIC_OK;
}

int pic_get_tuning(pic_ctx* ctx, int* tile, int* halo, int* splits, int* sort_every) {
  if (!ctx) return fail(PIC_ERR_ARG, "NULL context");
  Ctx& x = ctx->impl;
  if (x.tune.empty()) return fail(PIC_ERR_STATE, "context not allocated yet");
  if (tile) *tile = x.tune[0].tile;
  if (halo) *halo = x.tune[0].halo;
  if (splits) *splits = x.tune[0].per_tile ? 0 : x.tune[0].splits;
  if (sort_every) *sort_every = x.sort_every;
  return PIC_OK;
}

int pic_sort_fused(pic_ctx* ctx, int* fused) {
  if (!ctx || !fused) return fail(PIC_ERR_ARG, "NULL argument");
  Ctx& x = ctx->impl;
  if (x.tune.empty()) return fail(PIC_ERR_STATE, "context not allocated yet");
  *fused = x.fused ? 1 : 0;
  return PIC_OK;
}

int pic_set_tuning(pic_ctx* ctx, int tile, int halo, int splits, int sort_every) {
  CTX_CHECK(ctx);
  Ctx& x = ctx->impl;
  if (!x.allocated) return fail(PIC_ERR_STATE, "context not allocated yet");
  if (sort_every < 0 || splits < 0 || halo < 0 || tile < 0 || splits > 64)
    return fail(PIC_ERR_ARG, "tuning values must be >= 0");
  for (int r = 0; r < x.nruns; ++r) {
    if (tile && (!is_pow2(tile) || tile > std::min(x.cfg[r].nx, x.cfg[r].ny) / 2 || tile > 64))
      return fail(PIC_ERR_ARG, "tile must be a power of two <= min(nx, ny)/2 and <= 64");
  }
  PIC_HIP(hipStreamSynchronize(x.stream));
  for (int r = 0; r < x.nruns; ++r) x.tune[r] = default_tuning(x.cfg[r], tile, halo, splits);
  if (sort_every) x.sort_every = sort_every;
  int rc = x.build_items();
  if (rc) return rc;
  // offsets must describe the current order for the new tiling: sort now
  bool any = false;
  for (int r = 0; r < x.nruns; ++r) any |= x.has_particles[r] != 0;
  if (any) {
    x.launch_sort_stage();
    x.steps_since_sort = 0;
    PIC_HIP(hipStreamSynchronize(x.stream));
  }
  return x.rebalance_push_items();
}

// Project autotuner: times (tile, halo, sort cadence) candidates on the live state and
// keeps the fastest.  The particle state, charge, counters and diagnostics are restored
// afterwards, so tuning does not advance the simulation.
int pic_autotune(pic_ctx* ctx, int trial_steps, double* best_ms) {
  CTX_CHECK(ctx);
  Ctx& x = ctx->impl;
  if (trial_steps < 1) return fail(PIC_ERR_ARG, "trial_steps must be >= 1");
  for (int r = 0; r < x.nruns; ++r)
    if (!x.has_particles[r]) return fail(PIC_ERR_STATE, "run has no particles");
  PIC_HIP(hipStreamSynchronize(x.stream));
  // snapshot
  const size_t pb = x.rsize * x.ptotal;
  std::vector<void*> snap;
  auto keep = [&](const void* src, size_t bytes) -> void* {
    void* d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) return nullptr;
    if (hipMemcpy(d, src, bytes, hipMemcpyDeviceToDevice) != hipSuccess) return nullptr;
    snap.push_back(d);
    return d;
  };
  const int c0 = x.cur;
  void* sx = keep(x.px[c0], pb);
  void* sy = keep(x.py[c0], pb);
  void* svx = keep(x.pvx[c0], pb);
  void* svy = keep(x.pvy[c0], pb);
  void* sid = keep(x.pid[c0], sizeof(int) * x.ptotal);
  void* srho = keep(x.rho, sizeof(unsigned long long) * 2 * x.gsum);
  void* srlo = x.fp64 ? keep(x.rho_lo, sizeof(unsigned long long) * 2 * x.gsum) : (void*)1;
  void* sst = keep(x.d_st, sizeof(RunState) * x.nruns);
  void* sdg = keep(x.diag, sizeof(double) * (size_t)x.nruns * PIC_DIAG_CAP * 8);
  void* scc = keep(x.d_coll_counts, sizeof(long long) * 4 * x.nruns);
  void* sgm = x.any_dsmc ? keep(x.gmax, sizeof(float) * x.ctotal) : (void*)1;
  auto release = [&]() { for (void* p : snap) (void)hipFree(p); };
  if (!sx || !sy || !svx || !svy || !sid || !srho || !srlo || !sst || !sdg || !scc || !sgm) {
    release();
    return fail(PIC_ERR_OOM, "autotune snapshot allocation failed");
  }
  const std::vector<Tuning> orig = x.tune;
  const int orig_se = x.sort_every;
  auto restore = [&]() -> int {
    const int c = x.cur;
    PIC_HIP(hipMemcpy(x.px[c], sx, pb, hipMemcpyDeviceToDevice));
    PIC_HIP(hipMemcpy(x.py[c], sy, pb, hipMemcpyDeviceToDevice));
    PIC_HIP(hipMemcpy(x.pvx[c], svx, pb, hipMemcpyDeviceToDevice));
    PIC_HIP(hipMemcpy(x.pvy[c], svy, pb, hipMemcpyDeviceToDevice));
    PIC_HIP(hipMemcpy(x.pid[c], sid, sizeof(int) * x.ptotal, hipMemcpyDeviceToDevice));
    PIC_HIP(hipMemcpy(x.rho, srho, sizeof(unsigned long long) * 2 * x.gsum, hipMemcpyDeviceToDevice));
    if (x.fp64)
      PIC_HIP(hipMemcpy(x.rho_lo, srlo, sizeof(unsigned long long) * 2 * x.gsum, hipMemcpyDeviceToDevice));
    PIC_HIP(hipMemcpy(x.d_st, sst, sizeof(RunState) * x.nruns, hipMemcpyDeviceToDevice));
    PIC_HIP(hipMemcpy(x.diag, sdg, sizeof(double) * (size_t)x.nruns * PIC_DIAG_CAP * 8, hipMemcpyDeviceToDevice));
    PIC_HIP(hipMemcpy(x.d_coll_counts, scc, sizeof(long long) * 4 * x.nruns, hipMemcpyDeviceToDevice));
    if (x.any_dsmc) PIC_HIP(hipMemcpy(x.gmax, sgm, sizeof(float) * x.ctotal, hipMemcpyDeviceToDevice));
    return PIC_OK;
  };
  int mn = PIC_FFT_MAX;
  for (auto& c : x.cfg) mn = std::min(mn, std::min(c.nx, c.ny));
  double best = 1e30;
  int bt = 0, bh = 0, bs = 0;
  const int tiles[3] = {8, 16, 32}, halos[2] = {8, 16}, cad[5] = {1, 2, 4, 6, 8};
  // splits: 0 = sized by particles per tile (the only choice for ragged batches); the
  // larger counts fill the GPU on small grids (a 64^2 run has 16 tiles of 16^2 cells)
  const int splitc[6] = {0, 1, 2, 4, 8, 16};
  int rc = PIC_OK;
  int bsp = 0;
  for (int ti : tiles) {
    if (ti > mn || ti > 64) continue;
    for (int h : halos) {
      for (int se : cad)
      for (int sp : splitc) {
        if (x.any_dsmc && se != 1) continue;
        if (sp && x.nruns > 1) continue;
        // at least ~1k particles per workgroup for the larger split counts
        if (sp > 2 && (long long)x.cfg[0].capacity <
                          1024LL * sp * (x.cfg[0].nx / ti) * (x.cfg[0].ny / ti))
          continue;
        rc = restore();
        if (rc) break;
        for (int r = 0; r < x.nruns; ++r) x.tune[r] = default_tuning(x.cfg[r], ti, h, sp);
        x.sort_every = se;
        if ((rc = x.build_items())) break;
        x.launch_sort_stage();
        x.steps_since_sort = 0;
        if ((rc = x.rebalance_push_items())) break;
        // one warm cycle (captures graphs), then the timed cycles
        const int ml = x.macro_len();
        if ((rc = x.step(ml))) break;
        const int n = std::max(trial_steps / ml, 1) * ml;
        PIC_HIP(hipEventRecord(x.ev[0], x.stream));
        if ((rc = x.step(n))) break;
        PIC_HIP(hipEventRecord(x.ev[1], x.stream));
        PIC_HIP(hipEventSynchronize(x.ev[1]));
        float ms = 0.f;
        PIC_HIP(hipEventElapsedTime(&ms, x.ev[0], x.ev[1]));
        const double per = ms / n;
        if (per < best) { best = per; bt = ti; bh = h; bs = se; bsp = sp; }
      }
      if (rc) break;
    }
    if (rc) break;
  }
  int rc2 = restore();
  release();
  if (rc) { x.tune = orig; x.sort_every = orig_se; x.build_items(); return rc; }
  if (rc2) return rc2;
  if (best_ms) *best_ms = best;
  return pic_set_tuning(ctx, bt, bh, bsp, bs);
}

int pic_stage_times(pic_ctx* ctx, int nsteps, double* out) {
  CTX_CHECK(ctx);
  Ctx& x = ctx->impl;
  if (!out || nsteps < 1) return fail(PIC_ERR_ARG, "bad argument");
  for (int r = 0; r < x.nruns; ++r)
    if (!x.has_particles[r]) return fail(PIC_ERR_STATE, "run has no particles");
  double acc[PIC_NSTAGES] = {0, 0, 0, 0, 0, 0};
  for (int s = 0; s < nsteps; ++s) {
    x.launch_one_step(true, x.ev);
    PIC_HIP(hipEventRecord(x.ev[6], x.stream));
    PIC_HIP(hipEventSynchronize(x.ev[6]));
    // event order: 0 rows 1 cols 2 field 3 push 7 finish + fused-sort scans 4 collide 5
    // separate sort 6
    const int from[7] = {0, 1, 2, 3, 7, 4, 5}, to[7] = {1, 2, 3, 7, 4, 5, 6};
    const int stage[7] = {PIC_STAGE_FFT_ROWS, PIC_STAGE_FFT_COLS, PIC_STAGE_FFT_FIELD,
                          PIC_STAGE_PUSH_DEPOSIT, PIC_STAGE_SORT, PIC_STAGE_COLLIDE,
                          PIC_STAGE_SORT};
    for (int k = 0; k < 7; ++k) {
      float ms = 0.f;
      PIC_HIP(hipEventElapsedTime(&ms, x.ev[from[k]], x.ev[to[k]]));
      acc[stage[k]] += ms;
    }
  }
  for (int k = 0; k < PIC_NSTAGES; ++k) out[k] = acc[k] / nsteps;
  return PIC_OK;
}

int pic_run_compat(const double* x, const double* y, const double* vx, const double* vy,
                   double* x_out, double* y_out, int64_t n, int nsteps, const pic_config* cfg,
                   double* elapsed_s) {
  if (!x || !y || !vx || !vy || !x_out || !y_out || !cfg || nsteps < 0)
    return fail(PIC_ERR_ARG, "bad argument to pic_run_compat");
  pic_config c = *cfg;
  c.capacity = std::max<int64_t>(c.capacity, n);
  pic_ctx* ctx = nullptr;
  int rc = pic_create(&c, &ctx);
  if (rc) return rc;
  std::unique_ptr<pic_ctx, int (*)(pic_ctx*)> guard(ctx, pic_destroy);
  rc = pic_set_particles(ctx, 0, x, y, vx, vy, n);
  if (rc) return rc;
  Ctx& e = ctx->impl;
  PIC_HIP(hipEventRecord(e.ev[0], e.stream));
  rc = pic_step(ctx, nsteps);
  if (rc) return rc;
  PIC_HIP(hipEventRecord(e.ev[1], e.stream));
  PIC_HIP(hipEventSynchronize(e.ev[1]));
  float ms = 0.f;
  PIC_HIP(hipEventElapsedTime(&ms, e.ev[0], e.ev[1]));
  if (elapsed_s) *elapsed_s = ms * 1e-3;
  std::vector<double> tvx(n), tvy(n);
  return pic_get_particles(ctx, 0, x_out, y_out, tvx.data(), tvy.data(), n, 1);
}

}  // extern "C"
