// launch.h -- host-side launchers of the HIP kernels (internal to libpicamd.so).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pic_device.h"

// Item lists (device, built once per context):
//   chunk items  int2 {run, start}: PIC_CHUNK particles or PIC_CELL_CHUNK cells
//   tile items   int4 {run, tx, ty, split}; tx < 0 marks a tail chunk (start in .z)
//   row items    int2 {run, row}
#define PIC_CELL_CHUNK 2048
// cell sort: cells per item of the within-cell sort (one wave per cell, four waves per
// workgroup), and the slices of a particle chunk item that the rank, scatter and apply
// kernels give one workgroup each (more workgroups than chunks: the blob-like runs of the
// collision stages need the latency hiding)
#define PIC_SEG_CELLS 4
#define PIC_SORT_SLICES 4

struct SortBufs {
  int* keys;        // [ptotal]
  int* rank;        // [ptotal]
  int* perm;        // [ptotal]  run-local source index of sorted slot
  int* counts;      // [ctotal]
  int* offsets;     // [ctotal]
  int* chunk_sums;  // [n cell chunk items]
  int* big_list;    // [ptotal] cells needing the workgroup segment sort (run, cell) pairs
  int* big_count;   // [1]
  unsigned long long* scratch;  // [2*ptotal] power-of-two padded segments
  const int2* segitems;         // {run, first cell} items of 64 cells (segment sort)
  int nsegitems;
  // tile sort of the step
  int* table;                   // [sum nbuckets*nchunks] bucket-major count table
  int* tchunk_sums;             // [n table-scan items]
  const int2* tscan_items;      // {run, first entry} items of PIC_CELL_CHUNK table entries
  int ntscan_items;
  int* toff;                    // [sum (nbuckets+1)] tile offsets written by the sort
  int max_buckets;
  // sort fused into the push
  int* ftable;                  // [sum fchunks*nbuckets] row-major counts, then slots
  int* fchunk_sums;             // [n fused table-scan items][max_buckets] group sums
  const int2* fscan_items;      // {run, row group} items of FSCAN_ROWS table rows
  int nfscan_items;
};

template <typename R>
struct PartBufs {
  R *x, *y, *vx, *vy;
  int* id;
};

// ---- particle kernels (k_particles.hip)
template <typename R>
void launch_sort(hipStream_t s, const RunDesc* runs, RunState* st, int nruns,
                 PartBufs<R> src, PartBufs<R> dst, SortBufs sb,
                 const int2* pitems, int npitems, const int2* citems, int ncitems,
                 long long ctotal, bool move_data);
// stable tile (bucket) sort of the step; writes sb.toff and RunState::nsorted
// scan of the fused sort's count table and the new tile offsets (toff_new, nsorted)
void launch_fused_scan(hipStream_t s, const RunDesc* runs, RunState* st, int nruns,
                       SortBufs sb, int* toff_new);
template <typename R>
void launch_tile_sort(hipStream_t s, const RunDesc* runs, RunState* st, int nruns,
                      PartBufs<R> src, PartBufs<R> dst, SortBufs sb, const int2* pitems,
                      int npitems);
// op: keys given (run 0 layout), produce perm/counts/offsets
void launch_sort_keys_only(hipStream_t s, const RunDesc* runs, RunState* st, const int* keys,
                           SortBufs sb, const int2* pitems, int npitems, const int2* citems,
                           int ncitems, long long ctotal);

#define FSCAN_ROWS 64  // rows (push items) per fused count-table scan item
#define PUSH_NT 512            // threads per push workgroup
#define PUSH_NW (PUSH_NT / 64)  // waves per push workgroup

template <typename R>
struct PushArgs {
  const RunDesc* runs;
  RunState* st;
  R *x, *y, *vx, *vy;
  const R* E;
  unsigned long long* rho;      // [2][gsum] (hi word for FP64)
  unsigned long long* rho_lo;   // [2][gsum] (FP64 only)
  long long gsum;
  const int* off;
  const int4* items;
  int nitems;
  int nruns;
  double* partials;             // [nitems*3] per-workgroup kinetic energy, px, py
  int mode;                     // 0: gather+push+deposit, 1: deposit only
  int lds_bytes;
  // sort fused into the push (mode 0, tile items): 0 off; 1 count the tiles of the new
  // positions per item into its row of ftable; 2 scatter the particles tile-major (by their
  // positions at the start of this step) into dst, with their ids
  int sort_mode;
  int* ftable;
  PartBufs<R> dst;
  const int* id;                // ids of the source buffer
};
template <typename R>
void launch_push(hipStream_t s, const PushArgs<R>& a);
// per run: diagnostics row of the step from the field and push partials, then step + 1
void launch_step_finish(hipStream_t s, const RunDesc* runs, RunState* st, int nruns,
                        const double* ppush, const double* prows, double* diag);

template <typename R>
void launch_half_kick_out(hipStream_t s, const RunDesc* runs, RunState* st, int run,
                          PartBufs<R> p, const R* E, R* out_vx, R* out_vy,
                          const int2* pitems, int npitems);
template <typename R>
void launch_unpermute(hipStream_t s, const RunDesc* runs, RunState* st, int run,
                      PartBufs<R> p, const R* vxs, const R* vys, R* ox, R* oy, R* ovx, R* ovy,
                      const int2* pitems, int npitems);
template <typename R>
void launch_check_finite(hipStream_t s, const RunDesc* runs, RunState* st, PartBufs<R> p,
                         const int2* pitems, int npitems, int* flag);

// ops on plain device arrays (run 0 descriptor)
template <typename R>
void launch_op_cell_keys(hipStream_t s, const R* x, const R* y, int64_t n, int nx, int ny,
                         double lx, double ly, int* keys);
template <typename R>
void launch_op_gather_push(hipStream_t s, R* x, R* y, R* vx, R* vy, int64_t n, const double* ex,
                           const double* ey, int nx, int ny, double lx, double ly, double dt,
                           double qm, double c, double* ke_partials, int nblocks);
template <typename R>
void launch_op_deposit(hipStream_t s, const R* x, const R* y, int64_t n, int nx, int ny,
                       double lx, double ly, unsigned long long* hi, unsigned long long* lo);
void launch_op_fixed_to_double(hipStream_t s, const unsigned long long* hi,
                               const unsigned long long* lo, long long G, int fp64,
                               double scale, double* out);
void launch_op_validate_keys(hipStream_t s, const int* keys, int64_t n, int ncells, int* flag);
void launch_fill_i32(hipStream_t s, int* p, long long n, int v);
void launch_fill_u64(hipStream_t s, unsigned long long* p, long long n);
void launch_reduce_sum(hipStream_t s, const double* parts, int n, double* out);

// ---- field kernels (k_field.hip)
// Row passes work on groups of consecutive rows (at most about 4096 complex values of LDS),
// so the transposed spectra S[k][x] are read and written in contiguous runs along x.
// Forward rows are packed in pairs (two real rows = one complex FFT).
__host__ __device__ inline int fft_row_pairs_per_block(int nx, int ny) {
  int p = 4096 / ny;
  p = p < 1 ? 1 : (p > 8 ? 8 : p);
  return p > nx / 2 ? nx / 2 : p;
}
__host__ __device__ inline int fft_rows_per_block(int nx, int ny) {
  int r = 4096 / ny;
  r = r < 1 ? 1 : (r > 16 ? 16 : r);
  return r > nx ? nx : r;
}
// Group sizes used by a launch: about one workgroup per CU (256) over all runs' rows, never
// above the per-grid limits.  A single large grid keeps one pair / row per workgroup (its
// rows alone fill the chip); many small grids get grouped rows.
inline int pow2_floor(long long v) {
  int p = 1;
  while ((long long)p * 2 <= v) p *= 2;
  return p;
}
inline void fft_groups(long long rows_total, int nx, int ny, int* pp, int* rf) {
  const long long t = rows_total / 256, tp = rows_total / 512;
  const int gr = t >= 1 ? pow2_floor(t) : 1, gp = tp >= 1 ? pow2_floor(tp) : 1;
  *rf = fft_rows_per_block(nx, ny) < gr ? fft_rows_per_block(nx, ny) : gr;
  *pp = fft_row_pairs_per_block(nx, ny) < gp ? fft_row_pairs_per_block(nx, ny) : gp;
}

struct FieldArgs {
  const RunDesc* runs;
  RunState* st;
  const unsigned long long* rho;     // [2][gsum]
  unsigned long long* rho_w;         // same buffer, written (zeroing of the next parity)
  const unsigned long long* rho_lo;  // FP64 only
  unsigned long long* rho_lo_w;
  long long gsum;
  double2* S;    // rho_hat, then phi_hat
  double2* Sx;   // Ex_hat
  double2* Sy;   // Ey_hat
  void* E;       // padded (Ex,Ey) pairs, float or double
  double* phi;   // optional node potential (may be NULL)
  const double2* tw;
  const int2* row_pair_items; int n_row_pair_items;
  const int2* col_items; int n_col_items;
  const int2* row_items; int n_row_items;
  double* partials;                  // [n_row_items] per-row-group field energy (may be NULL)
  double* diag;
  int nruns;
  int fp64;                          // particle precision (E type, rho words)
  int want_phi;
  int max_nx, max_ny;                // largest grid of the context (LDS sizing)
  int row_lds;                       // complex values per row-pass workgroup: the largest
                                     // max(fft_pp, fft_rf) * ny over the runs
  const double* rho_in;              // op mode: FP64 density input instead of rho (may be NULL)
  int clear_read;                    // 1: zero the charge words that are read (one buffer);
                                     // 0: zero the other parity's (the next push's target)
  // End-of-step work of the previous step run by extra workgroups of these kernels
  // (step_tail.h).  fold_finish: the rows pass writes the previous step's diagnostics row
  // (nruns workgroups after its row items), and the column pass advances the step counter
  // (so every kernel of this solve reads the counter one behind and adds fold_finish).
  // fold_scan: the rows, column and field passes run the fused sort's count-table scan
  // (fscan_part, fscan_base, fscan_rows) after their own items.
  int fold_finish;
  const double* ppush;               // push partials of the previous step (fold_finish)
  int fold_scan;
  int* ftable;
  int* fchunk_sums;
  int fstride;
  const int2* fscan_items; int nfscan_items;
  int* ftoff;                        // tile offsets of the order the next push writes
};
void launch_field(hipStream_t s, const FieldArgs& a, hipEvent_t* ev /* 4 events or NULL */);
void launch_op_efield(hipStream_t s, const double* phi, double* ex, double* ey, int nx, int ny,
                      double lx, double ly);
// copy padded E (float/double pairs) of one run out to two FP64 node arrays
void launch_unpad_E(hipStream_t s, const void* E, int fp64, int nx, int ny, double* ex, double* ey);
// rho (fixed point, parity of current step) of one run to FP64 density incl. background
void launch_rho_to_double(hipStream_t s, const RunDesc* runs, RunState* st, int run,
                          const unsigned long long* rho, const unsigned long long* rho_lo,
                          long long gsum, int fp64, double background, double* out);

// ---- collision kernels (k_collide.hip)
struct MccArgs {
  float *x, *y, *vx, *vy;
  int* id;
  RunState* st;       // np / next_id / err of the run (device)
  long long pbase;    // particle base of the run
  int cap;
  double v_unit2_half_me_over_e;  // E[eV] = this * |v|^2
  double ng_vunit_dtphys;         // nu*dt = this * sigma(E) * |v|
  double p_coll;                  // 1 - exp(-nu_max dt)
  double nu_max_dt;               // nu_max * dt_phys
  double mass_ratio2;             // 2 m_e / M
  uint64_t seed;
  long long* counts;              // [4]
  int* birth_count;
  int* birth_parent;              // [cap]
  float* birth_buf;               // [4*cap]
  int from_state_step;            // 1: take the step index from st->step
  int finish_after;               // 1: the step's end (counter advance) runs after this
                                  // stage, so st->step is this step's index (else its + 1)
  long long step;
  double* diag;                   // the run's diagnostics rows (may be NULL): the step's
                                  // collisions are added to column PIC_DIAG_NCOLL
  int* coll_list;                 // slots that pass the null-collision test, a region of
                                  // per_block entries per workgroup (mcc_list_layout)
  int* coll_count;                // [blocks] entries of each region
  int* coll_partials;             // [blocks][4] collision counts of each k_mcc workgroup
  int* count_table;               // sort fused into the push, after its count step: the
                                  // count table, to which the births add their tiles (in
                                  // the row of the tail item that will push them); else NULL
};
// MCC collision list layout for n_upper particles: the null-test grid and the list region
// of each workgroup (its grid-stride share); the list needs blocks * per_block entries
#define PIC_MCC_BLOCKS 2048
void mcc_list_layout(long long n_upper, int* blocks, int* per_block);
void launch_mcc(hipStream_t s, const MccArgs& a, int n_upper);
void launch_mcc_births(hipStream_t s, const MccArgs& a, unsigned long long* scratch,
                       unsigned long long* rho, long long gsum, const RunDesc* runs, int run,
                       int n_upper);
struct DsmcArgs {
  float *vx, *vy;
  const int* offsets;             // [ncells + 1] cell ranges in the sort's numbering
  float* gmax;                    // [ncells] per-cell maximum relative speed, row-major cell
  int ncells;
  int tile_shift;                 // offsets in the engine's tile-major numbering with tiles of
                                  // 1 << tile_shift cells a side; < 0: row-major already
  int tiles_y, ny;
  double coeff;
  uint64_t seed;
  long long step;
  const RunState* st;             // may be NULL (op mode uses `step`)
  int finish_after;               // 1: the step's end runs after this stage (st->step is this
                                  // step's index; else its + 1)
  long long* ncoll;
  double* diag;                   // the run's diagnostics rows (may be NULL)
};
void launch_dsmc(hipStream_t s, const DsmcArgs& a);

// ---- the reference's own collide arithmetic on an injected uniform stream (k_refcollide.hip)
#define REF_CHAIN_CHUNK 1024  // stream positions per chunk of the consumer-start resolution
struct RefChainBufs {
  int2* maps;            // [7 * nchunks] (exit offset, consumers) per chunk and entry offset
  int* entry;            // [nchunks]
  long long* base;       // [nchunks] first consumer starting in the chunk
  long long* total;      // [1] consumers the stream starts
  long long* start;      // [m] stream position of each consumer
  long long* consumed;   // [1] uniforms taken by the m consumers, -1 if the stream is short
};
void launch_ref_chain(hipStream_t s, const double* u, long long nu, double prob, long long m,
                      RefChainBufs b);
void launch_mcc_reference(hipStream_t s, double* vx, double* vy, long long n, double prob,
                          const double* u, const long long* start);
void launch_ref_cell_keys(hipStream_t s, const double* x, const double* y, long long n, int nx,
                          int ny, double dx, double dy, int* keys, int* bad);
void launch_ref_nearest(hipStream_t s, const double* x, const double* y, long long n,
                        const int* keys, const int* perm, const int* offsets, int* target,
                        int* drawkey);
void launch_dsmc_reference(hipStream_t s, double* vx, double* vy, long long n, long long m,
                           double prob, const double* u, const long long* start,
                           const int* consumers, const int* target, double* m1, double* m2,
                           int* self, int* lastw, unsigned long long* ncoll);
