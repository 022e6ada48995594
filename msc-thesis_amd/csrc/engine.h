// engine.h -- host-side engine context (internal to libpicamd.so).
#pragma once
#include <string>
#include <vector>
#include "../../include/pic_capi.h"
#include "launch.h"

namespace pic {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define PIC_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return ::pic::fail(e_ == hipErrorOutOfMemory ? PIC_ERR_OOM : PIC_ERR_HIP,          \
                         std::string(#call) + ": " + hipGetErrorString(e_));            \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc((void**)p, count * sizeof(T));
}

int validate_config(const pic_config& c);
double mcc_nu_max(const pic_config& c);

struct Tuning {
  int tile, halo, splits;
  // splits per tile from each tile's particle count (DSMC runs asked for splits 0: the
  // blob-like densities of the collision workloads leave uniform splits unbalanced)
  bool per_tile = false;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int nruns = 0;
  int fp64 = 0;
  size_t rsize = 4;  // bytes per particle real
  std::vector<pic_config> cfg;
  std::vector<RunDesc> rd;
  std::vector<Tuning> tune;
  std::vector<int> has_particles;
  std::vector<double> numax_dt;     // MCC nu_max * dt_phys * dt per run
  int sort_every = 4;
  int steps_since_sort = 0;
  bool use_graphs = true;
  int max_nx = 0, max_ny = 0;
  long long ptotal = 0, gsum = 0, esum = 0, ssum = 0, ctotal = 0;
  bool allocated = false;
  bool any_mcc = false, any_dsmc = false;

  RunDesc* d_runs = nullptr;
  RunState* d_st = nullptr;
  void* px[2] = {nullptr, nullptr};
  void* py[2] = {nullptr, nullptr};
  void* pvx[2] = {nullptr, nullptr};
  void* pvy[2] = {nullptr, nullptr};
  int* pid[2] = {nullptr, nullptr};
  int cur = 0;
  unsigned long long* rho = nullptr;     // [2*gsum]
  unsigned long long* rho_lo = nullptr;  // [2*gsum] (FP64)
  void* E = nullptr;
  double2 *S = nullptr, *Sx = nullptr, *Sy = nullptr;
  double* phi = nullptr;
  double2* tw = nullptr;
  SortBufs sb{};
  int* toffs[2] = {nullptr, nullptr};  // tile offsets of the order in buffer cur / 1 - cur
  long long toffsum = 0;               // entries of each toffs[] array
  int* collide_count_table = nullptr;  // the count table while a count step's collision
                                       // stage runs (its births are counted; else NULL)
  bool fused = false;                  // sort fused into the push (see build_items)
  std::vector<int2> pitems, citems, rowpair, cols, rows, segitems, tscan_items, fscan_items;
  std::vector<int4> push_items;
  int2 *d_pitems = nullptr, *d_citems = nullptr, *d_rowpair = nullptr, *d_cols = nullptr,
       *d_rows = nullptr, *d_segitems = nullptr, *d_tscan_items = nullptr,
       *d_fscan_items = nullptr;
  int4* d_push = nullptr;
  double* partials_push = nullptr;
  double* partials_rows = nullptr;
  double* diag = nullptr;
  int lds_push = 0;
  // collisions
  long long* d_coll_counts = nullptr;  // [nruns*4]
  int* birth_count = nullptr;
  int* coll_list = nullptr;   // MCC: slots that pass the null-collision test
  int* coll_count = nullptr;
  int* coll_partials = nullptr;
  int* birth_parent = nullptr;
  float* birth_buf = nullptr;
  float* gmax = nullptr;               // per cell (ctotal)
  // output scratch
  int* flag = nullptr;
  // graphs
  hipGraphExec_t macro[2] = {nullptr, nullptr};
  int macro_k[2] = {0, 0};
  int macro_flip[2] = {0, 0};  // 1: the sequence ends in the other particle buffer
  hipEvent_t ev[8] = {};

  ~Ctx();
  int alloc();
  int build_items();
  void push_items_for(int r, const int* toff_h);
  int rebalance_push_items();
  int upload_runs();
  void free_graphs();

  template <typename R> PartBufs<R> bufs(int which) const {
    PartBufs<R> b;
    b.x = (R*)px[which]; b.y = (R*)py[which]; b.vx = (R*)pvx[which]; b.vy = (R*)pvy[which];
    b.id = pid[which];
    return b;
  }
  FieldArgs field_args(int want_phi) const;
  template <typename R> PushArgs<R> push_args(int mode, int sort_mode = 0) const;

  // launches (no synchronisation)
  int launch_field_stage(int want_phi, hipEvent_t* ev4 = nullptr);
  int launch_push_stage(int mode, int sort_mode = 0);
  int launch_sort_stage();
  int launch_collide_stage();
  double numax_dt_of(int r);
  // one step; ev (7 events or NULL) brackets the stages for pic_stage_times.  defer_tail:
  // the step's end-of-step work (diagnostics row and step counter; the count-table scan on
  // a count step) is left to the next step's field kernels (fold_finish / fold_scan of
  // FieldArgs), which must then be launched next, by launch_one_step
  int launch_one_step(bool allow_sort, hipEvent_t* ev = nullptr, bool defer_tail = false);
  bool fin_pending = false;   // the previous step's finish is deferred
  bool scan_pending = false;  // the previous step's count-table scan is deferred
  int* scan_toff = nullptr;   // its tile offsets output
  int step(int nsteps);
  int capture_macro(int which);
  // steps per captured sequence: a sort cycle, or (DSMC: a cell sort every step) eight steps
  int macro_len() const { return any_dsmc ? 8 : sort_every; }
  // a step's end runs after its collision stage (any collisions in the context)
  bool finish_late() const { return any_mcc || any_dsmc; }
};

}  // namespace pic

struct pic_ctx {
  pic::Ctx impl;
};

namespace pic {
// State restored with the particles (checkpoint restart): staggered velocities, step and id
// counters, and the macro-particle charge fixed when the run started.
struct Restore {
  long long step;
  int vsync;
  long long next_id;
  double qp;
};
// Uploads n particles into a run, sorts them and deposits their charge; restore == NULL
// starts a fresh run (synchronous velocities, step 0).
int set_particles_impl(pic_ctx* ctx, int run, const void* px, const void* py, const void* pvx,
                       const void* pvy, int64_t n, bool device_src, const Restore* restore);
}  // namespace pic
