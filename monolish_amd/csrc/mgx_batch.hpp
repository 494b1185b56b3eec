// mgx_batch.hpp -- multi-system fused Krylov engine (host side + descriptors).
#pragma once

#include <deque>
#include <memory>
#include <unordered_map>
#include <vector>

#include "mgx_internal.hpp"

namespace mgx {

// one system of a fused launch; vectors are indexed by global row
// (row0 + local row), matrix entries by the system's own row_ptr
template <class T> struct SysDesc {
  const int32_t *rp, *ci;
  const T *val;
  const int32_t *chunks, *tile_chunk;
  const T *M;  // Jacobi 1/diag or nullptr
  T *x;
  const T *b;
  T *r, *z, *q, *pa, *pb;
  T *P;  // per-unit partials of the reductions
  T *S;  // stage-2 lane sums (2 * kRedGroups)
  T *r0, *s, *shat, *phat, *t, *va, *vb;  // BiCGSTAB
  SolveState *st;
  uint32_t *tick;  // kTickWords arrival counters of this system's reductions
  double *rhist;
  int64_t rhist_len;
  int64_t row0, n, ntiles;
  // CG only: 1 = split iteration for large systems (p = z + beta p; q = A p
  // by the plan-based SpMV; <p, q>; update), 0 = two fused kernels
  int32_t split;
  int32_t vec;  // row0 even and x, b 16-byte aligned: vector loads allowed
  int32_t method;  // kMethodCG or kMethodBiCGSTAB
  // 1 = gather-once form: CG (two-kernel form) forms p = z + beta p in k_cg_p
  // before its SpMV, BiCGSTAB forms p, phat and s, shat in k_bcg_p / k_bcg_s,
  // so each SpMV gathers one vector per entry (instead of 2 or 3-4)
  int32_t once;
  // split form only: its SpMV also writes the <p, q> tile partials (sliced
  // layout, no long rows), so the separate stage-3 kernel skips it
  int32_t pq_fused;
  // CG, two-kernel form, at most kRedGroups tiles: no arrival chain; the
  // kernel after a reduction's producer reads all its tile partials and
  // finishes it in every block (the canonical order), the system's first
  // block writing the solve state (see k_cg_spmv, k_cg_p, k_cg_update)
  int32_t lf;
  // 1 = the matrix's sliced layout (soff, sci, sval: the CSR's sliced copy of
  // these values, no rows above the long-row threshold) is there: the fused
  // SpMV stages walk it one lane per row instead of staging products in LDS
  // (same bits)
  int32_t sell;
  const uint8_t *slen;
  const int64_t *soff;
  const int32_t *sci;
  const T *sval;
  double tol_scale;  // set-up: tol = tol_scale * nrm2(b) when nonzero
  // FP32 split-form CG inside the mixed-precision solver only (pq_fused):
  // 1 = the iteration's reductions multiply and add in FP64 (P holds double
  // partials), and q = A p adds dcorr[i] * p[i] per row (dcorr: the FP64
  // diagonal minus its FP32 rounding)
  int32_t acc64;
  const T *dcorr;
};

struct TileRef {
  int32_t slot, tile;
};

enum { kMethodCG = 0, kMethodBiCGSTAB = 1 };

// CG systems of at least this many rows run the split iteration: launch
// overhead is small against their SpMV, and the plan-based SpMV streams
// better than the fused quarter blocks, which gather two vectors
constexpr int64_t kSplitRows = 1 << 19;
// matrices whose 64 consecutive entries gather from more than this many
// distinct 128-byte lines of x (sampled at upload) use the gather-once form
constexpr double kScatteredLines = 24.0;

// one CG iteration of the listed systems (split-form systems are not
// listed): the fused SpMV stage (when a listed system uses the two-kernel
// form) and the update stage; gather-once systems need batch_launch_cg_p
// first
template <class T>
void batch_launch_cg_iter(Ctx *c, const SysDesc<T> *sys, const TileRef *tiles, int64_t nlaunch,
                          int32_t *stopped, int lds_entries, bool any_fused);
template <class T>
void batch_launch_cg_p(Ctx *c, const SysDesc<T> *sys, const TileRef *tiles, int64_t nlaunch,
                       int32_t *stopped);
// split form, one system per launch (descriptor by value, kSplitTiles tiles
// per block): stage 0 p = z + beta p, 1 the <p,q> tile partials (when its
// SpMV does not form them), 2 the update with the <z,r>, <r,r> partials
constexpr int kSplitTiles = 2;
template <class T> void batch_launch_split(Ctx *c, const SysDesc<T> &d, int stage);
// split form: stage 2 of <p,q> (stage 0, after the SpMV) or of <z,r>,
// <r,r> with the convergence test (stage 1, after the update), one block
// per listed split system
template <class T>
void batch_launch_cg_finish(Ctx *c, const SysDesc<T> *sys, const int32_t *slots, int nslots,
                            int stage, int32_t *stopped);
// one BiCGSTAB iteration; any_once: a listed system uses the gather-once
// form (p, phat and s, shat formed by streaming passes before the SpMVs)
template <class T>
void batch_launch_bicgstab_iter(Ctx *c, const SysDesc<T> *sys, const TileRef *tiles,
                                int64_t nlaunch, int32_t *stopped, int lds_entries,
                                bool any_once);

// grid barrier state of the persistent kernels; zero before each launch
struct GridSync {
  uint64_t group[64];
  uint64_t top;
  uint64_t release;
  uint32_t err;  // a block timed out waiting
  uint32_t pad;
};
// blocks of the persistent CG kernel that fit on the device at once
template <class T> int batch_cg_persist_blocks(Ctx *c, int lds_entries);
// up to max_iters iterations of the single system sys[0] in one cooperative
// launch of `grid` blocks (grid <= batch_cg_persist_blocks)
template <class T>
void batch_launch_cg_persist(Ctx *c, const SysDesc<T> *sys, GridSync *gs, int grid,
                             int32_t *stopped, int lds_entries, int max_iters);

// set-up of newly admitted systems: scatter their staged descriptors and
// states into their slots, then (over the list of their tiles) M = 1/diag,
// r = b - A x with z = M r or r0 = r, the relative tolerances, <z,r> or
// <r,r0>, and the BiCGSTAB first-iteration breakdown test
template <class T>
void batch_launch_setup(Ctx *c, const SysDesc<T> *stage_d, const SolveState *stage_s,
                        const int32_t *stage_slots, int nstage, SysDesc<T> *sys,
                        SolveState *states, const TileRef *tiles, int64_t ntiles, bool any_m,
                        bool any_rel, bool any_bcg, int lds_entries, int32_t *stopped);

// device memory pool with reuse by size (no hipMalloc on the steady path)
class DevicePool {
 public:
  explicit DevicePool(size_t budget) : budget_(budget) {}
  ~DevicePool();
  void *get(size_t bytes);  // nullptr when the budget would be exceeded
  void put(void *p);
  // Makes the next get() of each listed size (one entry per call) a hit:
  // the blocks not already cached come from ONE hipMalloc'd slab, carved
  // into the free lists (a batch of thousands of systems then makes one
  // allocation instead of one per work vector).  Carved blocks return to
  // the free lists like any other and are released with their slab when
  // the pool is destroyed.  False (nothing allocated) past the budget.
  bool reserve(const std::vector<size_t> &bytes);
  size_t in_use() const { return in_use_; }
  size_t budget() const { return budget_; }
  static size_t round(size_t bytes);

 private:
  size_t budget_, in_use_ = 0, reserved_ = 0;
  std::unordered_map<size_t, std::vector<void *>> free_;  // by rounded size
  std::unordered_map<void *, size_t> live_;
  std::vector<void *> slabs_;
  std::unordered_set<void *> carved_;  // blocks inside a slab (never freed alone)
};

struct SolveCfg {
  int method = kMethodCG;
  int precond = MGX_PRECOND_JACOBI;
  double tol = 1e-8;  // absolute, or relative to ||b|| when tol_relative
  int tol_relative = 0;
  int64_t miniter = 0, maxiter = INT64_MAX / 4;
  double *rhist = nullptr;  // device array or nullptr
  int64_t rhist_len = 0;
  // CG: gather-once form for any matrix (multi-system launches, where the
  // extra p pass costs less than the second gather); otherwise only for
  // matrices with scattered columns.  BiCGSTAB always uses it.
  int gather_once = 0;
  // mixed-precision inner CG (FP32 systems in the split form with the fused
  // <p, q>; ignored otherwise): FP64 reductions, and the FP32 diagonal
  // corrected by dcorr (float per row) in q = A p
  int acc64 = 0;
  const void *dcorr = nullptr;
};

// FP32-refined solves: one system's step of iterative refinement
struct IrSys {
  const double *r;  // FP64 residual (an FP64 engine's work vector)
  float *r32, *d32;
  double *x;
  double scale;  // 1 / ||r||
  double rn;     // ||r||
  int64_t n;
};
// over tiles (slot = index into sys): r32 = (float)(r / ||r||), d32 = 0
void launch_ir_scale(Ctx *c, const IrSys *sys, const TileRef *tiles, int64_t ntiles);
// over tiles: x += ||r|| d32
void launch_ir_update(Ctx *c, const IrSys *sys, const TileRef *tiles, int64_t ntiles);

// A set of systems iterated together.  add() uploads nothing: the matrix and
// x/b are device resident; it allocates the work vectors from the pool and
// records the system.  The set-up of all systems added since the last launch
// runs in a few shared launches ahead of the next step() (or
// run_persistent()).  step() queues `iters` fused iterations of all running
// systems; poll() reads which systems have stopped.
template <class T> class Engine {
 public:
  Engine(Ctx *c, DevicePool *pool, int capacity);
  ~Engine();
  // returns the slot, or -1 when the pool budget or capacity is exhausted
  int add(Csr *A, const void *val, T *x, const T *b, const SolveCfg &cfg);
  size_t bytes_needed(const Csr *A, int method) const;
  // the device blocks add() takes from the pool for this system (for
  // DevicePool::reserve ahead of a batch)
  static void sizes_of(const Csr *A, const SolveCfg &cfg, std::vector<size_t> &out);
  void step(int iters);
  int running() const { return host_running_; }
  // synchronises; returns slots that stopped since the last poll
  std::vector<int> poll();
  SolveState state(int slot);  // valid after poll() reported the slot
  // continue a stopped slot from the state s (its vectors as they are now)
  void resume(int slot, const SolveState &s);
  // the same, with the slot's new state already written to its device
  // state (status running, fresh and ia as resume sets them) by work
  // queued on the context's stream
  void resume_device(int slot);
  // the slot's device state (for kernels that write it)
  SolveState *device_state(int slot) { return d_states_ + slot; }
  void release(int slot);
  int active() const { return (int)active_.size(); }
  // run the set-up of the systems added since the last launch now
  void prepare() { flush_pending(); }
  // copy every slot's device state to the host (state() then reads it)
  void fetch_states();
  const SysDesc<T> &desc(int slot) const { return h_sys_[slot]; }
  // the one running system is a CG system of the two-kernel form, and the
  // context allows the persistent kernel
  bool persistent_ok(int slot) const;
  // iterate that system to its end with the persistent kernel
  void run_persistent(int slot);

 private:
  void rebuild_lists();
  void launch_iters(int iters);
  void flush_pending();  // set-up of the systems added since the last launch
  hipGraphExec_t gexec_ = nullptr;
  int giters_ = 0;
  bool gstale_ = true;
  Ctx *c_;
  DevicePool *pool_;
  int cap_;
  SysDesc<T> *d_sys_ = nullptr;
  SolveState *d_states_ = nullptr;
  int32_t *d_stopped_ = nullptr;
  int32_t host_stopped_ = 0;
  TileRef *d_tiles_[2] = {nullptr, nullptr};
  int64_t tiles_cap_[2] = {0, 0};
  int64_t ntiles_[2] = {0, 0};
  bool dirty_ = true;
  // the running slots the tile lists were last built for; listed_ok_ false
  // after an add() or release() (a slot's descriptor may differ)
  std::vector<int> listed_;
  bool listed_ok_ = false;
  std::vector<SysDesc<T>> h_sys_;
  std::vector<int> method_;
  std::vector<int> free_slots_;
  std::vector<int> active_;
  std::vector<std::vector<void *>> bufs_;
  std::vector<SolveState> h_states_;
  std::vector<char> reported_;
  int host_running_ = 0;
  int lds_entries_ = 0;
  GridSync *d_gs_ = nullptr;      // grid barrier of the persistent kernel
  uint32_t *d_ticks_ = nullptr;   // kTickWords arrival counters per slot
  std::vector<int> pending_;      // slots added since the last launch
  std::vector<char> rel_;         // per slot: relative tolerance
  std::vector<char> h_stage_;     // staged descriptors / states / slots
  void *d_stage_ = nullptr;
  size_t stage_cap_ = 0;
  TileRef *d_ptiles_ = nullptr;   // tiles of the pending systems
  int64_t ptiles_cap_ = 0;
  bool any_fused_ = false;        // a listed CG system uses the two-kernel form
  bool any_once_ = false;         // ... in its gather-once variant
  bool any_bcg_once_ = false;     // a listed BiCGSTAB system in the gather-once form
  std::vector<int> split_slots_;  // listed CG systems of the split form
  int32_t *d_split_slots_ = nullptr;
  int64_t split_cap_ = 0;
  std::vector<Csr *> csr_;        // per slot: matrix and values (split SpMV)
  std::vector<const void *> val_;
};

// the context's pool and single-solve engines (created on first use, freed
// by ctx_free_engines at mgx_finalize)
DevicePool &ctx_pool(Ctx *c);
template <class T> Engine<T> &ctx_engine(Ctx *c);
void ctx_free_engines(Ctx *c);

}  // namespace mgx
