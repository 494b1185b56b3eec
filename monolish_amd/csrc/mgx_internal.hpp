// mgx_internal.hpp -- shared definitions of the gfx950 CRS / Krylov library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/monolish_amd.h"

namespace mgx {

// ---------------------------------------------------------------- errors --
void set_error(const char *fmt, ...);

struct Error {
  int code;
};

#define MGX_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::mgx::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),  \
                       __FILE__, __LINE__);                                    \
      throw ::mgx::Error{e_ == hipErrorOutOfMemory ? MGX_ERR_NOMEM             \
                                                   : MGX_ERR_HIP};             \
    }                                                                          \
  } while (0)

#define MGX_REQUIRE(cond, ...)                                                 \
  do {                                                                         \
    if (!(cond)) {                                                             \
      ::mgx::set_error(__VA_ARGS__);                                           \
      throw ::mgx::Error{MGX_ERR_INVALID};                                     \
    }                                                                          \
  } while (0)

// every extern "C" entry point runs its body through guard(): exceptions
// become status codes
template <class F> int guard(F &&f) {
  try {
    return f();
  } catch (const Error &e) {
    return e.code;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed");
    return MGX_ERR_NOMEM;
  } catch (const std::exception &e) {
    set_error("%s", e.what());
    return MGX_ERR_INVALID;
  }
}

// ------------------------------------------------------- device constants --
constexpr int kBlock = 256;      // threads per block for every kernel
constexpr int kTile = 512;       // canonical reduction tile (elements)
constexpr int kMaxRemapRun = 1024;  // longest run of blocks per XCD in a plan
constexpr int kQuarter = 128;    // rows of one wave's share of a tile
constexpr int kFusedNnz = 2048;  // entry budget of a fused-SpMV chunk
constexpr int kStatRunning = 1;  // solver status while iterating

// arrival counters of the cross-block reductions (mgx_device.hpp): groups
// of blocks, supergroups of groups, one top counter; an array of kTickWords
// zeros per reduction user: [0] top, [1, 1 + kRedSuper) supergroups, then
// the groups (last_block_arrives uses [0, 1 + 64))
constexpr uint32_t kTicketGroups = 64;  // last_block_arrives
constexpr int kRedGroups = 256;         // arrive_reduce: one group per stage-2 lane
constexpr int kRedSuper = 16;
constexpr int kTickWords = 1 + kRedSuper + kRedGroups;

// device-resident solver scalars (one per solve / per batched system)
struct SolveState {
  double alpha, beta, zr, pq, rr, resid;  // CG scalars (FP64 storage)
  double rho, rho_old, omega, tt, ts, vr0;  // BiCGSTAB scalars
  double tol, bnorm;
  int64_t iter, miniter, maxiter;
  int32_t status, first;
  // CG systems whose kernels finish the reductions themselves (SysDesc::lf):
  // zr of iteration i in zrs[i & 1]; ia = iterations completed (written by
  // the update kernel), ib = the running iteration (written by the kernel
  // that starts it); fresh: 1 = start (p = z), 2 = resumed with beta and zr
  // given, 0 = continuing (beta from the last update's partials)
  double zrs[2];
  int64_t ia, ib;
  int32_t fresh, pad_;
};

// ------------------------------------------------------------ handles --
struct Ctx;

struct Vec {
  Ctx *ctx = nullptr;
  int64_t n = 0;
  mgx_dtype dtype = MGX_F64;
  void *d = nullptr;
  bool owner = true;
  size_t elem() const { return dtype == MGX_F64 ? 8 : 4; }
};

// schedule of a triangular sweep (SOR, IC / ILU factorisation and solves):
// rows grouped by level, level l = rows[ptr[l], ptr[l+1])
struct SorLevels {
  int32_t nlev = 0, max_width = 0;
  int32_t *d_rows = nullptr, *d_ptr = nullptr;
  std::vector<int32_t> ptr;
};

// plan of a triangular sweep (SOR sweeps, IC / ILU solves): rows in level
// order, rec[k] = {row, first, end, 0} of its dependency entries, packed in
// the row's order: dcol = the column read, dent = the entry (position in the
// values / factor) whose coefficient multiplies it
struct TriPlan {
  bool built = false;
  SorLevels lev;  // d_rows = the rows in level order
  int4 *d_rec = nullptr;
  // dpos: level-order position of the row read, or -1 when the sweep's LDS
  // copy of it may already be overwritten
  int32_t *d_dcol = nullptr, *d_dent = nullptr, *d_dpos = nullptr;
  int64_t ndeps = 0;
  // blocks of whole levels staged through LDS: {first level, first
  // position, first dependency entry}, nblk + 1 entries (0: not staged)
  int4 *d_blk = nullptr;
  int32_t nblk = 0;
  // dependency-driven sweep across the chip (ticket: its work counter)
  bool flow = false;
  uint32_t *d_ticket = nullptr;
};
// plans of a matrix: entries left of the diagonal; left of it up to the
// first other entry (the SOR solver's sweep); right of it; rows of L^T (IC,
// built by the IC set-up)
enum { kTriLower = 0, kTriLowerBreak = 1, kTriUpper = 2, kTriLt = 3 };
// how a row ends: out = acc * sc; the SOR backward form (acc in double
// from 0, out -= acc * sc); out = acc / sc; out = acc
enum { kTriMul = 0, kTriSorBack = 1, kTriDiv = 2, kTriUnit = 3 };
// a plan with its coefficients and per-row scale gathered in level order
struct TriOp {
  const TriPlan *plan = nullptr;
  void *coef = nullptr, *sc = nullptr;
  void *ginit = nullptr;  // staged and flow plans: the initial values in level order
  int fin = kTriMul;
};

// structure-only data of IC(0) / ILU(0) (built on first use): the checks on
// the pattern and the diagonal position of each row (the L^T plan goes to
// Csr::tri[kTriLt])
struct IluStruct {
  bool built = false;
  int32_t *d_diag = nullptr;
  int32_t max_len = 0;  // longest row (entries)
};

struct SpmvPlan {
  int2 *d_blocks = nullptr;  // nblocks+1 descriptors (first row | long flag, first entry)
  int64_t nblocks = 0;
  int64_t nlong = 0;
  int32_t max_row = 0, min_row = 0;
  int32_t nnz_max = 0;  // entry budget of a stream block (LDS products)
  int32_t rows_max = 0;  // row budget of a stream block
  int32_t long_row = 0;  // rows longer than this get a block of their own
  int32_t remap = 1;     // block order: 0 launch order, 1 one range per XCD,
                         // c > 1 runs of c blocks dealt to the XCDs in turn
  int32_t walk = 0;      // stream blocks: 0 product-staged, 1 row walk
  double mean_row = 0;
  double tuned_us = 0;   // SpMV time measured by the autotuner (0: not tuned)
  // column-panel mode (0: off): the short rows' entries are also stored
  // panel-major (panel k holds columns [k*pwidth, (k+1)*pwidth)), each pass
  // gathers x from one panel and continues the rows' sums in y; long rows
  // keep their own pass over the long blocks of d_blocks
  int32_t npanels = 0;
  int64_t pwidth = 0;
  int32_t *d_prp = nullptr;    // npanels x (rows + 1) row pointers into pci/pval
  int32_t *d_pci = nullptr;
  void *d_pval = nullptr;
  int2 *d_pblocks = nullptr;   // stream-block descriptors of all panels
  std::vector<int64_t> pblock_off;  // panel k: descriptors [off[k], off[k+1]) (+1 end)
  int32_t *d_lrows = nullptr;  // the long rows (nlong of them)
  // sliced mode (0: off): the short rows' entries are also stored in slices
  // of 64 rows and w-entry tuples, entry j of a slice's row l at
  // soff[slice] + ((j / w) * 64 + l) * w + j % w, w = 16 / sizeof(value):
  // `sl` for A's values, `sl32` (w = 4) for the FP32 copy of an FP64
  // matrix; one lane per row; long rows keep the block-tree path
  struct Sliced {
    int w = 0;
    int64_t entries = 0;
    int64_t *d_soff = nullptr;
    int32_t *d_sci = nullptr;
    void *d_val = nullptr;
    // row lengths, min(len, 255): the kernels read row_ptr only for rows
    // of 255 entries or more
    uint8_t *d_len = nullptr;
  };
  int32_t sell = 0;
  Sliced sl, sl32;
  std::vector<int64_t> soff32;  // host offsets of sl32, kept until it is built
  int32_t *d_slrows = nullptr;
  int64_t nslong = 0;
  // L2-panel mode (0: off): a block owns lrows consecutive rows and keeps
  // their sums in LDS while it walks lpanels column panels narrow enough for
  // one XCD's L2; segment (block, panel) lists the block's short-row entries
  // of that panel grouped by their rank within the row's run in the panel
  // (rank 0 of every row, then rank 1, ...), each group in row order, so a
  // barrier between groups adds every row's products in storage order.
  // Entry index = (row - first row of the block) << 20 | (col - panel start).
  int32_t lpanels = 0, lrows = 0;
  int64_t lpw = 0, lnblk = 0;
  int32_t *d_lsegoff = nullptr;  // lnblk * lpanels + 1 entry offsets
  int32_t *d_lrgoff = nullptr;   // lnblk * lpanels + 1 offsets into d_lrbound
  int32_t *d_lrbound = nullptr;  // end of each rank group, from its segment start
  uint32_t *d_lidx = nullptr;
  void *d_lval = nullptr;
  uint32_t *d_llong = nullptr;   // bit per row: a long row (written by the long-row kernel)
  int32_t *d_llrows = nullptr;   // the long rows
  int64_t nllong = 0;
};

struct Csr {
  Ctx *ctx = nullptr;
  int64_t rows = 0, cols = 0, nnz = 0;
  mgx_dtype dtype = MGX_F64;
  int32_t *rp = nullptr;   // effective (non-decreasing) row_ptr on device
  int32_t *ci = nullptr;
  void *val = nullptr;
  std::vector<int32_t> rp_raw;  // kept only when it differs from rp
  SpmvPlan plan;
  // fused-solver plan: row chunks that never cross a 512-row tile; the sign
  // bit marks a long row; tile k owns chunks [tile_chunk[k], tile_chunk[k+1])
  int32_t *d_chunks = nullptr;
  int32_t *d_tile_chunk = nullptr;
  int32_t fused_nnz = 0;     // entry budget of a chunk
  float *val32 = nullptr;    // FP32 shadow values (mixed precision)
  // per row: the FP64 diagonal entries minus their FP32 roundings (FP64
  // matrices, built with val32)
  float *dcorr32 = nullptr;
  Csr *trans = nullptr;      // A^T, built on first matvec_T
  TriPlan tri[4];            // triangular sweep plans (kTriLower ... kTriLt)
  IluStruct ilu;             // IC / ILU structure
  // distinct 128-byte lines of x gathered by 64 consecutive entries
  // (sampled average; 0 when not measured): high for scattered columns
  double gather_lines = 0;
  // the fused solvers' own sliced copies of the values (esl) and of their
  // FP32 copy (esl32), apart from the SpMV plan: offsets computed at upload
  // for square matrices whose rows all hold at most kEngineSellRow entries
  // and whose slices pad by at most 30 % (empty otherwise), copies built on
  // a solver's first use while the plan has no sliced copy of its own
  std::vector<int64_t> esoff, esoff32;
  SpmvPlan::Sliced esl, esl32;
  size_t elem() const { return dtype == MGX_F64 ? 8 : 4; }
};
constexpr int32_t kEngineSellRow = 256;

struct Ctx {
  int device = 0;
  int32_t batch_iters_per_poll = 0;  // schedule of the last batched solve
  hipStream_t stream = nullptr;      // where every call queues its work
  hipStream_t own_stream = nullptr;  // the library's stream (stream unless replaced)
  int num_cus = 0;
  size_t hbm_bytes = 0;
  char name[256] = {0};
  // reduction scratch
  void *d_partials = nullptr;
  size_t partials_bytes = 0;
  SolveState *d_state = nullptr;     // scratch solve state for BLAS calls
  uint32_t *d_tick = nullptr;        // arrival counters of the BLAS reductions
  double *h_pinned = nullptr;        // pinned host scalars (64 doubles)
  hipEvent_t ev = nullptr;
  // SpMV tuning
  int spmv_rows = 256, spmv_nnz = 2048, spmv_long = 256, xcd_remap = 1;
  int use_graphs = 0;  // replay fused iterations from a captured hipGraph
  int use_persistent = 0;  // single small CG solves in one cooperative launch
                           // (> 1: also the largest grid it may use)
  // fused-engine state owned by the context (engine.hip)
  void *pool = nullptr;   // DevicePool for solver work vectors
  void *eng64 = nullptr;  // Engine<double> for single fused solves
  void *eng32 = nullptr;  // Engine<float>
  // workspace for solver vectors
  std::vector<std::pair<void *, size_t>> ws_free;
  std::vector<std::pair<void *, size_t>> ws_used;
  // vectors and matrices created in this context and not destroyed yet;
  // mgx_finalize releases them
  std::mutex objs_mu;
  std::unordered_set<Vec *> live_vecs;
  std::unordered_set<Csr *> live_csrs;
};

// ------------------------------------------------------- host helpers --
void *dev_alloc(size_t bytes);
void dev_free(void *p);
void *ws_get(Ctx *c, size_t bytes);   // pooled device scratch
void ws_release_all(Ctx *c);
void ws_release_to(Ctx *c, size_t mark);  // release buffers taken after mark

inline void activate(Ctx *c) {
  static thread_local int cur = -1;
  if (cur != c->device) {
    MGX_HIP(hipSetDevice(c->device));
    cur = c->device;
  }
}
void ensure_partials(Ctx *c, size_t bytes);
// looks up the device address of the allocation-free dot's static scratch
// on `device` (which must be current); called once per context set-up
void noalloc_scratch_init(int device);

__host__ __device__ inline int64_t ntiles_of(int64_t n) { return (n + kTile - 1) / kTile; }

// kernel launchers (defined in the .hip files)
void launch_fill(Ctx *c, mgx_dtype dt, void *y, int64_t n, double v);
void launch_axpy(Ctx *c, mgx_dtype dt, int64_t n, double a, const void *x,
                 void *y);
void launch_xpay(Ctx *c, mgx_dtype dt, int64_t n, double a, const void *x,
                 void *y);
void launch_axpyz(Ctx *c, mgx_dtype dt, int64_t n, double a, const void *x,
                  const void *y, void *z);
void launch_scal(Ctx *c, mgx_dtype dt, int64_t n, double a, void *x);
void launch_copy(Ctx *c, mgx_dtype dt, int64_t n, const void *x, void *y);
// op: 0 add, 1 sub, 2 mul
void launch_vv(Ctx *c, mgx_dtype dt, int op, int64_t n, const void *a,
               const void *b, void *y);
void launch_reciprocal(Ctx *c, mgx_dtype dt, int64_t n, const void *a,
                       void *y);
void launch_convert(Ctx *c, mgx_dtype from, mgx_dtype to, int64_t n,
                    const void *x, void *y);
// dot into a device double slot (result converted to double exactly); sq:
// its square root in the vectors' precision (nrm2)
void launch_dot(Ctx *c, mgx_dtype dt, int64_t n, const void *x, const void *y,
                double *d_out, uint32_t *d_ticket, bool sq = false);
double dot_host(Ctx *c, mgx_dtype dt, int64_t n, const void *x,
                const void *y);
// d_out[r] = <x_r, x_r> for the m rows x_r = X + r ld, each the canonical
// dot of launch_dot; partials: m (ntiles(n) + 2 kRedGroups) values of dt,
// ticks: m kTickWords zeroed counters (zero again afterwards)
void launch_mnrm2sq(Ctx *c, mgx_dtype dt, int64_t n, int m, const void *X, int64_t ld,
                    void *partials, uint32_t *ticks, double *d_out);

void spmv_plan_build(Ctx *c, Csr *A, const int32_t *rp_host);
// plan with explicit budgets (the context's defaults otherwise)
void spmv_plan_build_with(Csr *A, const int32_t *rp_host, int rows_max, int nnz_max,
                          int long_row, int remap, int walk);
// live: optional device status word; the launch does nothing unless it
// reads kStatRunning (device-driven solver loops)
void launch_spmv(Csr *A, const void *val, mgx_dtype dt, const void *x,
                 void *y, double alpha, double beta, bool ab,
                 const int32_t *live = nullptr);
void launch_diag(Csr *A, void *d);
void spmv_autotune(Csr *A, int reps, mgx_spmv_tuning *out);
// build / drop the column-panel mode of A's plan (false: not applicable, e.g.
// columns not ascending within a row)
bool spmv_panels_build(Csr *A, int npanels);
void spmv_panels_free(Csr *A);
bool spmv_sell_build(Csr *A, const int32_t *rp_host);
// build / drop the L2-panel mode (false: not applicable -- columns not
// ascending within a short row, or a panel wider than 2^20 columns)
bool spmv_lpanel_build(Csr *A, int npanels);
void spmv_lpanel_free(Csr *A);
// host arrays of the L2-panel layout (built from host CRS arrays of es-byte
// values); false when not applicable
struct LpanelLayout {
  int P = 0, R = 0;
  int64_t pw = 0, nblk = 0;
  std::vector<int32_t> segoff, rgoff, rbound, lrows;
  std::vector<uint32_t> idx, lng;
  std::vector<char> val;
};
bool lpanel_layout(const int32_t *rp, const int32_t *ci, const char *val, size_t es, int64_t n,
                   int64_t cols, int long_t, int npanels, LpanelLayout &L);
void spmv_sell_free(Csr *A);
void spmv_sell_values(Csr *A);  // refresh the sliced values after a change
void launch_diag_v(Csr *A, const void *val, mgx_dtype dt, void *d);
// fused-solver row chunks (built at upload for square matrices, from the
// host row_ptr when given, else from a copy of the device one)
void fused_plan_build(Csr *A, const int32_t *rp_host = nullptr);
// col_ind / val of nnz entries plus 4 zeroed padding entries
void alloc_entries(Csr *A, int64_t nnz);
// average distinct 128-byte x lines per 64 consecutive entries of col_ind
// (host array), over up to 64 windows spread across it
double sample_gather_lines(const int32_t *ci, int64_t nnz, size_t es);
void csr_free(Csr *A);
// A^T built on the device (stable sort of the entries by column)
Csr *csr_transpose(Csr *A);
// schedule of any row structure (ptr, col; ent = entry positions, nullptr =
// the identity): filter 2 rows depend on their entries right of the
// diagonal, else on those left of it (filter 1 stops at the first other one)
void build_tri_plan(const std::vector<int32_t> &ptr, const std::vector<int32_t> &col,
                    const std::vector<int32_t> *ent, int64_t n, int filter, TriPlan &P);
// plan kTriLower / kTriLowerBreak / kTriUpper of A's pattern, built on first
// use from the device row_ptr / col_ind
const TriPlan &tri_plan(Csr *A, int which);
// gather coef[d] = vals[dent[d]] and sc[k] = scale[scale_pos ? scale_pos[row]
// : row] for the plan (buffers from the context's scratch)
TriOp tri_prepare(Csr *A, const TriPlan &P, mgx_dtype dt, int fin, const void *vals,
                  const void *scale, const int32_t *scale_pos);
// one sweep: out_row = fin(in_row - sum coef * out_col) in level order
void tri_run(Csr *A, const TriOp &op, mgx_dtype dt, const void *in, void *out);
void levels_free(SorLevels &L);
void tri_free(Csr *A);
// throws MGX_ERR_TIMEOUT (and clears the flag) when a dependency-driven sweep
// of one of A's plans gave up since the last check
void tri_check_stall(Csr *A);
// IC(0) (kind MGX_PRECOND_IC) / ILU(0) (MGX_PRECOND_ILU) of A's pattern:
// f (A->nnz entries of dt) holds a copy of the values on entry and the
// factor on return; throws on unsorted columns, a missing diagonal or a
// zero (IC: non-positive) pivot, naming the row.  Builds the structure
// (diagonal positions, the kTriLt plan for IC) on first use
void ilu_factor(Csr *A, mgx_dtype dt, int kind, void *f);
const int32_t *ilu_diag(Csr *A);
void ilu_free(Csr *A);
// LOBPCG: out (k (k + 1) doubles) = the lower triangles of Sa = WXP wxp^T
// and Sb = wxp wxp^T, row-major by pair (i, j), i >= j (Sa first); part
// holds ceil(n / 128) * k (k + 1) doubles of tile partials
void lobpcg_gram(Ctx *c, mgx_dtype dt, int64_t n, int k, const void *wxp, const void *WXP,
                 double *part, double *out);
// rows m..2m-1 (x) and 2m..3m-1 (p) of wxp and WXP from the previous basis
// wxp_p, WXP_p with coef (m rows of `stride` coefficients, device); has_p:
// the previous basis holds p
// Y_r = A X_r for m <= 10 rows (stride n) in one pass over A, when A's rows
// have at most 256 entries (then the sums equal the SpMV's); false otherwise
bool lobpcg_spmm(Csr *A, mgx_dtype dt, int m, const void *X, void *Y);
// rows r of A (and of B when not null), r < m <= 10, scaled by (T)s[r]
void lobpcg_scale_rows(Ctx *c, mgx_dtype dt, int64_t n, int m, const double *s, void *A,
                       void *B);
// pair r <= 10: x_r *= s[r], X_r *= s[r], w_r = nlam[r] x_r + X_r, then
// w_r = M o w_r when M (Jacobi) is not null; factors rounded to dt
void lobpcg_xw(Ctx *c, mgx_dtype dt, int64_t n, int m, const double *s, const double *nlam,
               void *x, void *X, void *w, const void *M);
// q = A p in the sliced layout with the tile partials of <p, q> written to
// P (the CG engine's split-form stage 3), for matrices whose values val
// have a sliced layout and no long rows; false (nothing launched) otherwise
bool spmv_sell_pq_ok(Csr *A, const void *val, mgx_dtype dt);
// the sliced copy of the values val (A's or their FP32 copy) for the fused
// solvers: the plan's when it has one and no rows above the long-row
// threshold, else the engine's own (built here on first use); nullptr when
// A has neither
const SpmvPlan::Sliced *spmv_sell_layout(Csr *A, const void *val, mgx_dtype dt);
// the same lookup without building anything
const SpmvPlan::Sliced *spmv_sell_find(const Csr *A, const void *val, mgx_dtype dt);
// at upload: the engine copies' offsets (Csr::esoff, esoff32)
void spmv_engine_offsets(Csr *A, const int32_t *rp_host);
// after new values: rewrite the engine copies that exist
void spmv_engine_values(Csr *A);
void spmv_engine_free(Csr *A);
// q = A p in the sliced layout with the <p, q> tile partials into P; FP32
// with acc64: partials in FP64 (P of doubles) and, when dcorr is given,
// q[i] += dcorr[i] * p[i]
bool spmv_sell_pq(Csr *A, const void *val, mgx_dtype dt, const void *p, void *q, void *P,
                  const int32_t *live, bool acc64 = false, const float *dcorr = nullptr);
void lobpcg_combine(Ctx *c, mgx_dtype dt, int64_t n, int m, int stride, bool has_p,
                    const void *coef, const void *wxp_p, const void *WXP_p, void *wxp,
                    void *WXP);
// FP32 shadow of an FP64 matrix's values (mixed precision), made once
void ensure_val32(Csr *A);
// (re)computes A->dcorr32 from A's FP64 values
void diag_corr_build(Csr *A);

double seconds_now();

}  // namespace mgx

struct mgx_ctx : mgx::Ctx {};
struct mgx_vec : mgx::Vec {};
struct mgx_csr : mgx::Csr {};
