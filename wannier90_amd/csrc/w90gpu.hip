// libw90gpu: MI355X wannierisation (spread minimisation without
// disentanglement) behind the C ABI of include/w90gpu.h.  Device code is in
// w90gpu_kernels.h; this file holds the host side.
//
// Reference: src/wannierise.F90 (wann_main :48-1072), src/overlap.F90
// (overlap_project :758-952), src/library_interface.F90 (w90_wannierise
// :464-530).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "w90gpu_ctx.h"
#include "w90gpu_kernels.h"
#include "w90gpu_mfma.h"


// Device memory of the allocation-free entry point w90gpu_omega_noalloc:
// one static array, which every GPU the library runs on holds until the
// process ends.  Host code never uses the name itself (on the host it is a
// stand-in, not a device address): w90gpu_create looks up the address of
// each device's copy once (scratch_address) and the entry point uses that.
constexpr size_t kScratchBytes = W90GPU_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) char scratch[kScratchBytes];

// Host side of libw90gpu: the engine that runs a set of problems on one
// device, and the C ABI of include/w90gpu.h on top of it.
//
// Engine.  Problems live in one device pool (sub-allocated first-fit, so
// admission is a question of free bytes).  Each problem has a slot in a
// device array of Prob; the problems that are running form a membership
// epoch, for which the host builds one item list (problem, k) sorted by
// kernel class and storage precision, and one problem list.  An iteration
// is the fixed launch sequence of wann_main over these lists; a block of
// graph_block iterations is captured once per epoch into a HIP graph and
// replayed, with or without the second-step kernels (lean blocks while no
// problem takes parabolic steps).  After each block a status kernel copies
// done/hold flags to pinned host memory; two blocks are kept in flight, so
// the host reads the flags of block j while block j+1 runs.  Finished
// problems are retired (final state gathered, unitarity checked), held
// problems get their FP64 rebuild of M (a new epoch) or, if a lean block
// held them for a parabolic step, that step.  A session is an engine with
// one problem.

struct Member {
  int id = -1;
  int slot = -1;
  int nw = 0, nk = 0, nn = 0;
  int cls = 0;     // kernel class by num_wann
  int rk = 0;      // rotation kernel (RotKernel)
  bool tune_pending = false;  // rk to be measured at the start of the first solve
  int phase = 1;   // 0: mixed bulk, 1: FP64 (host view)
  int stage = 0;   // 0 idle, 1 queued, 2 running, 3 finished
  w90gpu_control ctl{};
  double wbtot = 0.0;
  size_t off = 0, bytes = 0;  // sub-allocation in the pool
  Prob p{};
  State* st_init = nullptr;   // device template of the loop state
  cplx* M0 = nullptr;
  cplx* U0 = nullptr;
  w90gpu_result res{};
  double t_start = 0.0, t_end = 0.0;
};

// kernel classes by num_wann: the wavefront-per-neighbour MFMA rotation up
// to 16; the one-workgroup-per-block MFMA rotation (32 x 32 tile, T in LDS)
// with the per-k LDS retraction and gradient up to 32; the same rotation
// with a 64 x 64 (or, up to 48 bands, 48 x 48) tile, the MFMA retraction
// and the LDS gradient up to 64; the two-launch MFMA rotation (tiles over
// panels) with the global-memory retraction and the tiled gradient up to
// W90GPU_MAX_WANN
constexpr int kClasses = 4;
constexpr int kLdsMaxWann = 32;
// iterations per captured block (W90GPU_GRAPH_BLOCK overrides it, 1-64);
// the solve is polled once per block
constexpr int kGraphBlock = 16;

int wann_class(int nw) { return nw <= 16 ? 0 : nw <= kLdsMaxWann ? 1 : nw <= W90_STEP_N ? 2 : 3; }

// Rotation kernels by tile shape.  Each class above 16 bands has a default
// and candidates of other tile shapes:
//   17-32: fused32 (32 x 32 tile, T in LDS) | fused64 (64 x 64 tile)
//   33-64: fused64 | direct64 (operands straight into MFMA fragments) |
//          fused48 (a 48 x 48 tile over three wavefronts, up to 48 bands)
//   65-256: gemm64 (two launches over 64 x 64 panels) | gemm32 (32 x 32
//           panels: less padding for 65-96 bands, lower reuse per load)
// The choice is made per (num_wann, nntot) by the autotuner (rot_choice,
// tune_member) and fixed for the life of the process, so a problem is
// rotated by the same kernel alone and in a batch.
enum RotKernel { RK_W16 = 0, RK_FUSED32, RK_FUSED64, RK_DIRECT64, RK_GEMM64, RK_GEMM32, RK_FUSED48, kRotKernels };
const char* const kRotName[kRotKernels] = {"w16", "fused32", "fused64", "direct64", "gemm64", "gemm32", "fused48"};
// output tile edge (the initial reduction k_reduce_tiles uses the same tiles)
constexpr int kRotEdge[kRotKernels] = {16, 32, 64, 64, 64, 32, 48};

// Launch groups: (class, rotation kernel) pairs in class order.  Item lists
// are sorted by class, then storage precision, then group, so that one
// (class, precision) is a contiguous range for the gradient kernels and one
// (group, precision) a contiguous range for the rotation.
constexpr int kGroups = 8;
constexpr int kGroupClass[kGroups] = {0, 1, 1, 2, 2, 2, 3, 3};
constexpr int kGroupRot[kGroups] = {RK_W16,      RK_FUSED32, RK_FUSED64, RK_FUSED64,
                                    RK_DIRECT64, RK_FUSED48, RK_GEMM64,  RK_GEMM32};

int group_of(int cls, int rk) {
  for (int g = 0; g < kGroups; ++g)
    if (kGroupClass[g] == cls && kGroupRot[g] == rk) return g;
  return -1;
}

// Whether kernel rk rotates num_wann nw: it serves nw's class, and the one
// tile of fused48 covers the block.
bool rot_serves(int rk, int nw) {
  return rk >= 0 && group_of(wann_class(nw), rk) >= 0 && (rk != RK_FUSED48 || nw <= 48);
}
int default_rot(int cls) { return cls == 0 ? RK_W16 : cls == 1 ? RK_FUSED32 : cls == 2 ? RK_FUSED64 : RK_GEMM64; }
int first_group(int cls) {
  int g = 0;
  while (kGroupClass[g] != cls) ++g;
  return g;
}
int last_group(int cls) {
  int g = kGroups - 1;
  while (kGroupClass[g] != cls) --g;
  return g;
}

// Tiles per block M(k,b) of a rotation kernel: the per-(k,b) partial sums
// of its epilogue.
int rot_panels(int rk, int nw) {
  if (rk == RK_W16) return 1;
  const int t = (nw + kRotEdge[rk] - 1) / kRotEdge[rk];
  return t * t;
}
static_assert(((W90GPU_MAX_WANN + 31) / 32) * ((W90GPU_MAX_WANN + 31) / 32) <= W90_MAX_PANELS,
              "k_kpart stages every tile sum of a block");

// partial-sum slots per block in a problem's arena: the most that any
// candidate of its class writes
int max_panels(int nw) {
  const int c = wann_class(nw);
  int mx = 1;
  for (int g = first_group(c); g <= last_group(c); ++g)
    if (rot_serves(kGroupRot[g], nw)) mx = std::max(mx, rot_panels(kGroupRot[g], nw));
  return mx;
}

// A contiguous range of an item list with the launch shape its kernels need.
struct Seg {
  int off = 0, n = 0, nnmax = 0, npanmax = 0;
  int nwmin = 0, nwmax = 0;  // num_wann range of the items (0, 0 when empty)
};

// Device lists of one launch sequence: items sorted by (class, precision,
// group), problem indices for the per-problem kernels.
struct Lists {
  const Item* items = nullptr;
  const int* plist = nullptr;
  int np = 0;
  Seg seg[kGroups][2];  // [group][0: FP64, 1: FP32]
  bool any_second = false;
  bool any_group = false;  // some problem uses the two-level omega reduction
};

// Lists over n items of one group and precision (the segments before it
// empty at 0, those after it empty at n).
void one_group_lists(Lists& L, int g0, int q0, int n, int nn, int npan, int nw) {
  for (int c = 0; c < kClasses; ++c)
    for (int q = 0; q < 2; ++q)
      for (int g = first_group(c); g <= last_group(c); ++g) {
        const bool after = c > kGroupClass[g0] || (c == kGroupClass[g0] && (q > q0 || (q == q0 && g > g0)));
        L.seg[g][q] = Seg{after ? n : 0, 0, 0, 0};
      }
  L.seg[g0][q0] = Seg{0, n, nn, npan, nw, nw};
}

struct Pool {
  char* base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> bytes
  size_t used = 0;
  bool alloc(size_t bytes, size_t* off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes;
        const size_t o = it->first + bytes;
        free_.erase(it);
        if (rest) free_[o] = rest;
        used += bytes;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    used -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Engine {
  w90gpu_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;   // compute
  hipStream_t copy = nullptr;     // uploads of admitted problems
  Pool pool;
  std::vector<Member*> slots;     // by slot; nullptr = free
  std::map<int, Member*> by_id;
  int next_id = 0;
  // device tables
  Prob* d_probs = nullptr;
  int cap_probs = 0;
  Item* d_items = nullptr;        // run lists
  int* d_plist = nullptr;
  Item* d_aux_items = nullptr;    // start / rebuild / retire lists
  int* d_aux_plist = nullptr;
  int cap_items = 0, cap_plist = 0;
  Status* h_status = nullptr;     // [2][cap_plist] pinned
  Status* hd_status = nullptr;    // device address of h_status: k_status writes it directly
  Final* d_final = nullptr;       // [cap_plist]
  Final* h_final = nullptr;       // pinned
  // run lists of the current epoch
  long epoch = 0, lists_epoch = -1;
  Lists run;
  std::vector<int> run_slots;
  int epoch_blocks = 0;
  // graphs of one block of the epoch: with the second-step kernels (full)
  // and without them (lean, used while no problem takes parabolic steps)
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec_lean = nullptr;
  hipGraph_t graph_lean = nullptr;
  int clean_blocks = 0;           // processed blocks in a row without a parabolic step
  std::vector<std::pair<hipGraph_t, hipGraphExec_t>> retired;
  bool graph_failed = false;
  bool use_graph = true;
  bool use_lean = true;           // lean blocks allowed (W90GPU_NO_LEAN unset)
  int graph_block = kGraphBlock;  // iterations per block
  hipEvent_t blk_ev[2] = {nullptr, nullptr};
  std::vector<int> blk_slots[2];
  hipEvent_t t0 = nullptr, t1 = nullptr;
  std::vector<int> finished_ids;  // retired since the last report
};

struct w90gpu_session {
  Engine* eng = nullptr;
  Member* m = nullptr;
  bool loaded = false;
  w90gpu_result last{};
};

struct w90gpu_batch {
  Engine* eng = nullptr;
};

namespace {

thread_local std::string g_noctx_err;

void set_error_v(w90gpu_ctx* ctx, const char* fmt, va_list ap) {
  char buf[512];
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  if (ctx) ctx->err = buf;
  else g_noctx_err = buf;
}

void set_error(w90gpu_ctx* ctx, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  set_error_v(ctx, fmt, ap);
  va_end(ap);
}

#define HIPCHK(ctx, call)                                                                         \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) {                                                                       \
      set_error((ctx), "HIP error %s at %s:%d (%s)", hipGetErrorString(e_), __FILE__, __LINE__, \
                #call);                                                                           \
      return W90GPU_ERR_FATAL;                                                                    \
    }                                                                                             \
  } while (0)

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Arena of one problem: offsets in a fixed order that member_bind mirrors.
struct Layout {
  size_t off[40];
  int n = 0;
  size_t total = 0;
  void add(size_t bytes) {
    off[n++] = total;
    total += align_up(bytes);
  }
};

size_t problem_layout(int nw, int nk, int nn, int num_iter, Layout& lay) {
  const size_t nw2 = (size_t)nw * nw;
  const size_t mb = nw2 * nn * nk * sizeof(cplx);
  const size_t ub = nw2 * nk * sizeof(cplx);
  const size_t lnb = (size_t)nk * nn * nw * sizeof(double);
  const size_t psz = 8 * (size_t)nw + 8;
  lay.add(mb);   // M[0]
  lay.add(mb);   // M[1] (holds the two FP32 buffers of the mixed bulk phase)
  lay.add(mb);   // M0
  lay.add(ub);   // U[0]
  lay.add(ub);   // U[1]
  lay.add(ub);   // U0
  lay.add(ub);   // G
  lay.add(ub);   // D
  lay.add(ub);   // E
  lay.add(ub);   // W1
  lay.add(ub);   // W2
  lay.add(lnb);  // ln[0]
  lay.add(lnb);  // ln[1]
  lay.add(lnb);  // d2
  lay.add(wann_class(nw) >= 3 ? mb : 256);                  // Tm (two-stage rotation scratch)
  lay.add((size_t)nk * nn * max_panels(nw) * sizeof(double));  // part_od
  lay.add((size_t)nk * nn * max_panels(nw) * sizeof(double));  // part_all
  lay.add((size_t)nk * psz * sizeof(double));               // kpart
  lay.add((size_t)W90_MAX_GROUPS * psz * sizeof(double));   // gpart
  lay.add((size_t)nk * sizeof(double));                     // part_gg
  lay.add((size_t)nk * sizeof(double));                     // part_gd
  lay.add(3 * (size_t)nw * sizeof(double));                 // rave[0]
  lay.add(3 * (size_t)nw * sizeof(double));                 // rave[1]
  lay.add((size_t)nw * sizeof(double));                     // r2ave[0]
  lay.add((size_t)nw * sizeof(double));                     // r2ave[1]
  lay.add((size_t)nk * sizeof(double));                     // unit
  lay.add(4 * (size_t)(num_iter + 2) * sizeof(double));     // log
  lay.add(sizeof(State));                                   // state
  lay.add(sizeof(State));                                   // state template
  lay.add((size_t)nk * nn * sizeof(int));                   // nnl
  lay.add((size_t)nk * nn * 3 * sizeof(double));            // bk
  lay.add((size_t)nn * sizeof(double));                     // wb
  return lay.total;
}

int validate(w90gpu_ctx* ctx, const w90gpu_problem* prob, const w90gpu_control* ctl) {
  if (!prob || !ctl) {
    set_error(ctx, "null problem or control");
    return W90GPU_ERR_INPUT;
  }
  if (prob->num_wann < 1 || prob->num_wann > W90GPU_MAX_WANN) {
    set_error(ctx, "num_wann=%d outside [1, %d]", prob->num_wann, W90GPU_MAX_WANN);
    return W90GPU_ERR_INPUT;
  }
  if (prob->num_kpts < 1 || prob->nntot < 1 || prob->nntot > W90GPU_NUM_NNMAX) {
    set_error(ctx, "invalid num_kpts=%d or nntot=%d", prob->num_kpts, prob->nntot);
    return W90GPU_ERR_INPUT;
  }
  if (!prob->nnlist || !prob->bk || !prob->wb || !(prob->wbtot > 0.0)) {
    set_error(ctx, "missing k-mesh tables (nnlist, bk, wb, wbtot)");
    return W90GPU_ERR_INPUT;
  }
  for (size_t i = 0; i < (size_t)prob->num_kpts * prob->nntot; ++i) {
    if (prob->nnlist[i] < 1 || prob->nnlist[i] > prob->num_kpts) {
      set_error(ctx, "nnlist entry %d out of range [1, %d]", prob->nnlist[i], prob->num_kpts);
      return W90GPU_ERR_INPUT;
    }
  }
  if (ctl->precond || ctl->guiding_centres || ctl->selective_loc || ctl->conv_noise_amp > 0.0) {
    set_error(ctx, "precond, guiding centres, selective localisation and conv_noise are not on this path");
    return W90GPU_ERR_INPUT;
  }
  if (ctl->num_iter < 0) {
    set_error(ctx, "num_iter < 0");
    return W90GPU_ERR_INPUT;
  }
  if (ctl->precision != W90GPU_PREC_FP64 && ctl->precision != W90GPU_PREC_MIXED) {
    set_error(ctx, "unknown precision %d", ctl->precision);
    return W90GPU_ERR_INPUT;
  }
  return W90GPU_OK;
}

// ------------------------------------------------------------ engine setup

int grow_tables(Engine* E, int nprob, int nitems) {
  w90gpu_ctx* ctx = E->ctx;
  if (nprob > E->cap_probs) {
    const int cap = std::max(nprob, 2 * E->cap_probs);
    Prob* np_ = nullptr;
    HIPCHK(ctx, hipMalloc(&np_, sizeof(Prob) * cap));
    if (E->d_probs) {
      HIPCHK(ctx, hipMemcpyAsync(np_, E->d_probs, sizeof(Prob) * E->cap_probs, hipMemcpyDeviceToDevice, E->stream));
      HIPCHK(ctx, hipStreamSynchronize(E->stream));
      HIPCHK(ctx, hipFree(E->d_probs));
    }
    E->d_probs = np_;
    E->cap_probs = cap;
    E->epoch++;  // captured graphs hold the old table address
  }
  if (nprob > E->cap_plist) {
    const int cap = std::max(nprob, 2 * E->cap_plist);
    HIPCHK(ctx, hipStreamSynchronize(E->stream));
    for (void* q : {(void*)E->d_plist, (void*)E->d_aux_plist, (void*)E->d_final})
      if (q) HIPCHK(ctx, hipFree(q));
    if (E->h_status) HIPCHK(ctx, hipHostFree(E->h_status));
    if (E->h_final) HIPCHK(ctx, hipHostFree(E->h_final));
    E->d_plist = E->d_aux_plist = nullptr;
    E->hd_status = nullptr;
    E->d_final = nullptr;
    E->h_status = nullptr;
    E->h_final = nullptr;
    HIPCHK(ctx, hipMalloc(&E->d_plist, sizeof(int) * cap));
    HIPCHK(ctx, hipMalloc(&E->d_aux_plist, sizeof(int) * cap));
    HIPCHK(ctx, hipMalloc(&E->d_final, sizeof(Final) * cap));
    HIPCHK(ctx, hipHostMalloc((void**)&E->h_status, sizeof(Status) * 2 * cap, hipHostMallocDefault));
    HIPCHK(ctx, hipHostGetDevicePointer((void**)&E->hd_status, E->h_status, 0));
    HIPCHK(ctx, hipHostMalloc((void**)&E->h_final, sizeof(Final) * cap, hipHostMallocDefault));
    E->cap_plist = cap;
    E->epoch++;  // run lists must be uploaded again
  }
  if (nitems > E->cap_items) {
    const int cap = std::max(nitems, 2 * E->cap_items);
    HIPCHK(ctx, hipStreamSynchronize(E->stream));
    if (E->d_items) HIPCHK(ctx, hipFree(E->d_items));
    if (E->d_aux_items) HIPCHK(ctx, hipFree(E->d_aux_items));
    E->d_items = E->d_aux_items = nullptr;
    HIPCHK(ctx, hipMalloc(&E->d_items, sizeof(Item) * cap));
    HIPCHK(ctx, hipMalloc(&E->d_aux_items, sizeof(Item) * cap));
    E->cap_items = cap;
    E->epoch++;
  }
  return W90GPU_OK;
}

void drop_graph(Engine* E) {
  if (E->gexec || E->graph) E->retired.emplace_back(E->graph, E->gexec);
  if (E->gexec_lean || E->graph_lean) E->retired.emplace_back(E->graph_lean, E->gexec_lean);
  E->graph = nullptr;
  E->gexec = nullptr;
  E->graph_lean = nullptr;
  E->gexec_lean = nullptr;
  E->epoch_blocks = 0;
}

// Graphs of past epochs are destroyed once the stream has drained.
void reap_graphs(Engine* E) {
  for (auto& g : E->retired) {
    if (g.second) (void)hipGraphExecDestroy(g.second);
    if (g.first) (void)hipGraphDestroy(g.first);
  }
  E->retired.clear();
}

void engine_destroy(Engine* E);

void tune_refresh();

int engine_create(w90gpu_ctx* ctx, size_t pool_bytes, Engine** out) {
  tune_refresh();
  *out = nullptr;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  Engine* E = new (std::nothrow) Engine();
  if (!E) return W90GPU_ERR_ALLOC;
  E->ctx = ctx;
  auto fail = [&](int code) {
    // partial engine: release what was made
    if (E->stream) (void)hipStreamDestroy(E->stream);
    if (E->copy) (void)hipStreamDestroy(E->copy);
    if (E->pool.base) (void)hipFree(E->pool.base);
    delete E;
    return code;
  };
  if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&E->copy, hipStreamNonBlocking) != hipSuccess) {
    set_error(ctx, "hipStreamCreate failed");
    return fail(W90GPU_ERR_FATAL);
  }
  pool_bytes = align_up(std::max<size_t>(pool_bytes, 256));
  hipError_t e = hipMalloc((void**)&E->pool.base, pool_bytes);
  if (e != hipSuccess) {
    E->pool.base = nullptr;
    (void)hipGetLastError();
    set_error(ctx, "hipMalloc of a %zu-byte pool failed: %s", pool_bytes, hipGetErrorString(e));
    return fail(W90GPU_ERR_ALLOC);
  }
  E->pool.size = pool_bytes;
  E->pool.free_[0] = pool_bytes;
  if (hipEventCreate(&E->t0) != hipSuccess || hipEventCreate(&E->t1) != hipSuccess ||
      hipEventCreateWithFlags(&E->blk_ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&E->blk_ev[1], hipEventDisableTiming) != hipSuccess) {
    set_error(ctx, "hipEventCreate failed");
    return fail(W90GPU_ERR_FATAL);
  }
  E->use_graph = std::getenv("W90GPU_NO_GRAPH") == nullptr;
  E->use_lean = std::getenv("W90GPU_NO_LEAN") == nullptr;
  if (const char* gb = std::getenv("W90GPU_GRAPH_BLOCK")) E->graph_block = std::min(64, std::max(1, std::atoi(gb)));
  if (grow_tables(E, 4, 64)) {
    engine_destroy(E);
    return W90GPU_ERR_FATAL;
  }
  *out = E;
  return W90GPU_OK;
}

void engine_destroy(Engine* E) {
  if (!E) return;
  if (E->stream) (void)hipStreamSynchronize(E->stream);
  if (E->copy) (void)hipStreamSynchronize(E->copy);
  drop_graph(E);
  reap_graphs(E);
  for (auto& kv : E->by_id) delete kv.second;
  for (hipEvent_t ev : {E->t0, E->t1, E->blk_ev[0], E->blk_ev[1]})
    if (ev) (void)hipEventDestroy(ev);
  for (void* q : {(void*)E->d_probs, (void*)E->d_items, (void*)E->d_plist, (void*)E->d_aux_items,
                  (void*)E->d_aux_plist, (void*)E->d_final, (void*)E->pool.base})
    if (q) (void)hipFree(q);
  if (E->h_status) (void)hipHostFree(E->h_status);
  if (E->h_final) (void)hipHostFree(E->h_final);
  if (E->stream) (void)hipStreamDestroy(E->stream);
  if (E->copy) (void)hipStreamDestroy(E->copy);
  delete E;
}

// Point the Prob of a member at its arena.
void member_bind(Member* m, char* a, const Layout& lay) {
  int i = 0;
  Prob& p = m->p;
  p.nw = m->nw;
  p.nk = m->nk;
  p.nn = m->nn;
  p.nw2 = m->nw * m->nw;
  p.npan = rot_panels(m->rk, m->nw);
  p.psz = 8 * m->nw + 8;
  // two reduction levels only where one workgroup summing all per-k
  // vectors would be slow (over 128 Ki values)
  // meshes above 64 k-points reduce in two levels (32 k-ranges, then their
  // sums): one workgroup alone reads the nk vectors at the bandwidth of one
  // CU
  p.ngroup = (m->nk <= 64) ? 0 : std::min(m->nk, W90_MAX_GROUPS);
  p.M[0] = (cplx*)(a + lay.off[i++]);
  p.M[1] = (cplx*)(a + lay.off[i++]);
  m->M0 = (cplx*)(a + lay.off[i++]);
  p.M0 = m->M0;
  const size_t mcount = (size_t)p.nw2 * m->nn * m->nk;
  p.Mf[0] = (float2*)p.M[1];
  p.Mf[1] = p.Mf[0] + mcount;
  p.U[0] = (cplx*)(a + lay.off[i++]);
  p.U[1] = (cplx*)(a + lay.off[i++]);
  m->U0 = (cplx*)(a + lay.off[i++]);
  p.U0 = m->U0;
  p.G = (cplx*)(a + lay.off[i++]);
  p.D = (cplx*)(a + lay.off[i++]);
  p.E = (cplx*)(a + lay.off[i++]);
  p.W1 = (cplx*)(a + lay.off[i++]);
  p.W2 = (cplx*)(a + lay.off[i++]);
  p.ET = p.W1;
  p.ln[0] = (double*)(a + lay.off[i++]);
  p.ln[1] = (double*)(a + lay.off[i++]);
  p.d2 = (double*)(a + lay.off[i++]);
  p.Tm = (void*)(a + lay.off[i++]);
  p.part_od = (double*)(a + lay.off[i++]);
  p.part_all = (double*)(a + lay.off[i++]);
  p.kpart = (double*)(a + lay.off[i++]);
  p.gpart = (double*)(a + lay.off[i++]);
  p.part_gg = (double*)(a + lay.off[i++]);
  p.part_gd = (double*)(a + lay.off[i++]);
  p.rave[0] = (double*)(a + lay.off[i++]);
  p.rave[1] = (double*)(a + lay.off[i++]);
  p.r2ave[0] = (double*)(a + lay.off[i++]);
  p.r2ave[1] = (double*)(a + lay.off[i++]);
  p.unit = (double*)(a + lay.off[i++]);
  p.log = (double*)(a + lay.off[i++]);
  p.st = (State*)(a + lay.off[i++]);
  m->st_init = (State*)(a + lay.off[i++]);
  p.st_init = m->st_init;
  p.nnl = (const int*)(a + lay.off[i++]);
  p.bk = (const double*)(a + lay.off[i++]);
  p.wb = (const double*)(a + lay.off[i++]);
}

State state_template(const Member* m, int first_pass, double omega_invariant) {
  State s;
  std::memset(&s, 0, sizeof(s));
  s.num_iter = m->ctl.num_iter;
  s.num_cg_steps = m->ctl.num_cg_steps;
  s.conv_window = m->ctl.conv_window;
  s.lfixstep = m->ctl.lfixstep;
  s.conv_tol = m->ctl.conv_tol;
  s.trial_step = m->ctl.trial_step;
  s.fixed_step = m->ctl.fixed_step;
  s.wbtot = m->wbtot;
  s.first_pass = first_pass;
  s.omega_invariant = omega_invariant;
  s.dir_sign = 1.0;
  s.phase = (m->ctl.precision == W90GPU_PREC_MIXED) ? 0 : 1;
  return s;
}

int upload_state_template(Engine* E, Member* m, int first_pass, double omega_invariant) {
  const State s = state_template(m, first_pass, omega_invariant);
  HIPCHK(E->ctx, hipMemcpyAsync(m->st_init, &s, sizeof(State), hipMemcpyHostToDevice, E->stream));
  HIPCHK(E->ctx, hipStreamSynchronize(E->stream));
  return W90GPU_OK;
}

// ------------------------------------------------------------- autotuner
//
// Rotation tile shape per (num_wann, nntot).  Sources, first match wins:
//   W90GPU_ROT_KERNEL=<name>  forces a kernel wherever it serves the class;
//   choices already made in this process (w90gpu_tune_set, measurements),
//     or read from W90GPU_TUNE_FILE;
//   with W90GPU_TUNE=1, a measurement at the start of the first solve of
//     the key (tune_member), appended to W90GPU_TUNE_FILE if set;
//   otherwise the table measured on MI355X (w90gpu_tune_table.h), and the
//     class default where the table has no row.
// A measured choice is not revised within a process (w90gpu_tune_set may
// replace any choice); a member keeps the kernel it was admitted with.

#include "w90gpu_tune_table.h"

struct TuneCache {
  std::mutex mu;
  bool online = false;
  int forced = -1;
  std::string file;
  std::map<std::pair<int, int>, int> chosen;  // (num_wann, nntot) -> RotKernel
};

TuneCache& tune_cache() {
  static TuneCache c;
  return c;
}

int rot_by_name(const char* name) {
  for (int r = 0; r < kRotKernels; ++r)
    if (std::strcmp(name, kRotName[r]) == 0) return r;
  return -1;
}

// Environment (read at every context and engine creation, so that a
// process may turn the online tuner on and off) and the tuning file (read
// when its path changes; its choices do not replace ones already made).
void tune_refresh() {
  TuneCache& tc = tune_cache();
  std::lock_guard<std::mutex> lock(tc.mu);
  const char* on = std::getenv("W90GPU_TUNE");
  tc.online = on && std::strcmp(on, "0") != 0;
  const char* fk = std::getenv("W90GPU_ROT_KERNEL");
  tc.forced = fk ? rot_by_name(fk) : -1;
  const char* f = std::getenv("W90GPU_TUNE_FILE");
  const std::string path = f ? f : "";
  if (path == tc.file) return;
  tc.file = path;
  if (path.empty()) return;
  // lines "num_wann nntot kernel [name=microseconds ...]"; '#' starts a comment
  if (FILE* fp = std::fopen(path.c_str(), "r")) {
    char line[512];
    while (std::fgets(line, sizeof(line), fp)) {
      int nw = 0, nn = 0;
      char name[32] = {0};
      if (line[0] == '#' || std::sscanf(line, "%d %d %31s", &nw, &nn, name) != 3) continue;
      const int r = rot_by_name(name);
      if (r >= 0 && nw >= 1 && nw <= W90GPU_MAX_WANN && rot_serves(r, nw))
        tc.chosen.emplace(std::make_pair(nw, nn), r);
    }
    std::fclose(fp);
  }
}

// The rotation kernel of a problem.  *pending (when given) is set if the
// key is to be measured at the start of its first solve; until then the
// class default stands.
int rot_choice(int nw, int nn, bool* pending) {
  TuneCache& tc = tune_cache();
  std::lock_guard<std::mutex> lock(tc.mu);
  const int c = wann_class(nw);
  if (pending) *pending = false;
  if (tc.forced >= 0 && rot_serves(tc.forced, nw)) return tc.forced;
  auto it = tc.chosen.find({nw, nn});
  if (it != tc.chosen.end()) return it->second;
  if (tc.online) {
    if (pending && first_group(c) != last_group(c)) *pending = true;
    return default_rot(c);
  }
  for (const TuneRow& r : kTuneTable)
    if (nw >= r.nw_lo && nw <= r.nw_hi && nn >= r.nn_lo && nn <= r.nn_hi && rot_serves(r.rk, nw)) return r.rk;
  return default_rot(c);
}

// Admit a problem: pool bytes, a slot, its k-mesh tables and state template.
// M0/U0 are uploaded separately (member_upload).
int engine_add(Engine* E, const w90gpu_problem* prob, const w90gpu_control* ctl, Member** out) {
  w90gpu_ctx* ctx = E->ctx;
  int rc = validate(ctx, prob, ctl);
  if (rc) return rc;
  Layout lay;
  const size_t bytes = problem_layout(prob->num_wann, prob->num_kpts, prob->nntot, ctl->num_iter, lay);
  size_t off = 0;
  if (!E->pool.alloc(bytes, &off)) {
    set_error(ctx, "no room: problem needs %zu bytes, pool has %zu of %zu free", bytes,
              E->pool.size - E->pool.used, E->pool.size);
    return W90GPU_ERR_ALLOC;
  }
  Member* m = new (std::nothrow) Member();
  if (!m) {
    E->pool.release(off, bytes);
    return W90GPU_ERR_ALLOC;
  }
  m->nw = prob->num_wann;
  m->nk = prob->num_kpts;
  m->nn = prob->nntot;
  m->cls = wann_class(m->nw);
  m->rk = rot_choice(m->nw, m->nn, &m->tune_pending);
  m->ctl = *ctl;
  m->wbtot = prob->wbtot;
  m->off = off;
  m->bytes = bytes;
  m->phase = (ctl->precision == W90GPU_PREC_MIXED) ? 0 : 1;
  member_bind(m, E->pool.base + off, lay);
  int slot = -1;
  for (size_t i = 0; i < E->slots.size(); ++i)
    if (!E->slots[i]) {
      slot = (int)i;
      break;
    }
  if (slot < 0) {
    slot = (int)E->slots.size();
    E->slots.push_back(nullptr);
  }
  auto fail = [&](int code) {
    E->pool.release(off, bytes);
    delete m;
    return code;
  };
  if (grow_tables(E, (int)E->slots.size(), 0)) return fail(W90GPU_ERR_FATAL);
  // k-mesh tables: nnlist (nk,nn) Fortran -> [k][b] 0-based
  std::vector<int> hn((size_t)m->nk * m->nn);
  for (int k = 0; k < m->nk; ++k)
    for (int b = 0; b < m->nn; ++b) hn[(size_t)k * m->nn + b] = prob->nnlist[k + (size_t)m->nk * b] - 1;
  const State s = state_template(m, 1, 0.0);
  if (hipMemcpyAsync((void*)m->p.nnl, hn.data(), hn.size() * sizeof(int), hipMemcpyHostToDevice, E->stream) !=
          hipSuccess ||
      hipMemcpyAsync((void*)m->p.bk, prob->bk, (size_t)m->nk * m->nn * 3 * sizeof(double), hipMemcpyHostToDevice,
                     E->stream) != hipSuccess ||
      hipMemcpyAsync((void*)m->p.wb, prob->wb, (size_t)m->nn * sizeof(double), hipMemcpyHostToDevice,
                     E->stream) != hipSuccess ||
      hipMemcpyAsync(m->st_init, &s, sizeof(State), hipMemcpyHostToDevice, E->stream) != hipSuccess ||
      hipMemcpyAsync(m->p.st, &s, sizeof(State), hipMemcpyHostToDevice, E->stream) != hipSuccess ||
      hipMemcpyAsync(E->d_probs + slot, &m->p, sizeof(Prob), hipMemcpyHostToDevice, E->stream) != hipSuccess ||
      hipStreamSynchronize(E->stream) != hipSuccess) {
    set_error(ctx, "upload of k-mesh tables failed");
    return fail(W90GPU_ERR_FATAL);
  }
  m->slot = slot;
  m->id = E->next_id++;
  E->slots[slot] = m;
  E->by_id[m->id] = m;
  *out = m;
  return W90GPU_OK;
}

// Upload M0/U0 of a member from the host on the copy stream; the call
// returns once the caller's arrays have been read.
int member_upload(Engine* E, Member* m, const double* m_matrix, const double* u_matrix) {
  const size_t mb = (size_t)m->nw * m->nw * m->nn * m->nk * sizeof(cplx);
  const size_t ub = (size_t)m->nw * m->nw * m->nk * sizeof(cplx);
  HIPCHK(E->ctx, hipMemcpyAsync(m->M0, m_matrix, mb, hipMemcpyHostToDevice, E->copy));
  HIPCHK(E->ctx, hipMemcpyAsync(m->U0, u_matrix, ub, hipMemcpyHostToDevice, E->copy));
  HIPCHK(E->ctx, hipStreamSynchronize(E->copy));
  return W90GPU_OK;
}

void engine_remove(Engine* E, Member* m) {
  E->pool.release(m->off, m->bytes);
  E->slots[m->slot] = nullptr;
  E->by_id.erase(m->id);
  if (m->stage == 2) E->epoch++;
  delete m;
}

// ------------------------------------------------------------- lists

// Build lists for a set of members: items sorted by (class, precision,
// group).  prec_of(m) gives 0 (FP64) or 1 (FP32) for each member.
template <class F>
void build_lists(const std::vector<Member*>& ms, F prec_of, std::vector<Item>& items, std::vector<int>& plist,
                 Lists& L) {
  items.clear();
  plist.clear();
  L = Lists{};
  for (int c = 0; c < kClasses; ++c)
    for (int q = 0; q < 2; ++q)
      for (int g = first_group(c); g <= last_group(c); ++g) {
        Seg& s = L.seg[g][q];
        s.off = (int)items.size();
        for (Member* m : ms) {
          if (m->cls != c || m->rk != kGroupRot[g] || prec_of(m) != q) continue;
          for (int k = 0; k < m->nk; ++k) items.push_back(Item{m->slot, k});
          s.nnmax = std::max(s.nnmax, m->nn);
          s.npanmax = std::max(s.npanmax, rot_panels(m->rk, m->nw));
          s.nwmin = s.nwmin ? std::min(s.nwmin, m->nw) : m->nw;
          s.nwmax = std::max(s.nwmax, m->nw);
        }
        s.n = (int)items.size() - s.off;
      }
  for (Member* m : ms) {
    plist.push_back(m->slot);
    if (!m->ctl.lfixstep) L.any_second = true;
    if (m->p.ngroup > 0) L.any_group = true;
  }
  L.np = (int)plist.size();
}

int upload_lists(Engine* E, const std::vector<Item>& items, const std::vector<int>& plist, Item* d_items,
                 int* d_plist, Lists& L) {
  if (!items.empty())
    HIPCHK(E->ctx, hipMemcpyAsync(d_items, items.data(), sizeof(Item) * items.size(), hipMemcpyHostToDevice,
                                  E->stream));
  if (!plist.empty())
    HIPCHK(E->ctx, hipMemcpyAsync(d_plist, plist.data(), sizeof(int) * plist.size(), hipMemcpyHostToDevice,
                                  E->stream));
  L.items = d_items;
  L.plist = d_plist;
  return W90GPU_OK;
}

// Aux lists (start, rebuild, retire, host-launched second steps) for a set
// of members, uploaded in stream order.  Every member in FP64 storage, or
// with by_phase in the storage of its current phase (FP32 in the mixed
// bulk phase), as in the run lists.
int aux_lists(Engine* E, const std::vector<Member*>& ms, Lists& L, bool by_phase = false) {
  int nitems = 0;
  for (Member* m : ms) nitems += m->nk;
  int rc = grow_tables(E, (int)E->slots.size(), nitems);
  if (rc) return rc;
  std::vector<Item> items;
  std::vector<int> plist;
  if (by_phase)
    build_lists(ms, [](const Member* m) { return m->phase == 0 ? 1 : 0; }, items, plist, L);
  else
    build_lists(ms, [](const Member*) { return 0; }, items, plist, L);
  // pageable sources are staged by the copy call itself; keep them alive
  // until the stream reaches the copy
  rc = upload_lists(E, items, plist, E->d_aux_items, E->d_aux_plist, L);
  if (rc) return rc;
  HIPCHK(E->ctx, hipStreamSynchronize(E->stream));
  return W90GPU_OK;
}

// ---------------------------------------------------------- launches

// A range of whole classes [c0, c1] of the item list (both precisions).
Seg class_range(const Lists& L, int c0, int c1) {
  Seg r;
  r.off = L.seg[first_group(c0)][0].off;
  const Seg& last = L.seg[last_group(c1)][1];
  r.n = last.off + last.n - r.off;
  for (int g = first_group(c0); g <= last_group(c1); ++g)
    for (int q = 0; q < 2; ++q) {
      const Seg& x = L.seg[g][q];
      if (x.n == 0) continue;
      r.nwmin = r.nwmin ? std::min(r.nwmin, x.nwmin) : x.nwmin;
      r.nwmax = std::max(r.nwmax, x.nwmax);
    }
  return r;
}

// The items of one class in one storage precision (every group), with the
// widest launch shape among them.
Seg class_prec_range(const Lists& L, int c, int q) {
  Seg r;
  r.off = L.seg[first_group(c)][q].off;
  const Seg& last = L.seg[last_group(c)][q];
  r.n = last.off + last.n - r.off;
  for (int g = first_group(c); g <= last_group(c); ++g) {
    r.nnmax = std::max(r.nnmax, L.seg[g][q].nnmax);
    r.npanmax = std::max(r.npanmax, L.seg[g][q].npanmax);
  }
  return r;
}

// One rotation kernel over a segment of items in storage precision T2; for
// MODE_INIT the tile reduction of the current M with the kernel's tiles.
template <class T2>
void launch_rot_kernel(hipStream_t s, const Prob* P, const Item* it, const Seg& g, int rk, int mode) {
  // MFMA GEMM tiles; gridDim.y covers b and the tiles of a block
  const dim3 g2(g.n, g.nnmax * g.npanmax);
  // one tile per block: both products in one launch, XCD-aware
  // one-dimensional grid (see k_rot_fused)
  const dim3 gb((unsigned)(8 * ((g.n + 7) / 8) * g.nnmax));
  if (rk == RK_W16) {
    hipLaunchKernelGGL(k_rotate_w16<T2>, dim3(g.n), dim3(64 * g.nnmax), rotate_w16_lds<T2>(g.nnmax), s, P, it, g.n,
                       mode);
    return;
  }
  if (mode == MODE_INIT) {
    // fused48 writes one tile per block, as the 64 tiles do up to 64 bands
    if (kRotEdge[rk] == 32) hipLaunchKernelGGL((k_reduce_tiles<T2, 32>), g2, dim3(256), 0, s, P, it, g.n, mode);
    else hipLaunchKernelGGL((k_reduce_tiles<T2, 64>), g2, dim3(256), 0, s, P, it, g.n, mode);
    return;
  }
  switch (rk) {
    case RK_FUSED32:
      hipLaunchKernelGGL((k_rot_fused<T2, 32>), gb, dim3(256), 0, s, P, it, g.n, g.nnmax, mode);
      break;
    case RK_FUSED64:
      hipLaunchKernelGGL((k_rot_fused<T2, 64>), gb, dim3(256), 0, s, P, it, g.n, g.nnmax, mode);
      break;
    case RK_FUSED48:
      hipLaunchKernelGGL((k_rot_fused<T2, 48>), gb, dim3(FusedWaves<48>::NT), 0, s, P, it, g.n, g.nnmax, mode);
      break;
    case RK_DIRECT64:
      hipLaunchKernelGGL(k_rot_direct<T2>, gb, dim3(256), 0, s, P, it, g.n, g.nnmax, mode);
      break;
    case RK_GEMM64:
      hipLaunchKernelGGL((k_rot_gemm<T2, 1, 64>), g2, dim3(256), 0, s, P, it, g.n, mode);
      hipLaunchKernelGGL((k_rot_gemm<T2, 2, 64>), g2, dim3(256), 0, s, P, it, g.n, mode);
      break;
    default:  // RK_GEMM32
      hipLaunchKernelGGL((k_rot_gemm<T2, 1, 32>), g2, dim3(256), 0, s, P, it, g.n, mode);
      hipLaunchKernelGGL((k_rot_gemm<T2, 2, 32>), g2, dim3(256), 0, s, P, it, g.n, mode);
      break;
  }
}

// The rotation of M and the reduction of its diagonal into the per-k
// partial sums of wann_omega, over every group and precision of the lists.
int launch_rotate(Engine* E, const Lists& L, int mode) {
  hipStream_t s = E->stream;
  const Prob* P = E->d_probs;
  for (int g = 0; g < kGroups; ++g)
    for (int q = 0; q < 2; ++q) {
      const Seg& sg = L.seg[g][q];
      if (sg.n == 0) continue;
      const Item* it = L.items + sg.off;
      if (q == 0) launch_rot_kernel<cplx>(s, P, it, sg, kGroupRot[g], mode);
      else launch_rot_kernel<float2>(s, P, it, sg, kGroupRot[g], mode);
      HIPCHK(E->ctx, hipGetLastError());
    }
  // per-k partial vectors of the MFMA classes
  const Seg gk = class_range(L, 1, kClasses - 1);
  if (gk.n) {
    hipLaunchKernelGGL(k_kpart, dim3(gk.n), dim3(W90_NT), 0, s, P, L.items + gk.off, gk.n, mode);
    HIPCHK(E->ctx, hipGetLastError());
  }
  return W90GPU_OK;
}

// The reduction of the per-k partial vectors over k, with the control
// logic that follows each wann_omega.
int launch_reduce(Engine* E, const Lists& L, int mode, int lean = 0) {
  if (L.any_group) {
    hipLaunchKernelGGL(k_omega_group, dim3(L.np * W90_MAX_GROUPS), dim3(W90_NT), 0, E->stream, E->d_probs,
                       L.plist, L.np, mode);
    HIPCHK(E->ctx, hipGetLastError());
  }
  hipLaunchKernelGGL(k_omega_final, dim3(L.np), dim3(W90_FINAL_NT), 0, E->stream, E->d_probs, L.plist, L.np, mode,
                     lean);
  HIPCHK(E->ctx, hipGetLastError());
  return W90GPU_OK;
}

int launch_omega(Engine* E, const Lists& L, int mode, int lean = 0) {
  int rc = launch_rotate(E, L, mode);
  if (rc) return rc;
  return launch_reduce(E, L, mode, lean);
}

int launch_step(Engine* E, const Lists& L, int mode) {
  hipStream_t s = E->stream;
  const Seg a = class_range(L, 0, 0), b = class_range(L, 1, 1), c = class_range(L, 2, 2),
            d = class_range(L, 3, 3);
  if (a.n)
    hipLaunchKernelGGL(k_step_k<16>, dim3(a.n), dim3(W90_NT), 0, s, E->d_probs, L.items + a.off, a.n, mode);
  if (b.n)
    hipLaunchKernelGGL(k_step_k<kLdsMaxWann>, dim3(b.n), dim3(W90_NT), 0, s, E->d_probs, L.items + b.off, b.n,
                       mode);
  // num_wann up to 48 on the 48 layout, above on the 64 one (each launch
  // covers the class; workgroups of the other's items return at once)
  if (c.n && c.nwmin <= 48)
    hipLaunchKernelGGL(k_step_mfma<48>, dim3(c.n), dim3(StepWaves<48>::NT), step_mfma_lds<48>(), s, E->d_probs,
                       L.items + c.off, c.n, mode);
  if (c.n && c.nwmax > 48)
    hipLaunchKernelGGL(k_step_mfma<64>, dim3(c.n), dim3(StepWaves<64>::NT), step_mfma_lds<64>(), s, E->d_probs,
                       L.items + c.off, c.n, mode);
  if (d.n)
    hipLaunchKernelGGL(k_step, dim3(d.n), dim3(W90_STEP_NT), 0, s, E->d_probs, L.items + d.off, d.n, mode);
  HIPCHK(E->ctx, hipGetLastError());
  return W90GPU_OK;
}

int launch_domega(Engine* E, const Lists& L, int mode) {
  hipStream_t s = E->stream;
  const Prob* P = E->d_probs;
  for (int c = 0; c < kClasses; ++c)
    for (int q = 0; q < 2; ++q) {
      const Seg g = class_prec_range(L, c, q);
      if (g.n == 0) continue;
      const Item* it = L.items + g.off;
      if (c == 0) {
        if (q == 0)
          hipLaunchKernelGGL((k_domega_k<16, cplx>), dim3(g.n), dim3(domega_nt<16>()), (domega_lds<16, cplx>()), s, P, it,
                             g.n, mode);
        else
          hipLaunchKernelGGL((k_domega_k<16, float2>), dim3(g.n), dim3(domega_nt<16>()), (domega_lds<16, float2>()), s, P,
                             it, g.n, mode);
      } else if (c == 1) {
        if (q == 0)
          hipLaunchKernelGGL((k_domega_k<kLdsMaxWann, cplx>), dim3(g.n), dim3(domega_nt<kLdsMaxWann>()),
                             (domega_lds<kLdsMaxWann, cplx>()), s, P, it,
                             g.n, mode);
        else
          hipLaunchKernelGGL((k_domega_k<kLdsMaxWann, float2>), dim3(g.n), dim3(domega_nt<kLdsMaxWann>()),
                             (domega_lds<kLdsMaxWann, float2>()), s, P, it,
                             g.n, mode);
      } else if (c == 2) {
        if (q == 0)
          hipLaunchKernelGGL((k_domega_w64<cplx>), dim3(g.n), dim3(domega_nt<W90_STEP_N>()),
                             (domega_lds<W90_STEP_N, cplx>()), s, P, it, g.n, mode);
        else
          hipLaunchKernelGGL((k_domega_w64<float2>), dim3(g.n), dim3(domega_nt<W90_STEP_N>()),
                             (domega_lds<W90_STEP_N, float2>()), s, P, it, g.n, mode);
      } else {
        if (q == 0)
          hipLaunchKernelGGL(k_domega<cplx>, dim3(g.n), dim3(1024), domega_t_lds<cplx>(), s, P, it, g.n, mode);
        else
          hipLaunchKernelGGL(k_domega<float2>, dim3(g.n), dim3(1024), domega_t_lds<float2>(), s, P, it, g.n, mode);
      }
      HIPCHK(E->ctx, hipGetLastError());
    }
  return W90GPU_OK;
}

int all_items(const Lists& L) {
  const Seg r = class_range(L, 0, kClasses - 1);
  return r.n;
}

// One iteration of wann_main over the lists (wannierise.F90:555-817).
// lean: without the second-step kernels; a problem whose line search asks
// for the parabolic step then holds until the host launches it
// (enqueue_seconds).
int enqueue_iteration(Engine* E, const Lists& L, bool lean = false) {
  int rc;
  if ((rc = launch_domega(E, L, MODE_TRIAL))) return rc;
  // the search direction is formed inside the trial-step kernels
  if ((rc = launch_step(E, L, MODE_TRIAL))) return rc;
  if ((rc = launch_omega(E, L, MODE_TRIAL, lean ? 1 : 0))) return rc;
  if (L.any_second && !lean) {
    if ((rc = launch_step(E, L, MODE_SECOND))) return rc;
    if ((rc = launch_omega(E, L, MODE_SECOND))) return rc;
  }
  return W90GPU_OK;
}

// Measure the rotation candidates of a member's class on its own data and
// fix the choice for its (num_wann, nntot).  Each candidate runs in
// MODE_BENCH (a trial rotation of the current buffer that writes only the
// other buffer, the per-(k,b) outputs and the partial sums, all rebuilt by
// the initial evaluation that follows) over the member's k items, repeated
// so that one launch holds at least kTuneItems items, as in a batch; five
// timed launches after one untimed, twice in alternating order, the
// smaller time kept.  A candidate replaces the class default only when it
// is at least 3% faster.  The member's state must be valid (cur 0 or 1):
// engine_add uploads the template, so it is from admission on.
constexpr int kTuneItems = 2048;

int tune_member(Engine* E, Member* m) {
  const int c = m->cls, nw = m->nw, nn = m->nn;
  const int reps = std::max(1, (kTuneItems + m->nk - 1) / m->nk);
  const int n = reps * m->nk;
  int rc = grow_tables(E, (int)E->slots.size(), n);
  if (rc) return rc;
  std::vector<Item> items((size_t)n);
  for (int r = 0; r < reps; ++r)
    for (int k = 0; k < m->nk; ++k) items[(size_t)r * m->nk + k] = Item{m->slot, k};
  HIPCHK(E->ctx, hipMemcpyAsync(E->d_aux_items, items.data(), sizeof(Item) * n, hipMemcpyHostToDevice, E->stream));
  std::vector<int> cand;
  for (int g = first_group(c); g <= last_group(c); ++g)
    if (rot_serves(kGroupRot[g], nw)) cand.push_back(kGroupRot[g]);
  double t[kRotKernels];
  for (double& x : t) x = 1e30;
  auto time_one = [&](int rk) -> int {
    m->p.npan = rot_panels(rk, nw);
    HIPCHK(E->ctx, hipMemcpyAsync(E->d_probs + m->slot, &m->p, sizeof(Prob), hipMemcpyHostToDevice, E->stream));
    Lists L;
    one_group_lists(L, group_of(c, rk), 0, n, nn, m->p.npan, nw);
    L.items = E->d_aux_items;
    int r = launch_rotate(E, L, MODE_BENCH);
    if (r) return r;
    HIPCHK(E->ctx, hipEventRecord(E->t0, E->stream));
    for (int i = 0; i < 5 && !r; ++i) r = launch_rotate(E, L, MODE_BENCH);
    if (r) return r;
    HIPCHK(E->ctx, hipEventRecord(E->t1, E->stream));
    HIPCHK(E->ctx, hipEventSynchronize(E->t1));
    float ms = 0.f;
    HIPCHK(E->ctx, hipEventElapsedTime(&ms, E->t0, E->t1));
    t[rk] = std::min(t[rk], 1e3 * ms / 5.0);
    return W90GPU_OK;
  };
  for (int pass = 0; pass < 2; ++pass)
    for (size_t i = 0; i < cand.size(); ++i)
      if ((rc = time_one(cand[pass ? cand.size() - 1 - i : i]))) return rc;
  const int d = default_rot(c);
  int pick = d;
  for (int rk : cand)
    if (t[rk] < t[pick]) pick = rk;
  if (pick != d && !(t[pick] <= 0.97 * t[d])) pick = d;
  m->rk = pick;
  m->p.npan = rot_panels(pick, nw);
  HIPCHK(E->ctx, hipMemcpyAsync(E->d_probs + m->slot, &m->p, sizeof(Prob), hipMemcpyHostToDevice, E->stream));
  HIPCHK(E->ctx, hipStreamSynchronize(E->stream));
  TuneCache& tc = tune_cache();
  std::lock_guard<std::mutex> lock(tc.mu);
  if (!tc.chosen.emplace(std::make_pair(nw, nn), pick).second) return W90GPU_OK;
  if (!tc.file.empty()) {
    if (FILE* fp = std::fopen(tc.file.c_str(), "a")) {
      std::fprintf(fp, "%d %d %s", nw, nn, kRotName[pick]);
      for (int rk : cand) std::fprintf(fp, " %s=%.2f", kRotName[rk], t[rk]);
      std::fprintf(fp, " items=%d\n", n);
      std::fclose(fp);
    }
  }
  return W90GPU_OK;
}

// Members whose key is still to be measured: the first member of each key
// is measured, the others take its choice.
int tune_pending(Engine* E, const std::vector<Member*>& ms) {
  for (Member* m : ms) {
    if (!m->tune_pending) continue;
    m->tune_pending = false;
    bool again = false;
    const int rk = rot_choice(m->nw, m->nn, &again);
    if (!again) {
      if (rk != m->rk) {
        m->rk = rk;
        m->p.npan = rot_panels(rk, m->nw);
        HIPCHK(E->ctx,
               hipMemcpyAsync(E->d_probs + m->slot, &m->p, sizeof(Prob), hipMemcpyHostToDevice, E->stream));
      }
      continue;
    }
    const int rc = tune_member(E, m);
    if (rc) return rc;
  }
  return W90GPU_OK;
}

// Start the queued members: state from the template, M and U from M0/U0,
// the initial wann_omega in FP64, and the FP32 copy of M for the mixed
// bulk phase.
int enqueue_start(Engine* E, const std::vector<Member*>& ms) {
  int rc = tune_pending(E, ms);
  if (rc) return rc;
  Lists L;
  rc = aux_lists(E, ms, L);
  if (rc) return rc;
  hipStream_t s = E->stream;
  hipLaunchKernelGGL(k_reset, dim3((L.np + 255) / 256), dim3(256), 0, s, E->d_probs, L.plist, L.np);
  HIPCHK(E->ctx, hipGetLastError());
  const int n = all_items(L);
  hipLaunchKernelGGL(k_load_initial, dim3(n), dim3(W90_NT), 0, s, E->d_probs, L.items, n);
  HIPCHK(E->ctx, hipGetLastError());
  if ((rc = launch_omega(E, L, MODE_INIT))) return rc;
  bool any_mixed = false;
  for (Member* m : ms) any_mixed |= (m->ctl.precision == W90GPU_PREC_MIXED);
  if (any_mixed) {
    hipLaunchKernelGGL(k_to_f32, dim3(n), dim3(W90_NT), 0, s, E->d_probs, L.items, n);
    HIPCHK(E->ctx, hipGetLastError());
  }
  const double t = now_s();
  for (Member* m : ms) {
    m->stage = 2;
    m->phase = (m->ctl.precision == W90GPU_PREC_MIXED) ? 0 : 1;
    m->t_start = t;
  }
  E->epoch++;
  return W90GPU_OK;
}

// FP64 rebuild of M for members whose mixed bulk phase has ended:
// E(k) = U0(k)^+ U(k), M(k,b) = E(k)^+ M0(k,b) E(k+b), and the wann_omega
// of it.
int enqueue_remix(Engine* E, const std::vector<Member*>& ms) {
  Lists L;
  int rc = aux_lists(E, ms, L);
  if (rc) return rc;
  const int n = all_items(L);
  hipLaunchKernelGGL(k_delta_u, dim3(n), dim3(W90_NT), 0, E->stream, E->d_probs, L.items, n);
  HIPCHK(E->ctx, hipGetLastError());
  if ((rc = launch_omega(E, L, MODE_REMIX))) return rc;
  for (Member* m : ms) m->phase = 1;
  E->epoch++;
  return W90GPU_OK;
}

// Retire finished members: unitarity check of U and the final state.
int retire(Engine* E, const std::vector<Member*>& ms) {
  Lists L;
  int rc = aux_lists(E, ms, L);
  if (rc) return rc;
  const int n = all_items(L);
  hipLaunchKernelGGL(k_unitarity, dim3(n), dim3(W90_NT), 0, E->stream, E->d_probs, L.items, n);
  HIPCHK(E->ctx, hipGetLastError());
  hipLaunchKernelGGL(k_gather_final, dim3(L.np), dim3(W90_NT), 0, E->stream, E->d_probs, L.plist, L.np,
                     E->d_final);
  HIPCHK(E->ctx, hipGetLastError());
  HIPCHK(E->ctx, hipMemcpyAsync(E->h_final, E->d_final, sizeof(Final) * L.np, hipMemcpyDeviceToHost, E->stream));
  HIPCHK(E->ctx, hipStreamSynchronize(E->stream));
  const double t = now_s();
  for (size_t i = 0; i < ms.size(); ++i) {
    Member* m = ms[i];
    const State& st = E->h_final[i].st;
    w90gpu_result r{};
    r.omega_total = st.wann.om_tot;
    r.omega_invariant = st.wann.om_i;
    r.omega_d = st.wann.om_d;
    r.omega_od = st.wann.om_od;
    r.omega_tilde = st.wann.om_d + st.wann.om_od;
    r.gcnorm1 = st.gcnorm1;
    r.delta_omega = st.wann.om_tot - st.old.om_tot;
    r.iterations = st.iter;
    r.converged = st.converged;
    r.bulk_iterations = (m->ctl.precision == W90GPU_PREC_MIXED) ? st.bulk_iter : 0;
    r.held_second_steps = st.nheld;
    const double mx = E->h_final[i].unit_max;
    r.unitarity_ok = (mx <= 1e-5) ? 1 : 0;
    r.status = r.unitarity_ok ? W90GPU_OK : W90GPU_ERR_FATAL;
    if (!r.unitarity_ok)
      set_error(E->ctx, "wann_check_unitarity: final U deviates from unitarity by %.3e", mx);
    m->t_end = t;
    r.solve_seconds = m->t_end - m->t_start;
    m->res = r;
    m->stage = 3;
    E->finished_ids.push_back(m->id);
  }
  E->epoch++;
  return W90GPU_OK;
}

int build_run_lists(Engine* E) {
  std::vector<Member*> ms;
  int nitems = 0;
  for (Member* m : E->slots)
    if (m && m->stage == 2) {
      ms.push_back(m);
      nitems += m->nk;
    }
  int rc = grow_tables(E, (int)E->slots.size(), nitems);
  if (rc) return rc;
  static thread_local std::vector<Item> items;
  static thread_local std::vector<int> plist;
  build_lists(ms, [](const Member* m) { return m->phase == 0 ? 1 : 0; }, items, plist, E->run);
  rc = upload_lists(E, items, plist, E->d_items, E->d_plist, E->run);
  if (rc) return rc;
  HIPCHK(E->ctx, hipStreamSynchronize(E->stream));
  E->run_slots = plist;
  drop_graph(E);
  E->lists_epoch = E->epoch;
  return W90GPU_OK;
}

// Capture graph_block iterations over the run lists.  The kernels read the
// loop state from device memory, so one graph serves every block of the
// epoch.  On a capture failure the engine keeps launching directly.
int capture_block(Engine* E, bool lean) {
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    E->graph_failed = true;
    return W90GPU_OK;
  }
  int rc = W90GPU_OK;
  for (int i = 0; i < E->graph_block && rc == W90GPU_OK; ++i) rc = enqueue_iteration(E, E->run, lean);
  const hipError_t e = hipStreamEndCapture(E->stream, &g);
  if (rc != W90GPU_OK || e != hipSuccess || !g) {
    (void)hipGetLastError();
    if (g) (void)hipGraphDestroy(g);
    E->graph_failed = true;
    return W90GPU_OK;
  }
  hipGraphExec_t ge = nullptr;
  if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipGraphDestroy(g);
    E->graph_failed = true;
    return W90GPU_OK;
  }
  if (lean) {
    E->graph_lean = g;
    E->gexec_lean = ge;
  } else {
    E->graph = g;
    E->gexec = ge;
  }
  return W90GPU_OK;
}

int enqueue_block(Engine* E, int slot) {
  int rc;
  // the first block of an epoch is launched directly; the graph is
  // captured when the membership has stayed the same for a block
  const bool graph = E->use_graph && !E->graph_failed && E->epoch_blocks >= 1;
  // without second-step kernels after a processed block in which no problem
  // took a parabolic step (a problem that asks for one then waits for the
  // host, which launches it and returns to full blocks)
  const bool lean = E->use_lean && E->run.any_second && E->clean_blocks >= 1;
  hipGraphExec_t& ge = lean ? E->gexec_lean : E->gexec;
  if (graph && !ge) {
    if ((rc = capture_block(E, lean))) return rc;
  }
  if (graph && ge) {
    HIPCHK(E->ctx, hipGraphLaunch(ge, E->stream));
  } else {
    for (int i = 0; i < E->graph_block; ++i)
      if ((rc = enqueue_iteration(E, E->run, lean))) return rc;
  }
  E->epoch_blocks++;
  const int np = E->run.np;
  // the status rows go straight to pinned host memory; the event after the
  // kernel makes them visible to the host
  hipLaunchKernelGGL(k_status, dim3((np + 255) / 256), dim3(256), 0, E->stream, E->d_probs, E->run.plist, np,
                     E->hd_status + (size_t)slot * E->cap_plist);
  HIPCHK(E->ctx, hipGetLastError());
  HIPCHK(E->ctx, hipEventRecord(E->blk_ev[slot], E->stream));
  E->blk_slots[slot] = E->run_slots;
  return W90GPU_OK;
}

// Read the status of a completed block: finished members are retired,
// members whose bulk phase ended get their FP64 rebuild queued.
// The parabolic steps found in blocks without second-step kernels: release
// the held problems and launch the step, its rotation and wann_omega for
// them (stream order puts this after the blocks already in flight, in which
// these problems only wait).
int enqueue_seconds(Engine* E, const std::vector<Member*>& ms) {
  Lists L;
  int rc = aux_lists(E, ms, L, true);
  if (rc) return rc;
  hipLaunchKernelGGL(k_release, dim3((L.np + 255) / 256), dim3(256), 0, E->stream, E->d_probs, L.plist, L.np);
  HIPCHK(E->ctx, hipGetLastError());
  if ((rc = launch_step(E, L, MODE_SECOND))) return rc;
  return launch_omega(E, L, MODE_SECOND);
}

int process_block(Engine* E, int slot, std::vector<Member*>& remix) {
  HIPCHK(E->ctx, hipEventSynchronize(E->blk_ev[slot]));
  const Status* st = E->h_status + (size_t)slot * E->cap_plist;

  std::vector<Member*> done, seconds;
  bool any_second = false;
  const std::vector<int>& sl = E->blk_slots[slot];
  for (size_t i = 0; i < sl.size(); ++i) {
    Member* m = E->slots[sl[i]];
    if (!m || m->stage != 2) continue;
    if (st[i].nsecond > 0) any_second = true;
    if (st[i].done) {
      done.push_back(m);
    } else if (st[i].second) {
      seconds.push_back(m);
    } else if (st[i].hold && m->phase == 0 &&
               std::find(remix.begin(), remix.end(), m) == remix.end()) {
      remix.push_back(m);
    }
  }
  E->clean_blocks = any_second ? 0 : E->clean_blocks + 1;
  int rc;
  if (!seconds.empty() && (rc = enqueue_seconds(E, seconds))) return rc;
  if (!done.empty()) return retire(E, done);
  return W90GPU_OK;
}

int running_count(const Engine* E) {
  int n = 0;
  for (const Member* m : E->slots)
    if (m && m->stage == 2) ++n;
  return n;
}

// Start queued members, then run blocks over the running ones until
// max_blocks blocks were issued (max_blocks < 0: until none is running).
// Returns with the stream drained.
int engine_advance(Engine* E, int max_blocks) {
  int rc;
  int issued = 0;
  int inflight = 0, head = 0;  // blocks in flight: slots head, head^1
  std::vector<Member*> remix;
  while (true) {
    std::vector<Member*> queued;
    for (Member* m : E->slots)
      if (m && m->stage == 1) queued.push_back(m);
    if (!queued.empty() && (rc = enqueue_start(E, queued))) return rc;
    if (!remix.empty()) {
      if ((rc = enqueue_remix(E, remix))) return rc;
      remix.clear();
    }
    if (running_count(E) == 0) break;
    if (max_blocks >= 0 && issued >= max_blocks) break;
    if (E->lists_epoch != E->epoch) {
      // a new epoch: let the blocks in flight finish with the old lists
      while (inflight > 0) {
        if ((rc = process_block(E, head, remix))) return rc;
        head ^= 1;
        --inflight;
      }
      if (!remix.empty()) continue;
      if (running_count(E) == 0) break;
      if ((rc = build_run_lists(E))) return rc;
    }
    const int slot = (head + inflight) & 1;
    if ((rc = enqueue_block(E, slot))) return rc;
    ++inflight;
    ++issued;
    if (inflight == 2) {
      if ((rc = process_block(E, head, remix))) return rc;
      head ^= 1;
      --inflight;
    }
  }
  while (inflight > 0) {
    if ((rc = process_block(E, head, remix))) return rc;
    head ^= 1;
    --inflight;
  }
  if (!remix.empty() && (rc = enqueue_remix(E, remix))) return rc;
  HIPCHK(E->ctx, hipStreamSynchronize(E->stream));
  reap_graphs(E);
  return W90GPU_OK;
}

int read_member_state(Engine* E, Member* m, State* out) {
  HIPCHK(E->ctx, hipMemcpyAsync(out, m->p.st, sizeof(State), hipMemcpyDeviceToHost, E->stream));
  HIPCHK(E->ctx, hipStreamSynchronize(E->stream));
  return W90GPU_OK;
}

int member_fetch(Engine* E, Member* m, double* m_matrix, double* u_matrix, double* centres, double* spreads) {
  State st;
  int rc = read_member_state(E, m, &st);
  if (rc) return rc;
  const int cur = st.cur;
  const size_t mb = (size_t)m->nw * m->nw * m->nn * m->nk * sizeof(cplx);
  const size_t ub = (size_t)m->nw * m->nw * m->nk * sizeof(cplx);
  if (m_matrix) {
    if (st.phase == 0 && st.iter > 0) {
      set_error(E->ctx, "M of a mixed solve is available once its FP64 phase has started");
      return W90GPU_ERR_INPUT;
    }
    HIPCHK(E->ctx, hipMemcpyAsync(m_matrix, m->p.M[cur], mb, hipMemcpyDeviceToHost, E->stream));
  }
  if (u_matrix) HIPCHK(E->ctx, hipMemcpyAsync(u_matrix, m->p.U[cur], ub, hipMemcpyDeviceToHost, E->stream));
  std::vector<double> rave(3 * m->nw), r2(m->nw);
  HIPCHK(E->ctx, hipMemcpyAsync(rave.data(), m->p.rave[cur], 3 * m->nw * sizeof(double), hipMemcpyDeviceToHost,
                                E->stream));
  HIPCHK(E->ctx, hipMemcpyAsync(r2.data(), m->p.r2ave[cur], m->nw * sizeof(double), hipMemcpyDeviceToHost,
                                E->stream));
  HIPCHK(E->ctx, hipStreamSynchronize(E->stream));
  if (centres) std::memcpy(centres, rave.data(), 3 * m->nw * sizeof(double));
  if (spreads)
    for (int n = 0; n < m->nw; ++n) {
      const double rx = rave[3 * n], ry = rave[3 * n + 1], rz = rave[3 * n + 2];
      spreads[n] = r2[n] - (rx * rx + ry * ry + rz * rz);
    }
  return W90GPU_OK;
}

// A one-member list of a member for the per-phase entry points.
int single_lists(Engine* E, Member* m, Lists& L) { return aux_lists(E, std::vector<Member*>{m}, L); }

// Address of `scratch` on `device`, looked up once per device (with the
// device current) and kept for the life of the process.
int scratch_address(int device, char** out) {
  static std::mutex mu;
  static std::map<int, char*> addr;
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(device);
  if (it != addr.end()) {
    *out = it->second;
    return W90GPU_OK;
  }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_error(nullptr, "hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(e));
    return W90GPU_ERR_FATAL;
  }
  addr[device] = (char*)p;
  *out = (char*)p;
  return W90GPU_OK;
}

// Tables of the allocation-free path, built on the device: the work items
// (problem 0, k), the problem list {0} and the neighbour table nnl[k][b]
// (0-based) from the caller's nnlist(nk, nn), 1-based, copied raw.
__global__ void k_noalloc_tables(Item* items, int* plist, int* nnl, const int* nnlist_raw, int nk, int nn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) plist[0] = 0;
  if (i < nk) items[i] = Item{0, i};
  if (i < nk * nn) {
    const int k = i / nn, b = i % nn;
    nnl[i] = nnlist_raw[k + (size_t)nk * b] - 1;
  }
}

// Offsets of the allocation-free path in the scratch array: the problem
// record, the items, the problem list, the raw nnlist, then the arena of
// problem_layout (num_iter 0, as w90gpu_omega uses).
struct ScratchLayout {
  size_t items, plist, raw, arena, total;
};

ScratchLayout scratch_layout(int nw, int nk, int nn, Layout& lay) {
  ScratchLayout s;
  s.items = align_up(sizeof(Prob));
  s.plist = s.items + align_up(sizeof(Item) * (size_t)nk);
  s.raw = s.plist + align_up(sizeof(int));
  s.arena = s.raw + align_up(sizeof(int) * (size_t)nk * nn);
  s.total = s.arena + problem_layout(nw, nk, nn, 0, lay);
  return s;
}

}  // namespace

// =============================================================== C ABI

extern "C" {

const char* w90gpu_version(void) { return "w90gpu 0.2 (gfx950)"; }

void w90gpu_control_defaults(w90gpu_control* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->num_iter = 100;
  c->num_cg_steps = 5;
  c->conv_window = -1;
  c->lfixstep = 0;
  c->conv_tol = 1.0e-10;
  c->trial_step = 2.0;
  c->fixed_step = -999.0;
  c->conv_noise_amp = -1.0;
  c->precision = W90GPU_PREC_FP64;
}

int w90gpu_create(int device, w90gpu_ctx** ctx) {
  if (!ctx) return W90GPU_ERR_INPUT;
  *ctx = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    set_error(nullptr, "no HIP device available (%s)", hipGetErrorString(e));
    return W90GPU_ERR_FATAL;
  }
  if (device < 0 || device >= ndev) {
    set_error(nullptr, "device %d out of range (%d devices)", device, ndev);
    return W90GPU_ERR_INPUT;
  }
  if (hipSetDevice(device) != hipSuccess) {
    set_error(nullptr, "hipSetDevice(%d) failed", device);
    return W90GPU_ERR_FATAL;
  }
  tune_refresh();
  // dynamic LDS above 64 KiB must be enabled per kernel
  const struct {
    const void* f;
    size_t bytes;
  } lds[] = {
      {(const void*)k_rotate_w16<cplx>, rotate_w16_lds<cplx>(W90GPU_NUM_NNMAX)},
      {(const void*)k_rotate_w16<float2>, rotate_w16_lds<float2>(W90GPU_NUM_NNMAX)},
      {(const void*)k_step_mfma<48>, step_mfma_lds<48>()},
      {(const void*)k_step_mfma<64>, step_mfma_lds<64>()},
      {(const void*)k_domega_k<16, cplx>, domega_lds<16, cplx>()},
      {(const void*)k_domega_k<16, float2>, domega_lds<16, float2>()},
      {(const void*)k_domega_k<kLdsMaxWann, cplx>, domega_lds<kLdsMaxWann, cplx>()},
      {(const void*)k_domega_k<kLdsMaxWann, float2>, domega_lds<kLdsMaxWann, float2>()},
      {(const void*)k_domega_w64<cplx>, domega_lds<W90_STEP_N, cplx>()},
      {(const void*)k_domega_w64<float2>, domega_lds<W90_STEP_N, float2>()},
      {(const void*)k_domega<cplx>, domega_t_lds<cplx>()},
      {(const void*)k_domega<float2>, domega_t_lds<float2>()},
  };
  for (const auto& l : lds) {
    if (hipFuncSetAttribute(l.f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l.bytes) != hipSuccess) {
      set_error(nullptr, "hipFuncSetAttribute(MaxDynamicSharedMemorySize=%zu) failed", l.bytes);
      return W90GPU_ERR_FATAL;
    }
  }
  char* sp = nullptr;
  if (scratch_address(device, &sp) != W90GPU_OK) return W90GPU_ERR_FATAL;
  w90gpu_ctx* c = new (std::nothrow) w90gpu_ctx();
  if (!c) return W90GPU_ERR_ALLOC;
  c->device = device;
  c->scratch = sp;
  *ctx = c;
  return W90GPU_OK;
}

int w90gpu_destroy(w90gpu_ctx* ctx) {
  delete ctx;
  return W90GPU_OK;
}

const char* w90gpu_last_error(const w90gpu_ctx* ctx) { return ctx ? ctx->err.c_str() : g_noctx_err.c_str(); }

void w90gpu_set_error(w90gpu_ctx* ctx, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  set_error_v(ctx, fmt, ap);
  va_end(ap);
}

size_t w90gpu_problem_bytes(int num_wann, int num_kpts, int nntot, int num_iter) {
  if (num_wann < 1 || num_kpts < 1 || nntot < 1 || num_iter < 0) return 0;
  Layout lay;
  return problem_layout(num_wann, num_kpts, nntot, num_iter, lay);
}

size_t w90gpu_session_bytes(int num_wann, int num_kpts, int nntot) {
  return w90gpu_problem_bytes(num_wann, num_kpts, nntot, 0);
}

// ------------------------------------------------- rotation tile shapes

int w90gpu_tune_get(int num_wann, int nntot, char* name, size_t len) {
  if (num_wann < 1 || num_wann > W90GPU_MAX_WANN || nntot < 1 || nntot > W90GPU_NUM_NNMAX || !name || len == 0) {
    set_error(nullptr, "w90gpu_tune_get: invalid num_wann=%d, nntot=%d or buffer", num_wann, nntot);
    return W90GPU_ERR_INPUT;
  }
  std::snprintf(name, len, "%s", kRotName[rot_choice(num_wann, nntot, nullptr)]);
  return W90GPU_OK;
}

int w90gpu_tune_set(int num_wann, int nntot, const char* name) {
  if (num_wann < 1 || num_wann > W90GPU_MAX_WANN || nntot < 1 || nntot > W90GPU_NUM_NNMAX) {
    set_error(nullptr, "w90gpu_tune_set: invalid num_wann=%d or nntot=%d", num_wann, nntot);
    return W90GPU_ERR_INPUT;
  }
  TuneCache& tc = tune_cache();
  std::lock_guard<std::mutex> lock(tc.mu);
  if (!name || !name[0]) {
    tc.chosen.erase({num_wann, nntot});
    return W90GPU_OK;
  }
  const int r = rot_by_name(name);
  if (!rot_serves(r, num_wann)) {
    set_error(nullptr, "w90gpu_tune_set: kernel '%s' does not rotate num_wann=%d", name, num_wann);
    return W90GPU_ERR_INPUT;
  }
  tc.chosen[{num_wann, nntot}] = r;
  return W90GPU_OK;
}

// ------------------------------------------------------------ sessions

int w90gpu_session_create(w90gpu_ctx* ctx, const w90gpu_problem* prob, const w90gpu_control* ctl,
                          w90gpu_session** out) {
  if (!ctx || !out) return W90GPU_ERR_INPUT;
  *out = nullptr;
  int rc = validate(ctx, prob, ctl);
  if (rc) return rc;
  Layout lay;
  const size_t bytes = problem_layout(prob->num_wann, prob->num_kpts, prob->nntot, ctl->num_iter, lay);
  Engine* E = nullptr;
  if ((rc = engine_create(ctx, bytes, &E))) return rc;
  Member* m = nullptr;
  if ((rc = engine_add(E, prob, ctl, &m))) {
    engine_destroy(E);
    return rc;
  }
  w90gpu_session* s = new (std::nothrow) w90gpu_session();
  if (!s) {
    engine_destroy(E);
    return W90GPU_ERR_ALLOC;
  }
  s->eng = E;
  s->m = m;
  *out = s;
  return W90GPU_OK;
}

int w90gpu_session_destroy(w90gpu_session* s) {
  if (!s) return W90GPU_OK;
  engine_destroy(s->eng);
  delete s;
  return W90GPU_OK;
}

int w90gpu_session_load(w90gpu_session* s, const double* m_matrix, const double* u_matrix) {
  if (!s || !m_matrix || !u_matrix) return W90GPU_ERR_INPUT;
  int rc = member_upload(s->eng, s->m, m_matrix, u_matrix);
  if (rc) return rc;
  s->loaded = true;
  return W90GPU_OK;
}

int w90gpu_session_launch(w90gpu_session* s) {
  if (!s) return W90GPU_ERR_INPUT;
  if (!s->loaded) {
    set_error(s->eng->ctx, "session has no loaded M/U");
    return W90GPU_ERR_INPUT;
  }
  Engine* E = s->eng;
  HIPCHK(E->ctx, hipEventRecord(E->t0, E->stream));
  s->m->stage = 1;
  std::vector<Member*> q{s->m};
  return enqueue_start(E, q);
}

int w90gpu_session_wait(w90gpu_session* s, w90gpu_result* r) {
  if (!s) return W90GPU_ERR_INPUT;
  Engine* E = s->eng;
  int rc = engine_advance(E, -1);
  if (rc) return rc;
  HIPCHK(E->ctx, hipEventRecord(E->t1, E->stream));
  HIPCHK(E->ctx, hipEventSynchronize(E->t1));
  float ms = 0.f;
  HIPCHK(E->ctx, hipEventElapsedTime(&ms, E->t0, E->t1));
  E->finished_ids.clear();
  if (s->m->stage != 3) {
    set_error(E->ctx, "solve did not finish");
    return W90GPU_ERR_FATAL;
  }
  w90gpu_result res = s->m->res;
  res.solve_seconds = ms * 1e-3;
  s->last = res;
  if (r) *r = res;
  return res.status;
}

int w90gpu_session_solve(w90gpu_session* s, w90gpu_result* r) {
  int rc = w90gpu_session_launch(s);
  if (rc) return rc;
  return w90gpu_session_wait(s, r);
}

int w90gpu_session_fetch(w90gpu_session* s, double* m_matrix, double* u_matrix, double* centres, double* spreads) {
  if (!s) return W90GPU_ERR_INPUT;
  return member_fetch(s->eng, s->m, m_matrix, u_matrix, centres, spreads);
}

int w90gpu_session_history(w90gpu_session* s, double* out, int rows) {
  if (!s || !out || rows < 0) return -W90GPU_ERR_INPUT;
  State st;
  if (read_member_state(s->eng, s->m, &st)) return -W90GPU_ERR_FATAL;
  const int n = std::min(rows, st.iter + 1);
  if (n <= 0) return 0;
  if (hipMemcpy(out, s->m->p.log, sizeof(double) * 4 * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error(s->eng->ctx, "history download failed");
    return -W90GPU_ERR_FATAL;
  }
  return n;
}

int w90gpu_session_phase_time(w90gpu_session* s, int phase, int reps, double* avg_ms) {
  if (!s || reps < 1 || !avg_ms) return W90GPU_ERR_INPUT;
  if (phase < W90GPU_PHASE_ROTATION || phase > W90GPU_PHASE_REDUCTION) return W90GPU_ERR_INPUT;
  Engine* E = s->eng;
  if (s->m->stage == 1 || !s->loaded) {
    set_error(E->ctx, "phase timing needs a solved session");
    return W90GPU_ERR_INPUT;
  }
  Lists L;
  int rc = single_lists(E, s->m, L);
  if (rc) return rc;
  // one phase on the final state, repeated: a rotation reads M(cur) and E
  // and rewrites the other M buffer, a gradient rewrites G and its norms, a
  // retraction reads D and U(cur) and rewrites E and the other U buffer, so
  // every repetition does the same work.  The launches are replayed from a
  // graph, as in the solve.
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  HIPCHK(E->ctx, hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
  for (int i = 0; i < reps && rc == W90GPU_OK; ++i) {
    if (phase == W90GPU_PHASE_ROTATION) rc = launch_rotate(E, L, MODE_BENCH);
    else if (phase == W90GPU_PHASE_GRADIENT) rc = launch_domega(E, L, MODE_BENCH);
    else if (phase == W90GPU_PHASE_RETRACTION) rc = launch_step(E, L, MODE_BENCH);
    else rc = launch_reduce(E, L, MODE_BENCH);
  }
  const hipError_t ec = hipStreamEndCapture(E->stream, &g);
  if (rc || ec != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    set_error(E->ctx, "capture of the phase timing failed");
    return W90GPU_ERR_FATAL;
  }
  auto done = [&](int code) {
    if (ge) (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    return code;
  };
  if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
    set_error(E->ctx, "instantiation of the phase timing graph failed");
    return done(W90GPU_ERR_FATAL);
  }
  float ms = 0.f;
  if (hipGraphLaunch(ge, E->stream) != hipSuccess ||  // warm-up
      hipEventRecord(E->t0, E->stream) != hipSuccess || hipGraphLaunch(ge, E->stream) != hipSuccess ||
      hipEventRecord(E->t1, E->stream) != hipSuccess || hipEventSynchronize(E->t1) != hipSuccess ||
      hipEventElapsedTime(&ms, E->t0, E->t1) != hipSuccess) {
    set_error(E->ctx, "phase timing failed");
    return done(W90GPU_ERR_FATAL);
  }
  *avg_ms = ms / reps;
  return done(W90GPU_OK);
}

int w90gpu_session_rotation_time(w90gpu_session* s, int reps, double* avg_ms) {
  return w90gpu_session_phase_time(s, W90GPU_PHASE_ROTATION, reps, avg_ms);
}

int w90gpu_wann_main(w90gpu_ctx* ctx, const w90gpu_problem* prob, const w90gpu_control* ctl, double* m_matrix,
                     double* u_matrix, double* centres, double* spreads, w90gpu_result* result) {
  if (!ctx) return W90GPU_ERR_INPUT;
  if (!m_matrix || !u_matrix) {
    set_error(ctx, "m_matrix and u_matrix must be associated");
    return W90GPU_ERR_INPUT;
  }
  w90gpu_session* s = nullptr;
  int rc = w90gpu_session_create(ctx, prob, ctl, &s);
  if (rc) return rc;
  rc = w90gpu_session_load(s, m_matrix, u_matrix);
  if (!rc) rc = w90gpu_session_solve(s, result);
  if (rc <= 0) {
    const int rc2 = w90gpu_session_fetch(s, m_matrix, u_matrix, centres, spreads);
    if (rc2) rc = rc2;
  }
  w90gpu_session_destroy(s);
  return rc;
}

// ------------------------------------------------------------- batches

int w90gpu_batch_create(w90gpu_ctx* ctx, size_t pool_bytes, w90gpu_batch** out) {
  if (!ctx || !out) return W90GPU_ERR_INPUT;
  *out = nullptr;
  Engine* E = nullptr;
  int rc = engine_create(ctx, pool_bytes, &E);
  if (rc) return rc;
  w90gpu_batch* b = new (std::nothrow) w90gpu_batch();
  if (!b) {
    engine_destroy(E);
    return W90GPU_ERR_ALLOC;
  }
  b->eng = E;
  *out = b;
  return W90GPU_OK;
}

int w90gpu_batch_destroy(w90gpu_batch* b) {
  if (!b) return W90GPU_OK;
  engine_destroy(b->eng);
  delete b;
  return W90GPU_OK;
}

int w90gpu_batch_add(w90gpu_batch* b, const w90gpu_problem* prob, const w90gpu_control* ctl,
                     const double* m_matrix, const double* u_matrix, int* id) {
  if (!b || !id || !m_matrix || !u_matrix) return W90GPU_ERR_INPUT;
  *id = -1;
  Member* m = nullptr;
  int rc = engine_add(b->eng, prob, ctl, &m);
  if (rc) return rc;
  if ((rc = member_upload(b->eng, m, m_matrix, u_matrix))) {
    engine_remove(b->eng, m);
    return rc;
  }
  m->stage = 1;
  *id = m->id;
  return W90GPU_OK;
}

int w90gpu_batch_step(w90gpu_batch* b, int max_blocks, int* finished, int cap, int* nfinished) {
  if (!b) return W90GPU_ERR_INPUT;
  Engine* E = b->eng;
  int rc = engine_advance(E, max_blocks < 0 ? 0 : max_blocks);
  if (rc) return rc;
  int n = 0;
  for (int id : E->finished_ids) {
    if (finished && n < cap) finished[n] = id;
    ++n;
  }
  if (nfinished) *nfinished = std::min(n, cap);
  if (n <= cap) E->finished_ids.clear();
  else E->finished_ids.erase(E->finished_ids.begin(), E->finished_ids.begin() + cap);
  return W90GPU_OK;
}

int w90gpu_batch_run(w90gpu_batch* b) {
  if (!b) return W90GPU_ERR_INPUT;
  int rc = engine_advance(b->eng, -1);
  b->eng->finished_ids.clear();
  return rc;
}

int w90gpu_batch_restart(w90gpu_batch* b) {
  if (!b) return W90GPU_ERR_INPUT;
  for (Member* m : b->eng->slots)
    if (m && m->stage != 2) m->stage = 1;
  return W90GPU_OK;
}

int w90gpu_batch_result(w90gpu_batch* b, int id, w90gpu_result* r) {
  if (!b || !r) return W90GPU_ERR_INPUT;
  auto it = b->eng->by_id.find(id);
  if (it == b->eng->by_id.end()) {
    set_error(b->eng->ctx, "no problem with id %d", id);
    return W90GPU_ERR_INPUT;
  }
  if (it->second->stage != 3) {
    set_error(b->eng->ctx, "problem %d has not finished", id);
    return W90GPU_ERR_INPUT;
  }
  *r = it->second->res;
  return r->status;
}

int w90gpu_batch_fetch(w90gpu_batch* b, int id, double* m_matrix, double* u_matrix, double* centres,
                       double* spreads) {
  if (!b) return W90GPU_ERR_INPUT;
  auto it = b->eng->by_id.find(id);
  if (it == b->eng->by_id.end()) {
    set_error(b->eng->ctx, "no problem with id %d", id);
    return W90GPU_ERR_INPUT;
  }
  return member_fetch(b->eng, it->second, m_matrix, u_matrix, centres, spreads);
}

int w90gpu_batch_remove(w90gpu_batch* b, int id) {
  if (!b) return W90GPU_ERR_INPUT;
  auto it = b->eng->by_id.find(id);
  if (it == b->eng->by_id.end()) {
    set_error(b->eng->ctx, "no problem with id %d", id);
    return W90GPU_ERR_INPUT;
  }
  if (it->second->stage == 2) {
    set_error(b->eng->ctx, "problem %d is running", id);
    return W90GPU_ERR_INPUT;
  }
  HIPCHK(b->eng->ctx, hipStreamSynchronize(b->eng->copy));
  engine_remove(b->eng, it->second);
  return W90GPU_OK;
}

int w90gpu_batch_info(w90gpu_batch* b, size_t* pool_bytes, size_t* used_bytes, int* running, int* queued) {
  if (!b) return W90GPU_ERR_INPUT;
  const Engine* E = b->eng;
  if (pool_bytes) *pool_bytes = E->pool.size;
  if (used_bytes) *used_bytes = E->pool.used;
  int nr = 0, nq = 0;
  for (const Member* m : E->slots) {
    if (!m) continue;
    nr += (m->stage == 2);
    nq += (m->stage == 1);
  }
  if (running) *running = nr;
  if (queued) *queued = nq;
  return W90GPU_OK;
}

// ---------------------------------------------------- per-phase entry points

int w90gpu_omega(w90gpu_ctx* ctx, const w90gpu_problem* prob, const double* m_matrix, int first_pass,
                 double omega_invariant, double* rave, double* r2ave, double* omega) {
  if (!ctx || !m_matrix) return W90GPU_ERR_INPUT;
  w90gpu_control ctl;
  w90gpu_control_defaults(&ctl);
  ctl.num_iter = 0;
  w90gpu_session* s = nullptr;
  int rc = w90gpu_session_create(ctx, prob, &ctl, &s);
  if (rc) return rc;
  Engine* E = s->eng;
  Member* m = s->m;
  const size_t mb = (size_t)m->nw * m->nw * m->nn * m->nk * sizeof(cplx);
  const size_t nd = (size_t)m->nw * m->nw * m->nk;
  auto run = [&]() -> int {
    int r = upload_state_template(E, m, first_pass ? 1 : 0, omega_invariant);
    if (r) return r;
    Lists L;
    if ((r = single_lists(E, m, L))) return r;
    HIPCHK(ctx, hipMemcpyAsync(m->p.M[0], m_matrix, mb, hipMemcpyHostToDevice, E->stream));
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, E->stream, E->d_probs, L.plist, 1);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>(1024, (nd + 255) / 256)), dim3(256), 0, E->stream,
                       m->p.U[0], nd);
    if ((r = launch_omega(E, L, MODE_INIT))) return r;
    State st;
    if ((r = read_member_state(E, m, &st))) return r;
    if (rave) HIPCHK(ctx, hipMemcpy(rave, m->p.rave[0], 3 * m->nw * sizeof(double), hipMemcpyDeviceToHost));
    if (r2ave) HIPCHK(ctx, hipMemcpy(r2ave, m->p.r2ave[0], m->nw * sizeof(double), hipMemcpyDeviceToHost));
    if (omega) {
      omega[0] = st.wann.om_i;
      omega[1] = st.wann.om_d;
      omega[2] = st.wann.om_od;
      omega[3] = st.wann.om_tot;
    }
    return W90GPU_OK;
  };
  rc = run();
  w90gpu_session_destroy(s);
  return rc;
}

// w90gpu_omega with every device buffer in the static scratch array: the
// same kernels and launch sequence (initial wann_omega, MODE_INIT), on the
// null stream, with no allocation.  Calls on one device share the array,
// so calls that overlap on one GPU race; every call rewrites what it reads
// (the loop state from its template, U[0] cleared as in w90gpu_omega).
int w90gpu_omega_noalloc(w90gpu_ctx* ctx, const w90gpu_problem* prob, const double* m_matrix, int first_pass,
                         double omega_invariant, double* rave, double* r2ave, double* omega) {
  if (!ctx || !m_matrix) return W90GPU_ERR_INPUT;
  if (!ctx->scratch) {
    set_error(ctx, "w90gpu_omega_noalloc: no device scratch on this context");
    return W90GPU_ERR_FATAL;
  }
  w90gpu_control ctl;
  w90gpu_control_defaults(&ctl);
  ctl.num_iter = 0;
  int rc = validate(ctx, prob, &ctl);
  if (rc) return rc;
  const int nw = prob->num_wann, nk = prob->num_kpts, nn = prob->nntot;
  Layout lay;
  const ScratchLayout sl = scratch_layout(nw, nk, nn, lay);
  if (sl.total > kScratchBytes) {
    set_error(ctx,
              "w90gpu_omega_noalloc: num_wann=%d, num_kpts=%d, nntot=%d needs %zu bytes of device scratch; the "
              "static scratch holds %zu (use w90gpu_omega)",
              nw, nk, nn, sl.total, kScratchBytes);
    return W90GPU_ERR_INPUT;
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  char* base = ctx->scratch;
  Member m;
  m.nw = nw;
  m.nk = nk;
  m.nn = nn;
  m.cls = wann_class(nw);
  m.rk = rot_choice(nw, nn, nullptr);
  m.ctl = ctl;
  m.wbtot = prob->wbtot;
  member_bind(&m, base + sl.arena, lay);
  const State st0 = state_template(&m, first_pass ? 1 : 0, omega_invariant);
  const size_t mb = (size_t)nw * nw * nn * nk * sizeof(cplx);
  const size_t nd = (size_t)nw * nw * nk;
  Item* items = (Item*)(base + sl.items);
  int* plist = (int*)(base + sl.plist);
  int* raw = (int*)(base + sl.raw);
  HIPCHK(ctx, hipMemcpy(base, &m.p, sizeof(Prob), hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(raw, prob->nnlist, sizeof(int) * (size_t)nk * nn, hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy((void*)m.p.bk, prob->bk, sizeof(double) * 3 * (size_t)nk * nn, hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy((void*)m.p.wb, prob->wb, sizeof(double) * (size_t)nn, hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(m.st_init, &st0, sizeof(State), hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(m.p.M[0], m_matrix, mb, hipMemcpyHostToDevice));
  const int nt = std::max(nk * nn, nk);
  hipLaunchKernelGGL(k_noalloc_tables, dim3((nt + 255) / 256), dim3(256), 0, nullptr, items, plist, (int*)m.p.nnl,
                     raw, nk, nn);
  HIPCHK(ctx, hipGetLastError());
  Engine E;
  E.ctx = ctx;
  E.stream = nullptr;
  E.d_probs = (Prob*)base;
  Lists L;
  L.items = items;
  L.plist = plist;
  L.np = 1;
  // the segments build_lists gives a single member: empty before and after
  // its group, all nk items in (group, FP64)
  one_group_lists(L, group_of(m.cls, m.rk), 0, nk, nn, rot_panels(m.rk, nw), nw);
  L.any_group = m.p.ngroup > 0;
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, nullptr, E.d_probs, L.plist, 1);
  hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>(1024, (nd + 255) / 256)), dim3(256), 0, nullptr,
                     m.p.U[0], nd);
  HIPCHK(ctx, hipGetLastError());
  if ((rc = launch_omega(&E, L, MODE_INIT))) return rc;
  State st;
  HIPCHK(ctx, hipMemcpy(&st, m.p.st, sizeof(State), hipMemcpyDeviceToHost));
  if (rave) HIPCHK(ctx, hipMemcpy(rave, m.p.rave[0], 3 * (size_t)nw * sizeof(double), hipMemcpyDeviceToHost));
  if (r2ave) HIPCHK(ctx, hipMemcpy(r2ave, m.p.r2ave[0], (size_t)nw * sizeof(double), hipMemcpyDeviceToHost));
  if (omega) {
    omega[0] = st.wann.om_i;
    omega[1] = st.wann.om_d;
    omega[2] = st.wann.om_od;
    omega[3] = st.wann.om_tot;
  }
  return W90GPU_OK;
}

int w90gpu_domega(w90gpu_ctx* ctx, const w90gpu_problem* prob, const double* m_matrix, double* cdodq) {
  if (!ctx || !m_matrix || !cdodq) return W90GPU_ERR_INPUT;
  w90gpu_control ctl;
  w90gpu_control_defaults(&ctl);
  ctl.num_iter = 1;
  w90gpu_session* s = nullptr;
  int rc = w90gpu_session_create(ctx, prob, &ctl, &s);
  if (rc) return rc;
  Engine* E = s->eng;
  Member* m = s->m;
  const size_t mb = (size_t)m->nw * m->nw * m->nn * m->nk * sizeof(cplx);
  const size_t ub = (size_t)m->nw * m->nw * m->nk * sizeof(cplx);
  const size_t nd = (size_t)m->nw * m->nw * m->nk;
  auto run = [&]() -> int {
    Lists L;
    int r = single_lists(E, m, L);
    if (r) return r;
    HIPCHK(ctx, hipMemcpyAsync(m->p.M[0], m_matrix, mb, hipMemcpyHostToDevice, E->stream));
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, E->stream, E->d_probs, L.plist, 1);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>(1024, (nd + 255) / 256)), dim3(256), 0, E->stream,
                       m->p.D, nd);
    if ((r = launch_omega(E, L, MODE_INIT))) return r;
    if ((r = launch_domega(E, L, MODE_TRIAL))) return r;
    HIPCHK(ctx, hipMemcpyAsync(cdodq, m->p.G, ub, hipMemcpyDeviceToHost, E->stream));
    HIPCHK(ctx, hipStreamSynchronize(E->stream));
    return W90GPU_OK;
  };
  rc = run();
  w90gpu_session_destroy(s);
  return rc;
}

int w90gpu_new_u_and_m(w90gpu_ctx* ctx, const w90gpu_problem* prob, const double* cdq, double* u_matrix,
                       double* m_matrix) {
  if (!ctx || !cdq || !u_matrix || !m_matrix) return W90GPU_ERR_INPUT;
  w90gpu_control ctl;
  w90gpu_control_defaults(&ctl);
  ctl.num_iter = 1;
  w90gpu_session* s = nullptr;
  int rc = w90gpu_session_create(ctx, prob, &ctl, &s);
  if (rc) return rc;
  Engine* E = s->eng;
  Member* m = s->m;
  const size_t mb = (size_t)m->nw * m->nw * m->nn * m->nk * sizeof(cplx);
  const size_t ub = (size_t)m->nw * m->nw * m->nk * sizeof(cplx);
  auto run = [&]() -> int {
    Lists L;
    int r = single_lists(E, m, L);
    if (r) return r;
    HIPCHK(ctx, hipMemcpyAsync(m->p.M[0], m_matrix, mb, hipMemcpyHostToDevice, E->stream));
    HIPCHK(ctx, hipMemcpyAsync(m->p.U[0], u_matrix, ub, hipMemcpyHostToDevice, E->stream));
    HIPCHK(ctx, hipMemcpyAsync(m->p.D, cdq, ub, hipMemcpyHostToDevice, E->stream));
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, E->stream, E->d_probs, L.plist, 1);
    hipLaunchKernelGGL(k_set_direct, dim3(1), dim3(64), 0, E->stream, E->d_probs + m->slot, 1.0);
    if ((r = launch_step(E, L, MODE_SECOND))) return r;
    if ((r = launch_rotate(E, L, MODE_SECOND))) return r;
    HIPCHK(ctx, hipMemcpyAsync(m_matrix, m->p.M[1], mb, hipMemcpyDeviceToHost, E->stream));
    HIPCHK(ctx, hipMemcpyAsync(u_matrix, m->p.U[1], ub, hipMemcpyDeviceToHost, E->stream));
    HIPCHK(ctx, hipStreamSynchronize(E->stream));
    return W90GPU_OK;
  };
  rc = run();
  w90gpu_session_destroy(s);
  return rc;
}

int w90gpu_project(w90gpu_ctx* ctx, const w90gpu_problem* prob, const double* a_matrix, double* u_matrix,
                   double* m_matrix) {
  if (!ctx || !a_matrix || !u_matrix || !m_matrix) return W90GPU_ERR_INPUT;
  w90gpu_control ctl;
  w90gpu_control_defaults(&ctl);
  ctl.num_iter = 1;
  w90gpu_session* s = nullptr;
  int rc = w90gpu_session_create(ctx, prob, &ctl, &s);
  if (rc) return rc;
  Engine* E = s->eng;
  Member* m = s->m;
  const size_t mb = (size_t)m->nw * m->nw * m->nn * m->nk * sizeof(cplx);
  const size_t ub = (size_t)m->nw * m->nw * m->nk * sizeof(cplx);
  auto run = [&]() -> int {
    Lists L;
    int r = single_lists(E, m, L);
    if (r) return r;
    const int n = all_items(L);
    // A is staged in the gradient buffer, which this phase does not use
    HIPCHK(ctx, hipMemcpyAsync(m->p.G, a_matrix, ub, hipMemcpyHostToDevice, E->stream));
    HIPCHK(ctx, hipMemcpyAsync(m->p.M[0], m_matrix, mb, hipMemcpyHostToDevice, E->stream));
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, E->stream, E->d_probs, L.plist, 1);
    hipLaunchKernelGGL(k_polar, dim3(n), dim3(W90_NT), 0, E->stream, E->d_probs, L.items, n);
    HIPCHK(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_transpose_e, dim3(n), dim3(W90_NT), 0, E->stream, E->d_probs, L.items, n);
    HIPCHK(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_set_direct, dim3(1), dim3(64), 0, E->stream, E->d_probs + m->slot, 1.0);
    if ((r = launch_rotate(E, L, MODE_SECOND))) return r;
    HIPCHK(ctx, hipMemcpyAsync(m->p.U[0], m->p.E, ub, hipMemcpyDeviceToDevice, E->stream));
    hipLaunchKernelGGL(k_unitarity, dim3(n), dim3(W90_NT), 0, E->stream, E->d_probs, L.items, n);
    HIPCHK(ctx, hipGetLastError());
    std::vector<double> dev(m->nk);
    HIPCHK(ctx, hipMemcpyAsync(dev.data(), m->p.unit, sizeof(double) * m->nk, hipMemcpyDeviceToHost, E->stream));
    HIPCHK(ctx, hipMemcpyAsync(u_matrix, m->p.U[0], ub, hipMemcpyDeviceToHost, E->stream));
    HIPCHK(ctx, hipMemcpyAsync(m_matrix, m->p.M[1], mb, hipMemcpyDeviceToHost, E->stream));
    HIPCHK(ctx, hipStreamSynchronize(E->stream));
    double mx = 0.0;
    for (double d : dev) mx = (d > mx || d != d) ? d : mx;
    if (!(mx <= 1e-5)) {
      set_error(ctx, "Error in unitarity of initial U in overlap_project (deviation %.3e)", mx);
      return W90GPU_ERR_FATAL;
    }
    return W90GPU_OK;
  };
  rc = run();
  w90gpu_session_destroy(s);
  return rc;
}

}  // extern "C"
