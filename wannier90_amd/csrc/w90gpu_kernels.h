// Device code of libw90gpu: the kernels of the spread minimisation
// (wann_main without disentanglement) on MI355X (gfx950).
//
// Every kernel works on a set of problems.  A problem's device state is a
// Prob (pointers into its HBM arena) in a device array; per-k kernels take
// a list of work items (problem, k) and run one workgroup per item,
// per-problem kernels take a list of problem indices and run one workgroup
// per problem.  One launch per phase therefore covers every active problem
// of a batch, and a single problem is a batch of one.  A problem whose loop
// has finished (State::done) or waits for the host (State::hold) makes its
// workgroups exit at once: the per-problem convergence mask.
//
// All scalar control of wann_main -- the Fletcher-Reeves coefficient and its
// resets, the uphill test, the parabolic line search and the convergence
// window -- runs in per-problem control kernels, so an iteration is a fixed
// sequence of launches with no host round trip.
//
// Storage precision.  The kernels that read or write M(k,b) are templates
// on the complex storage type T2: double2 for the FP64 path, float2 for the
// bulk phase of the mixed-precision path, whose rotation runs in FP32 with
// FP32 accumulation.  Everything else is FP64 in both: U, the step
// exp(dW) (so U stays unitary to FP64 rounding), Im ln M_nn, the centres,
// every spread sum and the gradient arithmetic.  The mixed bulk phase ends
// when Omega stops decreasing by more than its FP32 noise floor; M is then
// rebuilt in FP64 from the initial overlaps, M(k,b) = E(k)^+ M0(k,b)
// E(k+b) with E = U0^+ U (MODE_REMIX), and the FP64 loop polishes to
// conv_tol.  A bulk
// step that does not lower Omega beyond the FP32 noise floor is not taken.
//
// Reference: src/wannierise.F90 (wann_main :48-1072, internal_new_u_and_m
// :1599-1751, wann_omega :1979-2284, wann_domega :2287-2569,
// wann_check_unitarity :2594-2661), src/overlap.F90 (overlap_project
// :758-952).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/w90gpu.h"

typedef double2 cplx;

#define W90_NT 256          // threads per workgroup of the per-k kernels
#define W90_STEP_NT 1024    // threads of the global-memory retraction (above 64 bands)
#define W90_FINAL_NT 1024   // threads of the per-problem omega kernel
#define W90_MAX_GROUPS 32   // k-ranges of the first omega reduction level
#define W90_MIXED_RTOL 1e-8 // relative FP32 noise floor of Omega differences in the bulk phase
#ifndef W90_BULK_STALLS
#define W90_BULK_STALLS 2   // rejected bulk steps in a row that end the bulk phase
#endif

// --------------------------------------------------------------- helpers

__device__ __forceinline__ cplx c_make(double r, double i) { return make_double2(r, i); }
__device__ __forceinline__ cplx c_add(cplx a, cplx b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cplx c_sub(cplx a, cplx b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cplx c_scale(cplx a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ cplx c_conj(cplx a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ cplx c_mul(cplx a, cplx b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ cplx c_cmul(cplx a, cplx b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ cplx c_div(cplx a, cplx b) {
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ __forceinline__ double c_abs2(cplx a) { return a.x * a.x + a.y * a.y; }

// FP32 complex (storage of the mixed bulk phase)
__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <class T2>
__device__ __forceinline__ T2 c_zero() {
  T2 z;
  z.x = 0;
  z.y = 0;
  return z;
}
// storage -> FP64 and FP64 -> storage
__device__ __forceinline__ cplx c_wide(cplx a) { return a; }
__device__ __forceinline__ cplx c_wide(float2 a) { return make_double2((double)a.x, (double)a.y); }
// Streaming (non-temporal) store of a complex element.  Used for per-k
// blocks whose next readers are later kernels: the XCD's L2 does not keep
// them across the kernel boundary, so holding them there only lengthens the
// write-back at the end of the kernel.  Measured on the 16-band rotation
// (17.2 -> 14.9 us per launch, fcc 512-k mesh); the 17-64 band rotation
// keeps ordinary stores (the streaming form was slower there, 707 -> 726 us
// on the triclinic 64-band 800-k mesh).
template <class T2>
__device__ __forceinline__ void st_stream(T2* dst, T2 v) {
  __builtin_nontemporal_store(v.x, &dst->x);
  __builtin_nontemporal_store(v.y, &dst->y);
}
template <class T2>
__device__ __forceinline__ T2 c_narrow(cplx a);
template <>
__device__ __forceinline__ cplx c_narrow<cplx>(cplx a) {
  return a;
}
template <>
__device__ __forceinline__ float2 c_narrow<float2>(cplx a) {
  return make_float2((float)a.x, (float)a.y);
}

// Sum over the 64 lanes of a wavefront (butterfly; lane 0's value is used).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum of one double per thread (blockDim a multiple of 64):
// each wavefront's butterfly sum (lane 0), then the wavefront sums in
// wavefront order by every thread; deterministic, two barriers.
__device__ double block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nwv = blockDim.x >> 6;
  v = wave_sum(v);
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = 0.0;
  for (int i = 0; i < nwv; ++i) r += red[i];
  __syncthreads();
  return r;
}

// Two block sums in one pass (red holds 2 blockDim / 64 doubles): each value
// is summed in block_sum's order, with two barriers for both.
__device__ void block_sum2(double& a, double& b, double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nwv = blockDim.x >> 6;
  a = wave_sum(a);
  b = wave_sum(b);
  if (lane == 0) {
    red[w] = a;
    red[nwv + w] = b;
  }
  __syncthreads();
  double ra = 0.0, rb = 0.0;
  for (int i = 0; i < nwv; ++i) {
    ra += red[i];
    rb += red[nwv + i];
  }
  __syncthreads();
  a = ra;
  b = rb;
}

// Block-wide maximum (blockDim a multiple of 64), as block_sum.
__device__ double block_max(double v, double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nwv = blockDim.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = red[0];
  for (int i = 1; i < nwv; ++i) r = fmax(r, red[i]);
  __syncthreads();
  return r;
}

// --------------------------------------------------------------- state

struct Spread {
  double om_i, om_d, om_od, om_tot;
};

// MODE_BENCH: a trial-step rotation that runs whatever the loop state
// (timing of the rotation kernel on a solved problem)
enum Mode { MODE_INIT = 0, MODE_TRIAL = 1, MODE_SECOND = 2, MODE_REMIX = 3, MODE_BENCH = 4 };

// Per-problem solver state, resident in device memory.
struct State {
  // control (wann_control_type)
  int num_iter, num_cg_steps, conv_window, lfixstep;
  double conv_tol, trial_step, fixed_step, wbtot;
  // loop state of wann_main
  int iter;        // completed iterations
  int ncg_s[2];    // CG step counter, double-buffered by iteration parity (see search_control)
  int cur;         // buffer holding the current U, M
  int done;        // loop finished (converged or num_iter)
  int converged;
  int lquad;
  int second;      // parabolic step pending in this iteration
  int conv_count;
  int first_pass;
  int phase;       // 0: mixed bulk (FP32 M), 1: FP64
  int hold;        // waiting for the host: the FP64 rebuild of M after the
                   // mixed bulk phase, or (with second set) a parabolic step
                   // found in a block launched without second-step kernels
  int bulk_iter;   // iterations done in the bulk phase
  int stall;       // consecutive rejected bulk steps
  int nsecond;     // parabolic steps taken since the host last read the status
  int nheld;       // parabolic steps released to the host (k_release)
  double gcnorm0_s[2];  // previous squared gradient norm, double-buffered like ncg_s
  double gcnorm1, gcfac, doda0, alphamin, falphamin;
  double beta, dir_sign, scale;
  double omega_invariant;
  Spread wann, old, trial;
  int conv_run;    // iterations in a row, up to this one, with |dOmega| <= conv_tol
  int pad_run;
};

// Device view of one problem: sizes and pointers into its HBM arena.
struct Prob {
  int nw, nk, nn, npan;
  int psz, nw2, ngroup, pad;  // psz: per-k partial vector length; ngroup: first-level k-ranges
                              // (0: the final reduction sums the per-k vectors itself)
  // npan: tiles of the MFMA rotation per block M(k,b) (num_wann > 16)
  const int* nnl;     // [k][b] 0-based neighbour
  const double* bk;   // [k][b][3]
  const double* wb;   // [b]
  cplx* M[2];         // [k][b][nw*nw] column-major blocks, FP64
  float2* Mf[2];      // FP32 blocks of the mixed bulk phase (inside M[1])
  const cplx* M0;     // initial M of the problem
  cplx* U[2];         // [k][nw*nw]
  const cplx* U0;     // initial U
  cplx* G;            // gradient
  cplx* D;            // search direction (cdqkeep)
  cplx* E;            // exp(dW) per k
  cplx* W1;           // workspace per k
  cplx* W2;           // workspace per k
  cplx* ET;           // [k][nw*nw] exp(dW) transposed (row-major E), the operand layout of
                      // k_rot_direct; shares W1's memory (the retraction above 64 bands
                      // uses W1 as a workspace and its rotation reads E)
  double* ln[2];      // [k][b][n] Im ln M_nn of each buffer
  double* d2;         // [k][b][n] |M_nn|^2 (MFMA path)
  double* part_od;    // [k][b][tile] (MFMA path)
  double* part_all;   // [k][b][tile] (MFMA path)
  void* Tm;           // [k][b][nw*nw] E_k^+ M(k,b) of the MFMA path, storage precision
  double* kpart;      // [k][psz] per-k partial sums of wann_omega (DiagAcc layout)
  double* gpart;      // [ngroup][psz] sums over contiguous k-ranges
  double* part_gg;    // [k]
  double* part_gd;    // [k]
  double* rave[2];    // [3][nw] per buffer (Fortran rave(3,nw): ind fastest)
  double* r2ave[2];   // [nw]
  double* unit;       // [k] unitarity deviation
  double* log;        // [num_iter+1][4] om_tot, om_d, om_od, gcnorm1
  State* st;
  const State* st_init;  // state at the start of a solve (control + first_pass)
};

// A per-k work item.
struct Item {
  int prob;
  int k;
};

// Per-problem status the host reads after each block of iterations.
struct Status {
  int done, hold, iter, converged;
  int second;  // a parabolic step waits for the host (hold set with it)
  int nsecond; // parabolic steps taken during the block
  int pad[2];
};

// Final state of a problem gathered for the host.
struct Final {
  State st;
  double unit_max;
  double pad;
};

template <class T2>
__device__ __forceinline__ T2* mbuf(const Prob& p, int i);
template <>
__device__ __forceinline__ cplx* mbuf<cplx>(const Prob& p, int i) {
  return p.M[i];
}
template <>
__device__ __forceinline__ float2* mbuf<float2>(const Prob& p, int i) {
  return p.Mf[i];
}

// Whether a kernel of this mode works on the problem.  The flags are read
// unconditionally, so their loads are in flight together: the state was
// written by the kernel before, and a read behind each test would pay a
// memory latency of its own.
__device__ __forceinline__ bool gate(const State* st, int mode) {
  const int done = st->done, hold = st->hold, second = st->second;
  if (mode == MODE_BENCH) return true;
  if (mode == MODE_REMIX) return (hold != 0) & (done == 0);
  if (done | hold) return false;
  if (mode == MODE_SECOND) return second != 0;
  return true;
}

// gate() and the current buffer in one round of loads: the buffer, or -1
// when the kernel skips the problem (the early return depends on the
// buffer's value, so its load is issued with the flags').
__device__ __forceinline__ int gate_cur(const State* st, int mode) {
  const int cur = st->cur;
  return gate(st, mode) ? cur : -1;
}

// buffer written by a rotation of this mode: the next buffer for a step;
// the current one for the initial evaluation and for the FP64 rebuild
// (whose M then pairs with the current U)
__device__ __forceinline__ int dst_buf(int mode, int cur) {
  return (mode == MODE_INIT || mode == MODE_REMIX) ? cur : 1 - cur;
}

// --------------------------------------------------------- setup kernels

// Start a solve: the loop state of each listed problem from its template.
__global__ void k_reset(const Prob* probs, const int* plist, int np) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const Prob& p = probs[plist[i]];
  *p.st = *p.st_init;
}

// Direct-step state for internal_new_u_and_m as a standalone phase.
__global__ void k_set_direct(const Prob* probs, double scale) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  State* st = probs[0].st;
  st->done = 0;
  st->hold = 0;
  st->second = 1;
  st->cur = 0;
  st->scale = scale;
}

__global__ void k_zero(cplx* x, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] = c_make(0.0, 0.0);
}

// Per item: M[0](k) <- M0(k), U[0](k) <- U0(k), D(k) <- 0 (start of a solve).
__global__ void __launch_bounds__(W90_NT) k_load_initial(const Prob* probs, const Item* items, int nitems) {
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const size_t nw2 = p.nw2, mb = nw2 * p.nn;
  const cplx* m0 = p.M0 + (size_t)it.k * mb;
  cplx* m = p.M[0] + (size_t)it.k * mb;
  for (size_t e = threadIdx.x; e < mb; e += blockDim.x) m[e] = m0[e];
  const cplx* u0 = p.U0 + (size_t)it.k * nw2;
  cplx* u = p.U[0] + (size_t)it.k * nw2;
  cplx* d = p.D + (size_t)it.k * nw2;
  for (size_t e = threadIdx.x; e < nw2; e += blockDim.x) {
    u[e] = u0[e];
    d[e] = c_make(0.0, 0.0);
  }
}

// Per item: FP32 copy of the current FP64 M (entry to the mixed bulk phase).
__global__ void __launch_bounds__(W90_NT) k_to_f32(const Prob* probs, const Item* items, int nitems) {
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const State* st = p.st;
  if (st->done || st->phase != 0) return;
  const size_t mb = (size_t)p.nw2 * p.nn;
  const cplx* src = p.M[st->cur] + (size_t)it.k * mb;
  float2* dst = p.Mf[st->cur] + (size_t)it.k * mb;
  for (size_t e = threadIdx.x; e < mb; e += blockDim.x) dst[e] = c_narrow<float2>(src[e]);
}

// Per item of a problem on hold: E(k) = U0(k)^+ U(k), the rotation that
// takes the initial overlaps M0 to the gauge of the current U.
__global__ void __launch_bounds__(W90_NT) k_delta_u(const Prob* probs, const Item* items, int nitems);

// Status of each listed problem (one thread per problem).
// (one thread per problem; clears the count of parabolic steps).
__global__ void k_status(const Prob* probs, const int* plist, int np, Status* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  State* st = probs[plist[i]].st;
  Status s;
  s.done = st->done;
  s.hold = st->hold;
  s.iter = st->iter;
  s.converged = st->converged;
  s.second = (st->hold && st->second) ? 1 : 0;
  s.nsecond = st->nsecond;
  s.pad[0] = s.pad[1] = 0;
  out[i] = s;
  st->nsecond = 0;
}

// Release the problems listed whose hold is a pending parabolic step (found
// in a block without second-step kernels), so that the step can be
// launched.  Any other hold (the end of the mixed bulk phase) stays, and a
// problem whose step was already taken is left as it is, so a request
// repeated from a stale status does nothing.
__global__ void k_release(const Prob* probs, const int* plist, int np) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  State* st = probs[plist[i]].st;
  if (st->hold && st->second && !st->done) {
    st->hold = 0;
    st->nheld += 1;
  }
}

// Final state and largest unitarity deviation of each listed problem (one
// workgroup per problem; run after k_unitarity).
__global__ void __launch_bounds__(W90_NT) k_gather_final(const Prob* probs, const int* plist, int np,
                                                         Final* out) {
  if (blockIdx.x >= (unsigned)np) return;
  const Prob& p = probs[plist[blockIdx.x]];
  __shared__ double red[W90_NT];
  double mx = 0.0;
  for (int k = threadIdx.x; k < p.nk; k += blockDim.x) {
    const double v = p.unit[k];
    mx = (v != v) ? INFINITY : fmax(mx, v);
  }
  mx = block_max(mx, red);
  if (threadIdx.x == 0) {
    out[blockIdx.x].st = *p.st;
    out[blockIdx.x].unit_max = mx;
  }
}

// ------------------------------------------------- omega partial vectors

// Per-k partial sums of wann_omega.  For Wannier function n, over the b of
// one k:
//   [q*nw + n]     S_rave[q] = sum_b wb b_q ln_n              (q = 0,1,2)
//   [3*nw + n]     S_r2      = sum_b wb (1 - |M_nn|^2 + ln_n^2)
//   [4*nw + n]     A         = sum_b wb (ln_n + b.r0_n)^2
//   [(5+q)*nw + n] V[q]      = sum_b wb b_q (ln_n + b.r0_n)
// and the scalars
//   [8*nw + 0]     S_od      = sum_b wb sum_{m!=n} |M_mn|^2
//   [8*nw + 1]     S_i       = sum_b wb (nw - sum_mn |M_mn|^2)
//   [8*nw + 2..7]  B         = sum_b wb b b^T  (xx, yy, zz, xy, yz, zx)
// with ln_n = Im ln M_nn and r0 the centres of the source buffer.  With
// d = rave - r0, Omega_D = (1/nk) sum_n [A + 2 d.V + d^T B d] is exact
// algebra for (1/nk) sum wb (ln + b.rave)^2 (wannierise.F90:2260-2275), and
// r0 close to rave keeps it free of cancellation, so no second pass over ln
// is needed once rave is known.
struct DiagAcc {
  double r[3], r2, a, v[3];
};

__device__ __forceinline__ void diag_add(DiagAcc& d, double w, const double* bv, double l, double d2,
                                         double r0x, double r0y, double r0z) {
  d.r[0] += w * bv[0] * l;
  d.r[1] += w * bv[1] * l;
  d.r[2] += w * bv[2] * l;
  d.r2 += w * (1.0 - d2 + l * l);
  const double x = l + (bv[0] * r0x + bv[1] * r0y + bv[2] * r0z);
  d.a += w * x * x;
  d.v[0] += w * bv[0] * x;
  d.v[1] += w * bv[1] * x;
  d.v[2] += w * bv[2] * x;
}

__device__ __forceinline__ void diag_store(double* out, int nw, int n, const DiagAcc& d) {
  out[0 * nw + n] = d.r[0];
  out[1 * nw + n] = d.r[1];
  out[2 * nw + n] = d.r[2];
  out[3 * nw + n] = d.r2;
  out[4 * nw + n] = d.a;
  out[5 * nw + n] = d.v[0];
  out[6 * nw + n] = d.v[1];
  out[7 * nw + n] = d.v[2];
}

// Scalars of the per-k vector that depend only on the k-mesh: B = sum wb b b^T
// from the weights wb[b] and the vectors bk[3 b + i] of the k.
__device__ void kpart_bmat(const double* wb, const double* bk, int nn, double* out) {
  double b6[6] = {0, 0, 0, 0, 0, 0};
  for (int b = 0; b < nn; ++b) {
    const double w = wb[b];
    const double* bv = bk + b * 3;
    b6[0] += w * bv[0] * bv[0];
    b6[1] += w * bv[1] * bv[1];
    b6[2] += w * bv[2] * bv[2];
    b6[3] += w * bv[0] * bv[1];
    b6[4] += w * bv[1] * bv[2];
    b6[5] += w * bv[2] * bv[0];
  }
  for (int u = 0; u < 6; ++u) out[2 + u] = b6[u];
}

// The weights wb and the vectors bk(:,:,k) of one k staged in LDS by
// threads 0 .. 4 nn - 1 (one round of loads instead of a chain of dependent
// ones inside the loops over b); visible after the next barrier.
__device__ __forceinline__ void kmesh_stage(const Prob& p, int k, double* wbL, double* bkL) {
  const int t = threadIdx.x, nn = p.nn;
  if (t < nn) wbL[t] = p.wb[t];
  else if (t < 4 * nn) bkL[t - nn] = p.bk[(size_t)k * nn * 3 + (t - nn)];
}

// Reference centres r0 of the partial vectors: the centres of the source
// buffer (zero for MODE_INIT, where no earlier centres exist).
__device__ __forceinline__ void ref_centre(const Prob& p, int mode, int cur, int n, double* r0) {
  if (mode == MODE_INIT) {
    r0[0] = r0[1] = r0[2] = 0.0;
  } else {
    r0[0] = p.rave[cur][0 + 3 * n];
    r0[1] = p.rave[cur][1 + 3 * n];
    r0[2] = p.rave[cur][2 + 3 * n];
  }
}

// Source block of a rotation: M(k,b) of the current buffer, rotated by
// E = exp(dW); for MODE_REMIX the initial M0(k,b), rotated by
// E = U0^+ U (k_delta_u), since M0 is the overlap in the gauge of U0.
template <class T2>
__device__ __forceinline__ const T2* rot_src(const Prob& p, int mode, int cur) {
  if constexpr (std::is_same<T2, cplx>::value) {
    if (mode == MODE_REMIX) return p.M0;
  }
  return mbuf<T2>(p, cur);
}
__device__ __forceinline__ const cplx* rot_factor(const Prob& p, int, int) { return p.E; }

// --------------------------------------------------------- M rotation

// Tiles per block M(k,b) of the MFMA rotation at most: 32 x 32 panels over
// 256 bands.
#define W90_MAX_PANELS 64
static_assert(W90GPU_MAX_WANN <= W90_NT, "k_kpart: one thread per Wannier function");

// Per-k partial vector from the per-tile outputs of the MFMA rotation
// (num_wann above the LDS kernels): diagonal phases and |M_nn|^2, and the
// tile sums of |M_mn|^2.
__global__ void __launch_bounds__(W90_NT) k_kpart(const Prob* probs, const Item* items, int nitems, int mode) {
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const State* st = p.st;
  const int cur = gate_cur(st, mode);
  if (cur < 0) return;
  const int nw = p.nw, nn = p.nn;
  const int k = it.k;
  const int dst = dst_buf(mode, cur);
  const double* ln = p.ln[dst];
  double* out = p.kpart + (size_t)k * p.psz;
  const int npan = p.npan;
  // every input of the k is loaded in one round: the mesh tables and the
  // tile sums into LDS, the diagonal phases and moduli into registers
  __shared__ double wbL[W90GPU_NUM_NNMAX], bkL[3 * W90GPU_NUM_NNMAX];
  __shared__ double podL[W90GPU_NUM_NNMAX * W90_MAX_PANELS], pallL[W90GPU_NUM_NNMAX * W90_MAX_PANELS];
  kmesh_stage(p, k, wbL, bkL);
  for (int t = threadIdx.x; t < nn * npan; t += blockDim.x) {
    podL[t] = p.part_od[(size_t)k * nn * npan + t];
    pallL[t] = p.part_all[(size_t)k * nn * npan + t];
  }
  double lnv[W90GPU_NUM_NNMAX], d2v[W90GPU_NUM_NNMAX], r0[3];
  const int n0 = threadIdx.x;
  if (n0 < nw) {
    ref_centre(p, mode, cur, n0, r0);
#pragma unroll
    for (int b = 0; b < W90GPU_NUM_NNMAX; ++b) {
      if (b < nn) {
        const size_t kb = (size_t)k * nn + b;
        lnv[b] = ln[kb * nw + n0];
        d2v[b] = p.d2[kb * nw + n0];
      }
    }
  }
  __syncthreads();
  if (n0 < nw) {
    DiagAcc d = {};
#pragma unroll
    for (int b = 0; b < W90GPU_NUM_NNMAX; ++b)
      if (b < nn) diag_add(d, wbL[b], bkL + b * 3, lnv[b], d2v[b], r0[0], r0[1], r0[2]);
    diag_store(out, nw, n0, d);
  }
  if (threadIdx.x == 0) {
    double sod = 0.0, si = 0.0;
    for (int b = 0; b < nn; ++b) {
      double od = 0.0, al = 0.0;
      for (int q = 0; q < npan; ++q) {
        od += podL[b * npan + q];
        al += pallL[b * npan + q];
      }
      sod += wbL[b] * od;
      si += wbL[b] * ((double)nw - al);
    }
    out[8 * nw + 0] = sod;
    out[8 * nw + 1] = si;
    kpart_bmat(wbL, bkL, nn, out + 8 * nw);
  }
}

// C = A B (with ident: I + scale A B) of n x n column-major complex
// matrices in LDS with leading dimension n + 1 (n <= NWM, NWM a multiple of
// 16), on FP64 MFMA tiles; defined with the other matrix-core code.
template <int NWM>
__device__ void lds_gemm_mfma(cplx* C, const cplx* A, const cplx* B, int n, double scale = 1.0,
                              bool ident = false);

// ---------------------------------------------------------- omega / control

// internal_test_convergence (wannierise.F90:1076-1164) with the noise branch
// off (conv_noise_amp <= 0), then the loop-exit bookkeeping.
__device__ void conv_step(State* st, double* logrow, double gc) {
  const int iter = st->iter + 1;
  const double delta = st->wann.om_tot - st->old.om_tot;
  st->iter = iter;
  if (logrow) {
    logrow[0] = st->wann.om_tot;
    logrow[1] = st->wann.om_d;
    logrow[2] = st->wann.om_od;
    logrow[3] = gc;
  }
  // internal_test_convergence (wannierise.F90:1076-1164) keeps the last
  // conv_window changes of Omega and stops when none exceeds conv_tol after
  // conv_window tests: the same as a run of conv_window small changes in a
  // row, which needs no history of any length ("not greater", so a NaN
  // change counts as small, as it does there)
  if (st->conv_window > 1) {
    st->conv_count += 1;
    st->conv_run = (fabs(delta) > st->conv_tol) ? 0 : st->conv_run + 1;
    if (st->conv_count >= st->conv_window && st->conv_run >= st->conv_window) st->converged = 1;
  }
  if (st->converged || iter >= st->num_iter) st->done = 1;
}

// Acceptance of a step with spread sp into buffer dst (wannierise.F90
// :660-701), then the convergence test.  In the mixed bulk phase a step is
// taken only if it lowers Omega by more than the FP32 noise floor of Omega
// (W90_MIXED_RTOL relative, at least conv_tol), so a noise-level line
// search is never followed.  A step that does not is rejected and the
// conjugate gradient restarts from steepest descent; a second rejection in
// a row -- or the end of the iteration budget -- ends the bulk phase: the
// problem holds at its current point for the FP64 rebuild of M, from which
// the FP64 loop continues.
__device__ void accept_step(State* st, const Spread& sp, int dst, double* logrow) {
  if (st->phase == 0) {
    const double floor_tol = fmax(st->conv_tol, W90_MIXED_RTOL * fabs(st->wann.om_tot));
    const bool progress = sp.om_tot - st->wann.om_tot < -floor_tol;
    if (progress) {
      st->old = st->wann;
      st->wann = sp;
      st->cur = dst;
      st->stall = 0;
    } else {
      st->stall += 1;
    }
    st->iter += 1;
    st->bulk_iter = st->iter;
    if (!progress) st->ncg_s[st->iter & 1] = st->num_cg_steps;  // next search restarts CG
    if (logrow) {
      logrow[0] = st->wann.om_tot;
      logrow[1] = st->wann.om_d;
      logrow[2] = st->wann.om_od;
      logrow[3] = st->gcnorm1;
    }
    if (st->stall >= W90_BULK_STALLS || st->iter >= st->num_iter) st->hold = 1;
    return;
  }
  st->old = st->wann;
  st->wann = sp;
  st->cur = dst;
  conv_step(st, logrow, st->gcnorm1);
}

// First level of the omega reduction: workgroup (problem, g) sums the per-k
// vectors of a contiguous k-range, in k order (rows read coalesced, 16 rows
// loaded at once and then added in order).
__global__ void __launch_bounds__(W90_NT) k_omega_group(const Prob* probs, const int* plist, int np, int mode) {
  const int pi = blockIdx.x / W90_MAX_GROUPS;
  const int g = blockIdx.x % W90_MAX_GROUPS;
  if (pi >= np) return;
  const Prob& p = probs[plist[pi]];
  if (g >= p.ngroup || !gate(p.st, mode)) return;
  const int G = p.ngroup, psz = p.psz;
  const int k0 = (int)(((long)g * p.nk) / G), k1 = (int)(((long)(g + 1) * p.nk) / G);
  for (int v = threadIdx.x; v < psz; v += blockDim.x) {
    double s = 0.0;
    // chunks of 16 rows: the loads of a chunk are issued together (rows
    // past k1 read as zero), then added in row order
    for (int k = k0; k < k1; k += 16) {
      double x[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) x[u] = (k + u < k1) ? p.kpart[(size_t)(k + u) * psz + v] : 0.0;
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (k + u < k1) s += x[u];
    }
    p.gpart[(size_t)g * psz + v] = s;
  }
}

// Final reduction of wann_omega (one workgroup per problem) fused with the
// control logic that follows each wann_omega call in wann_main: initial
// state (:452-470), internal_optimal_step (:1528-1596) after the trial step
// and the acceptance of the trial or parabolic step (:660-701).  MODE_REMIX
// installs the FP64 rebuild of M at the end of the mixed bulk phase.
// Control logic of wann_main after a wann_omega evaluation, run by one
// thread on the workgroup's copy of the state: the initial state
// (:452-470), internal_optimal_step (:1528-1596) after the trial step, the
// acceptance of the trial or parabolic step (:660-701), and the FP64
// rebuild at the end of the mixed bulk phase.  sod, si, omd: the summed
// off-diagonal, invariant and Omega_D numerators.  lean: the launch
// sequence has no second-step kernels after this one, so a parabolic step
// holds the problem for the host.
__device__ void omega_control(State* st, const Prob& p, double sod, double si, double omd, int mode, int dst,
                              int lean) {
  const double rnk = (double)p.nk;
  Spread sp;
  sp.om_od = sod / rnk;
  sp.om_d = omd / rnk;
  if (mode == MODE_INIT && st->first_pass) {
    sp.om_i = si / rnk;
  } else {
    sp.om_i = st->omega_invariant;
  }
  sp.om_tot = sp.om_i + sp.om_d + sp.om_od;

  if (mode == MODE_INIT) {
    st->omega_invariant = sp.om_i;
    st->first_pass = 0;
    st->wann = sp;
    st->old.om_tot = 0.0;
    if (p.log) {
      p.log[0] = sp.om_tot;
      p.log[1] = sp.om_d;
      p.log[2] = sp.om_od;
      p.log[3] = 0.0;
    }
    if (st->num_iter <= 0) st->done = 1;
    return;
  }
  if (mode == MODE_REMIX) {
    // FP64 state from the rebuilt M; the FP64 loop restarts its conjugate
    // gradient and its convergence window from here
    st->wann = sp;
    st->cur = dst;
    st->phase = 1;
    st->hold = 0;
    st->ncg_s[st->iter & 1] = st->num_cg_steps;
    st->conv_count = 0;
    st->conv_run = 0;
    if (p.log) {
      double* row = p.log + 4 * (size_t)st->iter;
      row[0] = sp.om_tot;
      row[1] = sp.om_d;
      row[2] = sp.om_od;
    }
    if (st->iter >= st->num_iter) st->done = 1;
    return;
  }
  double* logrow = p.log ? p.log + 4 * (size_t)(st->iter + 1) : nullptr;
  if (mode == MODE_TRIAL) {
    st->trial = sp;
    if (st->lfixstep) {
      accept_step(st, sp, dst, logrow);
      return;
    }
    // internal_optimal_step
    const double trial_step = st->trial_step;
    double fac = st->trial.om_tot - st->wann.om_tot;
    double shift;
    if (fabs(fac) > 2.2250738585072014e-308) {  // tiny(1.0_dp)
      fac = 1.0 / fac;
      shift = 1.0;
    } else {
      fac = 1.0e6;
      shift = fac * st->trial.om_tot - fac * st->wann.om_tot;
    }
    const double eqb = fac * st->doda0;
    const double eqa = shift - eqb * trial_step;
    double alphamin, falphamin;
    int lquad;
    if (fabs(eqa / (fac * st->wann.om_tot)) > 2.220446049250313e-16) {  // epsilon(1.0_dp)
      lquad = 1;
      alphamin = -0.5 * eqb / eqa * (trial_step * trial_step);
      falphamin = st->wann.om_tot - 0.25 * eqb * eqb / (fac * eqa) * (trial_step * trial_step);
    } else {
      lquad = 0;
      alphamin = trial_step;
      falphamin = st->trial.om_tot;
    }
    if (st->doda0 * alphamin > 0.0) {
      lquad = 0;
      alphamin = trial_step;
      falphamin = st->trial.om_tot;
    }
    st->alphamin = alphamin;
    st->falphamin = falphamin;
    st->lquad = lquad;
    if (lquad) {
      st->second = 1;
      st->nsecond += 1;
      if (lean) st->hold = 1;
      st->scale = alphamin / (4.0 * st->wbtot);
    } else {
      accept_step(st, st->trial, dst, logrow);
    }
    return;
  }
  // MODE_SECOND: the parabolic step replaced the trial step
  st->second = 0;
  accept_step(st, sp, dst, logrow);
}

__global__ void __launch_bounds__(W90_FINAL_NT) k_omega_final(const Prob* probs, const int* plist, int np,
                                                              int mode, int lean) {
  if (blockIdx.x >= (unsigned)np) return;
  const Prob& p = probs[plist[blockIdx.x]];
  // the state is worked on in LDS: the control logic below is a chain of
  // dependent reads and writes that would each pay a global-memory latency.
  // Its words are requested first and stored to LDS after the row sums
  // below have issued their loads, so the two rounds of loads overlap (the
  // sums of a problem that the gate then skips are not used).
  __shared__ State sst;
  static_assert(sizeof(State) % sizeof(double) == 0, "State is copied as doubles");
  constexpr int NSD = sizeof(State) / sizeof(double);
  static_assert(NSD <= W90_FINAL_NT, "one state word per thread");
  const int tid = threadIdx.x;
  const double sword = tid < NSD ? reinterpret_cast<const double*>(p.st)[tid] : 0.0;
  const int nw = p.nw, nk = p.nk, psz = p.psz;
  const double rnk = (double)nk;
  __shared__ double red[W90_FINAL_NT];
  __shared__ double sums[8 * W90GPU_MAX_WANN + 8];
  __shared__ double sc[8];

  // sum of the group vectors (or, for problems without the group level,
  // of the per-k vectors): VB values per pass (coalesced along a row), J
  // lanes over rows, lanes combined in fixed order
  const int VB = min(psz, W90_FINAL_NT);
  const int J = W90_FINAL_NT / VB;
  const int G = p.ngroup > 0 ? p.ngroup : nk;
  const double* rows = p.ngroup > 0 ? p.gpart : p.kpart;
  for (int v0 = 0; v0 < psz; v0 += VB) {
    const int v = v0 + tid % VB;
    const int j = tid / VB;
    double s = 0.0;
    if (j < J && v < psz) {
      // rows j, j+J, ...: eight independent partial sums, combined in a
      // fixed order; four rounds of eight rows are loaded at once (32 loads
      // in flight) and added round by round, the same order as one round
      // at a time
      double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      int g = j;
      for (; g + 31 * J < G; g += 32 * J) {
        double x[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) x[u] = rows[(size_t)(g + u * J) * psz + v];
#pragma unroll
        for (int u = 0; u < 32; ++u) a[u % 8] += x[u];
      }
      for (; g + 7 * J < G; g += 8 * J) {
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] += rows[(size_t)(g + u * J) * psz + v];
      }
      for (int u = 0; g < G; g += J, ++u) a[u] += rows[(size_t)g * psz + v];
      s = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    }
    red[tid] = s;
    __syncthreads();
    if (tid < VB && v0 + tid < psz) {
      double t = 0.0;
      for (int jj = 0; jj < J; ++jj) t += red[tid + jj * VB];
      sums[v0 + tid] = t;
    }
    __syncthreads();
  }
  if (tid < NSD) reinterpret_cast<double*>(&sst)[tid] = sword;
  __syncthreads();
  State* st = &sst;
  const int cur = gate_cur(st, mode);
  if (cur < 0) return;
  const int dst = dst_buf(mode, cur);

  double omd_n = 0.0;
  for (int n = tid; n < nw; n += blockDim.x) {
    double q[8];
    for (int u = 0; u < 8; ++u) q[u] = sums[u * nw + n];
    double bb[6];
    for (int u = 0; u < 6; ++u) bb[u] = sums[8 * nw + 2 + u];
    // reference centres are read before rave[dst] is written (dst == cur
    // for the rebuild)
    double r0[3];
    ref_centre(p, mode, cur, n, r0);
    const double rx = -q[0] / rnk, ry = -q[1] / rnk, rz = -q[2] / rnk;
    p.rave[dst][0 + 3 * n] = rx;
    p.rave[dst][1 + 3 * n] = ry;
    p.rave[dst][2 + 3 * n] = rz;
    p.r2ave[dst][n] = q[3] / rnk;
    const double dx = rx - r0[0], dy = ry - r0[1], dz = rz - r0[2];
    const double dbd = bb[0] * dx * dx + bb[1] * dy * dy + bb[2] * dz * dz +
                       2.0 * (bb[3] * dx * dy + bb[4] * dy * dz + bb[5] * dz * dx);
    omd_n += q[4] + 2.0 * (dx * q[5] + dy * q[6] + dz * q[7]) + dbd;
  }
  // Omega_D summed over n in a fixed tree
  const double omd = block_sum(omd_n, red);
  if (threadIdx.x == 0) {
    sc[0] = sums[8 * nw + 0];
    sc[1] = sums[8 * nw + 1];
  }
  __syncthreads();
  if (mode == MODE_BENCH) return;  // timing of the reduction alone: the state stays as it is
  if (threadIdx.x == 0) omega_control(st, p, sc[0], sc[1], omd, mode, dst, lean);
  __syncthreads();
  for (int i = threadIdx.x; i < NSD; i += blockDim.x)
    reinterpret_cast<double*>(p.st)[i] = reinterpret_cast<const double*>(&sst)[i];
}


// ---------------------------------------------------------- gradient

// One (m, n) term of the non-SLWF gradient of wann_domega
// (wannierise.F90:2510-2528) for one block: A[R] - S[T] with
// R_mn = M_mn conj(M_nn), Rt_mn = M_mn / M_nn, qn = wb Im ln M_nn and
// rn = b.rave_n.
__device__ __forceinline__ cplx domega_term(cplx acc, double w, cplx mmn, cplx mnm, cplx mnn, cplx mmm, double qn,
                                            double qm, double rn, double rm) {
  const cplx mhalf_i = c_make(0.0, -0.5);
  const cplx cr_mn = c_mul(mmn, c_conj(mnn));
  const cplx cr_nm = c_mul(mnm, c_conj(mmm));
  const cplx crt_mn = c_div(mmn, mnn);
  const cplx crt_nm = c_div(mnm, mmm);
  // A[R] = (R - R^+)/2
  acc = c_add(acc, c_scale(c_sub(cr_mn, c_conj(cr_nm)), w * 0.5));
  // -S[T] = -(T + T^+)/2i, T_mn = Rt_mn q_n
  acc = c_sub(acc, c_mul(c_add(c_scale(crt_mn, qn), c_conj(c_scale(crt_nm, qm))), mhalf_i));
  acc = c_sub(acc, c_scale(c_mul(c_add(c_scale(crt_mn, rn), c_conj(c_scale(crt_nm, rm))), mhalf_i), w));
  return acc;
}

// The same term with the divisions taken out: rn = 1/M_nn and
// tn = wb Im ln M_nn + wb b.rave_n precomputed per (b, n), so that
// T_mn = M_mn rn tn.
__device__ __forceinline__ cplx domega_term_r(cplx acc, double w, cplx mmn, cplx mnm, cplx mnn, cplx mmm, cplx rn,
                                              cplx rm, double tn, double tm) {
  const cplx mhalf_i = c_make(0.0, -0.5);
  const cplx cr_mn = c_mul(mmn, c_conj(mnn));
  const cplx cr_nm = c_mul(mnm, c_conj(mmm));
  // A[R] = (R - R^+)/2
  acc = c_add(acc, c_scale(c_sub(cr_mn, c_conj(cr_nm)), w * 0.5));
  // -S[T] = -(T + T^+)/2i
  const cplx t_mn = c_scale(c_mul(mmn, rn), tn);
  const cplx t_nm = c_scale(c_mul(mnm, rm), tm);
  return c_sub(acc, c_mul(c_add(t_mn, c_conj(t_nm)), mhalf_i));
}

// wann_domega (wannierise.F90:2287-2569), non-SLWF gradient, above 64
// bands: one workgroup of 1024 threads per (problem, k), fused with the
// per-k partial sums Re<G,G> and Re<G,D_old> needed by
// internal_search_direction.  Element (m, n) needs M_mn and M_nm, so the
// block is taken in pairs of 32 x 32 tiles (I, J) and (J, I), I <= J: for
// each b both tiles are staged in LDS from coalesced loads (the next b's
// requested while this one is used), and each thread forms one element of
// each tile, its sum over b in b order.  The diagonals M_nn and the phases
// w_b Im ln M_nn of every b are staged once per k.  G is stored, and the
// per-k sums are then formed over G in element order e = t + 1024 i per
// thread, which makes them independent of the tiling.
constexpr int kDtTile = 32;
template <class T2>
__host__ __device__ constexpr size_t domega_t_lds() {
  return (sizeof(T2) + sizeof(double)) * W90GPU_NUM_NNMAX * W90GPU_MAX_WANN +
         sizeof(T2) * 2 * kDtTile * (kDtTile + 1);
}
template <class T2>
__global__ void __launch_bounds__(1024) k_domega(const Prob* probs, const Item* items, int nitems, int mode) {
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const State* st = p.st;
  const int cur = gate_cur(st, (mode == MODE_BENCH ? MODE_BENCH : MODE_TRIAL));
  if (cur < 0) return;
  constexpr int TS = kDtTile, LD = kDtTile + 1;
  const int nw = p.nw, nn = p.nn, nw2 = p.nw2;
  const int k = it.k, tid = threadIdx.x;
  const int lr = tid % TS, lc = tid / TS;
  const double* ln = p.ln[cur];
  const double* rave = p.rave[cur];
  const T2* Mk = mbuf<T2>(p, cur) + (size_t)k * nn * nw2;
  extern __shared__ double lds_dt[];
  T2* ta = reinterpret_cast<T2*>(lds_dt);                // tile (I, J), column-major
  T2* tb = ta + TS * LD;                                  // tile (J, I)
  T2* dg = tb + TS * LD;                                  // [b][n] M_nn
  double* qv = reinterpret_cast<double*>(dg + nn * nw);  // [b][n] w_b Im ln M_nn
  __shared__ double wbL[W90GPU_NUM_NNMAX], bkL[3 * W90GPU_NUM_NNMAX];
  __shared__ double red[1024];
  kmesh_stage(p, k, wbL, bkL);
  for (int t = tid; t < nn * nw; t += blockDim.x) {
    const int b = t / nw, n = t % nw;
    dg[t] = Mk[(size_t)b * nw2 + n + (size_t)n * nw];
    qv[t] = p.wb[b] * ln[((size_t)k * nn + b) * nw + n];
  }
  const double rnk = (double)p.nk;
  cplx* Gk = p.G + (size_t)k * nw2;
  const int nt = (nw + TS - 1) / TS;
  for (int J = 0, I = 0; J < nt;) {
    const bool diag = I == J;
    const int m1 = I * TS + lr, n1 = J * TS + lc;  // element of tile (I, J)
    const int m2 = J * TS + lr, n2 = I * TS + lc;  // element of tile (J, I)
    const bool v1 = m1 < nw && n1 < nw, v2 = !diag && m2 < nw && n2 < nw;
    T2 xa = {}, xb = {};
    if (v1) xa = Mk[m1 + (size_t)n1 * nw];
    if (v2) xb = Mk[m2 + (size_t)n2 * nw];
    cplx acc1 = c_make(0.0, 0.0), acc2 = acc1;
    for (int b = 0; b < nn; ++b) {
      __syncthreads();  // the tiles of b - 1 are used (and, for b = 0, the staging above is done)
      ta[lr + lc * LD] = xa;
      tb[lr + lc * LD] = xb;
      if (b + 1 < nn) {
        const T2* Mb1 = Mk + (size_t)(b + 1) * nw2;
        if (v1) xa = Mb1[m1 + (size_t)n1 * nw];
        if (v2) xb = Mb1[m2 + (size_t)n2 * nw];
      }
      __syncthreads();
      const double w = wbL[b];
      const double* bv = bkL + 3 * b;
      const T2* dgb = dg + b * nw;
      const double* qb = qv + b * nw;
      if (v1) {
        const double rn = bv[0] * rave[0 + 3 * n1] + bv[1] * rave[1 + 3 * n1] + bv[2] * rave[2 + 3 * n1];
        const double rm = bv[0] * rave[0 + 3 * m1] + bv[1] * rave[1 + 3 * m1] + bv[2] * rave[2 + 3 * m1];
        acc1 = domega_term(acc1, w, c_wide(ta[lr + lc * LD]), c_wide((diag ? ta : tb)[lc + lr * LD]),
                           c_wide(dgb[n1]), c_wide(dgb[m1]), qb[n1], qb[m1], rn, rm);
      }
      if (v2) {
        const double rn = bv[0] * rave[0 + 3 * n2] + bv[1] * rave[1 + 3 * n2] + bv[2] * rave[2 + 3 * n2];
        const double rm = bv[0] * rave[0 + 3 * m2] + bv[1] * rave[1 + 3 * m2] + bv[2] * rave[2 + 3 * m2];
        acc2 = domega_term(acc2, w, c_wide(tb[lr + lc * LD]), c_wide(ta[lc + lr * LD]), c_wide(dgb[n2]),
                           c_wide(dgb[m2]), qb[n2], qb[m2], rn, rm);
      }
    }
    if (v1) Gk[m1 + (size_t)n1 * nw] = c_scale(make_double2(acc1.x / rnk, acc1.y / rnk), 4.0);
    if (v2) Gk[m2 + (size_t)n2 * nw] = c_scale(make_double2(acc2.x / rnk, acc2.y / rnk), 4.0);
    if (++I > J) {
      ++J;
      I = 0;
    }
  }
  __syncthreads();  // G of the k is stored
  double gg = 0.0, gd = 0.0;
  for (int e = tid; e < nw2; e += blockDim.x) {
    const cplx g = Gk[e];
    gg += c_abs2(g);
    const cplx dold = p.D[(size_t)k * nw2 + e];
    gd += g.x * dold.x + g.y * dold.y;
  }
  block_sum2(gg, gd, red);
  if (tid == 0) {
    p.part_gg[k] = gg;
    p.part_gd[k] = gd;
  }
}

// Threads of k_domega_k: enough wavefronts per k to keep the block loads
// of the larger num_wann in flight.
template <int NWM>
__host__ __device__ constexpr int domega_nt() {
  return NWM > 32 ? 1024 : W90_NT;
}

// Blocks M(k,b) that k_domega_k holds in LDS at once: all of them for
// num_wann <= 16, four for <= 32, one above.
template <int NWM>
__host__ __device__ constexpr int domega_chunk() {
  return NWM <= 16 ? W90GPU_NUM_NNMAX : (NWM <= 32 ? 4 : 1);
}

// LDS bytes of k_domega_k (dynamic shared memory).
template <int NWM, class T2>
__host__ __device__ constexpr size_t domega_lds() {
  return sizeof(T2) * (size_t)NWM * (NWM + 1) * domega_chunk<NWM>() +
         sizeof(double) * (3 * W90GPU_NUM_NNMAX * NWM + domega_nt<NWM>());
}

// wann_domega with the blocks M(k,b) staged in LDS (num_wann <= NWM): the
// transposed and diagonal reads of the gradient formula hit LDS (padded
// leading dimension), and per (b, n) the phase wb (Im ln M_nn + b.rave_n)
// and the reciprocal 1/M_nn are precomputed (no division per element).  The blocks are loaded in chunks of
// domega_chunk<NWM>() (every load of a chunk in flight at once), the next
// chunk into registers while the current one is used; the sum over b runs
// in b order whatever the chunking.
template <int NWM, class T2>
__global__ void __launch_bounds__(domega_nt<NWM>()) k_domega_k(const Prob* probs, const Item* items, int nitems,
                                                                int mode) {
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const State* st = p.st;
  const int cur = gate_cur(st, (mode == MODE_BENCH ? MODE_BENCH : MODE_TRIAL));
  if (cur < 0) return;
  constexpr int NT = domega_nt<NWM>();
  constexpr int SL = (NWM * NWM + NT - 1) / NT;
  constexpr int CB = domega_chunk<NWM>();
  const int nw = p.nw, nn = p.nn, nw2 = p.nw2;
  const int ld = nw + 1;
  const int k = it.k;
  const int tid = threadIdx.x;
  extern __shared__ double lds_domega[];
  double* qt = lds_domega;                    // [b][n] wb (Im ln M_nn + b.rave_n)
  cplx* rec = reinterpret_cast<cplx*>(qt + W90GPU_NUM_NNMAX * NWM);  // [b][n] 1/M_nn
  double* red = qt + 3 * W90GPU_NUM_NNMAX * NWM;
  T2* Ms = reinterpret_cast<T2*>(red + NT);
  const double* rave = p.rave[cur];
  const T2* Mk = mbuf<T2>(p, cur) + (size_t)k * nn * nw2;
  // the first chunk of blocks is requested before the per-(b, n) inputs
  // below, so both rounds of loads are in flight together
  T2 nxt[CB][SL];
#pragma unroll
  for (int j = 0; j < CB; ++j)
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      const int e = tid + s * NT;
      T2 v = {};
      if (j < nn && e < nw2) v = Mk[(size_t)j * nw2 + e];
      nxt[j][s] = v;
    }
  for (int t = tid; t < nn * nw; t += NT) {
    const int b = t / nw, n = t % nw;
    const size_t kb = (size_t)k * nn + b;
    const double* bv = p.bk + kb * 3;
    const double w = p.wb[b];
    const double rbn = bv[0] * rave[0 + 3 * n] + bv[1] * rave[1 + 3 * n] + bv[2] * rave[2 + 3 * n];
    qt[t] = w * p.ln[cur][kb * nw + n] + w * rbn;
    const cplx d = c_wide(Mk[(size_t)b * nw2 + n + (size_t)n * nw]);
    const double inv = 1.0 / c_abs2(d);
    rec[t] = make_double2(d.x * inv, -d.y * inv);
  }
  cplx acc[SL];
#pragma unroll
  for (int s = 0; s < SL; ++s) acc[s] = c_make(0.0, 0.0);
  __shared__ double wL[W90GPU_NUM_NNMAX];
  if (tid < nn) wL[tid] = p.wb[tid];
  const int bs = NWM * (NWM + 1);  // LDS stride of the chunk's blocks
  for (int b0 = 0; b0 < nn; b0 += CB) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CB; ++j)
#pragma unroll
      for (int s = 0; s < SL; ++s) {
        const int e = tid + s * NT;
        if (b0 + j < nn && e < nw2) Ms[j * bs + (e % nw) + (e / nw) * ld] = nxt[j][s];
      }
    if (b0 + CB < nn) {
      const T2* src = Mk + (size_t)(b0 + CB) * nw2;
#pragma unroll
      for (int j = 0; j < CB; ++j)
#pragma unroll
        for (int s = 0; s < SL; ++s) {
          const int e = tid + s * NT;
          T2 v = {};
          if (b0 + CB + j < nn && e < nw2) v = src[(size_t)j * nw2 + e];
          nxt[j][s] = v;
        }
    }
    __syncthreads();
    for (int j = 0; j < CB && b0 + j < nn; ++j) {
      const int b = b0 + j;
      const T2* Mb = Ms + j * bs;
      const double w = wL[b];
#pragma unroll
      for (int s = 0; s < SL; ++s) {
        const int e = tid + s * NT;
        if (e < nw2) {
          const int m = e % nw, n = e / nw;
          acc[s] = domega_term_r(acc[s], w, c_wide(Mb[m + n * ld]), c_wide(Mb[n + m * ld]), c_wide(Mb[n + n * ld]),
                                 c_wide(Mb[m + m * ld]), rec[b * nw + n], rec[b * nw + m], qt[b * nw + n],
                                 qt[b * nw + m]);
        }
      }
    }
  }
  const double rnk = (double)p.nk;
  double gg = 0.0, gd = 0.0;
#pragma unroll
  for (int s = 0; s < SL; ++s) {
    const int e = tid + s * NT;
    if (e < nw2) {
      const cplx g = c_scale(make_double2(acc[s].x / rnk, acc[s].y / rnk), 4.0);
      st_stream(&p.G[(size_t)k * nw2 + e], g);
      gg += c_abs2(g);
      const cplx dold = p.D[(size_t)k * nw2 + e];
      gd += g.x * dold.x + g.y * dold.y;
    }
  }
  block_sum2(gg, gd, red);
  if (tid == 0) {
    p.part_gg[k] = gg;
    p.part_gd[k] = gd;
  }
}

// wann_domega for 32 < num_wann <= 64: the formula and the sum over b of
// k_domega_k, with one block in LDS and the next two requested ahead into
// registers (two block loads in flight while a block is reduced).  Thread
// t takes row m = t % 64 and columns n = t / 64 + 16 s (s < 4) of the block
// padded to 64 x 64, so its row terms (M_mm, 1/M_mm, the phase of row m)
// are read once per b and its element offsets are one base and a uniform
// stride; with this layout the two blocks fit the register budget of 1024
// threads (three do not).  For num_wann = 64 the thread's elements are
// those of k_domega_k, so the per-k sums Re<G,G> and Re<G,D_old> are formed
// in the same order too.
constexpr int kDw64Ahead = 2;
template <class T2>
__global__ void __launch_bounds__(1024) k_domega_w64(const Prob* probs, const Item* items, int nitems, int mode) {
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const State* st = p.st;
  const int cur = gate_cur(st, (mode == MODE_BENCH ? MODE_BENCH : MODE_TRIAL));
  if (cur < 0) return;
  constexpr int NWM = 64, NT = 1024, SL = NWM * NWM / NT, RS = NT / NWM, LD = NWM + 1;
  constexpr int AH = kDw64Ahead;
  static_assert(domega_nt<NWM>() == NT && domega_chunk<NWM>() == 1, "LDS layout of domega_lds<64>");
  const int nw = p.nw, nn = p.nn, nw2 = p.nw2;
  const int k = it.k;
  const int tid = threadIdx.x;
  const int m = tid % NWM, n0 = tid / NWM;
  extern __shared__ double lds_domega[];
  double* qt = lds_domega;                                           // [b][n]
  cplx* rec = reinterpret_cast<cplx*>(qt + W90GPU_NUM_NNMAX * NWM);  // [b][n]
  double* red = qt + 3 * W90GPU_NUM_NNMAX * NWM;
  T2* Ms = reinterpret_cast<T2*>(red + NT);
  const double* rave = p.rave[cur];
  const T2* Mk = mbuf<T2>(p, cur) + (size_t)k * nn * nw2;
  // element s: (m, n0 + RS s), at g0 + s RS nw in the block
  const int g0 = m + n0 * nw, gs = RS * nw;
  bool ev[SL];
#pragma unroll
  for (int s = 0; s < SL; ++s) ev[s] = m < nw && n0 + RS * s < nw;
  T2 nxt[AH][SL];
#pragma unroll
  for (int j = 0; j < AH; ++j)
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      T2 v = {};
      if (j < nn && ev[s]) v = Mk[(size_t)j * nw2 + g0 + s * gs];
      nxt[j][s] = v;
    }
  for (int t = tid; t < nn * nw; t += NT) {
    const int b = t / nw, n = t % nw;
    const size_t kb = (size_t)k * nn + b;
    const double* bv = p.bk + kb * 3;
    const double w = p.wb[b];
    const double rbn = bv[0] * rave[0 + 3 * n] + bv[1] * rave[1 + 3 * n] + bv[2] * rave[2 + 3 * n];
    qt[t] = w * p.ln[cur][kb * nw + n] + w * rbn;
    const cplx d = c_wide(Mk[(size_t)b * nw2 + n + (size_t)n * nw]);
    const double inv = 1.0 / c_abs2(d);
    rec[t] = make_double2(d.x * inv, -d.y * inv);
  }
  cplx acc[SL];
#pragma unroll
  for (int s = 0; s < SL; ++s) acc[s] = c_make(0.0, 0.0);
  __shared__ double wL[W90GPU_NUM_NNMAX];
  if (tid < nn) wL[tid] = p.wb[tid];
  for (int b0 = 0; b0 < nn; b0 += AH) {
#pragma unroll
    for (int j = 0; j < AH; ++j) {
      const int b = b0 + j;
      if (b < nn) {
        __syncthreads();
#pragma unroll
        for (int s = 0; s < SL; ++s)
          if (ev[s]) Ms[m + (n0 + RS * s) * LD] = nxt[j][s];
        if (b + AH < nn) {
          const T2* src = Mk + (size_t)(b + AH) * nw2;
#pragma unroll
          for (int s = 0; s < SL; ++s) {
            T2 v = {};
            if (ev[s]) v = src[g0 + s * gs];
            nxt[j][s] = v;
          }
        }
        __syncthreads();
        if (m < nw) {
          const double w = wL[b];
          const cplx mmm = c_wide(Ms[m * (LD + 1)]);
          const cplx rm = rec[b * nw + m];
          const double tm = qt[b * nw + m];
#pragma unroll
          for (int s = 0; s < SL; ++s) {
            const int n = n0 + RS * s;
            if (n < nw)
              acc[s] = domega_term_r(acc[s], w, c_wide(Ms[m + n * LD]), c_wide(Ms[n + m * LD]),
                                     c_wide(Ms[n * (LD + 1)]), mmm, rec[b * nw + n], rm, qt[b * nw + n], tm);
          }
        }
      }
    }
  }
  const double rnk = (double)p.nk;
  double gg = 0.0, gd = 0.0;
#pragma unroll
  for (int s = 0; s < SL; ++s) {
    if (ev[s]) {
      const cplx g = c_scale(make_double2(acc[s].x / rnk, acc[s].y / rnk), 4.0);
      st_stream(&p.G[(size_t)k * nw2 + g0 + s * gs], g);
      gg += c_abs2(g);
      const cplx dold = p.D[(size_t)k * nw2 + g0 + s * gs];
      gd += g.x * dold.x + g.y * dold.y;
    }
  }
  block_sum2(gg, gd, red);
  if (tid == 0) {
    p.part_gg[k] = gg;
    p.part_gd[k] = gd;
  }
}

// Scalar part of internal_search_direction (wannierise.F90:1371-1525) on
// a state copy: gg = Re<G,G>, gd = Re<G,D_old>.  Every workgroup of the
// trial-step kernel runs it on its own copy; the CG counter and the previous
// gradient norm are read from the slot of this iteration's parity and
// written to the other one, so the workgroup that stores the results never
// changes what the others read.
__device__ void search_control(State* st, double gg, double gd) {
  const int iter = st->iter + 1;
  const int sl = st->iter & 1, nx = sl ^ 1;
  int ncg = st->ncg_s[sl];
  st->second = 0;
  st->gcnorm1 = gg;
  double gcfac;
  if (iter == 1 || ncg >= st->num_cg_steps) {
    gcfac = 0.0;
    ncg = 0;
  } else {
    if (st->gcnorm0_s[sl] > 2.220446049250313e-16) {
      gcfac = gg / st->gcnorm0_s[sl];
      if (gcfac > 3.0) {
        gcfac = 0.0;
        ncg = 0;
      } else {
        ncg += 1;
      }
    } else {
      gcfac = 0.0;
      ncg = 0;
    }
  }
  st->gcnorm0_s[nx] = gg;
  const double w4 = 4.0 * st->wbtot;
  double doda0 = -(gg + gcfac * gd) / w4;
  double beta = gcfac, sign = 1.0;
  if (doda0 > 0.0) {
    if (ncg > 0) {
      beta = 0.0;
      ncg = 0;
      gcfac = 0.0;
      doda0 = -gg / w4;
      if (doda0 > 0.0) {
        sign = -1.0;
        doda0 = -doda0;
      }
    } else {
      sign = -sign;
      doda0 = -doda0;
    }
  }
  st->ncg_s[nx] = ncg;
  st->gcfac = gcfac;
  st->beta = beta;
  st->dir_sign = sign;
  st->doda0 = doda0;
  st->scale = (st->lfixstep ? st->fixed_step : st->trial_step) / w4;
}

// internal_search_direction (wannierise.F90:1371-1525) without
// preconditioning or noise, fused into the trial-step kernels: each
// workgroup sums Re<G,G> and Re<G,D_old> over k in a fixed order (every
// workgroup of a problem gets the same sums), runs the scalar part on a copy
// of the state in LDS and forms its D from the result; the workgroup of
// k = 0 stores the results.  Called by all threads of the block; blockDim
// must be W90_NT (the summation order depends on it).
__device__ void search_in_step(const Prob& p, int k, State* lst, double* red) {
  constexpr int NSD = sizeof(State) / sizeof(double);
  // the partial sums are requested first, the state copy overlaps them;
  // the k range is split over W90_NT lanes whatever the workgroup size
  // (lanes beyond add exact zeros), so every retraction kernel forms the
  // same sums
  double gg = 0.0, gd = 0.0;
  const int nsplit = min((int)blockDim.x, W90_NT);
  const int per = (p.nk + nsplit - 1) / nsplit;
  const int kb = threadIdx.x * per, ke = min(p.nk, kb + per);
  // chunks of 4: the loads of a chunk are issued together, then added in
  // k order
  for (int q = kb; q < ke; q += 4) {
    double xg[4], xd[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xg[u] = (q + u < ke) ? p.part_gg[q + u] : 0.0;
      xd[u] = (q + u < ke) ? p.part_gd[q + u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (q + u < ke) {
        gg += xg[u];
        gd += xd[u];
      }
  }
  for (int i = threadIdx.x; i < NSD; i += blockDim.x)
    reinterpret_cast<double*>(lst)[i] = reinterpret_cast<const double*>(p.st)[i];
  block_sum2(gg, gd, red);  // (its barriers also order the copy above)
  if (threadIdx.x == 0) {
    search_control(lst, gg, gd);
    if (k == 0) {
      State* st = p.st;
      const int nx = (lst->iter & 1) ^ 1;
      st->second = lst->second;
      st->gcnorm1 = lst->gcnorm1;
      st->ncg_s[nx] = lst->ncg_s[nx];
      st->gcnorm0_s[nx] = lst->gcnorm0_s[nx];
      st->gcfac = lst->gcfac;
      st->beta = lst->beta;
      st->dir_sign = lst->dir_sign;
      st->doda0 = lst->doda0;
      st->scale = lst->scale;
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------- retraction

// Degree and squarings of the Taylor approximant of exp(X) for ||X||_1 = nrm:
// X is scaled by 2^-sq to norm <= 1/8, and the degree m is the smallest with
// (nrm/2^sq)^(m+1)/(m+1)! < 1e-18, far below the FP64 rounding of the
// result (the exponential of an anti-Hermitian matrix is unitary, norm 1).
__device__ __forceinline__ void taylor_plan(double nrm, int* sq_out, int* m_out) {
  if (!isfinite(nrm)) nrm = 0.0;  // a non-finite step yields a non-finite U
  int sq = 0;
  while (nrm > 0.125 && sq < 64) {
    nrm *= 0.5;
    ++sq;
  }
  int m = 1;
  double term = nrm;
  while (m < 24) {
    term = term * nrm / (double)(m + 1);
    if (term < 1e-18) break;
    ++m;
  }
  *sq_out = sq;
  *m_out = m;
}

// C = A * B for n x n column-major complex matrices held by one workgroup.
__device__ void wg_gemm(cplx* C, const cplx* A, const cplx* B, int n) {
  const int n2 = n * n;
  for (int e = threadIdx.x; e < n2; e += blockDim.x) {
    const int i = e % n, j = e / n;
    const cplx* bc = B + (size_t)j * n;
    cplx acc = c_make(0.0, 0.0);
    for (int l = 0; l < n; ++l) acc = c_add(acc, c_mul(A[i + (size_t)l * n], bc[l]));
    C[e] = acc;
  }
  __syncthreads();
}

// C = A^+ B for n x n column-major complex matrices held by one workgroup.
__device__ void wg_gemm_cn(cplx* C, const cplx* A, const cplx* B, int n) {
  const int n2 = n * n;
  for (int e = threadIdx.x; e < n2; e += blockDim.x) {
    const int i = e % n, j = e / n;
    const cplx* ac = A + (size_t)i * n;
    const cplx* bc = B + (size_t)j * n;
    cplx acc = c_make(0.0, 0.0);
    for (int l = 0; l < n; ++l) acc = c_add(acc, c_cmul(ac[l], bc[l]));
    C[e] = acc;
  }
  __syncthreads();
}

// ET(k) = E(k)^T for one k by one workgroup (E written by this workgroup
// before; the caller synchronised it): coalesced writes, reads along rows.
__device__ void wg_transpose_e(const Prob& p, int k) {
  const int nw = p.nw;
  const size_t o = (size_t)k * p.nw2;
  for (int e = threadIdx.x; e < p.nw2; e += blockDim.x) {
    const int c = e % nw, r = e / nw;
    p.ET[o + c + (size_t)r * nw] = p.E[o + r + (size_t)c * nw];
  }
}

__global__ void __launch_bounds__(W90_NT) k_delta_u(const Prob* probs, const Item* items, int nitems) {
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const State* st = p.st;
  if (!gate(st, MODE_REMIX)) return;
  const size_t o = (size_t)it.k * p.nw2;
  wg_gemm_cn(p.E + o, p.U0 + o, p.U[st->cur] + o, p.nw);
  wg_transpose_e(p, it.k);
}

// ET = E^T for every item (E formed outside the retraction kernels: the
// polar factor of the projection).
__global__ void __launch_bounds__(W90_NT) k_transpose_e(const Prob* probs, const Item* items, int nitems) {
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  wg_transpose_e(probs[it.prob], it.k);
}

// C = A * B (with ident: I + scale A B) on FP64 MFMA for matrices in global
// memory (w90gpu_mfma.h).
__device__ void wg_gemm_mfma(cplx* C, const cplx* A, const cplx* B, int n, double scale = 1.0,
                             bool ident = false);

// Retraction of one (problem, k) with workspaces in global memory (num_wann
// above the LDS kernels): forms the search direction D (trial step only), X = scale * D,
// E = exp(X) by scaling and squaring of a Taylor polynomial, and U_new = U E.
// X is anti-Hermitian, so exp(X) is the same unitary that
// internal_new_u_and_m builds from the eigenvectors of i X
// (wannierise.F90:1662-1689, 1729-1734).
__global__ void __launch_bounds__(W90_STEP_NT) k_step(const Prob* probs, const Item* items, int nitems, int mode) {
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const State* st = p.st;
  const int cur = gate_cur(st, mode);
  if (cur < 0) return;
  const int nw = p.nw, nw2 = p.nw2;
  const int k = it.k;
  __shared__ double red[W90_NT];
  cplx* Dk = p.D + (size_t)k * nw2;
  cplx* X = p.W1 + (size_t)k * nw2;
  cplx* P = p.W2 + (size_t)k * nw2;
  cplx* Ek = p.E + (size_t)k * nw2;
  __shared__ State lst;
  const State* ss = st;
  if (mode == MODE_TRIAL) {
    search_in_step(p, k, &lst, red);
    ss = &lst;
  }
  const double scale = ss->scale;
  if (mode == MODE_TRIAL) {
    const cplx* Gk = p.G + (size_t)k * nw2;
    const double beta = ss->beta, sign = ss->dir_sign;
    for (int e = threadIdx.x; e < nw2; e += blockDim.x) {
      cplx d = c_scale(c_add(Gk[e], c_scale(Dk[e], beta)), sign);
      Dk[e] = d;
      X[e] = c_scale(d, scale);
    }
  } else {
    for (int e = threadIdx.x; e < nw2; e += blockDim.x) X[e] = c_scale(Dk[e], scale);
  }
  __syncthreads();
  double cs = 0.0;
  for (int j = threadIdx.x; j < nw; j += blockDim.x) {
    double s = 0.0;
    for (int i = 0; i < nw; ++i) s += sqrt(c_abs2(X[i + (size_t)j * nw]));
    cs = fmax(cs, s);
  }
  int sq, m;
  taylor_plan(block_max(cs, red), &sq, &m);
  if (sq > 0) {
    const double f = ldexp(1.0, -sq);
    for (int e = threadIdx.x; e < nw2; e += blockDim.x) X[e] = c_scale(X[e], f);
  }
  __syncthreads();
  const double invm = 1.0 / (double)m;
  for (int e = threadIdx.x; e < nw2; e += blockDim.x) {
    cplx v = c_scale(X[e], invm);
    if (e % nw == e / nw) v.x += 1.0;
    P[e] = v;
  }
  __syncthreads();
  // Horner: P <- I + X P / j, the scale and the identity applied as each
  // product is stored; P and E take turns holding the polynomial
  cplx* pc = P;
  cplx* pn = Ek;
  for (int j = m - 1; j >= 1; --j) {
    wg_gemm_mfma(pn, X, pc, nw, 1.0 / (double)j, true);
    cplx* t = pc;
    pc = pn;
    pn = t;
  }
  cplx* a = pc;
  cplx* b = X;
  for (int s = 0; s < sq; ++s) {
    wg_gemm_mfma(b, a, a, nw);
    cplx* t = a;
    a = b;
    b = t;
  }
  for (int e = threadIdx.x; e < nw2; e += blockDim.x) Ek[e] = a[e];
  __syncthreads();
  wg_gemm_mfma(p.U[1 - cur] + (size_t)k * nw2, p.U[cur] + (size_t)k * nw2, Ek, nw);
}

// Retraction of one (problem, k) with every matrix in LDS (num_wann <= NWM):
// forms D (trial step), X = scale D, E = exp(X) and U_new = U E.  Padded
// column-major LDS matrices (leading dimension nw + 1).
template <int NWM>
__global__ void __launch_bounds__(W90_NT) k_step_k(const Prob* probs, const Item* items, int nitems, int mode) {
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const State* st = p.st;
  const int cur = gate_cur(st, mode);
  if (cur < 0) return;
  const int nw = p.nw, nw2 = p.nw2;
  const int ld = nw + 1;
  const int k = it.k;
  const int tid = threadIdx.x;
  __shared__ cplx X[NWM * (NWM + 1)];
  __shared__ cplx P[NWM * (NWM + 1)];
  __shared__ cplx T[NWM * (NWM + 1)];
  __shared__ double red[W90_NT];
  cplx* Dk = p.D + (size_t)k * nw2;
  // G, D and U of the k are loaded before the search and the exponential,
  // so their latency overlaps that work
  constexpr int SL = (NWM * NWM + W90_NT - 1) / W90_NT;
  cplx gpre[SL], dpre[SL], upre[SL];
  {
    const cplx* Gk = p.G + (size_t)k * nw2;
    const cplx* Uc = p.U[cur] + (size_t)k * nw2;
#pragma unroll
    for (int i = 0; i < SL; ++i) {
      const int e = tid + i * W90_NT;
      cplx g = c_make(0.0, 0.0), d = g, u = g;
      if (e < nw2) {
        if (mode == MODE_TRIAL) g = Gk[e];
        d = Dk[e];
        u = Uc[e];
      }
      gpre[i] = g;
      dpre[i] = d;
      upre[i] = u;
    }
  }
  __shared__ State lst;
  const State* ss = st;
  if (mode == MODE_TRIAL) {
    search_in_step(p, k, &lst, red);
    ss = &lst;
  }
  const double scale = ss->scale;
  // X, and the moduli of its elements for the 1-norm (into T, free until
  // the Taylor products)
  double* absx = reinterpret_cast<double*>(T);
  if (mode == MODE_TRIAL) {
    const double beta = ss->beta, sign = ss->dir_sign;
#pragma unroll
    for (int i = 0; i < SL; ++i) {
      const int e = tid + i * W90_NT;
      if (e < nw2) {
        cplx d = c_scale(c_add(gpre[i], c_scale(dpre[i], beta)), sign);
        st_stream(&Dk[e], d);
        const int q = (e % nw) + (e / nw) * ld;
        const cplx x = c_scale(d, scale);
        X[q] = x;
        absx[q] = sqrt(c_abs2(x));
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < SL; ++i) {
      const int e = tid + i * W90_NT;
      if (e < nw2) {
        const int q = (e % nw) + (e / nw) * ld;
        const cplx x = c_scale(dpre[i], scale);
        X[q] = x;
        absx[q] = sqrt(c_abs2(x));
      }
    }
  }
  __syncthreads();
  // column j < nw <= 64 is summed by lane j of wavefront 0, whose maximum
  // is formed by lane exchange and passed on through LDS
  double cs = 0.0;
  if (tid < 64) {
    if (tid < nw) {
#pragma unroll 4
      for (int i = 0; i < nw; ++i) cs += absx[i + tid * ld];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cs = fmax(cs, __shfl_xor(cs, o));
    if (tid == 0) red[0] = cs;
  }
  __syncthreads();
  cs = red[0];
  int sq, m;
  taylor_plan(cs, &sq, &m);
  const double f = ldexp(1.0, -sq);
  const double invm = 1.0 / (double)m;
  for (int e = tid; e < nw2; e += W90_NT) {
    const int q = (e % nw) + (e / nw) * ld;
    const cplx x = c_scale(X[q], f);
    X[q] = x;
    cplx v = c_scale(x, invm);
    if (e % nw == e / nw) v.x += 1.0;
    P[q] = v;
  }
  __syncthreads();
  // Horner: P <- I + X P / j, the scale and the identity applied as each
  // product is stored; P and T take turns holding the polynomial
  cplx* pc = P;
  cplx* pn = T;
  for (int j = m - 1; j >= 1; --j) {
    lds_gemm_mfma<NWM>(pn, X, pc, nw, 1.0 / (double)j, true);
    cplx* t = pc;
    pc = pn;
    pn = t;
  }
  cplx* a = pc;
  cplx* b = pn;
  for (int s = 0; s < sq; ++s) {
    lds_gemm_mfma<NWM>(b, a, a, nw);
    cplx* t = a;
    a = b;
    b = t;
  }
  cplx* Ek = p.E + (size_t)k * nw2;
#pragma unroll
  for (int i = 0; i < SL; ++i) {
    const int e = tid + i * W90_NT;
    if (e < nw2) {
      const int q = (e % nw) + (e / nw) * ld;
      st_stream(&Ek[e], a[q]);
      X[q] = upre[i];
    }
  }
  __syncthreads();
  lds_gemm_mfma<NWM>(b, X, a, nw);
  cplx* Un = p.U[1 - cur] + (size_t)k * nw2;
  for (int e = tid; e < nw2; e += W90_NT) st_stream(&Un[e], b[(e % nw) + (e / nw) * ld]);
}

// ---------------------------------------------------------- projection

// Lowdin orthonormalisation of overlap_project (src/overlap.F90:758-952)
// for num_bands = num_wann: the unitary polar factor U = A (A^+ A)^-1/2 of
// A(k) (staged in G), which is the Z V^+ built there from the SVD
// A = Z S V^+.  Newton-Schulz iteration X <- X (3 - X^+ X)/2 from
// X = A/||A||_F (all singular values in (0,1]) converges to it for
// non-singular A, quadratically near the end: once X^+ X = 1 to 1e-12 one
// more step reaches rounding level and the iteration stops (at most 200
// steps).  Writes U(k) to E.
__global__ void __launch_bounds__(W90_NT) k_polar(const Prob* probs, const Item* items, int nitems) {
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const int nw = p.nw, nw2 = p.nw2;
  const int k = it.k;
  cplx* X = p.E + (size_t)k * nw2;
  cplx* Y = p.W1 + (size_t)k * nw2;
  cplx* Z = p.W2 + (size_t)k * nw2;
  __shared__ double red[W90_NT];
  const cplx* Ak = p.G + (size_t)k * nw2;
  double fro = 0.0;
  for (int e = threadIdx.x; e < nw2; e += blockDim.x) fro += c_abs2(Ak[e]);
  fro = sqrt(block_sum(fro, red));
  const double inv = (fro > 0.0 && isfinite(fro)) ? 1.0 / fro : 0.0;
  for (int e = threadIdx.x; e < nw2; e += blockDim.x) X[e] = c_scale(Ak[e], inv);
  __syncthreads();
  for (int iter = 0; iter < 200; ++iter) {
    wg_gemm_cn(Y, X, X, nw);
    double dev = 0.0;
    for (int e = threadIdx.x; e < nw2; e += blockDim.x) {
      cplx y = Y[e];
      if (e % nw == e / nw) y.x -= 1.0;
      dev = fmax(dev, sqrt(c_abs2(y)));
      // 3 - X^+ X, halved below
      Y[e] = c_make(-Y[e].x + ((e % nw == e / nw) ? 3.0 : 0.0), -Y[e].y);
    }
    dev = block_max(dev, red);
    if (!(dev == dev)) break;  // non-finite input: left to the unitarity check
    wg_gemm(Z, X, Y, nw);
    for (int e = threadIdx.x; e < nw2; e += blockDim.x) X[e] = c_scale(Z[e], 0.5);
    __syncthreads();
    if (dev < 1e-12) break;
  }
}

// wann_check_unitarity (wannierise.F90:2594-2661): per-k maximum deviation
// of U U^+ and U^+ U from the identity (NaN-propagating).
__global__ void __launch_bounds__(W90_NT) k_unitarity(const Prob* probs, const Item* items, int nitems) {
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const int nw = p.nw, nw2 = p.nw2;
  const int k = it.k;
  const cplx* U = p.U[p.st->cur] + (size_t)k * nw2;
  __shared__ double red[W90_NT];
  double dev = 0.0;
  for (int e = threadIdx.x; e < nw2; e += blockDim.x) {
    const int i = e % nw, j = e / nw;
    cplx c1 = c_make(0, 0), c2 = c_make(0, 0);
    for (int m = 0; m < nw; ++m) {
      c1 = c_add(c1, c_mul(U[i + m * nw], c_conj(U[j + m * nw])));
      c2 = c_add(c2, c_mul(U[m + j * nw], c_conj(U[m + i * nw])));
    }
    if (i == j) {
      c1.x -= 1.0;
      c2.x -= 1.0;
    }
    const double v = fmax(sqrt(c_abs2(c1)), sqrt(c_abs2(c2)));
    if (!(v <= dev)) dev = isnan(v) ? INFINITY : fmax(v, dev);
    if (isnan(c1.x) || isnan(c1.y) || isnan(c2.x) || isnan(c2.y)) dev = INFINITY;
  }
  dev = block_max(dev, red);
  if (threadIdx.x == 0) p.unit[k] = dev;
}
