// libw90gpu: disentanglement on MI355X (gfx950) -- dis_main for k-point
// meshes (src/disentangle.F90:31-284) and setup_m_loc (:286-370), behind
// w90gpu_disentangle of include/w90gpu.h.
//
// Every per-k step is one workgroup per k-point:
//   k_dis_project     dis_project (:1173-1429): U_opt(k) = the polar factor
//                     of the projections A restricted to the outer window;
//   k_dis_proj_froz   dis_proj_froz (:1431-1891): the frozen bands, and the
//                     leading eigenvectors of Q P Q for the rest;
//   k_dis_orthonorm   internal_check_orthonorm (:372-451);
//   k_dis_zmat        internal_zmatrix (:2922-2988) with the Omega_I terms
//                     of dis_extract (:2069-2104, :2229-2256), one
//                     workgroup per (k, b): C = M(k,b) U_opt(k+b) once, from
//                     which the term w_b C C^+ of Z(k) (non-frozen rows),
//                     the spread term w_b ||U_opt(k)^+ C||^2 and its frozen
//                     rows;
//   k_dis_update      one extraction step (:2029-2068, :2115-2160): Z (the
//                     terms summed over b) mixed with the previous one, its
//                     leading eigenvectors as the new non-frozen columns of
//                     U_opt(k), the eigenproblem warm-started from the
//                     previous iteration's eigenvectors;
//   k_dis_reduce      one workgroup: Omega_I and its estimate over k, the
//                     change, and the convergence window (:2866-2920);
//   k_dis_ham         the Hamiltonian in the optimal subspace and U_opt in
//                     its eigenbasis (:2574-2640);
//   k_dis_find_u      internal_find_u (:531-698): U(k) = the polar factor of
//                     U_opt(k)^+ A(k); then W(k) = U_opt(k) U(k);
//   k_dis_rotate      one workgroup per (k, b): M(k,b) <- W(k)^+ M(k,b)
//                     W(k+b), the dis_main rotation (:229-241) and
//                     setup_m_loc in one product.
// Hermitian eigenproblems (of Z, Q P Q and the subspace Hamiltonian) are
// solved in the workgroup by a two-sided cyclic Jacobi method with
// round-robin (tournament) ordering: n/2 disjoint complex rotations per
// round, applied to A (by 2 x 2 blocks) and to the basis in one parallel
// pass; the polar factors (of A and of U_opt^+ A) by the one-sided Jacobi
// SVD with the same ordering.  A workgroup's matrices live in LDS up to 64
// bands (two 64 x 65 complex blocks) and in its own area of global memory
// above that, up to 256 bands (the same code, instantiated for each).  The
// extraction only uses the spanned subspaces and the eigenvalue sums, which
// do not depend on the eigenvector phases; the final U_opt and U do, so they
// agree with the reference up to the phase of each Hamiltonian eigenvector
// (their product U_opt U, and M, agree to rounding).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/w90gpu.h"
#include "w90gpu_ctx.h"

namespace {

typedef double2 zd;

__device__ __forceinline__ zd zmk(double r, double i) { return make_double2(r, i); }
__device__ __forceinline__ zd zadd(zd a, zd b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ zd zsub(zd a, zd b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ zd zmul(zd a, zd b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// conj(a) * b
__device__ __forceinline__ zd zcmul(zd a, zd b) { return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }
// a * conj(b)
__device__ __forceinline__ zd zmulc(zd a, zd b) { return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
__device__ __forceinline__ zd zconj(zd a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ zd zscale(zd a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double zabs2(zd a) { return a.x * a.x + a.y * a.y; }
// a += b * c
__device__ __forceinline__ void zfma(zd& a, zd b, zd c) {
  a.x = fma(b.x, c.x, fma(-b.y, c.y, a.x));
  a.y = fma(b.x, c.y, fma(b.y, c.x, a.y));
}
// a += conj(b) * c
__device__ __forceinline__ void zcfma(zd& a, zd b, zd c) {
  a.x = fma(b.x, c.x, fma(b.y, c.y, a.x));
  a.y = fma(b.x, c.y, fma(-b.y, c.x, a.y));
}

constexpr int kNT = 256;             // threads per workgroup
constexpr int kMaxN = W90GPU_DIS_MAX_BANDS;
// Largest matrix order of the LDS kernels: two kLdsN x (kLdsN + 1) complex
// blocks fill 133 KB of a workgroup's 160 KB.  Above it the same kernels
// run with their blocks in a per-workgroup area of global memory
// (instantiated with N = kMaxN), launched in chunks of kGwsBlocks
// workgroups that share those areas.
constexpr int kLdsN = 64;
constexpr size_t kLdsBytes = 2 * (size_t)kLdsN * (kLdsN + 1) * sizeof(double2);
constexpr int kGwsBlocks = 1024;

// Window of one k: first band (0-based), size, frozen count and the
// window row of the first frozen band (the frozen bands are contiguous).
struct Win {
  int first, ndim, nfroz, fmin;
};

struct DisState {
  int iter;        // iterations done
  int done;        // loop finished (converged or num_iter)
  int converged;
  int num_iter;
  int conv_window;
  int run;         // trailing iterations with |delta| < conv_tol
  int sweeps_max;  // most Jacobi sweeps of one extraction eigenproblem
  unsigned long long sweeps_sum, solves;
  double mix, conv_tol;
  double womegai, womegai1, delta;
};

// error and diagnostic bits (DisProb::flags)
enum : int {
  F_PROJECT_UNIT = 1,    // dis_project: U_opt not unitary within 1e-5
  F_FROZ_EIG = 2,        // dis_proj_froz: eigenvalues of QPQ outside [0, 1]
  F_FROZ_FIX = 4,        // dis_proj_froz: ortho-fix found too few vectors
  F_ORTHONORM = 8,       // internal_check_orthonorm
  F_JACOBI = 16,         // an eigenproblem did not converge in kMaxSweeps
};

struct DisProb {
  int nb, nw, nk, nn, ld;  // ld: leading dimension of the workgroup matrices (nb + 1)
  int ntri;                // nb (nb + 1) / 2
  int ldc;                 // k_dis_zmat: LDS leading dimension (odd, >= every ndim)
  int nzmax, ldz;          // k_dis_update: largest ndim - nfroz, leading dimension (odd)
  double wbtot;
  const int* nnl;      // [k * nn + b], 0-based
  const double* wb;    // [nn]
  const Win* win;      // [nk]
  const double* eig;   // (nb, nk)
  const zd* M;         // (nb, nb, nn, nk), all bands
  const zd* A;         // (nb, nw, nk), all bands
  zd* Uopt;            // (nb, nw, nk), rows relative to the window
  zd* Zin;             // (nb, nb, nk): the mixed Z, packed upper triangle
  zd* Zpart;           // (ntri, nn, nk): w_b C C^+ per (k, b), packed upper triangle
  zd* Vprev;           // (nb, nb, nk): eigenvectors of the last Z
  zd* U;               // (nw, nw, nk)
  zd* W;               // (nb, nw, nk): U_opt U
  zd* Mw;              // (nw, nw, nn, nk)
  double* wk1;         // [nk] Omega_I estimate terms
  double* part;        // (2, nn, nk): Omega_I terms and frozen-row terms per (k, b)
  double* eigopt;      // (nb, nk)
  double* log;         // (3, num_iter)
  DisState* st;
  int* flags;
  zd* gws;            // global-memory kernels: gws_stride complex per workgroup
  size_t gws_stride;
  int blk0;           // global-memory kernels: first block of this launch
};

// The workgroup's matrix area (LDS, or its slice of gws above kLdsN) and
// its block index (the launch's offset added above kLdsN).
extern __shared__ zd dis_lds[];
template <int N>
__device__ __forceinline__ zd* dis_ws(const DisProb& p) {
  if constexpr (N > kLdsN) return p.gws + (size_t)blockIdx.x * p.gws_stride;
  else return dis_lds;
}
template <int N>
__device__ __forceinline__ int dis_blk(const DisProb& p) {
  if constexpr (N > kLdsN) return (int)blockIdx.x + p.blk0;
  else return (int)blockIdx.x;
}

// window row of the i-th non-frozen state
__device__ __forceinline__ int nf_row(const Win& w, int i) { return i < w.fmin ? i : i + w.nfroz; }

// Deterministic block reductions (wavefront butterflies, then the four
// wavefront results in order); red holds kNT / 64 doubles.
__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  for (int i = 0; i < kNT / 64; ++i) r += red[i];
  __syncthreads();
  return r;
}

// ------------------------------------------------------------ Jacobi

constexpr int kMaxSweeps = 40;

// A complex Jacobi rotation J on (p, q): J = [[c, s], [-conj(s), c]].
struct JRot {
  int p, q;
  double c;
  zd s;
};

template <int N>
struct JacobiLds {
  static constexpr int kN = N;
  JRot rot[N / 2];
  int count;
  double red[kNT / 64];
  double lam[N];
  int perm[N];
  int rank[N];
};

// Position (r, c), r <= c, of entry x = c (c + 1) / 2 + r of a packed upper
// triangle (x < 2^22: the single-precision root is corrected by one step).
__device__ __forceinline__ void tri_decode(int x, int& r, int& c) {
  int cc = (int)((sqrtf(8.0f * (float)x + 1.0f) - 1.0f) * 0.5f);
  if (cc * (cc + 1) / 2 > x) --cc;
  if ((cc + 1) * (cc + 2) / 2 <= x) ++cc;
  c = cc;
  r = x - cc * (cc + 1) / 2;
}

// Pair (p, q), p < q, of slot s in round r of the round-robin (tournament)
// schedule on m (even) indices: index 0 stays, the others turn one place
// per round, and slot s pairs the indices s and m - 1 - s positions away.
__device__ __forceinline__ void rr_pair(int s, int r, int m, int& p, int& q) {
  int a = 0;
  if (s > 0) {
    a = s - 1 + r;
    if (a >= m - 1) a -= m - 1;
    a += 1;
  }
  int b = m - 2 - s + r;
  if (b >= m - 1) b -= m - 1;
  b += 1;
  p = min(a, b);
  q = max(a, b);
}

// rows of E per lane and chunk in a rotation pass; n <= 64 takes one chunk
// (at least 8 lanes a slot), n <= 256 up to eight
constexpr int kRowsPerLane = 8;
template <int N>
constexpr int row_chunks() {
  return (N + (kNT / (N / 2)) * kRowsPerLane - 1) / ((kNT / (N / 2)) * kRowsPerLane);
}

// The rotation of pair (p, q) is skipped when |a_pq| <= 1e-15 sqrt(|a_pp
// a_qq|) (relative to the diagonal) or <= 1e-16 ||A||_F (without this floor
// the rounding noise among the zero eigenvalues of a rank-deficient matrix
// would be rotated forever).
__device__ __forceinline__ bool jacobi_rotates(double b2, double app, double aqq, double floor2) {
  return b2 > floor2 && b2 > 1e-30 * fabs(app * aqq);
}

// Whether some off-diagonal pair of A still rotates (all lanes, one pass).
__device__ __forceinline__ bool jacobi_pending(const zd* A, int n, int ld, double floor2) {
  int any = 0;
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int i = e % n, j = e / n;
    if (i < j && jacobi_rotates(zabs2(A[i + j * ld]), A[i + i * ld].x, A[j + j * ld].x, floor2)) any = 1;
  }
  return __syncthreads_or(any) != 0;
}

// The rotation that zeroes a_pq: with D = diag(1, a_pq/|a_pq|) the 2 x 2
// block D^+ A D is real, and its rotation by the smaller angle (tan 2t =
// 2|a_pq| / (a_qq - a_pp)) gives J = D R D^+:
//   t = sign(h) 2|a_pq| / (|h| + sqrt(h^2 + 4|a_pq|^2)), h = a_qq - a_pp,
//   c = 1 / sqrt(1 + t^2), s = t c a_pq / |a_pq|.
__device__ __forceinline__ JRot jacobi_rotation(int p, int q, const zd* A, int n, int ld, double floor2) {
  JRot R;
  R.p = p;
  R.q = q;
  R.c = 1.0;
  R.s = zmk(0.0, 0.0);
  if (q < n) {
    const zd apq = A[p + q * ld];
    const double app = A[p + p * ld].x, aqq = A[q + q * ld].x;
    const double b2 = zabs2(apq);
    if (jacobi_rotates(b2, app, aqq, floor2)) {
      const double rb = rsqrt(b2);
      const double h = aqq - app;
      const double den = fabs(h) + sqrt(fma(h, h, 4.0 * b2));
      const double tt = copysign(2.0 * b2 * rb, h) / den;
      const double c = rsqrt(fma(tt, tt, 1.0));
      const double sr = tt * c * rb;
      R.c = c;
      R.s = zmk(sr * apq.x, sr * apq.y);
    }
  }
  return R;
}

// Eigen-decomposition of the Hermitian n x n matrix A (LDS, column-major,
// leading dimension ld, stored whole) by the two-sided cyclic Jacobi method
// with round-robin (tournament) ordering: each round rotates n/2 disjoint
// pairs at once; the rotations go into J.rot, and one parallel pass then
// applies A <- J^+ A J by 2 x 2 blocks (one lane per block of the upper
// block triangle, which writes the block and its mirror) and E <- E J to
// the rows of E (LDS, leading dimension ld).  On return A is diagonal to the
// tolerance of jacobi_rotates and the columns of E (as it came in) times the
// rotations are the eigenvectors: E = I gives those of A, E = V for A = V^+
// Z V those of Z.  The sweeps end when no pair rotates (checked before each
// sweep in one parallel pass).  Returns the number of sweeps, or -1 if
// kMaxSweeps were not enough.  A warm start (a nearly diagonal A) leaves
// off-diagonal entries just below the skip bounds, so they are kept tight:
// the selected subspace moves by their size over the eigenvalue gap at the
// selection.
template <class JL>
__device__ __forceinline__ int jacobi_run(zd* A, zd* E, int n, int ld, JL& J) {
  constexpr int CH = row_chunks<JL::kN>();
  const int t = threadIdx.x;
  double fro = 0.0;
  for (int e = t; e < n * n; e += kNT) fro += zabs2(A[(e % n) + (e / n) * ld]);
  const double floor2 = block_sum(fro, J.red) * 1e-32;
  const int m = n + (n & 1), half = m / 2;
  const int nblk = half * (half + 1) / 2;
  // this lane's first block of the upper block triangle, and its share of
  // the rows of E: slot es (g lanes per slot), rows eg, eg + g, ...
  int bs1 = 0, bs2 = 0;
  if (t < nblk) tri_decode(t, bs1, bs2);
  const int g = kNT / half, es = t / g, eg = t % g;
  int sweeps = 0;
  while (jacobi_pending(A, n, ld, floor2)) {
    if (sweeps == kMaxSweeps) return -1;
    ++sweeps;
    for (int r = 0; r < m - 1; ++r) {
      if (t < half) {
        int p, q;
        rr_pair(t, r, m, p, q);
        J.rot[t] = jacobi_rotation(p, q, A, n, ld, floor2);
      }
      __syncthreads();
      for (int x = t; x < nblk; x += kNT) {
        int r1 = bs1, r2 = bs2;
        if (x != t) tri_decode(x, r1, r2);
        int p1, q1, p2, q2;
        rr_pair(r1, r, m, p1, q1);
        rr_pair(r2, r, m, p2, q2);
        const bool v1 = q1 < n, v2 = q2 < n;
        // the block's entries requested with the rotations
        const zd z0 = zmk(0.0, 0.0);
        const zd xpp = A[p1 + p2 * ld];
        const zd xpq = v2 ? A[p1 + q2 * ld] : z0;
        const zd xqp = v1 ? A[q1 + p2 * ld] : z0;
        const zd xqq = (v1 && v2) ? A[q1 + q2 * ld] : z0;
        const double c1 = J.rot[r1].c, c2 = J.rot[r2].c;
        const zd s1 = J.rot[r1].s, s2 = J.rot[r2].s;
        if (s1.x == 0.0 && s1.y == 0.0 && s2.x == 0.0 && s2.y == 0.0) continue;
        // Y = X J2, then J1^+ Y
        const zd s2c = zconj(s2), s1c = zconj(s1);
        const zd ypp = zsub(zscale(xpp, c2), zmul(s2c, xpq)), ypq = zadd(zmul(s2, xpp), zscale(xpq, c2));
        const zd yqp = zsub(zscale(xqp, c2), zmul(s2c, xqq)), yqq = zadd(zmul(s2, xqp), zscale(xqq, c2));
        const zd npp = zsub(zscale(ypp, c1), zmul(s1, yqp)), nqp = zadd(zmul(s1c, ypp), zscale(yqp, c1));
        const zd npq = zsub(zscale(ypq, c1), zmul(s1, yqq)), nqq = zadd(zmul(s1c, ypq), zscale(yqq, c1));
        if (r1 != r2) {
          A[p1 + p2 * ld] = npp;
          A[p2 + p1 * ld] = zconj(npp);
          if (v2) {
            A[p1 + q2 * ld] = npq;
            A[q2 + p1 * ld] = zconj(npq);
          }
          if (v1) {
            A[q1 + p2 * ld] = nqp;
            A[p2 + q1 * ld] = zconj(nqp);
          }
          if (v1 && v2) {
            A[q1 + q2 * ld] = nqq;
            A[q2 + q1 * ld] = zconj(nqq);
          }
        } else {
          A[p1 + p1 * ld] = zmk(npp.x, 0.0);
          if (v1) {
            A[q1 + q1 * ld] = zmk(nqq.x, 0.0);
            A[p1 + q1 * ld] = npq;
            A[q1 + p1 * ld] = zconj(npq);
          }
        }
      }
      if (es < half) {
        int p, q;
        rr_pair(es, r, m, p, q);
        if (q < n) {
#pragma unroll 1
          for (int ch = 0; ch < CH; ++ch) {
            const int row0 = eg + ch * kRowsPerLane * g;
            zd ep[kRowsPerLane], eq[kRowsPerLane];
#pragma unroll
            for (int j = 0; j < kRowsPerLane; ++j) {
              const int row = row0 + j * g;
              if (row < n) {
                ep[j] = E[row + p * ld];
                eq[j] = E[row + q * ld];
              }
            }
            const double c = J.rot[es].c;
            const zd sv = J.rot[es].s, sc = zconj(sv);
            if (sv.x != 0.0 || sv.y != 0.0) {
#pragma unroll
              for (int j = 0; j < kRowsPerLane; ++j) {
                const int row = row0 + j * g;
                if (row < n) {
                  E[row + p * ld] = zsub(zscale(ep[j], c), zmul(sc, eq[j]));
                  E[row + q * ld] = zadd(zmul(sv, ep[j]), zscale(eq[j], c));
                }
              }
            }
          }
        }
      }
      __syncthreads();
    }
  }
  return sweeps;
}

// jacobi_run from E = I into V (the eigenvectors of A itself).
template <class JL>
__device__ __forceinline__ int jacobi_eigh(zd* A, zd* V, int n, int ld, JL& J) {
  for (int e = threadIdx.x; e < n * n; e += kNT) {
    const int r = e % n, c = e / n;
    V[r + c * ld] = zmk(r == c ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  return jacobi_run(A, V, n, ld, J);
}

// ------------------------------------------------------------ products

// out(i, j, C(i, j)) for C = op(A) B (M x N, inner dimension K) by RT x CT
// register tiles, one tile per lane and pass: op(A)(i, l) = A[i + l * lda],
// or conj(A[l + i * lda]) when ADJ; B(l, j) = B[l + j * ldb].  Rows and
// columns past the edge are read clamped and not stored.
template <int RT, int CT, bool ADJ, int UNR = 2, class Out>
__device__ void zgemm_tiles(int M, int N, int K, const zd* A, int lda, const zd* B, int ldb, Out out) {
  const int tr = (M + RT - 1) / RT, tc = (N + CT - 1) / CT;
  for (int x = threadIdx.x; x < tr * tc; x += kNT) {
    const int i0 = (x % tr) * RT, j0 = (x / tr) * CT;
    int ii[RT], jj[CT];
#pragma unroll
    for (int a = 0; a < RT; ++a) ii[a] = min(i0 + a, M - 1);
#pragma unroll
    for (int b = 0; b < CT; ++b) jj[b] = min(j0 + b, N - 1);
    zd acc[RT][CT];
#pragma unroll
    for (int a = 0; a < RT; ++a)
#pragma unroll
      for (int b = 0; b < CT; ++b) acc[a][b] = zmk(0.0, 0.0);
#pragma unroll UNR
    for (int l = 0; l < K; ++l) {
      zd av[RT], bv[CT];
#pragma unroll
      for (int a = 0; a < RT; ++a) av[a] = ADJ ? A[l + ii[a] * lda] : A[ii[a] + l * lda];
#pragma unroll
      for (int b = 0; b < CT; ++b) bv[b] = B[l + jj[b] * ldb];
#pragma unroll
      for (int a = 0; a < RT; ++a)
#pragma unroll
        for (int b = 0; b < CT; ++b) {
          if (ADJ) zcfma(acc[a][b], av[a], bv[b]);
          else zfma(acc[a][b], av[a], bv[b]);
        }
    }
#pragma unroll
    for (int a = 0; a < RT; ++a)
#pragma unroll
      for (int b = 0; b < CT; ++b)
        if (i0 + a < M && j0 + b < N) out(i0 + a, j0 + b, acc[a][b]);
  }
}

// zgemm with tiles of at most 2 x 3 (several passes for large results):
// the register budget of k_dis_zmat, whose (k, b) workgroups wait on
// memory and gain from more of them per CU; UNR K steps unrolled (four
// when A is in global memory, to request its loads ahead).
template <bool ADJ, int UNR, class Out>
__device__ void zgemm_small(int M, int N, int K, const zd* A, int lda, const zd* B, int ldb, Out out) {
  if (((M + 1) / 2) * ((N + 1) / 2) <= kNT) zgemm_tiles<2, 2, ADJ, UNR>(M, N, K, A, lda, B, ldb, out);
  else zgemm_tiles<2, 3, ADJ, UNR>(M, N, K, A, lda, B, ldb, out);
}

// zgemm_tiles with the smallest tile that needs at most one pass (M, N <=
// 64 fit 4 x 4 tiles; larger results take several passes of them).
template <bool ADJ, class Out>
__device__ void zgemm(int M, int N, int K, const zd* A, int lda, const zd* B, int ldb, Out out) {
  auto fits = [&](int r, int c) { return ((M + r - 1) / r) * ((N + c - 1) / c) <= kNT; };
  if (fits(2, 2)) zgemm_tiles<2, 2, ADJ>(M, N, K, A, lda, B, ldb, out);
  else if (fits(2, 3)) zgemm_tiles<2, 3, ADJ>(M, N, K, A, lda, B, ldb, out);
  else if (fits(3, 3)) zgemm_tiles<3, 3, ADJ>(M, N, K, A, lda, B, ldb, out);
  else if (fits(3, 4)) zgemm_tiles<3, 4, ADJ>(M, N, K, A, lda, B, ldb, out);
  else zgemm_tiles<4, 4, ADJ>(M, N, K, A, lda, B, ldb, out);
}

// The upper triangle (i <= j) of a Hermitian N x N product A^+ B by T x T
// tiles of the upper block triangle: out(i, j, C(i, j)) for i <= j only.
template <int T, class Out>
__device__ void zgemm_herm_tiles(int N, int K, const zd* A, int lda, const zd* B, int ldb, Out out) {
  const int nt = (N + T - 1) / T;
  for (int x = threadIdx.x; x < nt * (nt + 1) / 2; x += kNT) {
    int ib, jb;
    tri_decode(x, ib, jb);
    const int i0 = ib * T, j0 = jb * T;
    int ii[T], jj[T];
#pragma unroll
    for (int a = 0; a < T; ++a) {
      ii[a] = min(i0 + a, N - 1);
      jj[a] = min(j0 + a, N - 1);
    }
    zd acc[T][T];
#pragma unroll
    for (int a = 0; a < T; ++a)
#pragma unroll
      for (int b = 0; b < T; ++b) acc[a][b] = zmk(0.0, 0.0);
#pragma unroll 2
    for (int l = 0; l < K; ++l) {
      zd av[T], bv[T];
#pragma unroll
      for (int a = 0; a < T; ++a) {
        av[a] = A[l + ii[a] * lda];
        bv[a] = B[l + jj[a] * ldb];
      }
#pragma unroll
      for (int a = 0; a < T; ++a)
#pragma unroll
        for (int b = 0; b < T; ++b) zcfma(acc[a][b], av[a], bv[b]);
    }
#pragma unroll
    for (int a = 0; a < T; ++a)
#pragma unroll
      for (int b = 0; b < T; ++b)
        if (i0 + a <= j0 + b && j0 + b < N) out(i0 + a, j0 + b, acc[a][b]);
  }
}

template <class Out>
__device__ void zgemm_herm(int N, int K, const zd* A, int lda, const zd* B, int ldb, Out out) {
  auto fits = [&](int tl) {
    const int nt = (N + tl - 1) / tl;
    return nt * (nt + 1) / 2 <= kNT;
  };
  if (fits(2)) zgemm_herm_tiles<2>(N, K, A, lda, B, ldb, out);
  else if (fits(3)) zgemm_herm_tiles<3>(N, K, A, lda, B, ldb, out);
  else zgemm_herm_tiles<4>(N, K, A, lda, B, ldb, out);
}

// Eigenvalues (the diagonal of A after jacobi_run) into J.lam, their
// ascending order into J.perm (ties by index) and its inverse into J.rank.
template <class JL>
__device__ __forceinline__ void eig_order(const zd* A, int n, int ld, JL& J) {
  const int t = threadIdx.x;
  if (t < n) J.lam[t] = A[t + t * ld].x;
  __syncthreads();
  if (t < n) {
    const double l = J.lam[t];
    int rank = 0;
    for (int j = 0; j < n; ++j) rank += (J.lam[j] < l || (J.lam[j] == l && j < t)) ? 1 : 0;
    J.perm[rank] = t;
    J.rank[t] = rank;
  }
  __syncthreads();
}

// Polar factor U = Z V^+ of the rows x nw matrix X (global, leading
// dimension ldx; rows >= nw), from its singular value decomposition by the
// one-sided (Hestenes) Jacobi method: the columns of G = X are rotated in
// pairs until they are orthogonal (|g_p^+ g_q| <= 1e-13 ||g_p|| ||g_q||, or
// one of the two below 1e-15 ||X||_F),
// G = X V, and Z = G S^-1 with S the column norms.  Working on X itself
// (not on X^+ X) keeps small singular values to relative accuracy.  Columns
// with a singular value at or below 1e-12 of the largest (a rank-deficient
// X, where the polar factor is not unique, as with the SVD) are completed
// by Gram-Schmidt from unit vectors.  U goes to global memory (leading
// dimension ldu).  LDS: G (rows x nw, leading dimension ld), V.
template <class JL>
__device__ __forceinline__ void polar_factor(const zd* X, int ldx, int rows, int nw, zd* U, int ldu, zd* G, zd* V,
                                             int ld, JL& J, int* flags) {
  const int t = threadIdx.x;
  double fro = 0.0;
  for (int e = t; e < rows * nw; e += kNT) {
    const int r = e % rows, c = e / rows;
    const zd x = X[r + (size_t)c * ldx];
    G[r + c * ld] = x;
    fro += zabs2(x);
  }
  // columns below 1e-15 ||X||_F are numerically zero: rotating their
  // rounding noise against the others would never end
  const double zero2 = 1e-30 * block_sum(fro, J.red);
  for (int e = t; e < nw * nw; e += kNT) {
    const int r = e % nw, c = e / nw;
    V[r + c * ld] = zmk(r == c ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  const int m = nw + (nw & 1), half = m / 2;
  int sweep = 0;
  for (; sweep < kMaxSweeps; ++sweep) {
    if (t == 0) J.count = 0;
    __syncthreads();
    for (int r = 0; r < m - 1; ++r) {
      // Gram entries of each pair, 8 lanes per pair (kNT / 8 pairs a pass)
      for (int pb = 0; pb < half; pb += kNT / 8) {
        const int pi = pb + (t >> 3), lane = t & 7;
        double al = 0.0, be = 0.0;
        zd ga = zmk(0.0, 0.0);
        int P = 0, Q = 0;
        bool valid = false;
        if (pi < half) {
          rr_pair(pi, r, m, P, Q);
          valid = Q < nw;
          if (valid)
            for (int row = lane; row < rows; row += 8) {
              const zd gp = G[row + P * ld], gq = G[row + Q * ld];
              al += zabs2(gp);
              be += zabs2(gq);
              zcfma(ga, gp, gq);
            }
        }
        for (int off = 1; off < 8; off <<= 1) {
          al += __shfl_xor(al, off, 64);
          be += __shfl_xor(be, off, 64);
          ga.x += __shfl_xor(ga.x, off, 64);
          ga.y += __shfl_xor(ga.y, off, 64);
        }
        if (pi < half && lane == 0) {
          JRot R;
          R.p = P;
          R.q = Q;
          R.c = 1.0;
          R.s = zmk(0.0, 0.0);
          const double b2 = zabs2(ga);
          if (valid && b2 > 0.0 && b2 > 1e-26 * al * be && fmin(al, be) > zero2) {
            const double bn = sqrt(b2);
            const double tau = (be - al) / (2.0 * bn);
            double tt;
            if (fabs(tau) > 1e150) tt = 0.5 / tau;
            else tt = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            const double c = 1.0 / sqrt(1.0 + tt * tt), s = tt * c;
            R.c = c;
            R.s = zmk(s * ga.x / bn, s * ga.y / bn);
            atomicAdd(&J.count, 1);
          }
          J.rot[pi] = R;
        }
      }
      __syncthreads();
      for (int e = t; e < half * (rows + nw); e += kNT) {
        const int i = e / (rows + nw), row = e % (rows + nw);
        const JRot R = J.rot[i];
        if (R.s.x == 0.0 && R.s.y == 0.0) continue;
        zd* M = row < rows ? G : V;
        const int rr = row < rows ? row : row - rows;
        const zd ap = M[rr + R.p * ld], aq = M[rr + R.q * ld];
        M[rr + R.p * ld] = zadd(zscale(ap, R.c), zscale(zmul(zconj(R.s), aq), -1.0));
        M[rr + R.q * ld] = zadd(zmul(R.s, ap), zscale(aq, R.c));
      }
      __syncthreads();
    }
    const int cnt = J.count;
    __syncthreads();
    if (cnt == 0) break;
  }
  if (sweep >= kMaxSweeps && t == 0) atomicOr(flags, F_JACOBI);
  if (t < nw) {
    double s2 = 0.0;
    for (int r = 0; r < rows; ++r) s2 += zabs2(G[r + t * ld]);
    J.lam[t] = sqrt(s2);
  }
  __syncthreads();
  double smax = 0.0;
  for (int i = 0; i < nw; ++i) smax = fmax(smax, J.lam[i]);
  auto good = [&](int i) { return J.lam[i] > 1e-12 * smax && J.lam[i] > 0.0; };
  for (int e = t; e < rows * nw; e += kNT) {
    const int r = e % rows, i = e / rows;
    G[r + i * ld] = good(i) ? zscale(G[r + i * ld], 1.0 / J.lam[i]) : zmk(0.0, 0.0);
  }
  __syncthreads();
  // completion of rank-deficient columns (sequential, rare)
  if (t == 0) {
    for (int i = 0; i < nw; ++i) {
      if (good(i)) continue;
      for (int e0 = 0; e0 < rows; ++e0) {
        for (int r = 0; r < rows; ++r) G[r + i * ld] = zmk(r == e0 ? 1.0 : 0.0, 0.0);
        for (int pass = 0; pass < 2; ++pass)
          for (int j = 0; j < nw; ++j) {
            if (j == i || !(good(j) || j < i)) continue;
            zd d = zmk(0.0, 0.0);
            for (int r = 0; r < rows; ++r) zcfma(d, G[r + j * ld], G[r + i * ld]);
            for (int r = 0; r < rows; ++r) zfma(G[r + i * ld], zscale(d, -1.0), G[r + j * ld]);
          }
        double nrm = 0.0;
        for (int r = 0; r < rows; ++r) nrm += zabs2(G[r + i * ld]);
        if (nrm > 0.25) {
          const double inv = 1.0 / sqrt(nrm);
          for (int r = 0; r < rows; ++r) G[r + i * ld] = zscale(G[r + i * ld], inv);
          break;
        }
      }
    }
  }
  __syncthreads();
  // U = Z V^+
  for (int e = t; e < rows * nw; e += kNT) {
    const int r = e % rows, j = e / rows;
    zd u = zmk(0.0, 0.0);
    for (int i = 0; i < nw; ++i) zfma(u, G[r + i * ld], zconj(V[j + i * ld]));
    U[r + (size_t)j * ldu] = u;
  }
  __syncthreads();
}

// ------------------------------------------------------------ kernels
//
// Each kernel with matrices per workgroup is a template on the largest
// matrix order N: N = kLdsN keeps them in LDS, N = kMaxN in the
// workgroup's area of global memory (dis_ws, dis_blk).

// dis_project: U_opt(k) = polar factor of A(window rows, :); its unitarity
// checked to 1e-5; rows below the window zero.
template <int N>
__global__ void __launch_bounds__(kNT) k_dis_project(DisProb p) {
  __shared__ JacobiLds<N> J;
  const int k = dis_blk<N>(p);
  const Win w = p.win[k];
  const int nb = p.nb, nw = p.nw, ld = p.ld;
  zd* S = dis_ws<N>(p);
  zd* V = S + nb * ld;
  zd* U = p.Uopt + (size_t)k * nb * nw;
  const zd* X = p.A + (size_t)k * nb * nw + w.first;
  polar_factor(X, nb, w.ndim, nw, U, nb, S, V, ld, J, p.flags);
  for (int e = threadIdx.x; e < nb * nw; e += kNT)
    if (e % nb >= w.ndim) U[e] = zmk(0.0, 0.0);
  // unitarity of U_opt (dis_project, :1384-1410)
  bool bad = false;
  for (int e = threadIdx.x; e < nw * nw; e += kNT) {
    const int i = e % nw, j = e / nw;
    zd s = zmk(0.0, 0.0);
    for (int m = 0; m < w.ndim; ++m) zcfma(s, U[m + (size_t)i * nb], U[m + (size_t)j * nb]);
    if (i == j) s.x -= 1.0;
    if (sqrt(zabs2(s)) > 1e-5) bad = true;
  }
  if (bad) atomicOr(p.flags, F_PROJECT_UNIT);
}

// dis_proj_froz for one k: frozen columns = unit vectors at the frozen
// bands; the others = the leading nw - nfroz eigenvectors of Q P Q, with
// P = U_opt U_opt^+ and Q the projector off the frozen bands, in ascending
// order of eigenvalue; when some of those eigenvalues are below 1e-8 the
// reference's ortho-fix picks, from the top down, eigenvectors orthogonal
// to the frozen columns.
template <int N>
__global__ void __launch_bounds__(kNT) k_dis_proj_froz(DisProb p) {
  __shared__ JacobiLds<N> J;
  __shared__ int vmap[N];
  __shared__ int fix_ok;
  const int k = dis_blk<N>(p);
  const Win w = p.win[k];
  const int nb = p.nb, nw = p.nw, ld = p.ld, nd = w.ndim, t = threadIdx.x;
  zd* Aq = dis_ws<N>(p);
  zd* V = Aq + nb * ld;
  zd* U = p.Uopt + (size_t)k * nb * nw;
  const int r = nw - w.nfroz;
  auto frozen = [&](int i) { return i >= w.fmin && i < w.fmin + w.nfroz; };
  if (r > 0) {
    for (int e = t; e < nd * nd; e += kNT) {
      const int i = e % nd, j = e / nd;
      if (i > j) continue;
      zd s = zmk(0.0, 0.0);
      if (!frozen(i) && !frozen(j))
        for (int l = 0; l < nw; ++l) zfma(s, U[i + (size_t)l * nb], zconj(U[j + (size_t)l * nb]));
      if (i == j) s.y = 0.0;
      Aq[i + j * ld] = s;
      Aq[j + i * ld] = zconj(s);
    }
    __syncthreads();
    if (jacobi_eigh(Aq, V, nd, ld, J) < 0 && t == 0) atomicOr(p.flags, F_JACOBI);
    eig_order(Aq, nd, ld, J);
    if (t == 0) {
      fix_ok = 1;
      for (int j = 0; j < nd; ++j)
        if (J.lam[j] < -1e-8 || J.lam[j] > 1.0 + 1e-8) atomicOr(p.flags, F_FROZ_EIG);
      int nzero = 0, goods = 0;
      for (int j = nd - 1; j >= nd - r; --j) {
        if (J.lam[J.perm[j]] < 1e-8) nzero++;
        else goods++;
      }
      if (nzero == 0) {
        for (int l = 0; l < r; ++l) vmap[l] = J.perm[nd - r + l];
      } else {
        // ortho-fix (:1780-1850): the good eigenvectors from the top, then
        // eigenvectors orthogonal to the frozen columns, searched from the top
        int used[N];
        for (int j = 0; j < nd; ++j) used[j] = 0;
        int cnt = 0;
        for (int j = nd - 1; j >= nd - goods; --j) {
          vmap[cnt++] = J.perm[j];
          used[j] = 1;
        }
        for (int f = 0; f < nzero; ++f) {
          int found = -1;
          for (int j = nd - 1; j >= 0 && found < 0; --j) {
            if (used[j]) continue;
            const int v = J.perm[j];
            // orthogonal to the first nfroz columns of U_opt as dis_project
            // left them (the unit frozen columns are set afterwards)
            bool take = true;
            for (int m = 0; m < w.nfroz && take; ++m) {
              zd c = zmk(0.0, 0.0);
              for (int j2 = 0; j2 < nd; ++j2) zcfma(c, U[j2 + (size_t)m * nb], V[j2 + v * ld]);
              if (sqrt(zabs2(c)) > 1e-8) take = false;
            }
            if (take) found = j;
          }
          if (found < 0) {
            fix_ok = 0;
            break;
          }
          used[found] = 1;
          vmap[cnt++] = J.perm[found];
        }
        // the fix-up columns follow the good ones in the reference's order:
        // good ones from the top down, then the picked ones
      }
      if (!fix_ok) atomicOr(p.flags, F_FROZ_FIX);
    }
    __syncthreads();
    if (fix_ok) {
      for (int e = t; e < nd * r; e += kNT) {
        const int i = e % nd, l = e / nd;
        U[i + (size_t)(w.nfroz + l) * nb] = V[i + vmap[l] * ld];
      }
    }
  }
  __syncthreads();
  for (int e = t; e < nb * w.nfroz; e += kNT) {
    const int i = e % nb, l = e / nb;
    U[i + (size_t)l * nb] = zmk(i == w.fmin + l ? 1.0 : 0.0, 0.0);
  }
}

// internal_check_orthonorm: the columns of U_opt(k) orthonormal to 1e-8.
__global__ void __launch_bounds__(kNT) k_dis_orthonorm(DisProb p) {
  const int k = blockIdx.x;
  const Win w = p.win[k];
  const int nb = p.nb, nw = p.nw;
  const zd* U = p.Uopt + (size_t)k * nb * nw;
  bool bad = false;
  for (int e = threadIdx.x; e < nw * nw; e += kNT) {
    const int l = e % nw, m = e / nw;
    if (m > l) continue;
    zd s = zmk(0.0, 0.0);
    for (int j = 0; j < w.ndim; ++j) zcfma(s, U[j + (size_t)m * nb], U[j + (size_t)l * nb]);
    if (l == m) s.x -= 1.0;
    if (sqrt(zabs2(s)) > 1e-8) bad = true;
  }
  if (bad) atomicOr(p.flags, F_ORTHONORM);
}

// internal_zmatrix and the Omega_I terms for one (k, b), one workgroup
// each, with U_opt of every k as it is now: C = M(k,b) U_opt(k+b) and
//   Zpart(k,b) = w_b C C^+ over the non-frozen rows (upper triangle,
//                packed), summed over b by k_dis_update;
//   part(0,k,b) = w_b ||U_opt(k)^+ C||^2, the Omega_I term, summed over b
//                by k_dis_reduce;
//   part(1,k,b) = the same over the frozen rows of U_opt(k)^+ C, the frozen
//                term of the next extraction step.
// LDS: C (ndim(k) x nw) and U (ndim(k+b) x nw, then U_opt(k)), leading
// dimension ldc; the products by register tiles.
template <int T>
__device__ void zpart_tiles(const zd* C, int ldc, const Win& w, int nd, int nw, double wbb, zd* Zp) {
  const int nt = (nd + T - 1) / T;
  for (int x = threadIdx.x; x < nt * (nt + 1) / 2; x += kNT) {
    int ib, jb;
    tri_decode(x, ib, jb);
    int pr[T], qr[T];
#pragma unroll
    for (int a = 0; a < T; ++a) {
      pr[a] = nf_row(w, min(ib * T + a, nd - 1));
      qr[a] = nf_row(w, min(jb * T + a, nd - 1));
    }
    zd acc[T][T];
#pragma unroll
    for (int a = 0; a < T; ++a)
#pragma unroll
      for (int b = 0; b < T; ++b) acc[a][b] = zmk(0.0, 0.0);
#pragma unroll 2
    for (int l = 0; l < nw; ++l) {
      zd av[T], bv[T];
#pragma unroll
      for (int a = 0; a < T; ++a) {
        av[a] = C[pr[a] + l * ldc];
        bv[a] = zconj(C[qr[a] + l * ldc]);
      }
#pragma unroll
      for (int a = 0; a < T; ++a)
#pragma unroll
        for (int b = 0; b < T; ++b) zfma(acc[a][b], av[a], bv[b]);
    }
#pragma unroll
    for (int a = 0; a < T; ++a)
#pragma unroll
      for (int b = 0; b < T; ++b) {
        const int i = ib * T + a, j = jb * T + b;
        if (i <= j && j < nd) Zp[j * (j + 1) / 2 + i] = zscale(acc[a][b], wbb);
      }
  }
}

template <int N>
__global__ void __launch_bounds__(kNT) k_dis_zmat(DisProb p) {
  __shared__ double red[2][kNT / 64];
  const DisState* st = p.st;
  if (st->done) return;
  const int nn = p.nn;
  const int blk = dis_blk<N>(p);
  const int k = blk / nn, b = blk % nn, t = threadIdx.x;
  const Win w = p.win[k];
  const int nb = p.nb, nw = p.nw, ldc = p.ldc;
  const int nd = w.ndim - w.nfroz;
  zd* C = dis_ws<N>(p);    // ndim(k) x nw
  zd* U2 = C + ldc * nw;   // ndim(k+b) x nw, then ndim(k) x nw
  const zd* Uk = p.Uopt + (size_t)k * nb * nw;
  const int k2 = p.nnl[k * nn + b];
  const Win w2 = p.win[k2];
  const zd* Ukb = p.Uopt + (size_t)k2 * nb * nw;
  for (int e = t; e < w2.ndim * nw; e += kNT) {
    const int l = e % w2.ndim, j = e / w2.ndim;
    U2[l + j * ldc] = Ukb[l + (size_t)j * nb];
  }
  __syncthreads();
  const zd* Mkb = p.M + ((size_t)k * nn + b) * nb * nb + w.first + (size_t)w2.first * nb;
  zgemm_small<false, 4>(w.ndim, nw, w2.ndim, Mkb, nb, U2, ldc, [&](int i, int j, zd v) { C[i + j * ldc] = v; });
  __syncthreads();
  for (int e = t; e < w.ndim * nw; e += kNT) {
    const int l = e % w.ndim, j = e / w.ndim;
    U2[l + j * ldc] = Uk[l + (size_t)j * nb];
  }
  const double wbb = p.wb[b];
  const size_t kb = (size_t)k * nn + b;
  if (nw > w.nfroz) {
    zd* Zp = p.Zpart + kb * p.ntri;
    zpart_tiles<2>(C, ldc, w, nd, nw, wbb, Zp);
  }
  __syncthreads();
  double r = 0.0, rf = 0.0;
  zgemm_small<true, 2>(nw, nw, w.ndim, U2, ldc, C, ldc, [&](int i, int, zd v) {
    const double a2 = zabs2(v);
    r += a2;
    if (i < w.nfroz) rf += a2;
  });
  // both sums in one pass: wavefront butterflies, then the waves in order
  for (int off = 32; off > 0; off >>= 1) {
    r += __shfl_xor(r, off, 64);
    rf += __shfl_xor(rf, off, 64);
  }
  if ((t & 63) == 0) {
    red[0][t >> 6] = r;
    red[1][t >> 6] = rf;
  }
  __syncthreads();
  if (t == 0) {
    double s0 = 0.0, s1 = 0.0;
    for (int i = 0; i < kNT / 64; ++i) {
      s0 += red[0][i];
      s1 += red[1][i];
    }
    p.part[2 * kb] = wbb * s0;
    p.part[2 * kb + 1] = wbb * s1;
  }
}

// One extraction step for one k: Z of the last Z pass (its partial sums
// over b) mixed with the previous Z (Zin, packed upper triangle; the first
// iteration takes it unmixed), its eigenvectors; the leading nw - nfroz
// replace the non-frozen columns of U_opt(k) (in ascending order of
// eigenvalue) and their eigenvalues are subtracted from the Omega_I
// estimate.  From the second iteration on, the eigenproblem starts from the
// eigenvectors of the previous one, E = Vprev: A = E^+ Z E is nearly
// diagonal and the Jacobi rotations go straight into E.  One Newton-Schulz
// step E <- E (3 I - E^+ E) / 2 then keeps the product of the iterations'
// rotations orthonormal to rounding (without it the basis drifts by ~n eps
// per iteration).  LDS: A and E (nd x nd, leading dimension ldz); Z E
// passes through the Vprev area of global memory.
template <int N>
__global__ void __launch_bounds__(kNT) k_dis_update(DisProb p) {
  __shared__ JacobiLds<N> J;
  DisState* st = p.st;
  if (st->done) return;
  const int k = dis_blk<N>(p), t = threadIdx.x;
  const Win w = p.win[k];
  const int nb = p.nb, nw = p.nw, ld = p.ldz, nn = p.nn;
  const int nd = w.ndim - w.nfroz, r = nw - w.nfroz;
  double fz = 0.0;
  if (t == 0)
    for (int b = 0; b < nn; ++b) fz += p.part[2 * ((size_t)k * nn + b) + 1];
  if (r <= 0) {
    if (t == 0) p.wk1[k] = (double)nw * p.wbtot - fz;
    return;
  }
  const bool first = st->iter == 0;
  const double mix = st->mix;
  zd* Zi = p.Zin + (size_t)k * nb * nb;
  zd* Vp = p.Vprev + (size_t)k * nb * nb;
  zd* A = dis_ws<N>(p);
  zd* E = A + ld * p.nzmax;
  const zd* Zp = p.Zpart + (size_t)k * nn * p.ntri;
  for (int e = t; e < nd * nd; e += kNT) {
    const int i = e % nd, j = e / nd;
    E[i + j * ld] = first ? zmk(i == j ? 1.0 : 0.0, 0.0) : Vp[i + (size_t)j * nb];
    if (i > j) continue;
    const int x = j * (j + 1) / 2 + i;
    zd z = zmk(0.0, 0.0);
    for (int b = 0; b < nn; ++b) z = zadd(z, Zp[(size_t)b * p.ntri + x]);
    const zd v = first ? z : zadd(zscale(z, mix), zscale(Zi[x], 1.0 - mix));
    Zi[x] = v;
    A[i + j * ld] = i == j ? zmk(v.x, 0.0) : v;
    A[j + i * ld] = zconj(v);
  }
  __syncthreads();
  if (!first) {
    // Z E into the Vprev area (E holds its contents), then A = E^+ (Z E)
    zgemm<false>(nd, nd, nd, A, ld, E, ld, [&](int i, int j, zd v) { Vp[i + (size_t)j * nb] = v; });
    __threadfence_block();
    __syncthreads();
    zgemm_herm(nd, nd, E, ld, Vp, nb, [&](int i, int j, zd v) {
      A[i + j * ld] = i == j ? zmk(v.x, 0.0) : v;
      A[j + i * ld] = zconj(v);
    });
    __syncthreads();
  }
  const int sweeps = jacobi_run(A, E, nd, ld, J);
  if (t == 0) {
    if (sweeps < 0) {
      atomicOr(p.flags, F_JACOBI);
    } else {
      atomicMax(&st->sweeps_max, sweeps);
      atomicAdd(&st->sweeps_sum, (unsigned long long)sweeps);
      atomicAdd(&st->solves, 1ull);
    }
  }
  eig_order(A, nd, ld, J);
  // E^+ E into the A area (the eigenvalues are in J.lam), then the
  // Newton-Schulz step to Vprev and the selected columns to U_opt
  zgemm_herm(nd, nd, E, ld, E, ld, [&](int i, int j, zd v) {
    A[i + j * ld] = i == j ? zmk(v.x, 0.0) : v;
    A[j + i * ld] = zconj(v);
  });
  __syncthreads();
  zd* U = p.Uopt + (size_t)k * nb * nw;
  const int lead = nd - r;
  zgemm<false>(nd, nd, nd, E, ld, A, ld, [&](int i, int j, zd v) {
    const zd s = zsub(zscale(E[i + j * ld], 1.5), zscale(v, 0.5));
    Vp[i + (size_t)j * nb] = s;
    const int c = J.rank[j] - lead;
    if (c >= 0) U[nf_row(w, i) + (size_t)(w.nfroz + c) * nb] = s;
  });
  // the frozen rows of the non-frozen columns
  for (int e = t; e < w.nfroz * r; e += kNT)
    U[w.fmin + e % w.nfroz + (size_t)(w.nfroz + e / w.nfroz) * nb] = zmk(0.0, 0.0);
  if (t == 0) {
    double s = (double)nw * p.wbtot;
    if (w.nfroz > 0) s -= fz;
    for (int l = 0; l < r; ++l) s -= J.lam[J.perm[lead + l]];
    p.wk1[k] = s;
  }
}

// Omega_I estimate and Omega_I over k (per k its terms over b in order,
// then a deterministic block sum), the change and the convergence window
// (internal_test_convergence: the last conv_window changes all below
// conv_tol, once conv_window iterations are done).
__global__ void __launch_bounds__(kNT) k_dis_reduce(DisProb p) {
  __shared__ double red[kNT / 64];
  DisState* st = p.st;
  if (st->done) return;
  const int t = threadIdx.x, nn = p.nn;
  double s1 = 0.0, s2 = 0.0;
  for (int k = t; k < p.nk; k += kNT) {
    s1 += p.wk1[k];
    double om = 0.0;
    for (int b = 0; b < nn; ++b) om += p.part[2 * ((size_t)k * nn + b)];
    s2 += (double)p.nw * p.wbtot - om;
  }
  s1 = block_sum(s1, red);
  s2 = block_sum(s2, red);
  if (t != 0) return;
  const double womegai1 = s1 / p.nk, womegai = s2 / p.nk;
  const double delta = womegai1 / womegai - 1.0;
  const int iter = st->iter + 1;
  double* row = p.log + 3 * (size_t)(iter - 1);
  row[0] = womegai1;
  row[1] = womegai;
  row[2] = delta;
  st->iter = iter;
  st->womegai1 = womegai1;
  st->womegai = womegai;
  st->delta = delta;
  st->run = (fabs(delta) < st->conv_tol) ? st->run + 1 : 0;
  if (iter >= st->conv_window && st->run >= st->conv_window) st->converged = 1;
  if (st->converged || iter >= st->num_iter) st->done = 1;
}

// The Hamiltonian in the optimal subspace, H = U_opt^+ diag(e) U_opt, its
// eigenvalues into eigopt(1:nw) and U_opt <- U_opt V (ascending order).
template <int N>
__global__ void __launch_bounds__(kNT) k_dis_ham(DisProb p) {
  __shared__ JacobiLds<N> J;
  const int k = dis_blk<N>(p), t = threadIdx.x;
  const Win w = p.win[k];
  const int nb = p.nb, nw = p.nw, ld = p.ld;
  zd* U = p.Uopt + (size_t)k * nb * nw;
  const double* e = p.eig + (size_t)k * nb + w.first;
  zd* H = dis_ws<N>(p);
  zd* V = H + nb * ld;
  for (int q = t; q < nw * nw; q += kNT) {
    const int i = q % nw, j = q / nw;
    if (i > j) continue;
    zd s = zmk(0.0, 0.0);
    for (int l = 0; l < w.ndim; ++l) zcfma(s, U[l + (size_t)i * nb], zscale(U[l + (size_t)j * nb], e[l]));
    if (i == j) s.y = 0.0;
    H[i + j * ld] = s;
    H[j + i * ld] = zconj(s);
  }
  __syncthreads();
  if (jacobi_eigh(H, V, nw, ld, J) < 0 && t == 0) atomicOr(p.flags, F_JACOBI);
  eig_order(H, nw, ld, J);
  if (t < nw) p.eigopt[(size_t)k * nb + t] = J.lam[J.perm[t]];
  // U_opt V into the H area, then back
  for (int q = t; q < w.ndim * nw; q += kNT) {
    const int i = q % w.ndim, j = q / w.ndim;
    zd s = zmk(0.0, 0.0);
    const int c = J.perm[j];
    for (int l = 0; l < nw; ++l) zfma(s, U[i + (size_t)l * nb], V[l + c * ld]);
    H[i + j * ld] = s;
  }
  __syncthreads();
  for (int q = t; q < w.ndim * nw; q += kNT) {
    const int i = q % w.ndim, j = q / w.ndim;
    U[i + (size_t)j * nb] = H[i + j * ld];
  }
}

// internal_find_u: U(k) = polar factor of X = U_opt(k)^+ A(k) (X staged in
// the Zin area, free after the extraction), then W(k) = U_opt(k) U(k).
template <int N>
__global__ void __launch_bounds__(kNT) k_dis_find_u(DisProb p) {
  __shared__ JacobiLds<N> J;
  const int k = dis_blk<N>(p), t = threadIdx.x;
  const Win w = p.win[k];
  const int nb = p.nb, nw = p.nw, ld = p.ld;
  const zd* Uo = p.Uopt + (size_t)k * nb * nw;
  const zd* A = p.A + (size_t)k * nb * nw + w.first;
  zd* X = p.Zin + (size_t)k * nb * nb;  // nw x nw, leading dimension nw
  for (int q = t; q < nw * nw; q += kNT) {
    const int i = q % nw, j = q / nw;
    zd s = zmk(0.0, 0.0);
    for (int l = 0; l < w.ndim; ++l) zcfma(s, Uo[l + (size_t)i * nb], A[l + (size_t)j * nb]);
    X[i + (size_t)j * nw] = s;
  }
  __threadfence_block();
  __syncthreads();
  zd* U = p.U + (size_t)k * nw * nw;
  zd* G = dis_ws<N>(p);
  polar_factor(X, nw, nw, nw, U, nw, G, G + nb * ld, ld, J, p.flags);
  __threadfence_block();
  __syncthreads();
  zd* W = p.W + (size_t)k * nb * nw;
  for (int q = t; q < w.ndim * nw; q += kNT) {
    const int i = q % w.ndim, j = q / w.ndim;
    zd s = zmk(0.0, 0.0);
    for (int l = 0; l < nw; ++l) zfma(s, Uo[i + (size_t)l * nb], U[l + (size_t)j * nw]);
    W[i + (size_t)j * nb] = s;
  }
}

// M(k,b) <- W(k)^+ M(k,b) W(k+b) (num_wann x num_wann), one workgroup per
// (k, b).
template <int N>
__global__ void __launch_bounds__(kNT) k_dis_rotate(DisProb p) {
  const int blk = dis_blk<N>(p);
  const int k = blk / p.nn, b = blk % p.nn, t = threadIdx.x;
  const int nb = p.nb, nw = p.nw, ld = p.ld, nn = p.nn;
  const Win w = p.win[k];
  const int k2 = p.nnl[k * nn + b];
  const Win w2 = p.win[k2];
  zd* P = dis_ws<N>(p);  // ndim(k) x nw
  zd* W2 = P + nb * ld;   // ndim(k+b) x nw
  const zd* Wk = p.W + (size_t)k * nb * nw;
  const zd* Wkb = p.W + (size_t)k2 * nb * nw;
  for (int e = t; e < w2.ndim * nw; e += kNT) {
    const int l = e % w2.ndim, j = e / w2.ndim;
    W2[l + j * ld] = Wkb[l + (size_t)j * nb];
  }
  __syncthreads();
  const zd* Mkb = p.M + ((size_t)k * nn + b) * nb * nb + w.first + (size_t)w2.first * nb;
  for (int e = t; e < w.ndim * nw; e += kNT) {
    const int i = e % w.ndim, j = e / w.ndim;
    zd c = zmk(0.0, 0.0);
    for (int l = 0; l < w2.ndim; ++l) zfma(c, Mkb[i + (size_t)l * nb], W2[l + j * ld]);
    P[i + j * ld] = c;
  }
  __syncthreads();
  zd* out = p.Mw + ((size_t)k * nn + b) * nw * nw;
  for (int e = t; e < nw * nw; e += kNT) {
    const int i = e % nw, j = e / nw;
    zd v = zmk(0.0, 0.0);
    for (int l = 0; l < w.ndim; ++l) zcfma(v, Wk[l + (size_t)i * nb], P[l + j * ld]);
    out[i + (size_t)j * nw] = v;
  }
}

__global__ void k_dis_init_state(DisState* st, DisState s) { *st = s; }

// ------------------------------------------------------------ host

struct Arena {
  std::vector<size_t> off;
  size_t total = 0;
  size_t add(size_t bytes) {
    const size_t o = total;
    total += (bytes + 255) & ~(size_t)255;
    off.push_back(o);
    return o;
  }
};

#define DCHK(call)                                                                                      \
  do {                                                                                                  \
    hipError_t e_ = (call);                                                                             \
    if (e_ != hipSuccess) {                                                                             \
      w90gpu_set_error(ctx, "HIP error %s at %s:%d (%s)", hipGetErrorString(e_), __FILE__, __LINE__, #call); \
      return W90GPU_ERR_FATAL;                                                                          \
    }                                                                                                   \
  } while (0)

// dis_windows (:851-1171) without dis_spheres, on the host (each k is a
// scan over its sorted eigenvalues).
int windows(w90gpu_ctx* ctx, int nb, int nw, int nk, const w90gpu_dis_control* c, const double* eig,
            std::vector<Win>& win, std::vector<double>& eigopt, bool* linner) {
  win.assign(nk, Win{0, 0, 0, 0});
  eigopt.assign(eig, eig + (size_t)nb * nk);
  *linner = false;
  for (int k = 0; k < nk; ++k) {
    const double* e = eig + (size_t)k * nb;
    if (e[0] > c->win_max || e[nb - 1] < c->win_min) {
      w90gpu_set_error(ctx, "dis_windows: The outer energy window contains no eigenvalues (k-point %d)", k + 1);
      return W90GPU_ERR_FATAL;
    }
    int imin = 0, imax = 0;
    for (int i = 1; i <= nb; ++i) {
      if (imin == 0) {
        if (e[i - 1] >= c->win_min && e[i - 1] <= c->win_max) imin = i;
        imax = i;
      }
      if (e[i - 1] <= c->win_max) imax = i;
    }
    const int ndim = imax - imin + 1;
    if (ndim < nw) {
      w90gpu_set_error(ctx,
                       "dis_windows: Energy window contains fewer states than number of target WFs (k-point %d, "
                       "ndimwin=%d, num_wann=%d)",
                       k + 1, ndim, nw);
      return W90GPU_ERR_FATAL;
    }
    int kmin = 0, kmax = -1;
    if (c->frozen_states) {
      for (int i = imin; i <= imax; ++i) {
        if (kmin == 0) {
          if (e[i - 1] >= c->froz_min && e[i - 1] <= c->froz_max) {
            kmin = i - imin + 1;
            kmax = i - imin + 1;
          }
        } else if (e[i - 1] <= c->froz_max) {
          kmax += 1;
        }
      }
    }
    const int nfroz = kmax - kmin + 1;
    if (nfroz > nw) {
      w90gpu_set_error(ctx,
                       "dis_windows: More states in the frozen window than target WFs (k-point %d, %d frozen, "
                       "num_wann=%d)",
                       k + 1, nfroz, nw);
      return W90GPU_ERR_FATAL;
    }
    if (nfroz > 0) *linner = true;
    win[k] = Win{imin - 1, ndim, nfroz, nfroz > 0 ? kmin - 1 : 0};
    double* eo = eigopt.data() + (size_t)k * nb;
    for (int q = 0; q < ndim; ++q) eo[q] = e[imin - 1 + q];
    for (int q = ndim; q < nb; ++q) eo[q] = 0.0;
  }
  return W90GPU_OK;
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// Once per process (calls from several host threads are serialised).
int set_lds_limits(w90gpu_ctx* ctx, size_t bytes) {
  static std::mutex mu;
  static bool done = false;
  std::lock_guard<std::mutex> lock(mu);
  if (done) return W90GPU_OK;
  const void* ks[] = {(const void*)k_dis_project<kLdsN>, (const void*)k_dis_proj_froz<kLdsN>,
                      (const void*)k_dis_zmat<kLdsN>,    (const void*)k_dis_update<kLdsN>,
                      (const void*)k_dis_ham<kLdsN>,     (const void*)k_dis_find_u<kLdsN>,
                      (const void*)k_dis_rotate<kLdsN>};
  for (const void* f : ks)
    DCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  done = true;
  return W90GPU_OK;
}

// A kernel over nblocks workgroups: the LDS instantiation in one launch,
// the global-memory one in launches of at most kGwsBlocks workgroups (each
// launch reuses the workgroup areas of gws; the stream orders them).
template <int N>
void launch_blocks(void (*kern)(DisProb), int nblocks, size_t lds, hipStream_t sm, DisProb p) {
  if constexpr (N <= kLdsN) {
    hipLaunchKernelGGL(kern, dim3(nblocks), dim3(kNT), lds, sm, p);
  } else {
    for (int b0 = 0; b0 < nblocks; b0 += kGwsBlocks) {
      p.blk0 = b0;
      hipLaunchKernelGGL(kern, dim3(std::min(kGwsBlocks, nblocks - b0)), dim3(kNT), 0, sm, p);
    }
  }
}

// The instantiation that holds a kernel's matrices: LDS when they fit.
template <template <int> class K>
void launch_sized(bool global, int nblocks, size_t lds, hipStream_t sm, const DisProb& p) {
  if (global) launch_blocks<kMaxN>(K<kMaxN>::fn, nblocks, 0, sm, p);
  else launch_blocks<kLdsN>(K<kLdsN>::fn, nblocks, lds, sm, p);
}
template <int N> struct KProject { static constexpr void (*fn)(DisProb) = k_dis_project<N>; };
template <int N> struct KProjFroz { static constexpr void (*fn)(DisProb) = k_dis_proj_froz<N>; };
template <int N> struct KZmat { static constexpr void (*fn)(DisProb) = k_dis_zmat<N>; };
template <int N> struct KUpdate { static constexpr void (*fn)(DisProb) = k_dis_update<N>; };
template <int N> struct KHam { static constexpr void (*fn)(DisProb) = k_dis_ham<N>; };
template <int N> struct KFindU { static constexpr void (*fn)(DisProb) = k_dis_find_u<N>; };
template <int N> struct KRotate { static constexpr void (*fn)(DisProb) = k_dis_rotate<N>; };

}  // namespace

extern "C" {

void w90gpu_dis_control_defaults(w90gpu_dis_control* c) {
  if (!c) return;
  c->win_min = -1.7976931348623157e308;
  c->win_max = 1.7976931348623157e308;
  c->froz_min = -1.7976931348623157e308;
  c->froz_max = 1.7976931348623157e308;
  c->frozen_states = 0;
  c->num_iter = 200;
  c->mix_ratio = 0.5;
  c->conv_tol = 1e-10;
  c->conv_window = 3;
}

int w90gpu_disentangle(w90gpu_ctx* ctx, const w90gpu_problem* prob, int num_bands, const w90gpu_dis_control* ctl,
                       const double* eigval, const double* a_matrix, const double* m_matrix_orig,
                       double* u_matrix_opt, double* u_matrix, double* m_matrix, double* eigval_opt,
                       int* nfirstwin, int* ndimwin, double* history, int history_rows, w90gpu_dis_result* result) {
  if (!ctx || !prob || !ctl || !eigval || !a_matrix || !m_matrix_orig || !u_matrix_opt || !u_matrix || !m_matrix) {
    w90gpu_set_error(ctx, "w90gpu_disentangle: null argument");
    return W90GPU_ERR_INPUT;
  }
  const int nb = num_bands, nw = prob->num_wann, nk = prob->num_kpts, nn = prob->nntot;
  if (nw < 1 || nb < nw || nb > W90GPU_DIS_MAX_BANDS || nk < 1 || nn < 1 || nn > W90GPU_NUM_NNMAX ||
      !prob->nnlist || !prob->wb) {
    w90gpu_set_error(ctx, "w90gpu_disentangle: invalid sizes num_bands=%d num_wann=%d num_kpts=%d nntot=%d (num_bands "
                          "<= %d)", nb, nw, nk, nn, W90GPU_DIS_MAX_BANDS);
    return W90GPU_ERR_INPUT;
  }
  if (ctl->num_iter < 0 || !(ctl->mix_ratio > 0.0 && ctl->mix_ratio <= 1.0) || ctl->conv_tol < 0.0 ||
      ctl->conv_window < 0) {
    w90gpu_set_error(ctx, "w90gpu_disentangle: invalid dis_num_iter, dis_mix_ratio, dis_conv_tol or "
                          "dis_conv_window");
    return W90GPU_ERR_INPUT;
  }
  const auto t_start = std::chrono::steady_clock::now();
  std::vector<Win> win;
  std::vector<double> eigopt;
  bool linner = false;
  int rc = windows(ctx, nb, nw, nk, ctl, eigval, win, eigopt, &linner);
  if (rc) return rc;
  DCHK(hipSetDevice(ctx->device));
  const int ld = nb + 1;
  const size_t lds = 2 * (size_t)nb * ld * sizeof(zd);
  // the per-iteration kernels size their LDS by the windows: k_dis_zmat
  // holds two ndim x nw blocks, k_dis_update two (ndim - nfroz)^2 blocks
  // (odd leading dimensions: rows and columns of complex doubles both
  // spread over the banks)
  int ndmax = 1, nzmax = 1;
  for (const Win& wk : win) {
    ndmax = std::max(ndmax, wk.ndim);
    if (nw > wk.nfroz) nzmax = std::max(nzmax, wk.ndim - wk.nfroz);
  }
  const int ldc = ndmax | 1, ldz = nzmax | 1;
  const size_t lds_zmat = 2 * (size_t)ldc * nw * sizeof(zd), lds_update = 2 * (size_t)ldz * nzmax * sizeof(zd);
  if ((rc = set_lds_limits(ctx, kLdsBytes))) return rc;
  // matrices beyond the LDS go to per-workgroup areas of global memory:
  // those of the one-time kernels (two nb x (nb + 1) blocks) above kLdsN
  // bands, those of the extraction kernels when their window blocks do not
  // fit
  const bool g_once = nb > kLdsN, g_zmat = lds_zmat > kLdsBytes, g_update = nzmax > kLdsN;
  const size_t gws_stride = 2 * (size_t)nb * ld;
  const size_t gws_blocks = (g_once || g_zmat || g_update) ? std::min<size_t>(kGwsBlocks, (size_t)nk * nn) : 0;

  const size_t zb = sizeof(zd);
  const int nlog = std::max(ctl->num_iter, 1);
  Arena ar;
  const size_t o_nnl = ar.add(sizeof(int) * nk * nn), o_wb = ar.add(sizeof(double) * nn),
               o_win = ar.add(sizeof(Win) * nk), o_eig = ar.add(sizeof(double) * nb * nk),
               o_M = ar.add(zb * nb * nb * nn * nk), o_A = ar.add(zb * nb * nw * nk),
               o_Uopt = ar.add(zb * nb * nw * nk), o_Zin = ar.add(zb * nb * nb * nk),
               o_Zpart = ar.add(zb * (nb * (nb + 1) / 2) * nn * nk), o_Vprev = ar.add(zb * nb * nb * nk),
               o_U = ar.add(zb * nw * nw * nk), o_W = ar.add(zb * nb * nw * nk),
               o_Mw = ar.add(zb * nw * nw * nn * nk), o_wk1 = ar.add(sizeof(double) * nk),
               o_part = ar.add(sizeof(double) * 2 * nn * nk),
               o_eigopt = ar.add(sizeof(double) * nb * nk), o_log = ar.add(sizeof(double) * 3 * nlog),
               o_st = ar.add(sizeof(DisState)), o_flags = ar.add(sizeof(int)),
               o_gws = ar.add(zb * gws_stride * gws_blocks);
  DevBuf buf;
  DCHK(hipMalloc(&buf.p, ar.total));
  char* base = (char*)buf.p;
  DCHK(hipMemset(base, 0, ar.total));
  std::vector<int> nnl((size_t)nk * nn);
  for (int k = 0; k < nk; ++k)
    for (int b = 0; b < nn; ++b) {
      const int v = prob->nnlist[k + (size_t)nk * b];
      if (v < 1 || v > nk) {
        w90gpu_set_error(ctx, "w90gpu_disentangle: nnlist(%d,%d)=%d out of range", k + 1, b + 1, v);
        return W90GPU_ERR_INPUT;
      }
      nnl[(size_t)k * nn + b] = v - 1;
    }
  DCHK(hipMemcpy(base + o_nnl, nnl.data(), sizeof(int) * nnl.size(), hipMemcpyHostToDevice));
  DCHK(hipMemcpy(base + o_wb, prob->wb, sizeof(double) * nn, hipMemcpyHostToDevice));
  DCHK(hipMemcpy(base + o_win, win.data(), sizeof(Win) * nk, hipMemcpyHostToDevice));
  DCHK(hipMemcpy(base + o_eig, eigval, sizeof(double) * nb * nk, hipMemcpyHostToDevice));
  DCHK(hipMemcpy(base + o_M, m_matrix_orig, zb * nb * nb * nn * nk, hipMemcpyHostToDevice));
  DCHK(hipMemcpy(base + o_A, a_matrix, zb * nb * nw * nk, hipMemcpyHostToDevice));
  DCHK(hipMemcpy(base + o_eigopt, eigopt.data(), sizeof(double) * nb * nk, hipMemcpyHostToDevice));

  DisProb p;
  p.nb = nb;
  p.nw = nw;
  p.nk = nk;
  p.nn = nn;
  p.ld = ld;
  p.wbtot = prob->wbtot;
  p.nnl = (const int*)(base + o_nnl);
  p.wb = (const double*)(base + o_wb);
  p.win = (const Win*)(base + o_win);
  p.eig = (const double*)(base + o_eig);
  p.M = (const zd*)(base + o_M);
  p.A = (const zd*)(base + o_A);
  p.Uopt = (zd*)(base + o_Uopt);
  p.Zin = (zd*)(base + o_Zin);
  p.Zpart = (zd*)(base + o_Zpart);
  p.Vprev = (zd*)(base + o_Vprev);
  p.ntri = nb * (nb + 1) / 2;
  p.ldc = ldc;
  p.nzmax = nzmax;
  p.ldz = ldz;
  p.U = (zd*)(base + o_U);
  p.W = (zd*)(base + o_W);
  p.Mw = (zd*)(base + o_Mw);
  p.wk1 = (double*)(base + o_wk1);
  p.part = (double*)(base + o_part);
  p.eigopt = (double*)(base + o_eigopt);
  p.log = (double*)(base + o_log);
  p.st = (DisState*)(base + o_st);
  p.flags = (int*)(base + o_flags);
  p.gws = (zd*)(base + o_gws);
  p.gws_stride = gws_stride;
  p.blk0 = 0;

  DisState s0;
  std::memset(&s0, 0, sizeof(s0));
  s0.num_iter = ctl->num_iter;
  s0.conv_window = ctl->conv_window;
  s0.mix = ctl->mix_ratio;
  s0.conv_tol = ctl->conv_tol;
  s0.done = ctl->num_iter == 0 ? 1 : 0;
  hipStream_t sm = nullptr;
  hipLaunchKernelGGL(k_dis_init_state, dim3(1), dim3(1), 0, sm, p.st, s0);
  launch_sized<KProject>(g_once, nk, lds, sm, p);
  if (linner) launch_sized<KProjFroz>(g_once, nk, lds, sm, p);
  hipLaunchKernelGGL(k_dis_orthonorm, dim3(nk), dim3(kNT), 0, sm, p);
  DCHK(hipGetLastError());
  int flags = 0;
  DCHK(hipMemcpy(&flags, p.flags, sizeof(int), hipMemcpyDeviceToHost));
  if (flags & F_PROJECT_UNIT) {
    w90gpu_set_error(ctx, "dis_project: Error in unitarity of initial U in dis_project");
    return W90GPU_ERR_FATAL;
  }
  if (flags & F_FROZ_EIG) {
    w90gpu_set_error(ctx, "dis_proj_frozen: error - Eigenvalues not between 0 and 1");
    return W90GPU_ERR_FATAL;
  }
  if (flags & F_FROZ_FIX) {
    w90gpu_set_error(ctx, "dis_proj_froz: Ortho-fix failed to find enough vectors");
    return W90GPU_ERR_FATAL;
  }
  if (flags & F_ORTHONORM) {
    w90gpu_set_error(ctx, "Error in dis_main: orthonormal error (the trial orbitals for disentanglement are not "
                          "orthonormal)");
    return W90GPU_ERR_FATAL;
  }
  // the extraction: Z of the initial U_opt, then (update, Z, reduce) per
  // iteration, polled every kPoll iterations
  constexpr int kPoll = 16;
  launch_sized<KZmat>(g_zmat, nk * nn, lds_zmat, sm, p);
  DisState st;
  int launched = 0;
  while (launched < ctl->num_iter) {
    const int n = std::min(kPoll, ctl->num_iter - launched);
    for (int i = 0; i < n; ++i) {
      launch_sized<KUpdate>(g_update, nk, lds_update, sm, p);
      launch_sized<KZmat>(g_zmat, nk * nn, lds_zmat, sm, p);
      hipLaunchKernelGGL(k_dis_reduce, dim3(1), dim3(kNT), 0, sm, p);
    }
    DCHK(hipGetLastError());
    launched += n;
    DCHK(hipMemcpy(&st, p.st, sizeof(DisState), hipMemcpyDeviceToHost));
    if (st.done) break;
  }
  launch_sized<KHam>(g_once, nk, lds, sm, p);
  launch_sized<KFindU>(g_once, nk, lds, sm, p);
  launch_sized<KRotate>(g_once, nk * nn, lds, sm, p);
  DCHK(hipGetLastError());
  DCHK(hipMemcpy(&st, p.st, sizeof(DisState), hipMemcpyDeviceToHost));
  DCHK(hipMemcpy(&flags, p.flags, sizeof(int), hipMemcpyDeviceToHost));
  if (flags & F_JACOBI) {
    w90gpu_set_error(ctx, "dis_extract: Jacobi eigensolver did not converge in %d sweeps", kMaxSweeps);
    return W90GPU_ERR_FATAL;
  }
  DCHK(hipMemcpy(u_matrix_opt, p.Uopt, zb * nb * nw * nk, hipMemcpyDeviceToHost));
  DCHK(hipMemcpy(u_matrix, p.U, zb * nw * nw * nk, hipMemcpyDeviceToHost));
  DCHK(hipMemcpy(m_matrix, p.Mw, zb * nw * nw * nn * nk, hipMemcpyDeviceToHost));
  if (eigval_opt) DCHK(hipMemcpy(eigval_opt, p.eigopt, sizeof(double) * nb * nk, hipMemcpyDeviceToHost));
  if (history && history_rows > 0 && st.iter > 0)
    DCHK(hipMemcpy(history, p.log, sizeof(double) * 3 * std::min(history_rows, st.iter), hipMemcpyDeviceToHost));
  for (int k = 0; k < nk; ++k) {
    if (nfirstwin) nfirstwin[k] = win[k].first + 1;
    if (ndimwin) ndimwin[k] = win[k].ndim;
  }
  if (result) {
    result->omega_invariant = st.womegai;
    result->sweeps_max = st.sweeps_max;
    result->sweeps_mean = st.solves ? (double)st.sweeps_sum / (double)st.solves : 0.0;
    result->iterations = st.iter;
    result->converged = st.converged;
    result->status = W90GPU_OK;
    result->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  }
  return W90GPU_OK;
}

}  // extern "C"
