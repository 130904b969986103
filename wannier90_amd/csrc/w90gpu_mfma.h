// Matrix-core (MFMA) kernels of libw90gpu for num_wann above the LDS
// kernels (num_wann > 32): the M rotation M(k,b) <- E_k^+ M(k,b) E_{k+b} as
// two batched complex GEMMs on v_mfma_*_16x16x4 tiles, FP64 for the plain
// path and FP32 with FP32 accumulation for the mixed bulk phase.
//
// A workgroup computes a GT x GT tile of one product for one (problem, k,
// b): gridDim.y covers b and the tiles.  Four wavefronts hold 32 x 32 of the
// tile each, as 2 x 2 fragments of 16 x 16 with the real and imaginary parts
// in separate accumulators: a complex product is four real MFMA products.
// K runs in chunks of GK staged in LDS with coalesced global loads (planar
// re/im, padded rows).  Fragments wholly outside num_wann are skipped.
//
//   stage 1: T(k,b) = E_k^+ src(k,b)      (scratch T, storage precision)
//   stage 2: dst(k,b) = T(k,b) E_{k+b}    with the omega partial sums of the
//            tile (Im ln M_nn, |M_nn|^2, sum |M_mn|^2, off-diagonal sum)
//
// MFMA 16x16x4 lane layout (measured on gfx950): lane l holds A[l%16][l/16]
// and B[l/16][l%16]; element i of its result is D[r][l%16] with
// r = 4*(l/16) + i for f32 and r = l/16 + 4*i for f64.
//
// Reference: src/wannierise.F90 internal_new_u_and_m :1737-1748 (the two
// utility_zgemm calls per (k,b)), wann_omega :2034-2090, :2214-2258.
#pragma once

#include "w90gpu_kernels.h"

#define W90_GT 64  // output tile edge of a workgroup
#define W90_GK 16  // K chunk staged in LDS

typedef double w90_d4 __attribute__((ext_vector_type(4)));
typedef float w90_f4 __attribute__((ext_vector_type(4)));

template <class R>
struct Mfma;
template <>
struct Mfma<double> {
  typedef w90_d4 acc;
  static __device__ __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row of result element i in lane group lk = l/16
  static __device__ __forceinline__ int row(int lk, int i) { return lk + 4 * i; }
};
template <>
struct Mfma<float> {
  typedef w90_f4 acc;
  static __device__ __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int i) { return 4 * lk + i; }
};

template <class T2>
struct Real;
template <>
struct Real<cplx> {
  typedef double type;
};
template <>
struct Real<float2> {
  typedef float type;
};

// tiles of W90_GT per problem
__host__ __device__ __forceinline__ int gemm_tiles(int nw) {
  const int t = (nw + W90_GT - 1) / W90_GT;
  return t * t;
}

template <class T2, int STAGE>
__global__ void __launch_bounds__(256) k_rot_gemm(const Prob* probs, const Item* items, int nitems, int mode) {
  typedef typename Real<T2>::type R;
  typedef typename Mfma<R>::acc Acc;
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const State* st = p.st;
  if (!gate(st, mode)) return;
  const int nw = p.nw, nn = p.nn;
  const int nt1 = (nw + W90_GT - 1) / W90_GT;
  const int ntl = nt1 * nt1;
  const int by = blockIdx.y;
  if (by >= nn * ntl) return;
  const int b = by / ntl, tile = by % ntl;
  const int row0 = (tile % nt1) * W90_GT, col0 = (tile / nt1) * W90_GT;
  const int k = it.k;
  const size_t kb = (size_t)k * nn + b;
  const size_t nw2 = p.nw2;
  const int cur = st->cur;
  const int dst = dst_buf(mode, cur);
  T2* Tm = reinterpret_cast<T2*>(p.Tm) + kb * nw2;
  const cplx* Ek = p.E + (size_t)k * nw2;
  const cplx* Ekb = p.E + (size_t)p.nnl[kb] * nw2;
  const T2* src = rot_src<T2>(p, mode, cur) + kb * nw2;

  __shared__ R As[2][W90_GK][W90_GT + 1];  // [re/im][k][row]
  __shared__ R Bs[2][W90_GK][W90_GT + 1];  // [re/im][k][col]
  __shared__ double red[256];

  const int tid = threadIdx.x;
  const int w = tid >> 6, lane = tid & 63;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  bool fv_r[2], fv_c[2];
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    fv_r[f] = row0 + wr + f * 16 < nw;
    fv_c[f] = col0 + wc + f * 16 < nw;
  }
  Acc Cr[2][2], Ci[2][2];
#pragma unroll
  for (int fi = 0; fi < 2; ++fi)
#pragma unroll
    for (int fj = 0; fj < 2; ++fj)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        Cr[fi][fj][i] = 0;
        Ci[fi][fj][i] = 0;
      }

  for (int kk = 0; kk < nw; kk += W90_GK) {
    __syncthreads();
    // A chunk: rows row0.., k = kk..kk+GK
    for (int e = tid; e < W90_GK * W90_GT; e += 256) {
      cplx v;
      int l, i;
      if (STAGE == 1) {
        // A = E_k^+: A[i][l] = conj(E_k[l + i nw]); threads run along l
        l = e % W90_GK;
        i = e / W90_GK;
        const int gi = row0 + i, gl = kk + l;
        v = (gi < nw && gl < nw) ? Ek[gl + (size_t)gi * nw] : c_make(0.0, 0.0);
        v.y = -v.y;
      } else {
        // A = T: A[i][l] = T[i + l nw]; threads run along i
        i = e % W90_GT;
        l = e / W90_GT;
        const int gi = row0 + i, gl = kk + l;
        v = (gi < nw && gl < nw) ? c_wide(Tm[gi + (size_t)gl * nw]) : c_make(0.0, 0.0);
      }
      As[0][l][i] = (R)v.x;
      As[1][l][i] = (R)v.y;
    }
    // B chunk: B[l][j] = Bsrc[l + j nw]; threads run along l
    for (int e = tid; e < W90_GK * W90_GT; e += 256) {
      const int l = e % W90_GK, j = e / W90_GK;
      const int gl = kk + l, gj = col0 + j;
      cplx v = c_make(0.0, 0.0);
      if (gl < nw && gj < nw) v = (STAGE == 1) ? c_wide(src[gl + (size_t)gj * nw]) : Ekb[gl + (size_t)gj * nw];
      Bs[0][l][j] = (R)v.x;
      Bs[1][l][j] = (R)v.y;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < W90_GK; ks += 4) {
      R ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        ar[f] = As[0][ks + lk][wr + f * 16 + lr];
        ai[f] = As[1][ks + lk][wr + f * 16 + lr];
        br[f] = Bs[0][ks + lk][wc + f * 16 + lr];
        bi[f] = Bs[1][ks + lk][wc + f * 16 + lr];
      }
#pragma unroll
      for (int fi = 0; fi < 2; ++fi) {
        if (!fv_r[fi]) continue;
#pragma unroll
        for (int fj = 0; fj < 2; ++fj) {
          if (!fv_c[fj]) continue;
          Cr[fi][fj] = Mfma<R>::mma(ar[fi], br[fj], Cr[fi][fj]);
          Cr[fi][fj] = Mfma<R>::mma(-ai[fi], bi[fj], Cr[fi][fj]);
          Ci[fi][fj] = Mfma<R>::mma(ar[fi], bi[fj], Ci[fi][fj]);
          Ci[fi][fj] = Mfma<R>::mma(ai[fi], br[fj], Ci[fi][fj]);
        }
      }
    }
  }

  double od = 0.0, all = 0.0;
  T2* out = (STAGE == 1) ? Tm : mbuf<T2>(p, dst) + kb * nw2;
#pragma unroll
  for (int fi = 0; fi < 2; ++fi)
#pragma unroll
    for (int fj = 0; fj < 2; ++fj)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = row0 + wr + fi * 16 + Mfma<R>::row(lk, i);
        const int c = col0 + wc + fj * 16 + lr;
        if (r < nw && c < nw) {
          T2 v;
          v.x = Cr[fi][fj][i];
          v.y = Ci[fi][fj][i];
          out[r + (size_t)c * nw] = v;
          if (STAGE == 2) {
            const cplx vw = c_wide(v);
            const double a2 = c_abs2(vw);
            all += a2;
            if (r != c) {
              od += a2;
            } else {
              p.ln[dst][kb * nw + r] = atan2(vw.y, vw.x);
              p.d2[kb * nw + r] = a2;
            }
          }
        }
      }
  if (STAGE == 2) {
    od = block_sum(od, red);
    all = block_sum(all, red);
    if (tid == 0) {
      p.part_od[kb * p.npan + tile] = od;
      p.part_all[kb * p.npan + tile] = all;
    }
  }
}

// Omega partial sums of the current M by GEMM tile (MODE_INIT of the MFMA
// class): the layout the stage-2 epilogue writes, without a rotation.
template <class T2>
__global__ void __launch_bounds__(256) k_reduce_tiles(const Prob* probs, const Item* items, int nitems, int mode) {
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const State* st = p.st;
  if (!gate(st, mode)) return;
  const int nw = p.nw, nn = p.nn;
  const int nt1 = (nw + W90_GT - 1) / W90_GT;
  const int ntl = nt1 * nt1;
  const int by = blockIdx.y;
  if (by >= nn * ntl) return;
  const int b = by / ntl, tile = by % ntl;
  const int row0 = (tile % nt1) * W90_GT, col0 = (tile / nt1) * W90_GT;
  const size_t kb = (size_t)it.k * nn + b;
  const int cur = st->cur;
  const int dst = dst_buf(mode, cur);
  const T2* m = mbuf<T2>(p, cur) + kb * p.nw2;
  __shared__ double red[256];
  double od = 0.0, all = 0.0;
  for (int e = threadIdx.x; e < W90_GT * W90_GT; e += blockDim.x) {
    const int r = row0 + e % W90_GT, c = col0 + e / W90_GT;
    if (r >= nw || c >= nw) continue;
    const cplx v = c_wide(m[r + (size_t)c * nw]);
    const double a2 = c_abs2(v);
    all += a2;
    if (r != c) {
      od += a2;
    } else {
      p.ln[dst][kb * nw + r] = atan2(v.y, v.x);
      p.d2[kb * nw + r] = a2;
    }
  }
  od = block_sum(od, red);
  all = block_sum(all, red);
  if (threadIdx.x == 0) {
    p.part_od[kb * p.npan + tile] = od;
    p.part_all[kb * p.npan + tile] = all;
  }
}

// ---------------------------------------------------------- retraction

// Square complex matrices of edge W90_STEP_N in LDS for k_step_mfma:
// planar re/im, column-major with leading dimension W90_STEP_N + 1 (MFMA
// fragment loads of A along a column and of B along a row are both
// conflict-free with this padding).
#define W90_STEP_N 64
#define W90_STEP_LD (W90_STEP_N + 1)
#define W90_STEP_MAT (W90_STEP_N * W90_STEP_LD)

// LDS bytes of k_step_mfma: X and P (re, im each) and a reduction buffer.
__host__ __device__ constexpr size_t step_mfma_lds() { return sizeof(double) * (4 * W90_STEP_MAT + 256); }

// C = A B for the 64 x 64 LDS matrices (A, B planar at Ar/Ai, Br/Bi): each
// wavefront accumulates its 32 x 32 quadrant as 2 x 2 FP64 MFMA fragments;
// fragments wholly outside n are skipped (their rows and columns are zero).
__device__ __forceinline__ void step_gemm(w90_d4 (&Cr)[2][2], w90_d4 (&Ci)[2][2], const double* Ar,
                                          const double* Ai, const double* Br, const double* Bi, int n) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int fi = 0; fi < 2; ++fi)
#pragma unroll
    for (int fj = 0; fj < 2; ++fj)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        Cr[fi][fj][i] = 0.0;
        Ci[fi][fj][i] = 0.0;
      }
  const bool vr0 = wr < n, vr1 = wr + 16 < n, vc0 = wc < n, vc1 = wc + 16 < n;
  for (int ks = 0; ks < n; ks += 4) {
    double ar[2], ai[2], br[2], bi[2];
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      const int ia = (ks + lk) * W90_STEP_LD + wr + f * 16 + lr;  // A[row][ks+lk]
      const int ib = (wc + f * 16 + lr) * W90_STEP_LD + ks + lk;  // B[ks+lk][col]
      ar[f] = Ar[ia];
      ai[f] = Ai[ia];
      br[f] = Br[ib];
      bi[f] = Bi[ib];
    }
#pragma unroll
    for (int fi = 0; fi < 2; ++fi) {
      if (!(fi ? vr1 : vr0)) continue;
#pragma unroll
      for (int fj = 0; fj < 2; ++fj) {
        if (!(fj ? vc1 : vc0)) continue;
        Cr[fi][fj] = Mfma<double>::mma(ar[fi], br[fj], Cr[fi][fj]);
        Cr[fi][fj] = Mfma<double>::mma(-ai[fi], bi[fj], Cr[fi][fj]);
        Ci[fi][fj] = Mfma<double>::mma(ar[fi], bi[fj], Ci[fi][fj]);
        Ci[fi][fj] = Mfma<double>::mma(ai[fi], br[fj], Ci[fi][fj]);
      }
    }
  }
}

// Position (row, column) of accumulator element i of fragment (fi, fj).
__device__ __forceinline__ void step_pos(int fi, int fj, int i, int* r, int* c) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  *r = (w >> 1) * 32 + fi * 16 + Mfma<double>::row(lane >> 4, i);
  *c = (w & 1) * 32 + fj * 16 + (lane & 15);
}

// Retraction of one (problem, k) for 32 < num_wann <= 64, every matrix in
// LDS and every product on FP64 MFMA: forms D (trial step), X = scale D,
// E = exp(X) by the scaling and squaring of the Taylor polynomial that
// k_step uses (same plan), and U_new = U E.  The exponential of the
// anti-Hermitian X is the unitary internal_new_u_and_m builds from the
// eigenvectors of i X (wannierise.F90:1662-1689, 1729-1734).
__global__ void __launch_bounds__(256) k_step_mfma(const Prob* probs, const Item* items, int nitems, int mode) {
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const State* st = p.st;
  if (!gate(st, mode)) return;
  const int nw = p.nw, nw2 = p.nw2;
  const int k = it.k;
  const int tid = threadIdx.x;
  const int cur = st->cur;
  extern __shared__ double lds_step[];
  double* Xr = lds_step;
  double* Xi = Xr + W90_STEP_MAT;
  double* Pr = Xi + W90_STEP_MAT;
  double* Pi = Pr + W90_STEP_MAT;
  double* red = Pi + W90_STEP_MAT;
  cplx* Dk = p.D + (size_t)k * nw2;
  const double scale = st->scale;
  // X = scale D (zero outside nw), D formed on the trial step
  for (int e = tid; e < W90_STEP_N * W90_STEP_N; e += 256) {
    const int r = e % W90_STEP_N, c = e / W90_STEP_N;
    cplx x = c_make(0.0, 0.0);
    if (r < nw && c < nw) {
      const int g = r + c * nw;
      cplx d = Dk[g];
      if (mode == MODE_TRIAL) {
        d = c_scale(c_add(p.G[(size_t)k * nw2 + g], c_scale(d, st->beta)), st->dir_sign);
        Dk[g] = d;
      }
      x = c_scale(d, scale);
    }
    Xr[c * W90_STEP_LD + r] = x.x;
    Xi[c * W90_STEP_LD + r] = x.y;
  }
  __syncthreads();
  double cs = 0.0;
  for (int c = tid; c < nw; c += 256) {
    double s = 0.0;
    for (int r = 0; r < nw; ++r) s += sqrt(c_abs2(c_make(Xr[c * W90_STEP_LD + r], Xi[c * W90_STEP_LD + r])));
    cs = fmax(cs, s);
  }
  int sq, m;
  taylor_plan(block_max(cs, red), &sq, &m);
  const double f = ldexp(1.0, -sq);
  const double invm = 1.0 / (double)m;
  for (int e = tid; e < W90_STEP_N * W90_STEP_N; e += 256) {
    const int r = e % W90_STEP_N, c = e / W90_STEP_N;
    const int q = c * W90_STEP_LD + r;
    const double xr = Xr[q] * f, xi = Xi[q] * f;
    Xr[q] = xr;
    Xi[q] = xi;
    Pr[q] = xr * invm + (r == c ? 1.0 : 0.0);
    Pi[q] = xi * invm;
  }
  __syncthreads();
  w90_d4 Cr[2][2], Ci[2][2];
  // Horner: P <- I + X P / j
  for (int j = m - 1; j >= 1; --j) {
    step_gemm(Cr, Ci, Xr, Xi, Pr, Pi, nw);
    __syncthreads();
    const double invj = 1.0 / (double)j;
#pragma unroll
    for (int fi = 0; fi < 2; ++fi)
#pragma unroll
      for (int fj = 0; fj < 2; ++fj)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          int r, c;
          step_pos(fi, fj, i, &r, &c);
          const int q = c * W90_STEP_LD + r;
          Pr[q] = Cr[fi][fj][i] * invj + (r == c ? 1.0 : 0.0);
          Pi[q] = Ci[fi][fj][i] * invj;
        }
    __syncthreads();
  }
  // squarings: P <- P P
  for (int s = 0; s < sq; ++s) {
    step_gemm(Cr, Ci, Pr, Pi, Pr, Pi, nw);
    __syncthreads();
#pragma unroll
    for (int fi = 0; fi < 2; ++fi)
#pragma unroll
      for (int fj = 0; fj < 2; ++fj)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          int r, c;
          step_pos(fi, fj, i, &r, &c);
          const int q = c * W90_STEP_LD + r;
          Pr[q] = Cr[fi][fj][i];
          Pi[q] = Ci[fi][fj][i];
        }
    __syncthreads();
  }
  // E = P to global; U into the X slot
  cplx* Ek = p.E + (size_t)k * nw2;
  const cplx* Uc = p.U[cur] + (size_t)k * nw2;
  for (int e = tid; e < nw2; e += 256) {
    const int r = e % nw, c = e / nw;
    const int q = c * W90_STEP_LD + r;
    Ek[e] = c_make(Pr[q], Pi[q]);
    const cplx u = Uc[e];
    Xr[q] = u.x;
    Xi[q] = u.y;
  }
  __syncthreads();
  step_gemm(Cr, Ci, Xr, Xi, Pr, Pi, nw);
  cplx* Un = p.U[1 - cur] + (size_t)k * nw2;
#pragma unroll
  for (int fi = 0; fi < 2; ++fi)
#pragma unroll
    for (int fj = 0; fj < 2; ++fj)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        int r, c;
        step_pos(fi, fj, i, &r, &c);
        if (r < nw && c < nw) Un[r + c * nw] = c_make(Cr[fi][fj][i], Ci[fi][fj][i]);
      }
}

// C = A B for n x n column-major complex matrices in global memory, held by
// one workgroup: each wavefront takes 16 x 16 output tiles in turn, two at
// a time (independent accumulator chains), on FP64 MFMA with the operand
// fragments loaded straight from memory (the matrices of one k stay in the
// caches: A fragments read along a column, B fragments reuse cache lines
// over consecutive K steps).  Used by the retraction above num_wann 64.
__device__ void wg_gemm_mfma(cplx* C, const cplx* A, const cplx* B, int n) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int nt = (n + 15) / 16;
  const int ntt = nt * nt;
  for (int t0 = 2 * w; t0 < ntt; t0 += 2 * nwv) {
    const bool two = t0 + 1 < ntt;
    int i0[2], j0[2];
    w90_d4 cr[2], ci[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int t = two ? t0 + u : t0;
      i0[u] = (t % nt) * 16;
      j0[u] = (t / nt) * 16;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        cr[u][i] = 0.0;
        ci[u][i] = 0.0;
      }
    }
    for (int ks = 0; ks < n; ks += 4) {
      const int kk = ks + lk;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int r = i0[u] + lr, c = j0[u] + lr;
        const cplx a = (r < n && kk < n) ? A[r + (size_t)kk * n] : c_make(0.0, 0.0);
        const cplx b = (kk < n && c < n) ? B[kk + (size_t)c * n] : c_make(0.0, 0.0);
        cr[u] = Mfma<double>::mma(a.x, b.x, cr[u]);
        cr[u] = Mfma<double>::mma(-a.y, b.y, cr[u]);
        ci[u] = Mfma<double>::mma(a.x, b.y, ci[u]);
        ci[u] = Mfma<double>::mma(a.y, b.x, ci[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (u == 1 && !two) break;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = i0[u] + Mfma<double>::row(lk, i), c = j0[u] + lr;
        if (r < n && c < n) C[r + (size_t)c * n] = c_make(cr[u][i], ci[u][i]);
      }
    }
  }
  __syncthreads();
}
