// Matrix-core (MFMA) kernels of libw90gpu: the M rotation
// M(k,b) <- E_k^+ M(k,b) E_{k+b} as two complex products on
// v_mfma_*_16x16x4 tiles (FP64 for the plain path, FP32 with FP32
// accumulation for the mixed bulk phase), and the products of the
// retraction above num_wann 16.
//
// num_wann <= 16 (k_rotate_w16): one 16 x 16 tile per block, one wavefront
// per neighbour b.  num_wann 17-64 (k_rot_fused): a workgroup computes both
// products of one (k, b) block as one GT x GT tile (GT = 32, 48 or 64), T
// kept in LDS between them.  num_wann 65-256 (k_rot_gemm): each product is
// its own launch over up to 4 x 4 tiles of 64 (or 8 x 8 of 32), T through
// a scratch buffer.  Four wavefronts hold a quadrant of a tile each (three
// hold a 48 x 16 strip of a 48 tile), as 16 x 16 fragments; a complex
// fragment product is three real MFMA products per K step (Cfrag, Gauss's
// form).  K runs in chunks staged in LDS with
// coalesced global loads (planar re/im, padded rows), the next chunk
// loaded into registers while the current one is multiplied.  Fragments
// wholly outside num_wann are skipped.  The second product's epilogue
// forms the omega partial sums of the tile (Im ln M_nn, |M_nn|^2,
// sum |M_mn|^2, off-diagonal sum).
//
// MFMA 16x16x4 lane layout (measured on gfx950): lane l holds A[l%16][l/16]
// and B[l/16][l%16]; element i of its result is D[r][l%16] with
// r = 4*(l/16) + i for f32 and r = l/16 + 4*i for f64.
//
// Reference: src/wannierise.F90 internal_new_u_and_m :1737-1748 (the two
// utility_zgemm calls per (k,b)), wann_omega :2034-2090, :2214-2258.
#pragma once

#include "w90gpu_kernels.h"

#ifndef W90_GK
#define W90_GK 16  // K chunk staged in LDS
#endif

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

// Complex 16 x 16 fragment product accumulated with three real MFMA
// products per K step (Gauss): P1 += ar br, P2 += ai bi,
// P3 += (ar + ai)(br + bi); the product is Re = P1 - P2,
// Im = P3 - P1 - P2.  Three MFMA instructions instead of four.
template <class R>
struct Cfrag {
  typename Mfma<R>::acc p1, p2, p3;
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      p1[i] = 0;
      p2[i] = 0;
      p3[i] = 0;
    }
  }
  __device__ __forceinline__ void mma(R ar, R ai, R br, R bi) {
    p1 = Mfma<R>::mma(ar, br, p1);
    p2 = Mfma<R>::mma(ai, bi, p2);
    p3 = Mfma<R>::mma(ar + ai, br + bi, p3);
  }
  __device__ __forceinline__ R re(int i) const { return p1[i] - p2[i]; }
  __device__ __forceinline__ R im(int i) const { return p3[i] - p1[i] - p2[i]; }
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


// C = A B for the LDS retraction (k_step_k): n x n column-major complex
// matrices in LDS, leading dimension n + 1, n <= NWM.  The product is
// (NWM/16)^2 tiles of 16 x 16; each (tile, real or imaginary part) is one
// job of a wavefront: K in steps of 4 on FP64 MFMA 16x16x4, operand
// fragments read from LDS (zero outside n), the result part written to its
// half of the complex elements of C.  With ident, C = I + scale A B (one
// Taylor step, applied as the product is stored).  C must not overlap A or
// B; the workgroup is synchronised before return.
template <int NWM>
__device__ void lds_gemm_mfma(cplx* C, const cplx* A, const cplx* B, int n, double scale, bool ident) {
  static_assert(NWM % 16 == 0, "whole 16 x 16 tiles");
  constexpr int TN = NWM / 16;
  constexpr int JOBS = 2 * TN * TN;
  // the job index lives in a scalar register, and operands outside n are
  // read through a clamped index and a select: no lane branches in the K
  // loop, so its LDS reads are issued together
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int ld = n + 1;
  const int nwaves = blockDim.x >> 6;
  double* Cd = reinterpret_cast<double*>(C);
  for (int job = wave; job < JOBS; job += nwaves) {
    const int part = job & 1;  // 0: real part, 1: imaginary part
    const int t = job >> 1;
    const int ti = t % TN, tj = t / TN;
    const int i = 16 * ti + lr, j = 16 * tj + lr;
    double ax[NWM / 4], ay[NWM / 4], bx[NWM / 4], by[NWM / 4];
#pragma unroll
    for (int s = 0; s < NWM / 4; ++s) {
      const int l = 4 * s + lk;
      const bool oka = i < n && l < n, okb = l < n && j < n;
      const cplx a = A[oka ? i + l * ld : 0];
      const cplx bv = B[okb ? l + j * ld : 0];
      ax[s] = oka ? a.x : 0.0;
      ay[s] = oka ? a.y : 0.0;
      bx[s] = okb ? bv.x : 0.0;
      by[s] = okb ? bv.y : 0.0;
    }
    Mfma<double>::acc acc = {0.0, 0.0, 0.0, 0.0};
    if (part == 0) {
#pragma unroll
      for (int s = 0; s < NWM / 4; ++s) {
        acc = Mfma<double>::mma(ax[s], bx[s], acc);
        acc = Mfma<double>::mma(-ay[s], by[s], acc);
      }
    } else {
#pragma unroll
      for (int s = 0; s < NWM / 4; ++s) {
        acc = Mfma<double>::mma(ax[s], by[s], acc);
        acc = Mfma<double>::mma(ay[s], bx[s], acc);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = 16 * ti + Mfma<double>::row(lk, q), c = 16 * tj + lr;
      if (r < n && c < n) {
        double v = acc[q];
        if (ident) {
          v = v * scale;
          if (part == 0 && r == c) v += 1.0;
        }
        Cd[2 * (r + c * ld) + part] = v;
      }
    }
  }
  __syncthreads();
}

// output tile edge of a workgroup: 32 up to num_wann 32, 64 above
__host__ __device__ __forceinline__ int gemm_edge(int nw) { return nw <= 32 ? 32 : 64; }

// GEMM tiles per block M(k,b)
__host__ __device__ __forceinline__ int gemm_tiles(int nw) {
  const int t = (nw + gemm_edge(nw) - 1) / gemm_edge(nw);
  return t * t;
}

// One GT x GT output tile (row0, col0) of one stage of the rotation of
// block (k, b) by the workgroup (256 threads): operands and scratch as in
// k_rot_gemm, LDS passed in.
template <class T2, int STAGE, int GT>
__device__ __forceinline__ void rot_gemm_tile(const Prob& p, int k, int b, int tile, int row0, int col0, int mode,
                                              int cur, typename Real<T2>::type (*As)[W90_GK][GT + 1],
                                              typename Real<T2>::type (*Bs)[W90_GK][GT + 1], double* red) {
  constexpr int FR = GT / 32;  // fragments per wavefront along rows and columns
  typedef typename Real<T2>::type R;
  const int nw = p.nw, nn = p.nn;
  const size_t kb = (size_t)k * nn + b;
  const size_t nw2 = p.nw2;
  const int dst = dst_buf(mode, cur);
  T2* Tm = reinterpret_cast<T2*>(p.Tm) + kb * nw2;
  const cplx* Ek = p.E + (size_t)k * nw2;
  const cplx* Ekb = p.E + (size_t)p.nnl[kb] * nw2;
  const T2* src = rot_src<T2>(p, mode, cur) + kb * nw2;

  const int tid = threadIdx.x;
  const int w = tid >> 6, lane = tid & 63;
  const int wr = (w >> 1) * (GT / 2), wc = (w & 1) * (GT / 2);
  const int lr = lane & 15, lk = lane >> 4;
  bool fv_r[FR], fv_c[FR];
#pragma unroll
  for (int f = 0; f < FR; ++f) {
    fv_r[f] = row0 + wr + f * 16 < nw;
    fv_c[f] = col0 + wc + f * 16 < nw;
  }
  Cfrag<R> C[FR][FR];
#pragma unroll
  for (int fi = 0; fi < FR; ++fi)
#pragma unroll
    for (int fj = 0; fj < FR; ++fj) C[fi][fj].zero();

  // K chunks are software-pipelined: the operands of chunk kk + GK are
  // loaded into registers before the products of chunk kk, and stored to
  // LDS after them.  Thread slot e = tid + 256 u of a chunk.
  constexpr int NE = W90_GK * GT / 256;
  static_assert(NE * 256 == W90_GK * GT, "whole chunk per thread slot");
  cplx pa[NE], pb[NE];
  auto load_chunk = [&](int kk) {
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      const int e = tid + 256 * u;
      cplx v = c_make(0.0, 0.0);
      if (STAGE == 1) {
        // A = E_k^+: A[i][l] = conj(E_k[l + i nw]); threads run along l
        const int gi = row0 + e / W90_GK, gl = kk + e % W90_GK;
        if (gi < nw && gl < nw) v = Ek[gl + (size_t)gi * nw];
        v.y = -v.y;
      } else {
        // A = T: A[i][l] = T[i + l nw]; threads run along i
        const int gi = row0 + e % GT, gl = kk + e / GT;
        if (gi < nw && gl < nw) v = c_wide(Tm[gi + (size_t)gl * nw]);
      }
      pa[u] = v;
      // B[l][j] = Bsrc[l + j nw]; threads run along l
      const int gl = kk + e % W90_GK, gj = col0 + e / W90_GK;
      cplx w2 = c_make(0.0, 0.0);
      if (gl < nw && gj < nw) w2 = (STAGE == 1) ? c_wide(src[gl + (size_t)gj * nw]) : Ekb[gl + (size_t)gj * nw];
      pb[u] = w2;
    }
  };
  load_chunk(0);
  for (int kk = 0; kk < nw; kk += W90_GK) {
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      const int e = tid + 256 * u;
      const int la = (STAGE == 1) ? e % W90_GK : e / GT, ia = (STAGE == 1) ? e / W90_GK : e % GT;
      As[0][la][ia] = (R)pa[u].x;
      As[1][la][ia] = (R)pa[u].y;
      const int lb = e % W90_GK, jb = e / W90_GK;
      Bs[0][lb][jb] = (R)pb[u].x;
      Bs[1][lb][jb] = (R)pb[u].y;
    }
    if (kk + W90_GK < nw) load_chunk(kk + W90_GK);
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < W90_GK; ks += 4) {
      R ar[FR], ai[FR], br[FR], bi[FR];
#pragma unroll
      for (int f = 0; f < FR; ++f) {
        ar[f] = As[0][ks + lk][wr + f * 16 + lr];
        ai[f] = As[1][ks + lk][wr + f * 16 + lr];
        br[f] = Bs[0][ks + lk][wc + f * 16 + lr];
        bi[f] = Bs[1][ks + lk][wc + f * 16 + lr];
      }
#pragma unroll
      for (int fi = 0; fi < FR; ++fi) {
        if (!fv_r[fi]) continue;
#pragma unroll
        for (int fj = 0; fj < FR; ++fj) {
          if (!fv_c[fj]) continue;
          C[fi][fj].mma(ar[fi], ai[fi], br[fj], bi[fj]);
        }
      }
    }
  }

  double od = 0.0, all = 0.0;
  T2* out = (STAGE == 1) ? Tm : mbuf<T2>(p, dst) + kb * nw2;
#pragma unroll
  for (int fi = 0; fi < FR; ++fi)
#pragma unroll
    for (int fj = 0; fj < FR; ++fj)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = row0 + wr + fi * 16 + Mfma<R>::row(lk, i);
        const int c = col0 + wc + fj * 16 + lr;
        if (r < nw && c < nw) {
          T2 v;
          v.x = C[fi][fj].re(i);
          v.y = C[fi][fj].im(i);
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
    block_sum2(od, all, red);
    if (tid == 0) {
      p.part_od[kb * p.npan + tile] = od;
      p.part_all[kb * p.npan + tile] = all;
    }
  }
}

// Two-stage rotation, one launch per stage: workgroup (item, b, tile).
template <class T2, int STAGE, int GT>
__global__ void __launch_bounds__(256, 2) k_rot_gemm(const Prob* probs, const Item* items, int nitems, int mode) {
  typedef typename Real<T2>::type R;
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const State* st = p.st;
  const int cur = gate_cur(st, mode);
  if (cur < 0) return;
  const int nw = p.nw, nn = p.nn;
  const int nt1 = (nw + GT - 1) / GT;
  const int ntl = nt1 * nt1;
  const int by = blockIdx.y;
  if (by >= nn * ntl) return;
  const int b = by / ntl, tile = by % ntl;
  const int row0 = (tile % nt1) * GT, col0 = (tile / nt1) * GT;
  __shared__ R As[2][W90_GK][GT + 1];  // [re/im][k][row]
  __shared__ R Bs[2][W90_GK][GT + 1];  // [re/im][k][col]
  __shared__ double red[256];
  rot_gemm_tile<T2, STAGE, GT>(p, it.k, b, tile, row0, col0, mode, cur, As, Bs, red);
}

// K chunk of the second product of k_rot_fused (T held in LDS).
#define W90_GK2 8

// Wavefront layout of k_rot_fused by tile edge GT: WR x WC wavefronts, each
// holding (GT / 16 WR) x (GT / 16 WC) fragments of 16 x 16.  A 64 or 32
// tile is split in quadrants over four wavefronts; a 48 tile (num_wann up
// to 48: 3 x 3 fragments, which four wavefronts cannot share evenly) in
// three column strips of 48 x 16 over three wavefronts.
template <int GT>
struct FusedWaves {
  static constexpr int WR = GT == 48 ? 1 : 2, WC = GT == 48 ? 3 : 2;
  static constexpr int NT = 64 * WR * WC;
  static constexpr int FRR = GT / (16 * WR), FRC = GT / (16 * WC);
  static constexpr int MINB = GT == 48 ? 3 : 2;  // workgroups per CU the registers allow
};

// LDS elements (storage-precision reals) of k_rot_fused: the staging of the
// first product (A and B chunks) and, after it, T with the B chunks of the
// second product share one buffer.
template <class R, int GT>
__host__ __device__ constexpr int rot_fused_elems() {
  return (4 * W90_GK * (GT + 1)) > (2 * GT * (GT + 1) + 2 * W90_GK2 * (GT + 1))
             ? 4 * W90_GK * (GT + 1)
             : 2 * GT * (GT + 1) + 2 * W90_GK2 * (GT + 1);
}

// Both products of the rotation of block (k, b) by one workgroup, when one
// GT x GT tile covers the block (num_wann <= GT).  T = E_k^+ src is formed
// as in k_rot_gemm stage 1, then written from the accumulators into LDS
// (planar, k-major: T[i][l] at Ts[p][l][i], zero outside num_wann) over the
// staging buffers of the first product; R = T E_{k+b} takes its A
// fragments from there and its B operand in chunks of W90_GK2 rows of
// E_{k+b}, loaded one chunk ahead.  The epilogue is that of stage 2.
template <class T2, int GT>
__device__ __forceinline__ void rot_fused_tile(const Prob& p, int k, int b, int mode, int cur,
                                               typename Real<T2>::type* lds, double* red) {
  typedef FusedWaves<GT> WL;
  constexpr int FRR = WL::FRR, FRC = WL::FRC, FRM = FRR > FRC ? FRR : FRC, NT = WL::NT;
  constexpr int LDA = GT + 1;
  typedef typename Real<T2>::type R;
  const int nw = p.nw, nn = p.nn;
  const size_t kb = (size_t)k * nn + b;
  const size_t nw2 = p.nw2;
  const int dst = dst_buf(mode, cur);
  const cplx* Ek = p.E + (size_t)k * nw2;
  const cplx* Ekb = p.E + (size_t)p.nnl[kb] * nw2;
  const T2* src = rot_src<T2>(p, mode, cur) + kb * nw2;
  T2* out = mbuf<T2>(p, dst) + kb * nw2;
  double* lnd = p.ln[dst] + kb * nw;
  double* d2d = p.d2 + kb * nw;
  double* pod = p.part_od + kb * p.npan;
  double* pall = p.part_all + kb * p.npan;
  R* As = lds;                     // [2][GK][LDA]   first product
  R* Bs = lds + 2 * W90_GK * LDA;  // [2][GK][LDA]   first product
  R* Ts = lds;                     // [2][GT][LDA]   T, k-major
  R* B2 = lds + 2 * GT * LDA;      // [2][GK2][LDA]  second product

  const int tid = threadIdx.x;
  const int w = tid >> 6, lane = tid & 63;
  const int wr = (w / WL::WC) * (GT / WL::WR), wc = (w % WL::WC) * (GT / WL::WC);
  const int lr = lane & 15, lk = lane >> 4;
  Cfrag<R> C[FRR][FRC];
#pragma unroll
  for (int fi = 0; fi < FRR; ++fi)
#pragma unroll
    for (int fj = 0; fj < FRC; ++fj) C[fi][fj].zero();

  // ---- T = E_k^+ src, K chunks of GK loaded one chunk ahead
  constexpr int NE = W90_GK * GT / NT;
  static_assert(NE * NT == W90_GK * GT, "whole chunk per thread slot");
  cplx pa[NE], pb[NE];
  auto load1 = [&](int kk) {
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      const int e = tid + NT * u;
      const int l = e % W90_GK, x = e / W90_GK;
      const int gl = kk + l;
      cplx v = c_make(0.0, 0.0), v2 = c_make(0.0, 0.0);
      if (x < nw && gl < nw) {
        v = Ek[gl + (size_t)x * nw];  // A[x][l] = conj(E_k[l + x nw])
        v2 = c_wide(src[gl + (size_t)x * nw]);  // B[l][x] = src[l + x nw]
      }
      v.y = -v.y;
      pa[u] = v;
      pb[u] = v2;
    }
  };
  load1(0);
  for (int kk = 0; kk < nw; kk += W90_GK) {
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      const int e = tid + NT * u;
      const int l = e % W90_GK, x = e / W90_GK;
      As[(0 * W90_GK + l) * LDA + x] = (R)pa[u].x;
      As[(1 * W90_GK + l) * LDA + x] = (R)pa[u].y;
      Bs[(0 * W90_GK + l) * LDA + x] = (R)pb[u].x;
      Bs[(1 * W90_GK + l) * LDA + x] = (R)pb[u].y;
    }
    if (kk + W90_GK < nw) load1(kk + W90_GK);
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < W90_GK; ks += 4) {
      R ar[FRR], ai[FRR], br[FRC], bi[FRC];
#pragma unroll
      for (int f = 0; f < FRM; ++f) {
        if (f < FRR) {
          ar[f] = As[(0 * W90_GK + ks + lk) * LDA + wr + f * 16 + lr];
          ai[f] = As[(1 * W90_GK + ks + lk) * LDA + wr + f * 16 + lr];
        }
        if (f < FRC) {
          br[f] = Bs[(0 * W90_GK + ks + lk) * LDA + wc + f * 16 + lr];
          bi[f] = Bs[(1 * W90_GK + ks + lk) * LDA + wc + f * 16 + lr];
        }
      }
#pragma unroll
      for (int fi = 0; fi < FRR; ++fi)
#pragma unroll
        for (int fj = 0; fj < FRC; ++fj) C[fi][fj].mma(ar[fi], ai[fi], br[fj], bi[fj]);
    }
  }
  // every wavefront is done with the staging buffers before T overwrites them
  __syncthreads();
#pragma unroll
  for (int fi = 0; fi < FRR; ++fi)
#pragma unroll
    for (int fj = 0; fj < FRC; ++fj)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = wr + fi * 16 + Mfma<R>::row(lk, i);
        const int c = wc + fj * 16 + lr;
        Ts[(0 * GT + c) * LDA + r] = C[fi][fj].re(i);
        Ts[(1 * GT + c) * LDA + r] = C[fi][fj].im(i);
      }
#pragma unroll
  for (int fi = 0; fi < FRR; ++fi)
#pragma unroll
    for (int fj = 0; fj < FRC; ++fj) C[fi][fj].zero();

  // ---- R = T E_{k+b}, K chunks of GK2 rows of E_{k+b} loaded one ahead
  constexpr int NE2 = W90_GK2 * GT / NT;
  static_assert(NE2 * NT == W90_GK2 * GT, "whole chunk per thread slot");
  cplx pc[NE2];
  auto load2 = [&](int kk) {
#pragma unroll
    for (int u = 0; u < NE2; ++u) {
      const int e = tid + NT * u;
      const int l = e % W90_GK2, j = e / W90_GK2;
      const int gl = kk + l;
      cplx v = c_make(0.0, 0.0);
      if (j < nw && gl < nw) v = Ekb[gl + (size_t)j * nw];
      pc[u] = v;
    }
  };
  load2(0);
  for (int kk = 0; kk < nw; kk += W90_GK2) {
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NE2; ++u) {
      const int e = tid + NT * u;
      const int l = e % W90_GK2, j = e / W90_GK2;
      B2[(0 * W90_GK2 + l) * LDA + j] = (R)pc[u].x;
      B2[(1 * W90_GK2 + l) * LDA + j] = (R)pc[u].y;
    }
    if (kk + W90_GK2 < nw) load2(kk + W90_GK2);
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < W90_GK2; ks += 4) {
      R ar[FRR], ai[FRR], br[FRC], bi[FRC];
#pragma unroll
      for (int f = 0; f < FRM; ++f) {
        if (f < FRR) {
          ar[f] = Ts[(0 * GT + kk + ks + lk) * LDA + wr + f * 16 + lr];
          ai[f] = Ts[(1 * GT + kk + ks + lk) * LDA + wr + f * 16 + lr];
        }
        if (f < FRC) {
          br[f] = B2[(0 * W90_GK2 + ks + lk) * LDA + wc + f * 16 + lr];
          bi[f] = B2[(1 * W90_GK2 + ks + lk) * LDA + wc + f * 16 + lr];
        }
      }
#pragma unroll
      for (int fi = 0; fi < FRR; ++fi)
#pragma unroll
        for (int fj = 0; fj < FRC; ++fj) C[fi][fj].mma(ar[fi], ai[fi], br[fj], bi[fj]);
    }
  }

  // ---- epilogue: R to the other buffer, diagonal phases and tile sums
  double od = 0.0, all = 0.0;
#pragma unroll
  for (int fi = 0; fi < FRR; ++fi)
#pragma unroll
    for (int fj = 0; fj < FRC; ++fj)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = wr + fi * 16 + Mfma<R>::row(lk, i);
        const int c = wc + fj * 16 + lr;
        if (r < nw && c < nw) {
          T2 v;
          v.x = C[fi][fj].re(i);
          v.y = C[fi][fj].im(i);
          out[r + (size_t)c * nw] = v;
          const cplx vw = c_wide(v);
          const double a2 = c_abs2(vw);
          all += a2;
          if (r != c) {
            od += a2;
          } else {
            lnd[r] = atan2(vw.y, vw.x);
            d2d[r] = a2;
          }
        }
      }
  block_sum2(od, all, red);
  if (tid == 0) {
    pod[0] = od;
    pall[0] = all;
  }
}

// The rotation for num_wann <= GT (one tile per block), both products in
// one workgroup per (item, b) with T in LDS (rot_fused_tile).  Workgroups
// are placed XCD-aware: a one-dimensional grid of 8 * ceil(nitems / 8) *
// nnmax workgroups, where workgroup w runs on XCD w % 8 (round-robin
// dispatch) and takes item 8 t + w % 8 and neighbour b with
// (t, b) = divmod(w / 8, nnmax): the nnmax workgroups of one k run back to
// back on one XCD and share E_k through its L2.
template <class T2, int GT>
__global__ void __launch_bounds__(FusedWaves<GT>::NT, FusedWaves<GT>::MINB) k_rot_fused(const Prob* probs, const Item* items, int nitems,
                                                                     int nnmax, int mode) {
  typedef typename Real<T2>::type R;
  const int w = blockIdx.x;
  const int r = w >> 3;
  const int item = 8 * (r / nnmax) + (w & 7);
  const int b = r % nnmax;
  if (item >= nitems) return;
  const Item it = items[item];
  const Prob& p = probs[it.prob];
  const State* st = p.st;
  const int cur = gate_cur(st, mode);
  if (cur < 0) return;
  if (b >= p.nn) return;
  __shared__ R lds[rot_fused_elems<R, GT>()];
  __shared__ double red[256];
  rot_fused_tile<T2, GT>(p, it.k, b, mode, cur, lds, red);
}

// ------------------------------------------------ rotation, direct operands

// K steps of global operands in flight per wavefront in k_rot_direct.
#ifndef W90_RD_PF
#define W90_RD_PF 4
#endif

// The rotation of block (k, b) for num_wann <= 64 by one workgroup, with
// every global operand loaded straight into MFMA fragments, coalesced:
// the products are taken in the order P = M(k,b) E_{k+b}, then
// R = E_k^+ P, so that M (column-major) is the A operand, whose fragment
// rows are consecutive in memory, and E enters through its transpose ET
// (rows of E consecutive: the B operand of the first product, the
// conjugated A operand of the second).  Operands of W90_RD_PF K steps are
// in flight ahead of the MFMA products; P goes to LDS once (planar,
// row-major), the only barrier between the products; R goes through the
// same LDS for column-major stores of whole blocks.  Rows and columns
// outside num_wann are read at clamped (valid) addresses and zeroed.
// Epilogue as k_rot_fused: Im ln M_nn, |M_nn|^2, sum |M_mn|^2 and the
// off-diagonal sum of the block.
template <class T2>
__device__ __forceinline__ void rot_direct_tile(const Prob& p, int k, int b, int mode, int cur,
                                                typename Real<T2>::type* Ps, double* red) {
  constexpr int GT = 64, LD = GT + 1, FR = 2, PF = W90_RD_PF;
  typedef typename Real<T2>::type R;
  const int nw = p.nw, nn = p.nn;
  const size_t kb = (size_t)k * nn + b;
  const size_t nw2 = p.nw2;
  const int dst = dst_buf(mode, cur);
  const cplx* ETk = p.ET + (size_t)k * nw2;
  const cplx* ETkb = p.ET + (size_t)p.nnl[kb] * nw2;
  const T2* src = rot_src<T2>(p, mode, cur) + kb * nw2;
  T2* out = mbuf<T2>(p, dst) + kb * nw2;
  double* lnd = p.ln[dst] + kb * nw;
  double* d2d = p.d2 + kb * nw;

  const int tid = threadIdx.x;
  const int w = tid >> 6, lane = tid & 63;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  const int nks = (nw + 3) >> 2;
  // this lane's A rows and B columns, clamped into the block
  int ri[FR], ci[FR];
  bool rv[FR], cv[FR];
#pragma unroll
  for (int f = 0; f < FR; ++f) {
    const int r = wr + 16 * f + lr, c = wc + 16 * f + lr;
    rv[f] = r < nw;
    cv[f] = c < nw;
    ri[f] = rv[f] ? r : 0;
    ci[f] = cv[f] ? c : 0;
  }
  Cfrag<R> C[FR][FR];
#pragma unroll
  for (int fi = 0; fi < FR; ++fi)
#pragma unroll
    for (int fj = 0; fj < FR; ++fj) C[fi][fj].zero();

  // ---- P = M E_{k+b}: A[r][l] = src[r + l nw], B[l][c] = ETkb[c + l nw]
  T2 qa[PF][FR];
  cplx qb[PF][FR];
#pragma unroll
  for (int j = 0; j < PF; ++j) {
    const int l = 4 * j + lk;
    const size_t lo = (size_t)(l < nw ? l : 0) * nw;
#pragma unroll
    for (int f = 0; f < FR; ++f) {
      qa[j][f] = src[ri[f] + lo];
      qb[j][f] = ETkb[ci[f] + lo];
    }
  }
  for (int s0 = 0; s0 < nks; s0 += PF) {
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      const int s = s0 + j;
      if (s < nks) {
        const bool lv = 4 * s + lk < nw;
        R ar[FR], ai[FR], br[FR], bi[FR];
#pragma unroll
        for (int f = 0; f < FR; ++f) {
          const bool va = lv && rv[f], vb = lv && cv[f];
          ar[f] = va ? (R)qa[j][f].x : (R)0;
          ai[f] = va ? (R)qa[j][f].y : (R)0;
          br[f] = vb ? (R)qb[j][f].x : (R)0;
          bi[f] = vb ? (R)qb[j][f].y : (R)0;
        }
        const int sn = s + PF;
        if (sn < nks) {
          const int l = 4 * sn + lk;
          const size_t lo = (size_t)(l < nw ? l : 0) * nw;
#pragma unroll
          for (int f = 0; f < FR; ++f) {
            qa[j][f] = src[ri[f] + lo];
            qb[j][f] = ETkb[ci[f] + lo];
          }
        }
#pragma unroll
        for (int fi = 0; fi < FR; ++fi)
#pragma unroll
          for (int fj = 0; fj < FR; ++fj) C[fi][fj].mma(ar[fi], ai[fi], br[fj], bi[fj]);
      }
    }
  }
  // P to LDS, row-major: P[r][c] at Ps[(part * GT + r) * LD + c]
#pragma unroll
  for (int fi = 0; fi < FR; ++fi)
#pragma unroll
    for (int fj = 0; fj < FR; ++fj)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = wr + fi * 16 + Mfma<R>::row(lk, i);
        const int c = wc + fj * 16 + lr;
        Ps[(0 * GT + r) * LD + c] = C[fi][fj].re(i);
        Ps[(1 * GT + r) * LD + c] = C[fi][fj].im(i);
      }
#pragma unroll
  for (int fi = 0; fi < FR; ++fi)
#pragma unroll
    for (int fj = 0; fj < FR; ++fj) C[fi][fj].zero();

  // ---- R = E_k^+ P: A[r][l] = conj(ETk[r + l nw]), B[l][c] = P[l][c]
  cplx qe[PF][FR];
#pragma unroll
  for (int j = 0; j < PF; ++j) {
    const int l = 4 * j + lk;
    const size_t lo = (size_t)(l < nw ? l : 0) * nw;
#pragma unroll
    for (int f = 0; f < FR; ++f) qe[j][f] = ETk[ri[f] + lo];
  }
  __syncthreads();
  for (int s0 = 0; s0 < nks; s0 += PF) {
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      const int s = s0 + j;
      if (s < nks) {
        const int l = 4 * s + lk;
        const bool lv = l < nw;
        R ar[FR], ai[FR], br[FR], bi[FR];
#pragma unroll
        for (int f = 0; f < FR; ++f) {
          const bool va = lv && rv[f];
          ar[f] = va ? (R)qe[j][f].x : (R)0;
          ai[f] = va ? -(R)qe[j][f].y : (R)0;
          br[f] = Ps[(0 * GT + l) * LD + wc + 16 * f + lr];
          bi[f] = Ps[(1 * GT + l) * LD + wc + 16 * f + lr];
        }
        const int sn = s + PF;
        if (sn < nks) {
          const int ln = 4 * sn + lk;
          const size_t lo = (size_t)(ln < nw ? ln : 0) * nw;
#pragma unroll
          for (int f = 0; f < FR; ++f) qe[j][f] = ETk[ri[f] + lo];
        }
#pragma unroll
        for (int fi = 0; fi < FR; ++fi)
#pragma unroll
          for (int fj = 0; fj < FR; ++fj) C[fi][fj].mma(ar[fi], ai[fi], br[fj], bi[fj]);
      }
    }
  }

  // ---- epilogue: phases, moduli and sums from the accumulators; R through
  // LDS (column-major) for whole-block stores
  double od = 0.0, all = 0.0;
#pragma unroll
  for (int fi = 0; fi < FR; ++fi)
#pragma unroll
    for (int fj = 0; fj < FR; ++fj)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = wr + fi * 16 + Mfma<R>::row(lk, i);
        const int c = wc + fj * 16 + lr;
        if (r < nw && c < nw) {
          T2 v;
          v.x = C[fi][fj].re(i);
          v.y = C[fi][fj].im(i);
          const cplx vw = c_wide(v);
          const double a2 = c_abs2(vw);
          all += a2;
          if (r != c) {
            od += a2;
          } else {
            lnd[r] = atan2(vw.y, vw.x);
            d2d[r] = a2;
          }
        }
      }
  __syncthreads();  // every wavefront has read its last B fragment of P
#pragma unroll
  for (int fi = 0; fi < FR; ++fi)
#pragma unroll
    for (int fj = 0; fj < FR; ++fj)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = wr + fi * 16 + Mfma<R>::row(lk, i);
        const int c = wc + fj * 16 + lr;
        Ps[(0 * GT + c) * LD + r] = C[fi][fj].re(i);
        Ps[(1 * GT + c) * LD + r] = C[fi][fj].im(i);
      }
  __syncthreads();
  for (int e = tid; e < (int)nw2; e += 256) {
    const int r = e % nw, c = e / nw;
    T2 v;
    v.x = Ps[(0 * GT + c) * LD + r];
    v.y = Ps[(1 * GT + c) * LD + r];
    out[e] = v;
  }
  block_sum2(od, all, red);
  if (tid == 0) {
    p.part_od[kb * p.npan] = od;
    p.part_all[kb * p.npan] = all;
  }
}

// The rotation for 32 < num_wann <= 64 with direct operand loads
// (rot_direct_tile); workgroup placement as k_rot_fused (the nnmax
// workgroups of one k back to back on one XCD, sharing ET_k in its L2).
template <class T2>
__global__ void __launch_bounds__(256, 2) k_rot_direct(const Prob* probs, const Item* items, int nitems,
                                                       int nnmax, int mode) {
  typedef typename Real<T2>::type R;
  const int w = blockIdx.x;
  const int r = w >> 3;
  const int item = 8 * (r / nnmax) + (w & 7);
  const int b = r % nnmax;
  if (item >= nitems) return;
  const Item it = items[item];
  const Prob& p = probs[it.prob];
  const State* st = p.st;
  const int cur = gate_cur(st, mode);
  if (cur < 0) return;
  if (b >= p.nn) return;
  __shared__ R Ps[2 * 64 * 65];
  __shared__ double red[256];
  rot_direct_tile<T2>(p, it.k, b, mode, cur, Ps, red);
}

// Omega partial sums of the current M by GEMM tile (MODE_INIT of the MFMA
// class): the layout the stage-2 epilogue writes, without a rotation.
template <class T2, int GT>
__global__ void __launch_bounds__(256) k_reduce_tiles(const Prob* probs, const Item* items, int nitems, int mode) {
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const State* st = p.st;
  const int cur = gate_cur(st, mode);
  if (cur < 0) return;
  const int nw = p.nw, nn = p.nn;
  const int nt1 = (nw + GT - 1) / GT;
  const int ntl = nt1 * nt1;
  const int by = blockIdx.y;
  if (by >= nn * ntl) return;
  const int b = by / ntl, tile = by % ntl;
  const int row0 = (tile % nt1) * GT, col0 = (tile / nt1) * GT;
  const size_t kb = (size_t)it.k * nn + b;
  const int dst = dst_buf(mode, cur);
  const T2* m = mbuf<T2>(p, cur) + kb * p.nw2;
  __shared__ double red[256];
  double od = 0.0, all = 0.0;
  for (int e = threadIdx.x; e < GT * GT; e += blockDim.x) {
    const int r = row0 + e % GT, c = col0 + e / GT;
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
  block_sum2(od, all, red);
  if (threadIdx.x == 0) {
    p.part_od[kb * p.npan + tile] = od;
    p.part_all[kb * p.npan + tile] = all;
  }
}

// ---------------------------------------------------------- retraction

// Square complex matrices of edge N in LDS for k_step_mfma<N> (N = 64 for
// num_wann 49-64, 48 up to 48): planar re/im, column-major with leading
// dimension N + 1 (MFMA fragment loads of A along a column and of B along
// a row are both conflict-free with this padding).  Wavefront layout: a 64
// matrix in 32 x 16 blocks over eight wavefronts (two per SIMD, so one
// covers the other's LDS waits; the matrices fill the CU's LDS, so there
// is no second workgroup), a 48 matrix in three 48 x 16 column strips over
// three (3 x 3 fragments do not split evenly over four), which leaves room
// in LDS for two workgroups per CU.
#define W90_STEP_N 64
template <int N>
struct StepWaves {
  static constexpr int LD = N + 1, MAT = N * LD;
  static constexpr int WR = N == 48 ? 1 : 2, WC = N == 48 ? 3 : 4;
  static constexpr int NT = 64 * WR * WC;
  static constexpr int FRR = N / (16 * WR), FRC = N / (16 * WC), FRM = FRR > FRC ? FRR : FRC;
  static constexpr int MINB = N == 48 ? 2 : 1;  // workgroups per CU (LDS: 75 KB at 48, 133 KB at 64)
};

// LDS bytes of k_step_mfma<N>: X and P (re, im each) and a reduction buffer.
template <int N>
__host__ __device__ constexpr size_t step_mfma_lds() {
  return sizeof(double) * (4 * StepWaves<N>::MAT + 256);
}

// C = A B for the 64 x 64 LDS matrices (planar at Ar/Ai, Br/Bi) over the
// first 4 KS columns of A (rows of B): each wavefront accumulates its 32 x 32 quadrant as 2 x 2 complex
// FP64 MFMA fragments (three real products per K step, Cfrag).  The rows
// and columns beyond num_wann are zero in every operand, so all fragments
// are formed and the K steps past num_wann (up to KS) add zeros.  The K
// steps are unrolled without branches and the operands of step ks + 1 are
// read from LDS before the products of step ks are issued: with one
// wavefront per SIMD nothing else covers the LDS latency.
template <int KS, int N>
__device__ __forceinline__ void step_gemm(Cfrag<double> (&C)[StepWaves<N>::FRR][StepWaves<N>::FRC], const double* Ar,
                                          const double* Ai, const double* Br, const double* Bi) {
  typedef StepWaves<N> SW;
  constexpr int FRR = SW::FRR, FRC = SW::FRC, LD = SW::LD;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = (w / SW::WC) * (N / SW::WR), wc = (w % SW::WC) * (N / SW::WC);
  const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int fi = 0; fi < FRR; ++fi)
#pragma unroll
    for (int fj = 0; fj < FRC; ++fj) C[fi][fj].zero();
  double ar[2][FRR], ai[2][FRR], br[2][FRC], bi[2][FRC];  // [buffer][fragment]
  auto load = [&](int ks, int buf) {
#pragma unroll
    for (int f = 0; f < SW::FRM; ++f) {
      if (f < FRR) {
        const int ia = (4 * ks + lk) * LD + wr + f * 16 + lr;  // A[row][4 ks + lk]
        ar[buf][f] = Ar[ia];
        ai[buf][f] = Ai[ia];
      }
      if (f < FRC) {
        const int ib = (wc + f * 16 + lr) * LD + 4 * ks + lk;  // B[4 ks + lk][col]
        br[buf][f] = Br[ib];
        bi[buf][f] = Bi[ib];
      }
    }
  };
  load(0, 0);
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const int cb = ks & 1;
    if (ks + 1 < KS) load(ks + 1, cb ^ 1);
#pragma unroll
    for (int fi = 0; fi < FRR; ++fi)
#pragma unroll
      for (int fj = 0; fj < FRC; ++fj) C[fi][fj].mma(ar[cb][fi], ai[cb][fi], br[cb][fj], bi[cb][fj]);
  }
}

// step_gemm over the K steps num_wann needs: 12 up to 48 bands, else 16.
// Each output fragment is the same MFMA sequence in either layout, so a
// product's result does not depend on N.
template <int N>
__device__ __forceinline__ void step_gemm_n(Cfrag<double> (&C)[StepWaves<N>::FRR][StepWaves<N>::FRC],
                                            const double* Ar, const double* Ai, const double* Br, const double* Bi,
                                            int n) {
  if (N > 48 && n > 48) step_gemm<16, N>(C, Ar, Ai, Br, Bi);
  else step_gemm<12, N>(C, Ar, Ai, Br, Bi);
}

// Position (row, column) of accumulator element i of fragment (fi, fj).
template <int N>
__device__ __forceinline__ void step_pos(int fi, int fj, int i, int* r, int* c) {
  typedef StepWaves<N> SW;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  *r = (w / SW::WC) * (N / SW::WR) + fi * 16 + Mfma<double>::row(lane >> 4, i);
  *c = (w % SW::WC) * (N / SW::WC) + fj * 16 + (lane & 15);
}

// Retraction of one (problem, k) for 32 < num_wann <= 64, every matrix in
// LDS and every product on FP64 MFMA: forms D (trial step), X = scale D,
// E = exp(X) by the scaling and squaring of the Taylor polynomial that
// k_step uses (same plan), and U_new = U E.  The exponential of the
// anti-Hermitian X is the unitary internal_new_u_and_m builds from the
// eigenvectors of i X (wannierise.F90:1662-1689, 1729-1734).  N = 48 takes
// the items of num_wann up to 48 and N = 64 the others (a workgroup whose
// item is the other's returns at once).
template <int N>
__global__ void __launch_bounds__(StepWaves<N>::NT, StepWaves<N>::MINB) k_step_mfma(const Prob* probs,
                                                                                   const Item* items, int nitems,
                                                                                   int mode) {
  typedef StepWaves<N> SW;
  constexpr int NT = SW::NT, LD = SW::LD, MAT = SW::MAT, FRR = SW::FRR, FRC = SW::FRC;
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  if ((p.nw <= 48) != (N == 48)) return;
  const State* st = p.st;
  const int cur = gate_cur(st, mode);
  if (cur < 0) return;
  const int nw = p.nw, nw2 = p.nw2;
  const int k = it.k;
  const int tid = threadIdx.x;
  extern __shared__ double lds_step[];
  double* Xr = lds_step;
  double* Xi = Xr + MAT;
  double* Pr = Xi + MAT;
  double* Pi = Pr + MAT;
  double* red = Pi + MAT;
  cplx* Dk = p.D + (size_t)k * nw2;
  // every G, D and U element of the thread is requested before the search
  // and the exponential (one workgroup fills a CU's LDS, so there are no
  // other wavefronts to cover the latency of loads issued one loop trip
  // at a time)
  constexpr int SL = N * N / NT;
  static_assert(SL * NT == N * N, "whole matrix per thread slot");
  cplx gpre[SL], dpre[SL], upre[SL];
  {
    const cplx* Gk = p.G + (size_t)k * nw2;
    const cplx* Uc = p.U[cur] + (size_t)k * nw2;
#pragma unroll
    for (int i = 0; i < SL; ++i) {
      const int e = tid + NT * i;
      const int r = e % N, c = e / N;
      cplx g = c_make(0.0, 0.0), d = g, u = g;
      if (r < nw && c < nw) {
        const int q = r + c * nw;
        if (mode == MODE_TRIAL) g = Gk[q];
        d = Dk[q];
        u = Uc[q];
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
  // X = scale D (zero outside nw), D formed on the trial step, and the
  // moduli of X for its 1-norm (into the P slot, free until the Taylor
  // polynomial); each column is then summed in row order by lane c of
  // wavefront 0, the maximum by lane exchange
#pragma unroll
  for (int i = 0; i < SL; ++i) {
    const int e = tid + NT * i;
    const int r = e % N, c = e / N;
    cplx x = c_make(0.0, 0.0);
    if (r < nw && c < nw) {
      cplx d = dpre[i];
      if (mode == MODE_TRIAL) {
        d = c_scale(c_add(gpre[i], c_scale(d, ss->beta)), ss->dir_sign);
        st_stream(&Dk[r + c * nw], d);
      }
      x = c_scale(d, scale);
    }
    Xr[c * LD + r] = x.x;
    Xi[c * LD + r] = x.y;
    Pr[c * LD + r] = sqrt(c_abs2(x));
  }
  __syncthreads();
  double cs = 0.0;
  if (tid < 64) {
    if (tid < nw) {
#pragma unroll 4
      for (int r = 0; r < nw; ++r) cs += Pr[tid * LD + r];
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
  for (int e = tid; e < N * N; e += NT) {
    const int r = e % N, c = e / N;
    const int q = c * LD + r;
    const double xr = Xr[q] * f, xi = Xi[q] * f;
    Xr[q] = xr;
    Xi[q] = xi;
    Pr[q] = xr * invm + (r == c ? 1.0 : 0.0);
    Pi[q] = xi * invm;
  }
  __syncthreads();
  Cfrag<double> C[FRR][FRC];
  // Horner: P <- I + X P / j
  for (int j = m - 1; j >= 1; --j) {
    step_gemm_n<N>(C, Xr, Xi, Pr, Pi, nw);
    __syncthreads();
    const double invj = 1.0 / (double)j;
#pragma unroll
    for (int fi = 0; fi < FRR; ++fi)
#pragma unroll
      for (int fj = 0; fj < FRC; ++fj)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          int r, c;
          step_pos<N>(fi, fj, i, &r, &c);
          const int q = c * LD + r;
          Pr[q] = C[fi][fj].re(i) * invj + (r == c ? 1.0 : 0.0);
          Pi[q] = C[fi][fj].im(i) * invj;
        }
    __syncthreads();
  }
  // squarings: P <- P P
  for (int s = 0; s < sq; ++s) {
    step_gemm_n<N>(C, Pr, Pi, Pr, Pi, nw);
    __syncthreads();
#pragma unroll
    for (int fi = 0; fi < FRR; ++fi)
#pragma unroll
      for (int fj = 0; fj < FRC; ++fj)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          int r, c;
          step_pos<N>(fi, fj, i, &r, &c);
          const int q = c * LD + r;
          Pr[q] = C[fi][fj].re(i);
          Pi[q] = C[fi][fj].im(i);
        }
    __syncthreads();
  }
  // E = P to global, and its transpose ET (rows of E, for the rotation's
  // operand loads); U into the X slot (zero outside nw)
  cplx* Ek = p.E + (size_t)k * nw2;
  cplx* ETk = p.ET + (size_t)k * nw2;
#pragma unroll
  for (int i = 0; i < SL; ++i) {
    const int e = tid + NT * i;
    const int r = e % N, c = e / N;
    const int q = c * LD + r;
    if (r < nw && c < nw) {
      st_stream(&Ek[r + c * nw], c_make(Pr[q], Pi[q]));
      const int qt = r * LD + c;  // E[c][r]
      st_stream(&ETk[r + c * nw], c_make(Pr[qt], Pi[qt]));
    }
    Xr[q] = upre[i].x;
    Xi[q] = upre[i].y;
  }
  __syncthreads();
  step_gemm_n<N>(C, Xr, Xi, Pr, Pi, nw);
  cplx* Un = p.U[1 - cur] + (size_t)k * nw2;
#pragma unroll
  for (int fi = 0; fi < FRR; ++fi)
#pragma unroll
    for (int fj = 0; fj < FRC; ++fj)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        int r, c;
        step_pos<N>(fi, fj, i, &r, &c);
        if (r < nw && c < nw) st_stream(&Un[r + c * nw], c_make(C[fi][fj].re(i), C[fi][fj].im(i)));
      }
}

// C = A B for n x n column-major complex matrices in global memory, held by
// one workgroup: each wavefront takes 16 x 16 output tiles in turn, two at
// a time (independent accumulator chains), on FP64 MFMA with the operand
// fragments loaded straight from memory (the matrices of one k stay in the
// caches: A fragments read along a column, B fragments reuse cache lines
// over consecutive K steps).  With ident, C = I + scale A B (one Taylor
// step, applied as the product is stored).  Used by the retraction above
// num_wann 64.
__device__ void wg_gemm_mfma(cplx* C, const cplx* A, const cplx* B, int n, double scale, bool ident) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int nt = (n + 15) / 16;
  const int ntt = nt * nt;
  for (int t0 = 2 * w; t0 < ntt; t0 += 2 * nwv) {
    const bool two = t0 + 1 < ntt;
    int i0[2], j0[2];
    Cfrag<double> cc[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int t = two ? t0 + u : t0;
      i0[u] = (t % nt) * 16;
      j0[u] = (t / nt) * 16;
      cc[u].zero();
    }
    for (int ks = 0; ks < n; ks += 4) {
      const int kk = ks + lk;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int r = i0[u] + lr, c = j0[u] + lr;
        const cplx a = (r < n && kk < n) ? A[r + (size_t)kk * n] : c_make(0.0, 0.0);
        const cplx b = (kk < n && c < n) ? B[kk + (size_t)c * n] : c_make(0.0, 0.0);
        cc[u].mma(a.x, a.y, b.x, b.y);
      }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (u == 1 && !two) break;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = i0[u] + Mfma<double>::row(lk, i), c = j0[u] + lr;
        if (r < n && c < n) {
          cplx v = c_make(cc[u].re(i), cc[u].im(i));
          if (ident) {
            v = c_scale(v, scale);
            if (r == c) v.x += 1.0;
          }
          C[r + (size_t)c * n] = v;
        }
      }
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------- small blocks

// LDS bytes of k_rotate_w16 for up to nn neighbours: one padded 16 x 17
// block per wavefront, the diagonal of each rotated block (complex value,
// phase, modulus), the per-wave sums and the k's mesh tables (wb, bk).
template <class T2>
__host__ __device__ constexpr size_t rotate_w16_lds(int nn) {
  return sizeof(T2) * 16 * 17 * (size_t)nn + sizeof(double) * (4 * 16 * (size_t)nn + 6 * (size_t)nn);
}

// M rotation with the omega partial sums for num_wann <= 16: one workgroup
// per (problem, k), one wavefront per neighbour b, each block a single
// 16 x 16 MFMA tile (zero-padded).  T = E_k^+ M(k,b) is formed from operand
// fragments loaded straight from memory, passed through the wavefront's LDS
// block to become the A operand of R = T E_{k+b}; R goes back through the
// same LDS block so that it is stored in column order (coalesced) and its
// diagonal is collected per b; the phases Im ln M_nn of all b are then
// formed in one pass over the workgroup, and threads n < num_wann sum them
// over b in order.  blockDim.x = 64 * (largest nntot of the launch); wavefronts beyond
// the problem's nntot only take part in the barriers, and every sum is
// formed per wavefront and then over b in order, so a problem's result does
// not depend on the other problems of the launch.  MODE_INIT only reduces.
// (internal_new_u_and_m :1737-1748; wann_omega :2034-2090, :2214-2275)
template <class T2>
__global__ void __launch_bounds__(64 * W90GPU_NUM_NNMAX) k_rotate_w16(const Prob* probs, const Item* items,
                                                                      int nitems, int mode) {
  typedef typename Real<T2>::type R;
  if (blockIdx.x >= (unsigned)nitems) return;
  const Item it = items[blockIdx.x];
  const Prob& p = probs[it.prob];
  const State* st = p.st;
  const int cur = gate_cur(st, mode);
  if (cur < 0) return;
  const int nw = p.nw, nw2 = p.nw2, nn = p.nn;
  const int nnl = blockDim.x >> 6;
  const int k = it.k;
  const int tid = threadIdx.x;
  const int b = tid >> 6;
  const bool active = b < nn;
  const int lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  const bool rot = (mode != MODE_INIT);
  const int dst = dst_buf(mode, cur);
  extern __shared__ double lds_w16[];
  T2* Tl = reinterpret_cast<T2*>(lds_w16) + (size_t)b * 16 * 17;  // this wave's block, column-major
  double* lnL = reinterpret_cast<double*>(reinterpret_cast<T2*>(lds_w16) + (size_t)nnl * 16 * 17);
  double* d2L = lnL + nnl * 16;
  double* wod = d2L + nnl * 16;
  double* wall = wod + nnl;
  double* wbL = wall + nnl;
  double* bkL = wbL + nnl;
  double* dRe = bkL + 3 * nnl;  // [b][16] diagonal of the rotated block
  double* dIm = dRe + nnl * 16;
  // mesh tables and the reference centres are loaded with the operands,
  // not inside the loops over b of the epilogue
  kmesh_stage(p, k, wbL, bkL);
  double r0[3];
  if (tid < nw) ref_centre(p, mode, cur, tid, r0);
  const size_t kb = (size_t)k * nn + (active ? b : 0);
  const T2* src = rot_src<T2>(p, mode, cur) + kb * nw2;
  const cplx* Ek = p.E + (size_t)k * nw2;
  const cplx* Ekb = p.E + (size_t)p.nnl[kb] * nw2;
  // output addresses are formed before the first store (the compiler
  // would otherwise reload the problem's pointers after each store)
  T2* out = mbuf<T2>(p, dst) + kb * nw2;
  double* lnd = p.ln[dst] + (size_t)k * nn * nw;
  double* outk = p.kpart + (size_t)k * p.psz;
  Cfrag<R> Racc;
  Racc.zero();
  // every operand fragment of both products is loaded up front, so the
  // global-memory latencies overlap
  cplx fa[4], fm[4], fb[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int l = 4 * s + lk;
    fa[s] = fm[s] = fb[s] = c_make(0.0, 0.0);
    if (active && rot && l < nw && lr < nw) {
      fa[s] = Ek[l + lr * nw];
      fm[s] = c_wide(src[l + lr * nw]);
      fb[s] = Ekb[l + lr * nw];
    }
  }
  if (active && rot) {
    // T = E_k^+ M: A[i][l] = conj(E_k[l + i nw]), B[l][j] = M[l + j nw]
    Cfrag<R> Tacc;
    Tacc.zero();
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const cplx a = fa[s], v = fm[s];
      Tacc.mma((R)a.x, (R)(-a.y), (R)v.x, (R)v.y);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      T2 t;
      t.x = Tacc.re(i);
      t.y = Tacc.im(i);
      Tl[Mfma<R>::row(lk, i) + lr * 17] = t;
    }
  }
  __syncthreads();
  if (active && rot) {
    // R = T E_{k+b}: A[i][l] = T[i][l] (LDS), B[l][j] = E_{k+b}[l + j nw]
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int l = 4 * s + lk;
      const T2 a = Tl[lr + l * 17];
      Racc.mma(a.x, a.y, (R)fb[s].x, (R)fb[s].y);
    }
    // R replaces T in the wavefront's block (every read of T above feeds
    // the products that produce R)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      T2 t;
      t.x = Racc.re(i);
      t.y = Racc.im(i);
      Tl[Mfma<R>::row(lk, i) + lr * 17] = t;
    }
  }
  __syncthreads();
  double s_od = 0.0, s_all = 0.0;
  if (active) {
    // element e = lane + 64 i of the block: consecutive lanes hold
    // consecutive rows of a column, so the stores are contiguous
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = lane + 64 * i;
      const int r = e & 15, c = e >> 4;
      if (r < nw && c < nw) {
        T2 v;
        if (rot) {
          v = Tl[r + c * 17];
          // streamed: the next readers of the block run on other XCDs
          st_stream(&out[r + c * nw], v);
        } else {
          v = src[r + c * nw];
        }
        const cplx vw = c_wide(v);
        const double a2 = c_abs2(vw);
        s_all += a2;
        if (r != c) {
          s_od += a2;
        } else {
          dRe[b * 16 + r] = vw.x;
          dIm[b * 16 + r] = vw.y;
        }
      }
    }
  }
  s_od = wave_sum(s_od);
  s_all = wave_sum(s_all);
  if (active && lane == 0) {
    wod[b] = s_od;
    wall[b] = s_all;
  }
  __syncthreads();
  // Im ln M_nn and |M_nn|^2 of every b in one pass
  for (int t = tid; t < nn * 16; t += blockDim.x) {
    const int bb = t >> 4, n = t & 15;
    if (n < nw) {
      const cplx vw = c_make(dRe[t], dIm[t]);
      const double l = atan2(vw.y, vw.x);
      lnd[bb * nw + n] = l;
      lnL[t] = l;
      d2L[t] = c_abs2(vw);
    }
  }
  __syncthreads();
  if (tid < nw) {
    const int n = tid;
    DiagAcc d = {};
    for (int bb = 0; bb < nn; ++bb)
      diag_add(d, wbL[bb], bkL + bb * 3, lnL[bb * 16 + n], d2L[bb * 16 + n], r0[0], r0[1], r0[2]);
    diag_store(outk, nw, n, d);
  }
  if (tid == 0) {
    double sod = 0.0, sall = 0.0, wsum = 0.0;
    for (int bb = 0; bb < nn; ++bb) {
      const double w = wbL[bb];
      sod += w * wod[bb];
      sall += w * wall[bb];
      wsum += w;
    }
    outk[8 * nw + 0] = sod;
    outk[8 * nw + 1] = (double)nw * wsum - sall;
    kpart_bmat(wbL, bkL, nn, outk + 8 * nw);
  }
}
