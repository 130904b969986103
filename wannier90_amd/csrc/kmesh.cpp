// Finite-difference k-mesh for the spread functional: shells of b-vectors,
// B1 weights, neighbour lists.  Host-side C++; its tables are uploaded once
// to HBM by the solver.
//
// Follows kmesh_get (src/kmesh.F90:61-740): supercell images sorted by
// distance with the reproducible tie rule of internal_maxloc
// (kmesh.F90:984-1038, 1851-1885), shell distances/multiplicities from a
// Gamma-centred mesh (kmesh.F90:130-162), automatic shell selection with a
// least-squares B1 solve (kmesh_shell_automatic, kmesh.F90:1107-1432) or a
// user shell_list (kmesh_shell_fixed), then per-k neighbour search in
// (shell, image, k') order (kmesh.F90:319-345) and the B1 check
// (kmesh.F90:461-487).  The search order is kept so that nnlist/nncell/bk
// come out in the same order as the reference (the order fixes the b index
// of every M(k,b) block in .mmn files).

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>
#include <algorithm>

#include "../../include/w90gpu.h"

namespace {

constexpr int kNSupcell = 5;                        // kmesh.F90:55
constexpr int kNImg = (2 * kNSupcell + 1) * (2 * kNSupcell + 1) * (2 * kNSupcell + 1);
constexpr int kMaxShells = 6;                       // types.F90:115
constexpr int kNumNNMax = W90GPU_NUM_NNMAX;         // types.F90:116
constexpr double kEps5 = 1e-5, kEps6 = 1e-6, kEps8 = 1e-8;

void set_err(char* msg, int len, const char* text) {
  if (msg && len > 0) {
    std::snprintf(msg, (size_t)len, "%s", text);
  }
}

// recip(i,:) = 2*pi * (a_j x a_k) / V, rows are vectors (utility.F90 recip_lattice_base)
double recip_lattice(const double a[3][3], double b[3][3]) {
  auto cross = [](const double* u, const double* v, double* w) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
  };
  double c0[3], c1[3], c2[3];
  cross(a[1], a[2], c0);
  cross(a[2], a[0], c1);
  cross(a[0], a[1], c2);
  const double vol = a[0][0] * c0[0] + a[0][1] * c0[1] + a[0][2] * c0[2];
  const double f = 2.0 * M_PI / vol;
  for (int j = 0; j < 3; ++j) {
    b[0][j] = f * c0[j];
    b[1][j] = f * c1[j];
    b[2][j] = f * c2[j];
  }
  return vol;
}

// One-sided Jacobi SVD of a 6 x n matrix (n <= 6).  Returns singular values
// and the pseudo-inverse solution w = V S^-1 U^T t.  Used in place of LAPACK
// dgesvd for the B1 least-squares system (kmesh.F90:1266-1313).
void small_svd_solve(int n, double A[6][kMaxShells], const double t[6], double sv[kMaxShells],
                     double w[kMaxShells]) {
  double V[kMaxShells][kMaxShells] = {};
  for (int i = 0; i < n; ++i) V[i][i] = 1.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p) {
      for (int q = p + 1; q < n; ++q) {
        double alpha = 0, beta = 0, gamma = 0;
        for (int r = 0; r < 6; ++r) {
          alpha += A[r][p] * A[r][p];
          beta += A[r][q] * A[r][q];
          gamma += A[r][p] * A[r][q];
        }
        if (std::fabs(gamma) <= 1e-300) continue;
        off = std::max(off, std::fabs(gamma) / std::sqrt(alpha * beta));
        const double zeta = (beta - alpha) / (2.0 * gamma);
        const double tt = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + tt * tt), s = c * tt;
        for (int r = 0; r < 6; ++r) {
          const double ap = A[r][p], aq = A[r][q];
          A[r][p] = c * ap - s * aq;
          A[r][q] = s * ap + c * aq;
        }
        for (int r = 0; r < n; ++r) {
          const double vp = V[r][p], vq = V[r][q];
          V[r][p] = c * vp - s * vq;
          V[r][q] = s * vp + c * vq;
        }
      }
    }
    if (off < 1e-15) break;
  }
  double coef[kMaxShells];
  for (int j = 0; j < n; ++j) {
    double nrm = 0;
    for (int r = 0; r < 6; ++r) nrm += A[r][j] * A[r][j];
    nrm = std::sqrt(nrm);
    sv[j] = nrm;
    double ut = 0;
    if (nrm > 0) {
      for (int r = 0; r < 6; ++r) ut += (A[r][j] / nrm) * t[r];
      coef[j] = ut / nrm;
    } else {
      coef[j] = 0;
    }
  }
  for (int i = 0; i < n; ++i) {
    double s = 0;
    for (int j = 0; j < n; ++j) s += V[i][j] * coef[j];
    w[i] = s;
  }
}

struct Mesh {
  double recip[3][3];
  std::vector<double> kcart;  // 3*nk
  int lmn[kNImg][3];
  double gimg[kNImg][3];      // lmn . recip
};

// kmesh_supercell_sort: ascending distance; within a degenerate group the
// image with the lowest original index is placed last (internal_maxloc).
void supercell_sort(Mesh& m) {
  int lmn[kNImg][3];
  std::vector<double> dist(kNImg);
  int c = 0;
  lmn[0][0] = lmn[0][1] = lmn[0][2] = 0;
  dist[0] = 0.0;
  for (int l = -kNSupcell; l <= kNSupcell; ++l)
    for (int mm = -kNSupcell; mm <= kNSupcell; ++mm)
      for (int n = -kNSupcell; n <= kNSupcell; ++n) {
        if (l == 0 && mm == 0 && n == 0) continue;
        ++c;
        lmn[c][0] = l; lmn[c][1] = mm; lmn[c][2] = n;
        double p[3];
        for (int j = 0; j < 3; ++j) p[j] = l * m.recip[0][j] + mm * m.recip[1][j] + n * m.recip[2][j];
        dist[c] = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
      }
  for (int slot = kNImg - 1; slot >= 0; --slot) {
    int g = 0;
    for (int i = 1; i < kNImg; ++i)
      if (dist[i] > dist[g]) g = i;  // Fortran maxloc: first maximum
    int pick = g;
    for (int i = 0; i < kNImg; ++i)
      if (i != g && std::fabs(dist[i] - dist[g]) < kEps8 && i < pick) pick = i;
    for (int j = 0; j < 3; ++j) m.lmn[slot][j] = lmn[pick][j];
    dist[pick] = -1.0;
  }
  for (int i = 0; i < kNImg; ++i)
    for (int j = 0; j < 3; ++j)
      m.gimg[i][j] = m.lmn[i][0] * m.recip[0][j] + m.lmn[i][1] * m.recip[1][j] + m.lmn[i][2] * m.recip[2][j];
}

// kmesh_get_bvectors with kpt = 1: the first `multi` vectors of a shell in
// (image, k') order.
int shell_bvectors(const Mesh& m, int nk, double shell_dist, double tol, int multi, double* bvec) {
  int nb = 0;
  const double* k0 = &m.kcart[0];
  for (int img = 0; img < kNImg && nb < multi; ++img) {
    for (int k2 = 0; k2 < nk; ++k2) {
      double v[3], d2 = 0;
      for (int j = 0; j < 3; ++j) {
        v[j] = m.gimg[img][j] + m.kcart[3 * k2 + j];
        d2 += (k0[j] - v[j]) * (k0[j] - v[j]);
      }
      const double dist = std::sqrt(d2);
      if (dist >= shell_dist * (1.0 - tol) && dist <= shell_dist * (1.0 + tol)) {
        for (int j = 0; j < 3; ++j) bvec[3 * nb + j] = v[j] - k0[j];
        ++nb;
      }
      if (nb == multi) break;
    }
  }
  return nb;
}

}  // namespace

extern "C" int w90gpu_kmesh_get(const double* real_lattice, int num_kpts, const double* kpt_latt,
                                int search_shells, double tol, int num_shells_in,
                                const int* shell_list_in, int skip_b1_tests, int* nntot_out,
                                int* nnlist, int* nncell, double* bk, double* wb, double* wbtot,
                                char* errmsg, int errlen) {
  if (num_kpts <= 0 || !real_lattice || !kpt_latt || !nntot_out || !nnlist || !nncell || !bk || !wb ||
      !wbtot) {
    set_err(errmsg, errlen, "w90gpu_kmesh_get: invalid arguments");
    return W90GPU_ERR_INPUT;
  }
  if (search_shells <= 0) search_shells = 36;  // types.F90:127
  if (tol <= 0) tol = 1e-6;                    // types.F90:128
  Mesh m;
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = real_lattice[i + 3 * j];  // real_lattice(i,j), row i = a_i
  const double vol = recip_lattice(a, m.recip);
  if (std::fabs(vol) < kEps5) {
    set_err(errmsg, errlen, "Found almost zero Volume in utility_recip_lattice");
    return W90GPU_ERR_FATAL;
  }
  const int nk = num_kpts;
  m.kcart.resize(3 * (size_t)nk);
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j < 3; ++j)
      m.kcart[3 * k + j] = kpt_latt[3 * k + 0] * m.recip[0][j] + kpt_latt[3 * k + 1] * m.recip[1][j] +
                           kpt_latt[3 * k + 2] * m.recip[2][j];
  supercell_sort(m);

  // shell distances and multiplicities (kmesh.F90:130-162)
  const double eta = 99999999.0;
  std::vector<double> dnn(search_shells);
  std::vector<int> multi(search_shells);
  double dnn0 = 0.0, dnn1 = eta;
  for (int s = 0; s < search_shells; ++s) {
    int counter = 0;
    for (int k = 0; k < nk; ++k) {
      for (int img = 0; img < kNImg; ++img) {
        double v[3];
        for (int j = 0; j < 3; ++j) v[j] = m.kcart[3 * k + j] + m.gimg[img][j];
        const double dist = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (dist > tol && dist > dnn0 + tol) {
          if (dist < dnn1 - tol) {
            dnn1 = dist;
            counter = 0;
          }
          if (dist > dnn1 - tol && dist < dnn1 + tol) ++counter;
        }
      }
    }
    dnn[s] = dnn1;
    multi[s] = counter;
    dnn0 = dnn1;
    dnn1 = eta;
  }

  int num_shells = 0;
  int shell_list[kMaxShells];
  double bweight[kMaxShells] = {};
  int maxmulti = 0;
  for (int s = 0; s < search_shells; ++s) maxmulti = std::max(maxmulti, multi[s]);
  std::vector<double> bvec((size_t)kMaxShells * 3 * std::max(maxmulti, 1));
  static const double target[6] = {1, 1, 1, 0, 0, 0};

  auto build_amat = [&](int ns, double A[6][kMaxShells]) {
    for (int r = 0; r < 6; ++r)
      for (int s = 0; s < kMaxShells; ++s) A[r][s] = 0.0;
    for (int s = 0; s < ns; ++s) {
      const int mu = multi[shell_list[s]];
      for (int b = 0; b < mu; ++b) {
        const double* v = &bvec[((size_t)s * maxmulti + b) * 3];
        A[0][s] += v[0] * v[0];
        A[1][s] += v[1] * v[1];
        A[2][s] += v[2] * v[2];
        A[3][s] += v[0] * v[1];
        A[4][s] += v[1] * v[2];
        A[5][s] += v[2] * v[0];
      }
    }
  };
  auto b1_ok = [&](int ns) {
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j) {
        double delta = 0;
        for (int s = 0; s < ns; ++s) {
          const int mu = multi[shell_list[s]];
          for (int b = 0; b < mu; ++b) {
            const double* v = &bvec[((size_t)s * maxmulti + b) * 3];
            delta += bweight[s] * v[i] * v[j];
          }
        }
        if (i == j && std::fabs(delta - 1.0) > tol) return false;
        if (i != j && std::fabs(delta) > tol) return false;
      }
    return true;
  };

  if (num_shells_in <= 0) {
    // kmesh_shell_automatic
    bool b1sat = false;
    for (int s = 0; s < search_shells; ++s) {
      const int cur = num_shells;
      if (cur >= kMaxShells) break;
      if (multi[s] <= 0) continue;
      double* bv = &bvec[(size_t)cur * maxmulti * 3];
      if (shell_bvectors(m, nk, dnn[s], tol, multi[s], bv) < multi[s]) {
        set_err(errmsg, errlen, "kmesh_get_bvector: Not enough bvectors found");
        return W90GPU_ERR_FATAL;
      }
      bool lpar = false;
      for (int bn = 0; bn < multi[s] && !lpar; ++bn)
        for (int ls = 0; ls < cur && !lpar; ++ls)
          for (int b = 0; b < multi[shell_list[ls]]; ++b) {
            const double* u = bv + 3 * bn;
            const double* v = &bvec[((size_t)ls * maxmulti + b) * 3];
            const double uv = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
            const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
            const double vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            if (std::fabs(std::fabs(uv / std::sqrt(uu * vv)) - 1.0) < kEps6) {
              lpar = true;
              break;
            }
          }
      if (lpar) continue;
      shell_list[num_shells++] = s;
      double A[6][kMaxShells], sv[kMaxShells], w[kMaxShells];
      build_amat(num_shells, A);
      small_svd_solve(num_shells, A, target, sv, w);
      bool small = false;
      for (int i = 0; i < num_shells; ++i)
        if (std::fabs(sv[i]) < kEps5) small = true;
      if (small) {
        if (num_shells == 1) {
          set_err(errmsg, errlen,
                  "kmesh_shell_automatic: Singular Value Decomposition has found a very small singular value");
          return W90GPU_ERR_FATAL;
        }
        --num_shells;
        continue;
      }
      for (int i = 0; i < num_shells; ++i) bweight[i] = w[i];
      b1sat = b1_ok(num_shells);
      if (b1sat) break;
    }
    if (!b1sat) {
      set_err(errmsg, errlen, "kmesh_get_automatic: unable to satisfy B1 with the searched shells");
      return W90GPU_ERR_FATAL;
    }
  } else {
    // kmesh_shell_fixed: user shell_list (1-based)
    if (num_shells_in > kMaxShells || !shell_list_in) {
      set_err(errmsg, errlen, "kmesh_shell_fixed: invalid shell_list");
      return W90GPU_ERR_INPUT;
    }
    for (int i = 0; i < num_shells_in; ++i) {
      const int s = shell_list_in[i] - 1;
      if (s < 0 || s >= search_shells) {
        set_err(errmsg, errlen, "kmesh_shell_fixed: shell_list entry out of range");
        return W90GPU_ERR_INPUT;
      }
      shell_list[num_shells] = s;
      if (shell_bvectors(m, nk, dnn[s], tol, multi[s], &bvec[(size_t)num_shells * maxmulti * 3]) < multi[s]) {
        set_err(errmsg, errlen, "kmesh_get_bvector: Not enough bvectors found");
        return W90GPU_ERR_FATAL;
      }
      ++num_shells;
    }
    double A[6][kMaxShells], sv[kMaxShells], w[kMaxShells];
    build_amat(num_shells, A);
    small_svd_solve(num_shells, A, target, sv, w);
    for (int i = 0; i < num_shells; ++i) {
      if (std::fabs(sv[i]) < kEps5) {
        set_err(errmsg, errlen, "kmesh_shell_fixed: Singular Value Decomposition has found a very small singular value");
        return W90GPU_ERR_FATAL;
      }
      bweight[i] = w[i];
    }
    if (!skip_b1_tests && !b1_ok(num_shells)) {
      set_err(errmsg, errlen, "kmesh_shell_fixed: B1 condition not satisfied by the given shells");
      return W90GPU_ERR_FATAL;
    }
  }

  int nntot = 0;
  for (int s = 0; s < num_shells; ++s) nntot += multi[shell_list[s]];
  if (nntot > kNumNNMax) {
    set_err(errmsg, errlen, "kmesh_get: something wrong, found too many nearest neighbours");
    return W90GPU_ERR_FATAL;
  }

  std::vector<double> wb_local(nntot);
  {
    int x = 0;
    for (int s = 0; s < num_shells; ++s)
      for (int b = 0; b < multi[shell_list[s]]; ++b) wb_local[x++] = bweight[s];
  }

  // neighbour lists (kmesh.F90:319-345)
  std::vector<double> bkl(3 * (size_t)nntot * nk);
  for (int k = 0; k < nk; ++k) {
    int x = 0;
    for (int s = 0; s < num_shells; ++s) {
      const int sh = shell_list[s];
      int found = 0;
      for (int img = 0; img < kNImg && found < multi[sh]; ++img) {
        for (int k2 = 0; k2 < nk; ++k2) {
          double v[3], d2 = 0;
          for (int j = 0; j < 3; ++j) {
            v[j] = m.gimg[img][j] + m.kcart[3 * k2 + j];
            const double dd = m.kcart[3 * k + j] - v[j];
            d2 += dd * dd;
          }
          const double dist = std::sqrt(d2);
          if (dist >= dnn[sh] * (1 - tol) && dist <= dnn[sh] * (1 + tol)) {
            nnlist[k + (size_t)nk * x] = k2 + 1;
            for (int j = 0; j < 3; ++j) {
              nncell[j + 3 * (k + (size_t)nk * x)] = m.lmn[img][j];
              bkl[j + 3 * (x + (size_t)nntot * k)] = v[j] - m.kcart[3 * k + j];
            }
            ++x;
            ++found;
          }
          if (found == multi[sh]) break;
        }
      }
      if (found != multi[sh]) {
        set_err(errmsg, errlen, "kmesh_get: failed to find all neighbours of a k-point");
        return W90GPU_ERR_FATAL;
      }
    }
  }

  // every k sees the same shell radii (kmesh.F90:377-392)
  for (int k = 0; k < nk; ++k)
    for (int x = 0; x < nntot; ++x) {
      double b1 = 0, bn = 0;
      for (int j = 0; j < 3; ++j) {
        b1 += bkl[j + 3 * x] * bkl[j + 3 * x];
        bn += bkl[j + 3 * (x + (size_t)nntot * k)] * bkl[j + 3 * (x + (size_t)nntot * k)];
      }
      if (std::fabs(std::sqrt(b1) - std::sqrt(bn)) > tol) {
        set_err(errmsg, errlen, "Non-symmetric k-point neighbours!");
        return W90GPU_ERR_FATAL;
      }
    }

  if (!skip_b1_tests) {
    for (int k = 0; k < nk; ++k)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double dd = 0;
          for (int x = 0; x < nntot; ++x)
            dd += wb_local[x] * bkl[i + 3 * (x + (size_t)nntot * k)] * bkl[j + 3 * (x + (size_t)nntot * k)];
          if ((i == j && std::fabs(dd - 1.0) > tol) || (i != j && std::fabs(dd) > tol)) {
            set_err(errmsg, errlen, "Error with B1 Condition");
            return W90GPU_ERR_FATAL;
          }
        }
  }

  double wt = 0.0;
  for (int x = 0; x < nntot; ++x) {
    wt += wb_local[x];
    wb[x] = wb_local[x];
  }
  *wbtot = wt;
  std::memcpy(bk, bkl.data(), sizeof(double) * bkl.size());
  *nntot_out = nntot;
  set_err(errmsg, errlen, "");
  return W90GPU_OK;
}
