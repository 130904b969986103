/*
 * CPU restatement of wannier90's disentanglement (dis_main for k-points,
 * without gamma-only, site symmetry or dis_spheres), followed by
 * setup_m_loc.  TEST INFRASTRUCTURE ONLY: the checker the HIP library is
 * compared against; nothing in the product links it.
 *
 * Each function follows the loop order of the reference routine it cites
 * and calls the same LAPACK/BLAS routines (zhpevx, zgesvd, zgemm from the
 * MKL in /opt/conda/lib), so that it reproduces the reference to rounding,
 * eigenvector phases included.
 *
 *   oracle_dis_windows      dis_windows              src/disentangle.F90:851-1171
 *   oracle_dis_project      dis_project              :1173-1429
 *   oracle_dis_proj_froz    dis_proj_froz            :1431-1891
 *   oracle_dis_orthonorm    internal_check_orthonorm :372-451
 *   (slimming of M)         internal_slim_m          :453-529
 *   oracle_dis_extract      dis_extract              :1893-2864 (with
 *                           internal_zmatrix :2922-2988 and
 *                           internal_test_convergence :2866-2920)
 *   oracle_dis_main         dis_main :31-284 (the M rotation :229-241,
 *                           internal_find_u :531-698), then setup_m_loc
 *                           :286-370
 *
 * Arrays are Fortran layout: eig(nb,nk), A and U_opt (nb,nw,nk), M
 * (nb,nb,nn,nk) on input, U (nw,nw,nk), M out (nw,nw,nn,nk), nnlist(nk,nn)
 * 1-based, wb(nn).  Window indices are 1-based as in the reference.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex zc;

extern void zgemm_(const char* ta, const char* tb, const int* m, const int* n, const int* k,
                   const zc* alpha, const zc* a, const int* lda, const zc* b, const int* ldb,
                   const zc* beta, zc* c, const int* ldc, size_t, size_t);
extern void zgesvd_(const char* jobu, const char* jobvt, const int* m, const int* n, zc* a,
                    const int* lda, double* s, zc* u, const int* ldu, zc* vt, const int* ldvt,
                    zc* work, const int* lwork, double* rwork, int* info, size_t, size_t);
extern void zhpevx_(const char* jobz, const char* range, const char* uplo, const int* n, zc* ap,
                    const double* vl, const double* vu, const int* il, const int* iu,
                    const double* abstol, int* m, double* w, zc* z, const int* ldz, zc* work,
                    double* rwork, int* iwork, int* ifail, int* info, size_t, size_t, size_t);

#define I2(i, j, n) ((size_t)(i) + (size_t)(j) * (size_t)(n))

/* error codes of the restatement (the reference's messages in comments) */
enum {
  DIS_OK = 0,
  DIS_EMPTY_WINDOW = 1,   /* dis_windows: The outer energy window contains no eigenvalues */
  DIS_FEW_STATES = 2,     /* dis_windows: Energy window contains fewer states than number of target WFs */
  DIS_MANY_FROZEN = 3,    /* dis_windows: More states in the frozen window than target WFs */
  DIS_PROJECT_UNIT = 4,   /* dis_project: Error in unitarity of initial U in dis_project */
  DIS_LAPACK = 5,         /* zgesvd / zhpevx failure */
  DIS_PROJ_FROZ = 6,      /* dis_proj_froz: error (QPQ not Hermitian, eigenvalues outside [0,1], ortho-fix) */
  DIS_ORTHONORM = 7       /* internal_check_orthonorm */
};

typedef struct {
  int nb, nw, nk, nn;
  const int* nnlist; /* (nk,nn) */
  const double* wb;  /* (nn) */
  double wbtot;
} DisProblem;

typedef struct {
  double win_min, win_max, froz_min, froz_max;
  int frozen_states;
  int num_iter;
  double mix_ratio, conv_tol;
  int conv_window;
} DisControl;

static const zc one = 1.0, zero = 0.0;

static int hpev_all(int n, zc* ap, double* w, zc* z, int ldz) {
  /* ZHPEVX('V','A','U', ...) as the reference calls it, all eigenpairs ascending */
  const double vl = 0.0, vu = 0.0, abstol = -1.0;
  const int il = 0, iu = 0;
  int m = 0, info = 0;
  zc* work = malloc(sizeof(zc) * 2 * (size_t)(n > 0 ? n : 1));
  double* rwork = malloc(sizeof(double) * 7 * (size_t)(n > 0 ? n : 1));
  int* iwork = malloc(sizeof(int) * 5 * (size_t)(n > 0 ? n : 1));
  int* ifail = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  zhpevx_("V", "A", "U", &n, ap, &vl, &vu, &il, &iu, &abstol, &m, w, z, &ldz, work, rwork, iwork, ifail,
          &info, 1, 1, 1);
  free(work);
  free(rwork);
  free(iwork);
  free(ifail);
  return info;
}

/* dis_windows (:851-1171) without dis_spheres.  Outputs 1-based nfirst,
 * ndim, nfroz, indxfroz(nb,nk), indxnfroz(nb,nk) relative to the window;
 * lfrozen(nb,nk) 0/1; eigval_opt(nb,nk) slimmed to the window. */
int oracle_dis_windows(const DisProblem* P, const DisControl* c, const double* eig, int* nfirst, int* ndim,
                       int* nfroz, int* indxfroz, int* indxnfroz, int* lfrozen, double* eigopt, int* linner) {
  const int nb = P->nb, nk = P->nk, nw = P->nw;
  memcpy(eigopt, eig, sizeof(double) * (size_t)nb * nk);
  *linner = 0;
  for (int k = 0; k < nk; ++k) {
    const double* e = eigopt + (size_t)k * nb;
    if (e[0] > c->win_max || e[nb - 1] < c->win_min) return DIS_EMPTY_WINDOW;
    int imin = 0, imax = 0;
    for (int i = 1; i <= nb; ++i) {
      if (imin == 0) {
        if (e[i - 1] >= c->win_min && e[i - 1] <= c->win_max) imin = i;
        imax = i;
      }
      if (e[i - 1] <= c->win_max) imax = i;
    }
    ndim[k] = imax - imin + 1;
    nfirst[k] = imin;
    if (ndim[k] < nw) return DIS_FEW_STATES;
    for (int i = 0; i < ndim[k]; ++i) lfrozen[I2(i, k, nb)] = 0;
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
    nfroz[k] = kmax - kmin + 1;
    if (nfroz[k] > nw) return DIS_MANY_FROZEN;
    if (nfroz[k] > 0) *linner = 1;
    for (int i = 1; i <= nfroz[k]; ++i) {
      indxfroz[I2(i - 1, k, nb)] = kmin + i - 1;
      lfrozen[I2(kmin + i - 2, k, nb)] = 1;
    }
    int i = 0;
    for (int j = 1; j <= ndim[k]; ++j)
      if (!lfrozen[I2(j - 1, k, nb)]) indxnfroz[I2(i++, k, nb)] = j;
    double* eo = eigopt + (size_t)k * nb;
    for (int q = 1; q <= ndim[k]; ++q) eo[q - 1] = eo[nfirst[k] + q - 2];
    for (int q = ndim[k] + 1; q <= nb; ++q) eo[q - 1] = 0.0;
  }
  return DIS_OK;
}

/* dis_project (:1173-1429): A slimmed to the window, U_opt = Z V^+ from its
 * SVD, A rebuilt as Z S V^+, and the unitarity check of U_opt (1e-5). */
int oracle_dis_project(const DisProblem* P, const int* nfirst, const int* ndim, zc* a, zc* uopt) {
  const int nb = P->nb, nw = P->nw, nk = P->nk;
  zc* cz = malloc(sizeof(zc) * nb * nb);
  zc* cv = malloc(sizeof(zc) * nb * nb);
  double* sv = malloc(sizeof(double) * nb);
  const int lwork = 4 * nb;
  zc* work = malloc(sizeof(zc) * lwork);
  double* rwork = malloc(sizeof(double) * 5 * nb);
  int rc = DIS_OK;
  for (int k = 0; k < nk && !rc; ++k) {
    zc* ak = a + (size_t)k * nb * nw;
    if (ndim[k] != nb) {
      for (int j = 0; j < nw; ++j) {
        for (int i = 0; i < ndim[k]; ++i) ak[I2(i, j, nb)] = ak[I2(nfirst[k] + i - 1, j, nb)];
        for (int i = ndim[k]; i < nb; ++i) ak[I2(i, j, nb)] = 0.0;
      }
    }
  }
  for (int k = 0; k < nk && !rc; ++k) {
    zc* ak = a + (size_t)k * nb * nw;
    zc* uk = uopt + (size_t)k * nb * nw;
    int info = 0, m = ndim[k], n = nw;
    zgesvd_("A", "A", &m, &n, ak, &nb, sv, cz, &nb, cv, &nb, work, &lwork, rwork, &info, 1, 1);
    if (info) {
      rc = DIS_LAPACK;
      break;
    }
    for (size_t q = 0; q < (size_t)nb * nw; ++q) uk[q] = 0.0, ak[q] = 0.0;
    for (int j = 0; j < nw; ++j)
      for (int i = 0; i < ndim[k]; ++i)
        for (int l = 0; l < nw; ++l) {
          uk[I2(i, j, nb)] += cz[I2(i, l, nb)] * cv[I2(l, j, nb)];
          ak[I2(i, j, nb)] += cz[I2(i, l, nb)] * sv[l] * cv[I2(l, j, nb)];
        }
    for (int i = 0; i < nw && !rc; ++i)
      for (int j = 0; j < nw; ++j) {
        zc t = 0.0;
        for (int q = 0; q < ndim[k]; ++q) t += uk[I2(q, j, nb)] * conj(uk[I2(q, i, nb)]);
        if ((i == j && cabs(t - 1.0) > 1e-5) || (i != j && cabs(t) > 1e-5)) {
          rc = DIS_PROJECT_UNIT;
          break;
        }
      }
  }
  free(cz);
  free(cv);
  free(sv);
  free(work);
  free(rwork);
  return rc;
}

/* dis_proj_froz (:1431-1891): the leading eigenvectors of Q P Q (P the
 * projector of the projected states, Q the complement of the frozen
 * states) fill the non-frozen columns; the frozen columns are unit vectors
 * at the frozen bands.  The ortho-fix for zero eigenvalues follows the
 * reference (eigenvectors orthogonal to the frozen columns of U_opt, taken
 * from the top down). */
int oracle_dis_proj_froz(const DisProblem* P, const int* ndim, const int* nfroz, const int* indxfroz,
                         const int* lfrozen, zc* uopt) {
  const int nb = P->nb, nw = P->nw, nk = P->nk;
  const double eps8 = 1e-8;
  zc* cps = malloc(sizeof(zc) * nb * nb);
  zc* cq = malloc(sizeof(zc) * nb * nb);
  zc* cpq = malloc(sizeof(zc) * nb * nb);
  zc* cqpq = malloc(sizeof(zc) * nb * nb);
  zc* cap = malloc(sizeof(zc) * nb * (nb + 1) / 2);
  zc* cz = malloc(sizeof(zc) * nb * nb);
  double* w = malloc(sizeof(double) * nb);
  int* vmap = malloc(sizeof(int) * nb);
  int rc = DIS_OK;
  for (int k = 0; k < nk && !rc; ++k) {
    zc* uk = uopt + (size_t)k * nb * nw;
    const int nd = ndim[k];
    if (nw > nfroz[k]) {
      for (size_t q = 0; q < (size_t)nb * nb; ++q) cq[q] = 0.0, cps[q] = 0.0;
      for (int n = 0; n < nd; ++n) {
        for (int m = 0; m < nd; ++m)
          for (int l = 0; l < nw; ++l) cps[I2(m, n, nb)] += uk[I2(m, l, nb)] * conj(uk[I2(n, l, nb)]);
        if (!lfrozen[I2(n, k, nb)]) cq[I2(n, n, nb)] = 1.0;
      }
      for (size_t q = 0; q < (size_t)nb * nb; ++q) cpq[q] = 0.0, cqpq[q] = 0.0;
      for (int n = 0; n < nd; ++n)
        for (int m = 0; m < nd; ++m)
          for (int l = 0; l < nd; ++l) cpq[I2(m, n, nb)] += cps[I2(m, l, nb)] * cq[I2(l, n, nb)];
      for (int n = 0; n < nd; ++n)
        for (int m = 0; m < nd; ++m)
          for (int l = 0; l < nd; ++l) cqpq[I2(m, n, nb)] += cq[I2(m, l, nb)] * cpq[I2(l, n, nb)];
      for (int n = 0; n < nd && !rc; ++n)
        for (int m = 0; m <= n; ++m)
          if (cabs(cqpq[I2(m, n, nb)] - conj(cqpq[I2(n, m, nb)])) > eps8) rc = DIS_PROJ_FROZ;
      if (rc) break;
      for (int n = 0; n < nd; ++n)
        for (int m = 0; m <= n; ++m) cap[m + (size_t)n * (n + 1) / 2] = cqpq[I2(m, n, nb)];
      int il = nd - (nw - nfroz[k]) + 1;
      const int iu = nd;
      if (hpev_all(nd, cap, w, cz, nb)) {
        rc = DIS_LAPACK;
        break;
      }
      for (int j = 0; j < nd; ++j)
        if (w[j] < -eps8 || w[j] > 1.0 + eps8) rc = DIS_PROJ_FROZ;
      if (rc) break;
      int nzero = 0, goods = 0;
      for (int j = nd; j >= nd - (nw - nfroz[k]) + 1; --j) {
        if (w[j - 1] < eps8) nzero++;
        else goods++;
      }
      if (nzero > 0) {
        for (int q = 0; q < nb; ++q) vmap[q] = 0;
        int counter = 1;
        for (int j = nd; j >= nd - goods + 1; --j) vmap[counter++ - 1] = j;
        for (int f = 1; f <= nzero; ++f) {
          for (int v = nd; v >= 1; --v) {
            int used = 0;
            for (int q = 0; q < nb; ++q)
              if (vmap[q] == v) used = 1;
            if (used) continue;
            int take = 1;
            for (int m = 0; m < nfroz[k]; ++m) {
              zc t = 0.0;
              for (int j = 0; j < nd; ++j) t += conj(uk[I2(j, m, nb)]) * cz[I2(j, v - 1, nb)];
              if (cabs(t) > eps8) take = 0;
            }
            if (take) {
              vmap[goods + f - 1] = v;
              break;
            }
          }
        }
        for (int l = 0; l < nw - nfroz[k]; ++l)
          if (vmap[l] == 0) rc = DIS_PROJ_FROZ;
        if (rc) break;
        counter = 1;
        for (int l = nfroz[k]; l < nw; ++l, ++counter)
          for (int i = 0; i < nd; ++i) uk[I2(i, l, nb)] = cz[I2(i, vmap[counter - 1] - 1, nb)];
      } else {
        for (int l = nfroz[k]; l < nw; ++l, ++il)
          for (int i = 0; i < nd; ++i) uk[I2(i, l, nb)] = cz[I2(i, il - 1, nb)];
        if (il - 1 != iu) {
          rc = DIS_PROJ_FROZ;
          break;
        }
      }
    }
    for (int l = 0; l < nfroz[k]; ++l) {
      for (int i = 0; i < nb; ++i) uk[I2(i, l, nb)] = 0.0;
      uk[I2(indxfroz[I2(l, k, nb)] - 1, l, nb)] = 1.0;
    }
  }
  free(cps);
  free(cq);
  free(cpq);
  free(cqpq);
  free(cap);
  free(cz);
  free(w);
  free(vmap);
  return rc;
}

/* internal_check_orthonorm (:372-451): columns of U_opt orthonormal within 1e-8 */
int oracle_dis_orthonorm(const DisProblem* P, const int* ndim, const zc* uopt) {
  const int nb = P->nb, nw = P->nw;
  for (int k = 0; k < P->nk; ++k) {
    const zc* uk = uopt + (size_t)k * nb * nw;
    for (int l = 0; l < nw; ++l)
      for (int m = 0; m <= l; ++m) {
        zc t = 0.0;
        for (int j = 0; j < ndim[k]; ++j) t += uk[I2(j, m, nb)] * conj(uk[I2(j, l, nb)]);
        if (l == m && cabs(t - 1.0) > 1e-8) return DIS_ORTHONORM;
        if (l != m && cabs(t) > 1e-8) return DIS_ORTHONORM;
      }
  }
  return DIS_OK;
}

/* internal_slim_m (:453-529): M(k,b) restricted to the windows of k and k+b,
 * moved to the top-left corner, zero elsewhere. */
static void slim_m(const DisProblem* P, const int* nfirst, const int* ndim, zc* m) {
  const int nb = P->nb, nn = P->nn;
  zc* t = malloc(sizeof(zc) * nb * nb);
  for (int k = 0; k < P->nk; ++k)
    for (int b = 0; b < nn; ++b) {
      const int k2 = P->nnlist[I2(k, b, P->nk)] - 1;
      zc* mk = m + ((size_t)k * nn + b) * nb * nb;
      for (int j = 0; j < ndim[k2]; ++j)
        for (int i = 0; i < ndim[k]; ++i) t[I2(i, j, nb)] = mk[I2(nfirst[k] + i - 1, nfirst[k2] + j - 1, nb)];
      for (size_t q = 0; q < (size_t)nb * nb; ++q) mk[q] = 0.0;
      for (int j = 0; j < ndim[k2]; ++j)
        for (int i = 0; i < ndim[k]; ++i) mk[I2(i, j, nb)] = t[I2(i, j, nb)];
    }
  free(t);
}

/* internal_zmatrix (:2922-2988) */
static void zmatrix(const DisProblem* P, const zc* m, const zc* uopt, const int* ndim, const int* nfroz,
                    const int* indxnfroz, int k, zc* cbw, zc* cmtrx) {
  const int nb = P->nb, nw = P->nw, nn = P->nn;
  for (size_t q = 0; q < (size_t)nb * nb; ++q) cmtrx[q] = 0.0;
  const int ndimk = ndim[k] - nfroz[k];
  for (int b = 0; b < nn; ++b) {
    const int k2 = P->nnlist[I2(k, b, P->nk)] - 1;
    zgemm_("N", "N", &nb, &nw, &ndim[k2], &one, m + ((size_t)k * nn + b) * nb * nb, &nb,
           uopt + (size_t)k2 * nb * nw, &nb, &zero, cbw, &nb, 1, 1);
    for (int n = 1; n <= ndimk; ++n) {
      const int q = indxnfroz[I2(n - 1, k, nb)];
      for (int mm = 1; mm <= n; ++mm) {
        const int p = indxnfroz[I2(mm - 1, k, nb)];
        zc cs = 0.0;
        for (int l = 0; l < nw; ++l) cs += cbw[I2(p - 1, l, nb)] * conj(cbw[I2(q - 1, l, nb)]);
        cmtrx[I2(mm - 1, n - 1, nb)] += P->wb[b] * cs;
        cmtrx[I2(n - 1, mm - 1, nb)] = conj(cmtrx[I2(mm - 1, n - 1, nb)]);
      }
    }
  }
}

/* sum_b w_b || U_opt(k)^+ M(k,b) U_opt(k+b) ||^2 over the first nrows rows */
static double uMu(const DisProblem* P, const zc* m, const zc* uopt, const int* ndim, int k, int nrows, zc* cwb,
                  zc* cww) {
  const int nb = P->nb, nw = P->nw, nn = P->nn;
  double s = 0.0;
  for (int b = 0; b < nn; ++b) {
    const int k2 = P->nnlist[I2(k, b, P->nk)] - 1;
    zgemm_("C", "N", &nrows, &ndim[k2], &ndim[k], &one, uopt + (size_t)k * nb * nw, &nb,
           m + ((size_t)k * nn + b) * nb * nb, &nb, &zero, cwb, &nw, 1, 1);
    zgemm_("N", "N", &nrows, &nw, &ndim[k2], &one, cwb, &nw, uopt + (size_t)k2 * nb * nw, &nb, &zero, cww, &nw,
           1, 1);
    double r = 0.0;
    for (int n = 0; n < nw; ++n)
      for (int mm = 0; mm < nrows; ++mm) {
        const zc v = cww[I2(mm, n, nw)];
        r += creal(v) * creal(v) + cimag(v) * cimag(v);
      }
    s += P->wb[b] * r;
  }
  return s;
}

/* dis_extract (:1893-2864), serial.  history(3, num_iter) receives
 * (Omega_I(i-1) estimate, Omega_I(i), delta) per iteration; returns the
 * number of iterations done (negative: error code).  On exit omega_i holds
 * the final Omega_I, eigopt(1:nw,k) the eigenvalues of the Hamiltonian in
 * the optimal subspace and U_opt its eigenvectors. */
int oracle_dis_extract(const DisProblem* P, const DisControl* c, const zc* m, zc* uopt, double* eigopt,
                       const int* ndim, const int* nfroz, const int* indxnfroz, double* omega_i,
                       double* history, int* converged) {
  const int nb = P->nb, nw = P->nw, nk = P->nk;
  zc* cwb = malloc(sizeof(zc) * nw * nb);
  zc* cww = malloc(sizeof(zc) * nw * nw);
  zc* cbw = malloc(sizeof(zc) * nb * nw);
  zc* cap = malloc(sizeof(zc) * nb * (nb + 1) / 2);
  zc* cz = malloc(sizeof(zc) * nb * nb);
  double* w = malloc(sizeof(double) * nb);
  zc* zin = malloc(sizeof(zc) * nb * nb * (size_t)nk);
  zc* zout = malloc(sizeof(zc) * nb * nb * (size_t)nk);
  zc* cham = malloc(sizeof(zc) * nb * nb);
  zc* ceamp = malloc(sizeof(zc) * nb * nw);
  double* wk1 = malloc(sizeof(double) * nk);
  const int cw = c->conv_window > 0 ? c->conv_window : 1;
  double* hist = calloc((size_t)cw, sizeof(double));
  int iter, rc = 0, done = 0;
  double womegai = 0.0;
  *converged = 0;
  for (iter = 1; iter <= c->num_iter; ++iter) {
    if (iter == 1) {
      for (int k = 0; k < nk; ++k)
        if (nw > nfroz[k]) zmatrix(P, m, uopt, ndim, nfroz, indxnfroz, k, cbw, zin + (size_t)k * nb * nb);
    } else {
      for (int k = 0; k < nk; ++k) {
        if (nw <= nfroz[k]) continue;
        const int nd = ndim[k] - nfroz[k];
        zc* zi = zin + (size_t)k * nb * nb;
        const zc* zo = zout + (size_t)k * nb * nb;
        for (int i = 0; i < nd; ++i)
          for (int j = 0; j <= i; ++j) {
            zi[I2(j, i, nb)] = c->mix_ratio * zo[I2(j, i, nb)] + (1.0 - c->mix_ratio) * zi[I2(j, i, nb)];
            zi[I2(i, j, nb)] = conj(zi[I2(j, i, nb)]);
          }
      }
    }
    double womegai1 = 0.0;
    for (int k = 0; k < nk; ++k) wk1[k] = nw * P->wbtot;
    for (int k = 0; k < nk; ++k)
      if (nfroz[k] > 0) wk1[k] -= uMu(P, m, uopt, ndim, k, nfroz[k], cwb, cww);
    for (int k = 0; k < nk; ++k) {
      if (nw > nfroz[k]) {
        const int nd = ndim[k] - nfroz[k];
        const zc* zi = zin + (size_t)k * nb * nb;
        for (int j = 1; j <= nd; ++j)
          for (int i = 1; i <= j; ++i) cap[i - 1 + (size_t)(j - 1) * j / 2] = zi[I2(i - 1, j - 1, nb)];
        if (hpev_all(nd, cap, w, cz, nb)) {
          rc = -DIS_LAPACK;
          goto out;
        }
        int mm = nfroz[k];
        zc* uk = uopt + (size_t)k * nb * nw;
        for (int j = ndim[k] - nw + 1; j <= nd; ++j) {
          mm++;
          wk1[k] -= w[j - 1];
          for (int i = 0; i < ndim[k]; ++i) uk[I2(i, mm - 1, nb)] = 0.0;
          for (int i = 1; i <= nd; ++i) uk[I2(indxnfroz[I2(i - 1, k, nb)] - 1, mm - 1, nb)] = cz[I2(i - 1, j - 1, nb)];
        }
      }
      womegai1 += wk1[k];
    }
    womegai1 /= nk;
    womegai = 0.0;
    for (int k = 0; k < nk; ++k) womegai += nw * P->wbtot - uMu(P, m, uopt, ndim, k, nw, cwb, cww);
    womegai /= nk;
    const double delta = womegai1 / womegai - 1.0;
    if (history) {
      history[3 * (iter - 1) + 0] = womegai1;
      history[3 * (iter - 1) + 1] = womegai;
      history[3 * (iter - 1) + 2] = delta;
    }
    for (int k = 0; k < nk; ++k)
      if (nw > nfroz[k]) zmatrix(P, m, uopt, ndim, nfroz, indxnfroz, k, cbw, zout + (size_t)k * nb * nb);
    /* internal_test_convergence */
    if (iter <= cw) {
      hist[iter - 1] = delta;
    } else {
      for (int q = 0; q < cw - 1; ++q) hist[q] = hist[q + 1];
      hist[cw - 1] = delta;
    }
    if (iter >= c->conv_window) {
      int all = 1;
      for (int q = 0; q < cw; ++q)
        if (!(fabs(hist[q]) < c->conv_tol)) all = 0;
      if (all) {
        *converged = 1;
        done = iter;
        break;
      }
    }
  }
  if (!done) done = c->num_iter;
  *omega_i = womegai;
  /* the Hamiltonian in the optimal subspace and its eigenvectors */
  for (int k = 0; k < nk; ++k) {
    zc* uk = uopt + (size_t)k * nb * nw;
    double* e = eigopt + (size_t)k * nb;
    for (int j = 0; j < nw; ++j)
      for (int i = 0; i < nw; ++i) {
        zc s = 0.0;
        for (int l = 0; l < ndim[k]; ++l) s += conj(uk[I2(l, i, nb)]) * uk[I2(l, j, nb)] * e[l];
        cham[I2(i, j, nb)] = s;
      }
    for (int j = 1; j <= nw; ++j)
      for (int i = 1; i <= j; ++i) cap[i - 1 + (size_t)(j - 1) * j / 2] = cham[I2(i - 1, j - 1, nb)];
    if (hpev_all(nw, cap, w, cz, nb)) {
      rc = -DIS_LAPACK;
      goto out;
    }
    for (int j = 0; j < nw; ++j) e[j] = w[j];
    for (int j = 0; j < nw; ++j)
      for (int i = 0; i < ndim[k]; ++i) {
        zc s = 0.0;
        for (int l = 0; l < nw; ++l) s += cz[I2(l, j, nb)] * uk[I2(i, l, nb)];
        ceamp[I2(i, j, nb)] = s;
      }
    for (int j = 0; j < nw; ++j)
      for (int i = 0; i < ndim[k]; ++i) uk[I2(i, j, nb)] = ceamp[I2(i, j, nb)];
  }
  rc = done;
out:
  free(cwb);
  free(cww);
  free(cbw);
  free(cap);
  free(cz);
  free(w);
  free(zin);
  free(zout);
  free(cham);
  free(ceamp);
  free(wk1);
  free(hist);
  return rc;
}

/* dis_main (:31-284) with setup_m_loc (:286-370).  Inputs: eig(nb,nk),
 * a(nb,nw,nk) (the projections, as the reference passes them in
 * u_matrix_opt), m(nb,nb,nn,nk), overwritten.  Outputs: uopt(nb,nw,nk),
 * u(nw,nw,nk), mw(nw,nw,nn,nk) = U^+ U_opt^+ M U_opt U (the wannierise
 * input), eigopt(nb,nk), nfirst/ndim (nk), omega_i, history(3,num_iter).
 * Returns the iterations done (>= 0) or -error code. */
int oracle_dis_main(const DisProblem* P, const DisControl* c, const double* eig, zc* a, zc* m, zc* uopt, zc* u,
                    zc* mw, double* eigopt, int* nfirst, int* ndim, int* nfroz_out, double* omega_i,
                    double* history, int* converged) {
  const int nb = P->nb, nw = P->nw, nk = P->nk, nn = P->nn;
  int* nfroz = nfroz_out;
  int* indxfroz = calloc((size_t)nb * nk, sizeof(int));
  int* indxnfroz = calloc((size_t)nb * nk, sizeof(int));
  int* lfrozen = calloc((size_t)nb * nk, sizeof(int));
  int linner = 0, rc, iters = 0;
  rc = oracle_dis_windows(P, c, eig, nfirst, ndim, nfroz, indxfroz, indxnfroz, lfrozen, eigopt, &linner);
  if (rc) goto out;
  rc = oracle_dis_project(P, nfirst, ndim, a, uopt);
  if (rc) goto out;
  if (linner) {
    rc = oracle_dis_proj_froz(P, ndim, nfroz, indxfroz, lfrozen, uopt);
    if (rc) goto out;
  }
  rc = oracle_dis_orthonorm(P, ndim, uopt);
  if (rc) goto out;
  slim_m(P, nfirst, ndim, m);
  iters = oracle_dis_extract(P, c, m, uopt, eigopt, ndim, nfroz, indxnfroz, omega_i, history, converged);
  if (iters < 0) {
    rc = -iters;
    goto out;
  }
  {
    /* overlaps between the optimal subspaces, in place in the top-left
       num_wann x num_wann corner of each block */
    zc* cwb = malloc(sizeof(zc) * nw * nb);
    zc* cww = malloc(sizeof(zc) * nw * nw);
    for (int k = 0; k < nk; ++k)
      for (int b = 0; b < nn; ++b) {
        const int k2 = P->nnlist[I2(k, b, nk)] - 1;
        zc* mk = m + ((size_t)k * nn + b) * nb * nb;
        zgemm_("C", "N", &nw, &ndim[k2], &ndim[k], &one, uopt + (size_t)k * nb * nw, &nb, mk, &nb, &zero, cwb,
               &nw, 1, 1);
        zgemm_("N", "N", &nw, &nw, &ndim[k2], &one, cwb, &nw, uopt + (size_t)k2 * nb * nw, &nb, &zero, cww, &nw,
               1, 1);
        for (int j = 0; j < nw; ++j)
          for (int i = 0; i < nw; ++i) mk[I2(i, j, nb)] = cww[I2(i, j, nw)];
      }
    /* internal_find_u: U = Z V^+ from the SVD of U_opt^+ A */
    zc* caa = malloc(sizeof(zc) * nw * nw);
    zc* cz = malloc(sizeof(zc) * nw * nw);
    zc* cv = malloc(sizeof(zc) * nw * nw);
    double* sv = malloc(sizeof(double) * nw);
    const int lwork = 4 * nw;
    zc* work = malloc(sizeof(zc) * lwork);
    double* rwork = malloc(sizeof(double) * 5 * nw);
    for (int k = 0; k < nk && !rc; ++k) {
      int info = 0;
      zgemm_("C", "N", &nw, &nw, &ndim[k], &one, uopt + (size_t)k * nb * nw, &nb, a + (size_t)k * nb * nw, &nb,
             &zero, caa, &nw, 1, 1);
      zgesvd_("A", "A", &nw, &nw, caa, &nw, sv, cz, &nw, cv, &nw, work, &lwork, rwork, &info, 1, 1);
      if (info) rc = DIS_LAPACK;
      zgemm_("N", "N", &nw, &nw, &nw, &one, cz, &nw, cv, &nw, &zero, u + (size_t)k * nw * nw, &nw, 1, 1);
    }
    /* zero the rows of U_opt below each window */
    for (int k = 0; k < nk; ++k)
      for (int j = 0; j < nw; ++j)
        for (int i = ndim[k]; i < nb; ++i) uopt[(size_t)k * nb * nw + I2(i, j, nb)] = 0.0;
    /* setup_m_loc: M in the gauge of U, num_wann x num_wann blocks */
    for (int k = 0; k < nk && !rc; ++k)
      for (int b = 0; b < nn; ++b) {
        const int k2 = P->nnlist[I2(k, b, nk)] - 1;
        zgemm_("C", "N", &nw, &nw, &nw, &one, u + (size_t)k * nw * nw, &nw, m + ((size_t)k * nn + b) * nb * nb,
               &nb, &zero, cwb, &nw, 1, 1);
        zgemm_("N", "N", &nw, &nw, &nw, &one, cwb, &nw, u + (size_t)k2 * nw * nw, &nw, &zero,
               mw + ((size_t)k * nn + b) * nw * nw, &nw, 1, 1);
      }
    free(cwb);
    free(cww);
    free(caa);
    free(cz);
    free(cv);
    free(sv);
    free(work);
    free(rwork);
  }
out:
  free(indxfroz);
  free(indxnfroz);
  free(lfrozen);
  return rc ? -rc : iters;
}
