/*
 * CPU restatement of wannier90's spread minimisation (wann_main without
 * disentanglement, non-gamma, no preconditioner / site symmetry / SLWF /
 * guiding centres / noise).  TEST INFRASTRUCTURE ONLY: this is the checker
 * the HIP library is compared against; nothing in the product links it.
 *
 * Every function follows the loop order of the reference routine it cites,
 * and uses the same third-party LAPACK/BLAS calls (zheev, zgemm, zdotc from
 * the MKL in /opt/conda/lib, GNU interface layer), so that it reproduces
 * the reference trajectory to rounding.
 *
 *   oracle_omega            wann_omega          src/wannierise.F90:1979-2284
 *   oracle_domega           wann_domega         src/wannierise.F90:2287-2569
 *   oracle_search_direction internal_search_direction  :1371-1525
 *   oracle_optimal_step     internal_optimal_step      :1528-1596
 *   oracle_new_u_and_m      internal_new_u_and_m       :1599-1751
 *   oracle_test_convergence internal_test_convergence  :1076-1164
 *   oracle_wann_main        wann_main                  :48-1072
 *   oracle_check_unitarity  wann_check_unitarity       :2594-2661
 *   oracle_lowdin           overlap_project            src/overlap.F90:758-952
 *
 * Arrays are Fortran layout: M(nw,nw,nn,nk), U(nw,nw,nk), cdq(nw,nw,nk),
 * nnlist(nk,nn) 1-based, bk(3,nn,nk), wb(nn), rave(3,nw).
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex zc;

/* MKL, GNU Fortran interface (hidden string-length arguments last) */
extern void zheev_(const char* jobz, const char* uplo, const int* n, zc* a, const int* lda, double* w,
                   zc* work, const int* lwork, double* rwork, int* info, size_t, size_t);
extern void zgemm_(const char* ta, const char* tb, const int* m, const int* n, const int* k,
                   const zc* alpha, const zc* a, const int* lda, const zc* b, const int* ldb,
                   const zc* beta, zc* c, const int* ldc, size_t, size_t);
extern zc zdotc_(const int* n, const zc* x, const int* incx, const zc* y, const int* incy);
extern void zgesvd_(const char* jobu, const char* jobvt, const int* m, const int* n, zc* a,
                    const int* lda, double* s, zc* u, const int* ldu, zc* vt, const int* ldvt,
                    zc* work, const int* lwork, double* rwork, int* info, size_t, size_t);

#define IDX2(i, j, n) ((size_t)(i) + (size_t)(j) * (n))

typedef struct {
  int nw, nk, nn;
  const int* nnlist;  /* (nk,nn) */
  const double* bk;   /* (3,nn,nk) */
  const double* wb;   /* (nn) */
  double wbtot;
} oracle_problem;

typedef struct {
  int num_iter, num_cg_steps, conv_window, lfixstep;
  double conv_tol, trial_step, fixed_step;
} oracle_control;

typedef struct {
  double om_i, om_d, om_od, om_tot;
} spread_t;

static inline const zc* mblk(const zc* m, const oracle_problem* p, int b, int k) {
  return m + ((size_t)k * p->nn + b) * (size_t)p->nw * p->nw;
}
static inline double bkv(const oracle_problem* p, int ind, int b, int k) {
  return p->bk[ind + 3 * ((size_t)b + (size_t)p->nn * k)];
}
static inline int nnl(const oracle_problem* p, int k, int b) {
  return p->nnlist[k + (size_t)p->nk * b] - 1;
}

/* utility_zgemm: c = op(a) op(b), n x n (src/utility.F90:65-96) */
static void zgemm_n(zc* c, const zc* a, char ta, const zc* b, char tb, int n) {
  const zc one = 1.0, zero = 0.0;
  zgemm_(&ta, &tb, &n, &n, &n, &one, a, &n, b, &n, &zero, c, &n, 1, 1);
}

/* wann_omega, non-SLWF branch, csheet = 1, sheet = 0 */
void oracle_omega(const oracle_problem* p, const double* mr, int first_pass, double omega_invariant,
                  double* rave, double* r2ave, double* rave2, double* ln_tmp, double* om_out) {
  const zc* m = (const zc*)mr;
  const int nw = p->nw, nn = p->nn, nk = p->nk;
  for (int k = 0; k < nk; ++k)
    for (int b = 0; b < nn; ++b)
      for (int n = 0; n < nw; ++n) {
        const zc z = mblk(m, p, b, k)[IDX2(n, n, nw)];
        ln_tmp[n + (size_t)nw * (b + (size_t)nn * k)] = atan2(cimag(z), creal(z));
      }
  for (int i = 0; i < 3 * nw; ++i) rave[i] = 0.0;
  for (int iw = 0; iw < nw; ++iw)
    for (int ind = 0; ind < 3; ++ind)
      for (int k = 0; k < nk; ++k)
        for (int b = 0; b < nn; ++b)
          rave[ind + 3 * iw] += p->wb[b] * bkv(p, ind, b, k) * ln_tmp[iw + (size_t)nw * (b + (size_t)nn * k)];
  for (int i = 0; i < 3 * nw; ++i) rave[i] = -rave[i] / (double)nk;
  for (int iw = 0; iw < nw; ++iw) {
    const double* r = rave + 3 * iw;
    rave2[iw] = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  }
  for (int iw = 0; iw < nw; ++iw) {
    r2ave[iw] = 0.0;
    for (int k = 0; k < nk; ++k)
      for (int b = 0; b < nn; ++b) {
        const zc z = mblk(m, p, b, k)[IDX2(iw, iw, nw)];
        const double mnn2 = creal(z * conj(z));
        const double l = ln_tmp[iw + (size_t)nw * (b + (size_t)nn * k)];
        r2ave[iw] += p->wb[b] * (1.0 - mnn2 + l * l);
      }
  }
  for (int iw = 0; iw < nw; ++iw) r2ave[iw] /= (double)nk;
  spread_t s;
  if (first_pass) {
    s.om_i = 0.0;
    for (int k = 0; k < nk; ++k)
      for (int b = 0; b < nn; ++b) {
        double summ = 0.0;
        const zc* blk = mblk(m, p, b, k);
        for (int mm = 0; mm < nw; ++mm)
          for (int n = 0; n < nw; ++n) {
            const zc z = blk[IDX2(n, mm, nw)];
            summ += creal(z * conj(z));
          }
        s.om_i += p->wb[b] * ((double)nw - summ);
      }
    s.om_i /= (double)nk;
  } else {
    s.om_i = omega_invariant;
  }
  s.om_od = 0.0;
  for (int k = 0; k < nk; ++k)
    for (int b = 0; b < nn; ++b) {
      const zc* blk = mblk(m, p, b, k);
      for (int mm = 0; mm < nw; ++mm)
        for (int n = 0; n < nw; ++n)
          if (mm != n) {
            const zc z = blk[IDX2(n, mm, nw)];
            s.om_od += p->wb[b] * creal(z * conj(z));
          }
    }
  s.om_od /= (double)nk;
  s.om_d = 0.0;
  for (int k = 0; k < nk; ++k)
    for (int b = 0; b < nn; ++b)
      for (int n = 0; n < nw; ++n) {
        const double brn = bkv(p, 0, b, k) * rave[0 + 3 * n] + bkv(p, 1, b, k) * rave[1 + 3 * n] +
                           bkv(p, 2, b, k) * rave[2 + 3 * n];
        const double x = ln_tmp[n + (size_t)nw * (b + (size_t)nn * k)] + brn;
        s.om_d += p->wb[b] * x * x;
      }
  s.om_d /= (double)nk;
  s.om_tot = s.om_i + s.om_d + s.om_od;
  om_out[0] = s.om_i;
  om_out[1] = s.om_d;
  om_out[2] = s.om_od;
  om_out[3] = s.om_tot;
}

/* wann_domega, non-SLWF gradient; cdodq (nw,nw,nk) complex */
void oracle_domega(const oracle_problem* p, const double* mr, double* cdodq_r, double* rave_out) {
  const zc* m = (const zc*)mr;
  zc* cdodq = (zc*)cdodq_r;
  const int nw = p->nw, nn = p->nn, nk = p->nk;
  double* ln = malloc(sizeof(double) * (size_t)nw * nn * nk);
  double* rnkb = malloc(sizeof(double) * (size_t)nw * nn * nk);
  double* rave = malloc(sizeof(double) * 3 * (size_t)nw);
  zc* cr = malloc(sizeof(zc) * (size_t)nw * nw);
  zc* crt = malloc(sizeof(zc) * (size_t)nw * nw);
  for (int k = 0; k < nk; ++k)
    for (int b = 0; b < nn; ++b)
      for (int n = 0; n < nw; ++n) {
        const zc z = mblk(m, p, b, k)[IDX2(n, n, nw)];
        ln[n + (size_t)nw * (b + (size_t)nn * k)] = p->wb[b] * atan2(cimag(z), creal(z));
      }
  for (int i = 0; i < 3 * nw; ++i) rave[i] = 0.0;
  for (int iw = 0; iw < nw; ++iw)
    for (int ind = 0; ind < 3; ++ind)
      for (int k = 0; k < nk; ++k)
        for (int b = 0; b < nn; ++b)
          rave[ind + 3 * iw] += bkv(p, ind, b, k) * ln[iw + (size_t)nw * (b + (size_t)nn * k)];
  for (int i = 0; i < 3 * nw; ++i) rave[i] = -rave[i] / (double)nk;
  if (rave_out) memcpy(rave_out, rave, sizeof(double) * 3 * nw);
  for (int k = 0; k < nk; ++k)
    for (int b = 0; b < nn; ++b)
      for (int n = 0; n < nw; ++n)
        rnkb[n + (size_t)nw * (b + (size_t)nn * k)] = bkv(p, 0, b, k) * rave[0 + 3 * n] +
                                                      bkv(p, 1, b, k) * rave[1 + 3 * n] +
                                                      bkv(p, 2, b, k) * rave[2 + 3 * n];
  memset(cdodq, 0, sizeof(zc) * (size_t)nw * nw * nk);
  const zc mhalf_i = CMPLX(0.0, -0.5);
  for (int k = 0; k < nk; ++k)
    for (int b = 0; b < nn; ++b) {
      const zc* blk = mblk(m, p, b, k);
      for (int n = 0; n < nw; ++n) {
        const zc mnn = blk[IDX2(n, n, nw)];
        for (int i = 0; i < nw; ++i) {
          crt[IDX2(i, n, nw)] = blk[IDX2(i, n, nw)] / mnn;
          cr[IDX2(i, n, nw)] = blk[IDX2(i, n, nw)] * conj(mnn);
        }
      }
      const double w = p->wb[b];
      const double* lnb = ln + (size_t)nw * (b + (size_t)nn * k);
      const double* rb = rnkb + (size_t)nw * (b + (size_t)nn * k);
      zc* g = cdodq + (size_t)k * nw * nw;
      for (int n = 0; n < nw; ++n)
        for (int mm = 0; mm < nw; ++mm) {
          zc v = g[IDX2(mm, n, nw)];
          v = v + w * 0.5 * (cr[IDX2(mm, n, nw)] - conj(cr[IDX2(n, mm, nw)]));
          v = v - (crt[IDX2(mm, n, nw)] * lnb[n] + conj(crt[IDX2(n, mm, nw)] * lnb[mm])) * mhalf_i;
          v = v - w * (crt[IDX2(mm, n, nw)] * rb[n] + conj(crt[IDX2(n, mm, nw)] * rb[mm])) * mhalf_i;
          g[IDX2(mm, n, nw)] = v;
        }
    }
  for (size_t i = 0; i < (size_t)nw * nw * nk; ++i) cdodq[i] = cdodq[i] / (double)nk * 4.0;
  free(ln);
  free(rnkb);
  free(rave);
  free(cr);
  free(crt);
}

/* internal_new_u_and_m: zheev of i*cdq, exp, U <- U E, M <- E^+ M E_{k+b} */
int oracle_new_u_and_m(const oracle_problem* p, double* cdq_r, double* u_r, double* m_r) {
  zc* cdq = (zc*)cdq_r;
  zc* u = (zc*)u_r;
  zc* m = (zc*)m_r;
  const int nw = p->nw, nn = p->nn, nk = p->nk;
  const size_t n2 = (size_t)nw * nw;
  zc* tmp = malloc(sizeof(zc) * n2);
  zc* cmtmp = malloc(sizeof(zc) * n2);
  double* evals = malloc(sizeof(double) * nw);
  int lwork = 4 * nw;
  zc* work = malloc(sizeof(zc) * lwork);
  double* rwork = malloc(sizeof(double) * (3 * nw > 2 ? 3 * nw - 2 : 1));
  int info = 0;
  for (int k = 0; k < nk; ++k) {
    zc* c = cdq + (size_t)k * n2;
    for (size_t i = 0; i < n2; ++i) tmp[i] = I * c[i];
    zheev_("V", "U", &nw, tmp, &nw, evals, work, &lwork, rwork, &info, 1, 1);
    if (info != 0) break;
    for (int i = 0; i < nw; ++i) {
      const zc ph = cexp(-I * evals[i]);
      for (int r = 0; r < nw; ++r) cmtmp[IDX2(r, i, nw)] = tmp[IDX2(r, i, nw)] * ph;
    }
    zgemm_n(c, cmtmp, 'N', tmp, 'C', nw);
  }
  if (info == 0) {
    for (int k = 0; k < nk; ++k) {
      zgemm_n(cmtmp, u + (size_t)k * n2, 'N', cdq + (size_t)k * n2, 'N', nw);
      memcpy(u + (size_t)k * n2, cmtmp, sizeof(zc) * n2);
    }
    for (int k = 0; k < nk; ++k)
      for (int b = 0; b < nn; ++b) {
        zc* blk = m + ((size_t)k * nn + b) * n2;
        zgemm_n(tmp, cdq + (size_t)k * n2, 'C', blk, 'N', nw);
        zgemm_n(cmtmp, tmp, 'N', cdq + (size_t)nnl(p, k, b) * n2, 'N', nw);
        memcpy(blk, cmtmp, sizeof(zc) * n2);
      }
  }
  free(tmp);
  free(cmtmp);
  free(evals);
  free(work);
  free(rwork);
  return info;
}

/* wann_check_unitarity: max deviation of U U^+ and U^+ U from 1 */
double oracle_check_unitarity(const oracle_problem* p, const double* u_r) {
  const zc* u = (const zc*)u_r;
  const int nw = p->nw;
  double dev = 0.0;
  for (int k = 0; k < p->nk; ++k) {
    const zc* a = u + (size_t)k * nw * nw;
    for (int i = 0; i < nw; ++i)
      for (int j = 0; j < nw; ++j) {
        zc c1 = 0, c2 = 0;
        for (int mm = 0; mm < nw; ++mm) {
          c1 += a[IDX2(i, mm, nw)] * conj(a[IDX2(j, mm, nw)]);
          c2 += a[IDX2(mm, j, nw)] * conj(a[IDX2(mm, i, nw)]);
        }
        if (i == j) {
          c1 -= 1.0;
          c2 -= 1.0;
        }
        dev = fmax(dev, fmax(cabs(c1), cabs(c2)));
      }
  }
  return dev;
}

/* overlap_project: U = Z V^+ from zgesvd of A, then M <- U^+(k) M U(k+b).
 * a (nw,nw,nk) complex in, u out, m rotated in place. */
int oracle_lowdin(const oracle_problem* p, const double* a_r, double* u_r, double* m_r) {
  const zc* a = (const zc*)a_r;
  zc* u = (zc*)u_r;
  zc* m = (zc*)m_r;
  const int nw = p->nw, nn = p->nn, nk = p->nk;
  const size_t n2 = (size_t)nw * nw;
  zc* tmp = malloc(sizeof(zc) * n2);
  zc* cz = malloc(sizeof(zc) * n2);
  zc* cv = malloc(sizeof(zc) * n2);
  double* sv = malloc(sizeof(double) * nw);
  double* rwork = malloc(sizeof(double) * 5 * nw);
  int lwork = 4 * nw;
  zc* work = malloc(sizeof(zc) * lwork);
  int info = 0;
  for (int k = 0; k < nk && info == 0; ++k) {
    memcpy(tmp, a + (size_t)k * n2, sizeof(zc) * n2);
    zgesvd_("A", "A", &nw, &nw, tmp, &nw, sv, cz, &nw, cv, &nw, work, &lwork, rwork, &info, 1, 1);
    zgemm_n(u + (size_t)k * n2, cz, 'N', cv, 'N', nw);
  }
  if (info == 0)
    for (int k = 0; k < nk; ++k)
      for (int b = 0; b < nn; ++b) {
        zc* blk = m + ((size_t)k * nn + b) * n2;
        zgemm_n(cv, u + (size_t)k * n2, 'C', blk, 'N', nw);
        zgemm_n(cz, cv, 'N', u + (size_t)nnl(p, k, b) * n2, 'N', nw);
        memcpy(blk, cz, sizeof(zc) * n2);
      }
  free(tmp);
  free(cz);
  free(cv);
  free(sv);
  free(rwork);
  free(work);
  return info;
}

/* internal_optimal_step */
static void optimal_step(const spread_t* wann, const spread_t* trial, double doda0, double trial_step,
                         double* alphamin, double* falphamin, int* lquad) {
  double fac = trial->om_tot - wann->om_tot, shift;
  if (fabs(fac) > DBL_MIN) {
    fac = 1.0 / fac;
    shift = 1.0;
  } else {
    fac = 1.0e6;
    shift = fac * trial->om_tot - fac * wann->om_tot;
  }
  const double eqb = fac * doda0;
  const double eqa = shift - eqb * trial_step;
  if (fabs(eqa / (fac * wann->om_tot)) > DBL_EPSILON) {
    *lquad = 1;
    *alphamin = -0.5 * eqb / eqa * (trial_step * trial_step);
    *falphamin = wann->om_tot - 0.25 * eqb * eqb / (fac * eqa) * (trial_step * trial_step);
  } else {
    *lquad = 0;
    *alphamin = trial_step;
    *falphamin = trial->om_tot;
  }
  if (doda0 * (*alphamin) > 0.0) {
    *lquad = 0;
    *alphamin = trial_step;
    *falphamin = trial->om_tot;
  }
}

/* wann_main: minimises in place.  history_out (num_iter+1, 4) receives per
 * iteration {Omega, Omega_D, Omega_OD, gcnorm1}; row 0 is the initial state.
 * stats: {iterations, converged}.  Returns 0, or the LAPACK info. */
int oracle_wann_main(const oracle_problem* p, const oracle_control* c, double* m_r, double* u_r,
                     double* centres, double* spreads, double* om_out, double* history_out,
                     int* stats) {
  const int nw = p->nw, nn = p->nn, nk = p->nk;
  const size_t n2 = (size_t)nw * nw;
  const size_t msz = n2 * nn * nk, usz = n2 * nk;
  zc* m = (zc*)m_r;
  zc* u = (zc*)u_r;
  zc* cdodq = calloc(usz, sizeof(zc));
  zc* cdq = calloc(usz, sizeof(zc));
  zc* cdqkeep = calloc(usz, sizeof(zc));
  zc* m0 = malloc(sizeof(zc) * msz);
  zc* u0 = malloc(sizeof(zc) * usz);
  double* ln = malloc(sizeof(double) * (size_t)nw * nn * nk);
  double* rave = malloc(sizeof(double) * 3 * nw);
  double* r2ave = malloc(sizeof(double) * nw);
  double* rave2 = malloc(sizeof(double) * nw);
  double* history = calloc(c->conv_window > 0 ? c->conv_window : 1, sizeof(double));
  double om[4];
  spread_t wann, old, trial;
  int info = 0;

  oracle_omega(p, m_r, 1, 0.0, rave, r2ave, rave2, ln, om);
  const double omega_invariant = om[0];
  wann.om_i = om[0];
  wann.om_d = om[1];
  wann.om_od = om[2];
  wann.om_tot = om[3];
  old.om_tot = 0.0;
  if (history_out) {
    history_out[0] = wann.om_tot;
    history_out[1] = wann.om_d;
    history_out[2] = wann.om_od;
    history_out[3] = 0.0;
  }
  int ncg = 0, conv_count = 0, lconverged = 0, lquad = 0, iter;
  double gcnorm0 = 0.0, gcnorm1 = 0.0, doda0 = 0.0, gcfac = 0.0, alphamin = 0.0, falphamin = 0.0;
  const int msize = (int)usz, one = 1;
  for (iter = 1; iter <= c->num_iter; ++iter) {
    oracle_domega(p, m_r, (double*)cdodq, NULL);
    /* internal_search_direction */
    gcnorm1 = creal(zdotc_(&msize, cdodq, &one, cdodq, &one));
    if (iter == 1 || ncg >= c->num_cg_steps) {
      gcfac = 0.0;
      ncg = 0;
    } else if (gcnorm0 > DBL_EPSILON) {
      gcfac = gcnorm1 / gcnorm0;
      if (gcfac > 3.0) {
        gcfac = 0.0;
        ncg = 0;
      } else {
        ncg++;
      }
    } else {
      gcfac = 0.0;
      ncg = 0;
    }
    gcnorm0 = gcnorm1;
    for (size_t i = 0; i < usz; ++i) cdq[i] = cdodq[i] + cdqkeep[i] * gcfac;
    doda0 = -creal(zdotc_(&msize, cdodq, &one, cdq, &one));
    doda0 = doda0 / (4.0 * p->wbtot);
    if (doda0 > 0.0) {
      if (ncg > 0) {
        memcpy(cdq, cdodq, sizeof(zc) * usz);
        ncg = 0;
        gcfac = 0.0;
        doda0 = -creal(zdotc_(&msize, cdodq, &one, cdq, &one));
        doda0 = doda0 / (4.0 * p->wbtot);
        if (doda0 > 0.0) {
          for (size_t i = 0; i < usz; ++i) cdq[i] = -cdq[i];
          doda0 = -doda0;
        }
      } else {
        for (size_t i = 0; i < usz; ++i) cdq[i] = -cdq[i];
        doda0 = -doda0;
      }
    }
    memcpy(cdqkeep, cdq, sizeof(zc) * usz);
    if (c->lfixstep) {
      alphamin = c->fixed_step;
    } else {
      const double f = c->trial_step / (4.0 * p->wbtot);
      for (size_t i = 0; i < usz; ++i) cdq[i] = cdqkeep[i] * f;
      memcpy(u0, u, sizeof(zc) * usz);
      memcpy(m0, m, sizeof(zc) * msz);
      if ((info = oracle_new_u_and_m(p, (double*)cdq, u_r, m_r))) break;
      oracle_omega(p, m_r, 0, omega_invariant, rave, r2ave, rave2, ln, om);
      trial.om_i = om[0];
      trial.om_d = om[1];
      trial.om_od = om[2];
      trial.om_tot = om[3];
      optimal_step(&wann, &trial, doda0, c->trial_step, &alphamin, &falphamin, &lquad);
    }
    if (c->lfixstep || lquad) {
      const double f = alphamin / (4.0 * p->wbtot);
      for (size_t i = 0; i < usz; ++i) cdq[i] = cdqkeep[i] * f;
      if (!c->lfixstep) {
        memcpy(u, u0, sizeof(zc) * usz);
        memcpy(m, m0, sizeof(zc) * msz);
      }
      if ((info = oracle_new_u_and_m(p, (double*)cdq, u_r, m_r))) break;
      old = wann;
      oracle_omega(p, m_r, 0, omega_invariant, rave, r2ave, rave2, ln, om);
      wann.om_i = om[0];
      wann.om_d = om[1];
      wann.om_od = om[2];
      wann.om_tot = om[3];
    } else {
      old = wann;
      wann = trial;
    }
    if (history_out) {
      history_out[4 * iter + 0] = wann.om_tot;
      history_out[4 * iter + 1] = wann.om_d;
      history_out[4 * iter + 2] = wann.om_od;
      history_out[4 * iter + 3] = gcnorm1;
    }
    /* internal_test_convergence (noise branch off) */
    if (c->conv_window > 1) {
      const double delta = wann.om_tot - old.om_tot;
      const int cw = c->conv_window;
      if (iter <= cw) {
        history[iter - 1] = delta;
      } else {
        for (int j = 0; j < cw - 1; ++j) history[j] = history[j + 1];
        history[cw - 1] = delta;
      }
      conv_count++;
      if (conv_count >= cw) {
        int ok = 1;
        for (int j = 0; j < cw; ++j)
          if (fabs(history[j]) > c->conv_tol) ok = 0;
        if (ok) lconverged = 1;
      }
    }
    if (lconverged) break;
  }
  if (iter > c->num_iter) iter = c->num_iter;
  if (centres) memcpy(centres, rave, sizeof(double) * 3 * nw);
  if (spreads)
    for (int n = 0; n < nw; ++n) spreads[n] = r2ave[n] - rave2[n];
  if (om_out) {
    om_out[0] = wann.om_i;
    om_out[1] = wann.om_d;
    om_out[2] = wann.om_od;
    om_out[3] = wann.om_tot;
  }
  if (stats) {
    stats[0] = iter;
    stats[1] = lconverged;
  }
  free(cdodq);
  free(cdq);
  free(cdqkeep);
  free(m0);
  free(u0);
  free(ln);
  free(rave);
  free(r2ave);
  free(rave2);
  free(history);
  return info;
}
