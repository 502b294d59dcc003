/* ORACLE — test infrastructure only.  A CPU restatement of asr's augmented-space recursion and KPM path, used by tests/ and by
 * the CPU-baseline leg of bench.py as the checker and the timed CPU baseline.  Nothing in the product path links or calls it.
 *
 * Each routine follows the operation order of the reference routine it names (file:line of mavensgroup/asr):
 *   orc_local_weights    src/asr_augcls.f90:362-377   local_weights
 *   orc_apply_rand       src/asr_augcls.f90:380-403   apply_rand      (at_link :239-248)
 *   orc_apply_h          src/asr_augcls.f90:449-489   apply_h         (btab built as :738-750)
 *   orc_augcls_apply     src/asr_augcls.f90:275-297   augcls_apply
 *   orc_rand_c           src/asr_augcls.f90:419-446   augcls_rand_c
 *   orc_project_c        src/asr_augcls.f90:329-348   augcls_project_c
 *   orc_lanczos(_ro)     src/reference/asr_recursion.f90:17-58   lanczos (reorth = .false. / .true.)
 *   orc_edc_chain        src/asr_edc.f90:75-164       edc_green without the adaptive depth (depth_tol = 0): a, b, m
 *   orc_hp_c             src/asr_augcls.f90:301-320   augcls_hp_c
 *   orc_ed3_kmul         src/asr_ed3.f90:99-129       ed3_kmul
 *   orc_ed3_start        src/asr_ed3.f90:131-144      ed3_start
 *   orc_ed3_chain        src/asr_ed3.f90:148-211      chain of route ed3 without the adaptive depth: a, b, c, m, closure
 *   orc_epi_start        src/asr_epi.f90:578-587      epi_start
 *   orc_epi_chain        src/asr_epi.f90:590-626      chain of the EPI blocks (route edc operator, full depth)
 *   orc_edc_chain_proj   src/asr_edc.f90:174-232      edc_green_proj: the chain and its projections on target states
 *   orc_real_apply       src/reference/asr_lmto_realization.f90:120-191   h_apply / real_apply (order 2), n species per site
 *   orc_kpm_moments      src/reference/asr_kpm.f90:39-65   kpm_moments (Rademacher vectors from a counter-based hash)
 *   orc_spectral_bounds  src/reference/asr_kpm.f90:14-36   Lanczos part (the eigenvalues are taken on the caller's side)
 * The chain batch routines run one chain per thread with a dynamic schedule, as chain_worker
 * (src/asr_class_scf.f90:1093) does.
 */
#include <complex.h>
#include <omp.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

typedef struct {
  int nn, nslot, nop, nsmax, ncls, nsp, ntype;
  const int *hopr, *hopg, *flipr, *flipg, *lst, *ncl, *ntyp, *stab_ptr, *stab, *pi, *nsc, *spc;
  const unsigned char *fz;
  const double *sq, *sg, *block, *block0, *mt;
} orc_op;

/* One chain's parameters on an operator: chi (nop), c/d/o/e (nsp x 9, row-major [species][orbital]), frozen species of the
 * start site (-1: none) and of a frozen pair partner (-1: none), mt (NULL: the operator's class tables).  fz[i] of the
 * operator: 1 = node on the start site, 2 = node on the pair partner (augcls_freeze_any), 0 = neither. */
typedef struct {
  const double *chi, *c, *d, *o, *e, *mt;
  int frozen, frozen2;
} orc_member;

/* the frozen species that applies at node i (-1: none): op%frozen(i) of the reference */
static inline int frozen_at(const orc_op *op, const orc_member *m, int i) {
  return op->fz[i] == 1 ? m->frozen : op->fz[i] == 2 ? m->frozen2 : -1;
}

static inline const double *mt_of(const orc_op *op, const orc_member *m) { return m->mt ? m->mt : op->mt; }

/* wt[kap][l][k][lam] = sum_s par[spc(s,kap)][lam] mt[kap][l][k][s] */
void orc_local_weights(const orc_op *op, const double *mt, const double *par, double *wt) {
  const int ns = op->nsmax;
  memset(wt, 0, sizeof(double) * op->ncls * ns * ns * 9);
  for (int kap = 0; kap < op->ncls; ++kap)
    for (int kk = 0; kk < op->nsc[kap]; ++kk)
      for (int k = 0; k < op->nsc[kap]; ++k)
        for (int s = 0; s < op->nsc[kap]; ++s)
          for (int lam = 0; lam < 9; ++lam)
            wt[((kap * ns + k) * ns + kk) * 9 + lam] +=
                par[op->spc[kap * ns + s] * 9 + lam] * mt[((kap * ns + k) * ns + kk) * ns + s];
}

/* w = X x in amplitude coordinates (apply_rand) */
void orc_apply_rand(const orc_op *op, const orc_member *m, const double *par, const double *x, double *w) {
  const int ns = op->nsmax;
  double *wt = (double *)malloc(sizeof(double) * op->ncls * ns * ns * 9);
  orc_local_weights(op, mt_of(op, m), par, wt);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < op->nn; ++i) {
    const int fsp = frozen_at(op, m, i);
    if (fsp >= 0) {
      for (int l = 0; l < 9; ++l) w[9 * i + l] = par[fsp * 9 + l] * x[9 * i + l];
      continue;
    }
    int kap = op->ncl[i], l0 = op->lst[i];
    double acc[9];
    for (int l = 0; l < 9; ++l) acc[l] = wt[((kap * ns + l0) * ns + l0) * 9 + l] * x[9 * i + l];
    for (int k = 0; k < op->nsc[kap]; ++k) {
      int r = op->flipr[i * ns + k];
      if (k == l0 || r < 0) continue;
      int g = op->flipg[i * ns + k];
      for (int l = 0; l < 9; ++l) {
        double tf = m->chi[g] * op->sg[g * 9 + l] * x[9 * r + op->pi[g * 9 + l]];
        acc[l] = acc[l] + wt[((kap * ns + l0) * ns + k) * 9 + l] * tf;
      }
    }
    for (int l = 0; l < 9; ++l) w[9 * i + l] = acc[l];
  }
  free(wt);
}

/* btab[t][g][k][l][l2] = sum_l' block[t][k][l][l'] sg(l', g) delta(l2, pi(l', g)) */
static double *make_btab(const orc_op *op) {
  const int ns = op->nslot, nop = op->nop;
  double *bt = (double *)calloc((size_t)op->ntype * nop * ns * 81, sizeof(double));
  for (int t = 0; t < op->ntype; ++t)
    for (int g = 0; g < nop; ++g)
      for (int k = 0; k < ns; ++k)
        for (int l = 0; l < 9; ++l)
          for (int jj = 0; jj < 9; ++jj)
            bt[((((size_t)t * nop + g) * ns + k) * 9 + l) * 9 + op->pi[g * 9 + jj]] +=
                op->block[(((size_t)t * ns + k) * 9 + l) * 9 + jj] * op->sg[g * 9 + jj];
  return bt;
}

/* hx = c x + D S D x in amplitude coordinates (apply_h); t1, t2, t3: work vectors of 9 nn */
void orc_apply_h(const orc_op *op, const orc_member *m, const double *btab, const double *x, double *hx, double *t1,
                 double *t2, double *t3) {
  const int ns = op->nslot, nop = op->nop;
  orc_apply_rand(op, m, m->d, x, t1);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < op->nn; ++i) {
    int t = op->ntyp[i];
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int l2 = 0; l2 < 9; ++l2) {
      double tl = t1[9 * i + l2];
      for (int l = 0; l < 9; ++l) acc[l] = acc[l] + op->block0[((size_t)t * 9 + l) * 9 + l2] * tl;
    }
    for (int k = 0; k < ns; ++k) {
      int r = op->hopr[i * ns + k];
      if (r < 0) continue;
      int g = op->hopg[i * ns + k];
      const double *bt = btab + (((size_t)t * nop + g) * ns + k) * 81;
      for (int l2 = 0; l2 < 9; ++l2) {
        double tl = m->chi[g] * t1[9 * r + l2];
        for (int l = 0; l < 9; ++l) acc[l] = acc[l] + bt[l * 9 + l2] * tl;
      }
    }
    for (int l = 0; l < 9; ++l) t2[9 * i + l] = acc[l];
  }
  orc_apply_rand(op, m, m->d, t2, t3);
  orc_apply_rand(op, m, m->c, x, hx);
  for (int i = 0; i < 9 * op->nn; ++i) hx[i] = hx[i] + t3[i];
}

/* y = H x in normalised coordinates, H = E_nu + h - h o h (augcls_apply).  wk: 8 work vectors of 9 nn. */
void orc_augcls_apply(const orc_op *op, const orc_member *m, const double *btab, const double *x, double *y, double *wk) {
  const size_t n = (size_t)9 * op->nn;
  double *w1 = wk, *w2 = wk + n, *w3 = wk + 2 * n, *w4 = wk + 3 * n, *w5 = wk + 4 * n, *w6 = wk + 5 * n, *w7 = wk + 6 * n,
         *w8 = wk + 7 * n;
  for (int i = 0; i < op->nn; ++i)
    for (int l = 0; l < 9; ++l) w7[9 * i + l] = x[9 * i + l] / op->sq[i];
  orc_apply_h(op, m, btab, w7, w1, w4, w5, w6);
  orc_apply_rand(op, m, m->o, w1, w2);
  orc_apply_h(op, m, btab, w2, w3, w4, w5, w6);
  orc_apply_rand(op, m, m->e, w7, w8);
  for (int i = 0; i < op->nn; ++i)
    for (int l = 0; l < 9; ++l) y[9 * i + l] = op->sq[i] * (w8[9 * i + l] + w1[9 * i + l] - w3[9 * i + l]);
}

/* Public wrapper: one application of H with its own work space. */
int orc_apply(const orc_op *op, const orc_member *m, const double *x, double *y) {
  double *btab = make_btab(op);
  double *wk = (double *)malloc(sizeof(double) * 8 * 9 * (size_t)op->nn);
  orc_augcls_apply(op, m, btab, x, y, wk);
  free(wk);
  free(btab);
  return 0;
}

/* w = X(par) x for complex parameters par (nsp x 9 complex) in normalised coordinates (augcls_rand_c) */
void orc_rand_c(const orc_op *op, const orc_member *m, const cplx *par, const cplx *x, cplx *w) {
  const int ns = op->nsmax;
  const size_t nw = (size_t)op->ncls * ns * ns * 9;
  double *pr = (double *)malloc(sizeof(double) * op->nsp * 9), *pim = (double *)malloc(sizeof(double) * op->nsp * 9);
  for (int q = 0; q < op->nsp * 9; ++q) { pr[q] = creal(par[q]); pim[q] = cimag(par[q]); }
  double *wr = (double *)malloc(sizeof(double) * nw), *wi = (double *)malloc(sizeof(double) * nw);
  orc_local_weights(op, mt_of(op, m), pr, wr);
  orc_local_weights(op, mt_of(op, m), pim, wi);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < op->nn; ++i) {
    cplx xi[9], acc[9];
    for (int l = 0; l < 9; ++l) xi[l] = x[9 * i + l] / op->sq[i];
    const int fsp = frozen_at(op, m, i);
    if (fsp >= 0) {
      for (int l = 0; l < 9; ++l) w[9 * i + l] = par[fsp * 9 + l] * x[9 * i + l];
      continue;
    }
    int kap = op->ncl[i], l0 = op->lst[i];
    for (int l = 0; l < 9; ++l) {
      size_t q = ((size_t)(kap * ns + l0) * ns + l0) * 9 + l;
      acc[l] = CMPLX(wr[q], wi[q]) * xi[l];
    }
    for (int k = 0; k < op->nsc[kap]; ++k) {
      int r = op->flipr[i * ns + k];
      if (k == l0 || r < 0) continue;
      int g = op->flipg[i * ns + k];
      for (int l = 0; l < 9; ++l) {
        cplx tf = m->chi[g] * op->sg[g * 9 + l] * x[9 * r + op->pi[g * 9 + l]];
        tf = tf / op->sq[r];
        size_t q = ((size_t)(kap * ns + l0) * ns + k) * 9 + l;
        acc[l] = acc[l] + CMPLX(wr[q], wi[q]) * tf;
      }
    }
    for (int l = 0; l < 9; ++l) w[9 * i + l] = op->sq[i] * acc[l];
  }
  free(pr); free(pim); free(wr); free(wi);
}

/* projection onto the symmetric subspace, in place (augcls_project_c) */
void orc_project_c(const orc_op *op, const orc_member *m, cplx *x) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < op->nn; ++i) {
    int ns = op->stab_ptr[i + 1] - op->stab_ptr[i];
    if (ns <= 1) continue;
    cplx acc[9] = {0};
    for (int j = op->stab_ptr[i]; j < op->stab_ptr[i + 1]; ++j) {
      int g = op->stab[j];
      for (int l = 0; l < 9; ++l) acc[l] = acc[l] + m->chi[g] * op->sg[g * 9 + l] * x[9 * i + op->pi[g * 9 + l]];
    }
    for (int l = 0; l < 9; ++l) x[9 * i + l] = acc[l] / (double)ns;
  }
}

static double dot(const double *a, const double *b, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* Real Lanczos from the unit vector at component iv (lanczos, src/reference/asr_recursion.f90:17-58); reorth != 0: full
 * reorthogonalisation, w = w - Q (Q^T w) twice against every stored vector q_1..q_i (classical Gram-Schmidt, as the
 * reference's matmul form).  a(0..n-1), b(0..n-1): b(n-1) is the norm of the last residual when n == nmax.  Returns n. */
int orc_lanczos_ro(const orc_op *op, const orc_member *m, int iv, int nmax, double tiny, int reorth, double *a, double *b) {
  const size_t dim = (size_t)9 * op->nn;
  double *btab = make_btab(op);
  double *u = (double *)calloc(dim, sizeof(double)), *up = (double *)calloc(dim, sizeof(double));
  double *w = (double *)malloc(sizeof(double) * dim), *wk = (double *)malloc(sizeof(double) * 8 * dim);
  double *qs = reorth ? (double *)malloc(sizeof(double) * dim * nmax) : NULL;
  double *cq = reorth ? (double *)malloc(sizeof(double) * nmax) : NULL;
  for (int i = 0; i < nmax; ++i) { a[i] = 0.0; b[i] = 0.0; }
  u[iv] = 1.0;
  double bn = 0.0;
  int n = 0;
  for (int i = 0; i < nmax; ++i) {
    if (reorth) memcpy(qs + (size_t)i * dim, u, sizeof(double) * dim);
    orc_augcls_apply(op, m, btab, u, w, wk);
    a[i] = dot(u, w, dim);
    for (size_t q = 0; q < dim; ++q) w[q] = w[q] - a[i] * u[q] - bn * up[q];
    if (reorth)
      for (int pass = 0; pass < 2; ++pass) {
        for (int j = 0; j <= i; ++j) cq[j] = dot(qs + (size_t)j * dim, w, dim);
        for (size_t q = 0; q < dim; ++q) {
          double t = 0.0;
          for (int j = 0; j <= i; ++j) t += qs[(size_t)j * dim + q] * cq[j];
          w[q] -= t;
        }
      }
    n = i + 1;
    bn = sqrt(dot(w, w, dim));
    if (i == nmax - 1) break;
    b[i] = bn;
    if (bn < tiny) break;
    for (size_t q = 0; q < dim; ++q) { up[q] = u[q]; u[q] = w[q] / bn; }
  }
  if (n == nmax) b[nmax - 1] = bn;
  free(u); free(up); free(w); free(wk); free(btab);
  free(qs); free(cq);
  return n;
}

int orc_lanczos(const orc_op *op, const orc_member *m, int iv, int nmax, double tiny, double *a, double *b) {
  return orc_lanczos_ro(op, m, iv, nmax, tiny, 0, a, b);
}

/* Complex-symmetric chain of route edc at one energy (edc_green, src/asr_edc.f90:75-164) with depth_tol = 0:
 * ep (nsp x 9 complex) = E - D_c (E - E_nu,s) / D_s.  Returns m; a, b (nlev complex) filled up to m. */
int orc_edc_chain(const orc_op *op, const orc_member *m, const cplx *ep, int iv, int nlev, int project, cplx *a, cplx *b) {
  const size_t dim = (size_t)9 * op->nn;
  double *btab = make_btab(op);
  cplx *u = (cplx *)calloc(dim, sizeof(cplx)), *up = (cplx *)calloc(dim, sizeof(cplx)), *w = (cplx *)malloc(sizeof(cplx) * dim);
  cplx *dg = (cplx *)malloc(sizeof(cplx) * dim);
  double *re = (double *)malloc(sizeof(double) * dim), *im = (double *)malloc(sizeof(double) * dim);
  double *sre = (double *)malloc(sizeof(double) * dim), *sim = (double *)malloc(sizeof(double) * dim);
  double *wk = (double *)malloc(sizeof(double) * 8 * dim);
  for (int k = 0; k < nlev; ++k) { a[k] = 0; b[k] = 0; }
  u[iv] = 1.0;
  int mm = 0;
  for (int k = 0; k < nlev; ++k) {
    for (size_t q = 0; q < dim; ++q) { re[q] = creal(u[q]); im[q] = cimag(u[q]); }
    orc_augcls_apply(op, m, btab, re, sre, wk);
    orc_augcls_apply(op, m, btab, im, sim, wk);
    orc_rand_c(op, m, ep, u, dg);
    for (size_t q = 0; q < dim; ++q) w[q] = CMPLX(sre[q], sim[q]) + dg[q];
    cplx s = 0;
    for (size_t q = 0; q < dim; ++q) s += u[q] * w[q];
    a[k] = s;
    mm = k + 1;
    if (k == nlev - 1) break;
    cplx ak = a[k], bk = (k > 0) ? b[k - 1] : 0;
    for (size_t q = 0; q < dim; ++q) {
      w[q] = w[q] - ak * u[q];
      if (k > 0) w[q] = w[q] - bk * up[q];
    }
    if (project) orc_project_c(op, m, w);
    s = 0;
    for (size_t q = 0; q < dim; ++q) s += w[q] * w[q];
    b[k] = csqrt(s);
    if (cabs(b[k]) < 1.0e-12) break;
    bk = b[k];
    for (size_t q = 0; q < dim; ++q) { up[q] = u[q]; u[q] = w[q] / bk; }
  }
  free(u); free(up); free(w); free(dg); free(re); free(im); free(sre); free(sim); free(wk); free(btab);
  return mm;
}

/* y = h x = sq (c x~ + D S D x~) for complex x in normalised coordinates: apply_h on the real and the imaginary part
 * (augcls_hp_c).  wk: 8 work vectors of 9 nn. */
void orc_hp_c(const orc_op *op, const orc_member *m, const double *btab, const cplx *x, cplx *y, double *wk) {
  const size_t n = (size_t)9 * op->nn;
  double *w1 = wk, *w2 = wk + n, *w4 = wk + 3 * n, *w5 = wk + 4 * n, *w6 = wk + 5 * n, *w7 = wk + 6 * n, *w8 = wk + 7 * n;
  for (int i = 0; i < op->nn; ++i)
    for (int l = 0; l < 9; ++l) {
      w7[9 * i + l] = creal(x[9 * i + l]) / op->sq[i];
      w8[9 * i + l] = cimag(x[9 * i + l]) / op->sq[i];
    }
  orc_apply_h(op, m, btab, w7, w1, w4, w5, w6);
  orc_apply_h(op, m, btab, w8, w2, w4, w5, w6);
  for (int i = 0; i < op->nn; ++i)
    for (int l = 0; l < 9; ++l) y[9 * i + l] = op->sq[i] * CMPLX(w1[9 * i + l], w2[9 * i + l]);
}

/* y = K' x = et x - h'x + u h'x + h'(u x - Qt h'x) (ed3_kmul); et, uu, qt: nsp x 9 complex species coefficients at the
 * chain's energy; the member carries h' (c = g, d = D_c^1/2, o = e = 0).  ct: 4 complex work vectors, wk: 8 real. */
void orc_ed3_kmul(const orc_op *op, const orc_member *m, const double *btab, const cplx *et, const cplx *uu, const cplx *qt,
                  const cplx *x, cplx *y, cplx *ct, double *wk) {
  const size_t n = (size_t)9 * op->nn;
  cplx *t1 = ct, *t2 = ct + n, *t3 = ct + 2 * n, *t4 = ct + 3 * n;
  orc_hp_c(op, m, btab, x, t1, wk);
  orc_rand_c(op, m, qt, t1, t2);
  orc_rand_c(op, m, uu, x, t3);
  for (size_t q = 0; q < n; ++q) t3[q] = t3[q] - t2[q];
  orc_hp_c(op, m, btab, t3, t4, wk);
  orc_rand_c(op, m, et, x, y);
  orc_rand_c(op, m, uu, t1, t2);
  for (size_t q = 0; q < n; ++q) y[q] = y[q] - t1[q] + t2[q] + t4[q];
}

/* Public wrapper of orc_ed3_kmul with its own work space. */
int orc_ed3_apply(const orc_op *op, const orc_member *m, const cplx *et, const cplx *uu, const cplx *qt, const cplx *x,
                  cplx *y) {
  const size_t n = (size_t)9 * op->nn;
  double *btab = make_btab(op);
  double *wk = (double *)malloc(sizeof(double) * 8 * n);
  cplx *ct = (cplx *)malloc(sizeof(cplx) * 4 * n);
  orc_ed3_kmul(op, m, btab, et, uu, qt, x, y, ct, wk);
  free(ct); free(wk); free(btab);
  return 0;
}

/* Start vectors of orbital iv (ed3_start): |v> = h'|u> (projected if project), |a> = |u> + oc |v>; norms nv, na
 * (sqrt of the real part of the bilinear square).  vv, av: 9 nn complex.  Returns 1 for a zero start vector. */
int orc_ed3_start(const orc_op *op, const orc_member *m, int iv, double oc, int project, cplx *vv, cplx *av, double *nv,
                  double *na) {
  const size_t n = (size_t)9 * op->nn;
  double *btab = make_btab(op);
  double *wk = (double *)malloc(sizeof(double) * 8 * n);
  cplx *s0 = (cplx *)calloc(n, sizeof(cplx));
  s0[iv] = 1.0;
  orc_hp_c(op, m, btab, s0, vv, wk);
  if (project) orc_project_c(op, m, vv);
  for (size_t q = 0; q < n; ++q) av[q] = s0[q] + oc * vv[q];
  cplx sv = 0, sa = 0;
  for (size_t q = 0; q < n; ++q) sv += vv[q] * vv[q];
  for (size_t q = 0; q < n; ++q) sa += av[q] * av[q];
  *nv = sqrt(creal(sv));
  *na = sqrt(creal(sa));
  free(s0); free(wk); free(btab);
  return (*nv < 1.0e-14 || *na < 1.0e-14) ? 1 : 0;
}

/* One chain of route ed3 (chain, src/asr_ed3.f90:148-211) on H_eff = z - K' from u0 (u0^T u0 = 1) without the adaptive
 * depth: a, b (nlev complex), c(k) = tv^T u_k (nlev complex); *closed = 1 when b_k < 1e-12 ended it.  Returns m. */
int orc_ed3_chain(const orc_op *op, const orc_member *m, const cplx *et, const cplx *uu, const cplx *qt, cplx z,
                  const cplx *u0, const cplx *tv, int nlev, int project, cplx *a, cplx *b, cplx *c, int *closed) {
  const size_t n = (size_t)9 * op->nn;
  double *btab = make_btab(op);
  double *wk = (double *)malloc(sizeof(double) * 8 * n);
  cplx *ct = (cplx *)malloc(sizeof(cplx) * 4 * n);
  cplx *u = (cplx *)malloc(sizeof(cplx) * n), *up = (cplx *)calloc(n, sizeof(cplx)), *w = (cplx *)malloc(sizeof(cplx) * n);
  for (int k = 0; k < nlev; ++k) { a[k] = 0; b[k] = 0; c[k] = 0; }
  memcpy(u, u0, sizeof(cplx) * n);
  int mm = 0;
  *closed = 0;
  for (int k = 0; k < nlev; ++k) {
    cplx s = 0;
    for (size_t q = 0; q < n; ++q) s += tv[q] * u[q];
    c[k] = s;
    orc_ed3_kmul(op, m, btab, et, uu, qt, u, w, ct, wk);
    for (size_t q = 0; q < n; ++q) w[q] = z * u[q] - w[q];
    s = 0;
    for (size_t q = 0; q < n; ++q) s += u[q] * w[q];
    a[k] = s;
    mm = k + 1;
    if (k == nlev - 1) break;
    cplx ak = a[k], bk = (k > 0) ? b[k - 1] : 0;
    for (size_t q = 0; q < n; ++q) {
      w[q] = w[q] - ak * u[q];
      if (k > 0) w[q] = w[q] - bk * up[q];
    }
    if (project) orc_project_c(op, m, w);
    s = 0;
    for (size_t q = 0; q < n; ++q) s += w[q] * w[q];
    b[k] = csqrt(s);
    if (cabs(b[k]) < 1.0e-12) { *closed = 1; break; }
    bk = b[k];
    for (size_t q = 0; q < n; ++q) { up[q] = u[q]; u[q] = w[q] / bk; }
  }
  free(u); free(up); free(w); free(ct); free(wk); free(btab);
  return mm;
}

/* Start vector of an EPI block (epi_start, src/asr_epi.f90:578-587): s0 = v9 on the start node, |v> = h'|s0> (projected
 * if project), nv = sqrt(Re v.v).  s0, vv: 9 nn complex.  Returns 1 for a zero |v>. */
int orc_epi_start(const orc_op *op, const orc_member *m, const double *v9, int project, cplx *s0, cplx *vv, double *nv) {
  const size_t n = (size_t)9 * op->nn;
  double *btab = make_btab(op);
  double *wk = (double *)malloc(sizeof(double) * 8 * n);
  for (size_t q = 0; q < n; ++q) s0[q] = 0;
  for (int l = 0; l < 9; ++l) s0[l] = v9[l];
  orc_hp_c(op, m, btab, s0, vv, wk);
  if (project) orc_project_c(op, m, vv);
  cplx sv = 0;
  for (size_t q = 0; q < n; ++q) sv += vv[q] * vv[q];
  *nv = sqrt(creal(sv));
  free(wk); free(btab);
  return *nv < 1.0e-14 ? 1 : 0;
}

/* One chain of an EPI block (chain, src/asr_epi.f90:590-626): complex-symmetric Lanczos on H(z) of route edc (augcls_apply
 * on the real and imaginary parts plus augcls_rand_c with ep = E - D_c (E - E_nu,s) / D_s) from u0 (u0^T u0 = 1), full depth
 * nlev: a, b, c(k) = tv^T u_k; *closed = 1 when b_k < 1e-12 ended it.  Returns m. */
int orc_epi_chain(const orc_op *op, const orc_member *m, const cplx *ep, const cplx *u0, const cplx *tv, int nlev, int project,
                  cplx *a, cplx *b, cplx *c, int *closed) {
  const size_t dim = (size_t)9 * op->nn;
  double *btab = make_btab(op);
  cplx *u = (cplx *)malloc(sizeof(cplx) * dim), *up = (cplx *)calloc(dim, sizeof(cplx)), *w = (cplx *)malloc(sizeof(cplx) * dim);
  cplx *dg = (cplx *)malloc(sizeof(cplx) * dim);
  double *re = (double *)malloc(sizeof(double) * dim), *im = (double *)malloc(sizeof(double) * dim);
  double *sre = (double *)malloc(sizeof(double) * dim), *sim = (double *)malloc(sizeof(double) * dim);
  double *wk = (double *)malloc(sizeof(double) * 8 * dim);
  for (int k = 0; k < nlev; ++k) { a[k] = 0; b[k] = 0; c[k] = 0; }
  memcpy(u, u0, sizeof(cplx) * dim);
  int mm = 0;
  *closed = 0;
  for (int k = 0; k < nlev; ++k) {
    cplx s = 0;
    for (size_t q = 0; q < dim; ++q) s += tv[q] * u[q];
    c[k] = s;
    for (size_t q = 0; q < dim; ++q) { re[q] = creal(u[q]); im[q] = cimag(u[q]); }
    orc_augcls_apply(op, m, btab, re, sre, wk);
    orc_augcls_apply(op, m, btab, im, sim, wk);
    orc_rand_c(op, m, ep, u, dg);
    for (size_t q = 0; q < dim; ++q) w[q] = CMPLX(sre[q], sim[q]) + dg[q];
    s = 0;
    for (size_t q = 0; q < dim; ++q) s += u[q] * w[q];
    a[k] = s;
    mm = k + 1;
    if (k == nlev - 1) break;
    for (size_t q = 0; q < dim; ++q) {
      w[q] = w[q] - a[k] * u[q];
      if (k > 0) w[q] = w[q] - b[k - 1] * up[q];
    }
    if (project) orc_project_c(op, m, w);
    s = 0;
    for (size_t q = 0; q < dim; ++q) s += w[q] * w[q];
    b[k] = csqrt(s);
    if (cabs(b[k]) < 1.0e-12) { *closed = 1; break; }
    for (size_t q = 0; q < dim; ++q) { up[q] = u[q]; u[q] = w[q] / b[k]; }
  }
  free(u); free(up); free(w); free(dg); free(re); free(im); free(sre); free(sim); free(wk); free(btab);
  return mm;
}

/* Projected chain of route edc (edc_green_proj, src/asr_edc.f90:174-232) without the fraction: the chain of edc_green run
 * all nlev steps (no depth test) and, at every level k before the operator, the projections of u_k on the target states
 *   cpr[k][t][l] = sum_ks tw[t][ks][l] amp_l(u_k, tnode[t][ks], tgop[t][ks]),  amp_l = chi(g) sg(l, g) u(9 r + pi(l, g)) / sq(r)
 * (augcls_amp_c; tnode < 0: no such node).  cpr: nlev x nt x 9 complex.  Returns m. */
int orc_edc_chain_proj(const orc_op *op, const orc_member *m, const cplx *ep, int iv, int nlev, int project, int nt,
                       const int *tnode, const int *tgop, const double *tw, cplx *a, cplx *b, cplx *cpr) {
  const size_t dim = (size_t)9 * op->nn;
  const int nsm = op->nsmax;
  double *btab = make_btab(op);
  cplx *u = (cplx *)calloc(dim, sizeof(cplx)), *up = (cplx *)calloc(dim, sizeof(cplx)), *w = (cplx *)malloc(sizeof(cplx) * dim);
  cplx *dg = (cplx *)malloc(sizeof(cplx) * dim);
  double *re = (double *)malloc(sizeof(double) * dim), *im = (double *)malloc(sizeof(double) * dim);
  double *sre = (double *)malloc(sizeof(double) * dim), *sim = (double *)malloc(sizeof(double) * dim);
  double *wk = (double *)malloc(sizeof(double) * 8 * dim);
  for (int k = 0; k < nlev; ++k) { a[k] = 0; b[k] = 0; }
  for (size_t q = 0; q < (size_t)nlev * nt * 9; ++q) cpr[q] = 0;
  u[iv] = 1.0;
  int mm = 0;
  for (int k = 0; k < nlev; ++k) {
    for (int t = 0; t < nt; ++t)
      for (int ks = 0; ks < nsm; ++ks) {
        const int r = tnode[t * nsm + ks], g = tgop[t * nsm + ks];
        if (r < 0) continue;
        for (int l = 0; l < 9; ++l) {
          const cplx amp = m->chi[g] * op->sg[g * 9 + l] * u[9 * (size_t)r + op->pi[g * 9 + l]] / op->sq[r];
          cpr[((size_t)k * nt + t) * 9 + l] += tw[((size_t)t * nsm + ks) * 9 + l] * amp;
        }
      }
    for (size_t q = 0; q < dim; ++q) { re[q] = creal(u[q]); im[q] = cimag(u[q]); }
    orc_augcls_apply(op, m, btab, re, sre, wk);
    orc_augcls_apply(op, m, btab, im, sim, wk);
    orc_rand_c(op, m, ep, u, dg);
    for (size_t q = 0; q < dim; ++q) w[q] = CMPLX(sre[q], sim[q]) + dg[q];
    cplx s = 0;
    for (size_t q = 0; q < dim; ++q) s += u[q] * w[q];
    a[k] = s;
    mm = k + 1;
    if (k == nlev - 1) break;
    for (size_t q = 0; q < dim; ++q) {
      w[q] = w[q] - a[k] * u[q];
      if (k > 0) w[q] = w[q] - b[k - 1] * up[q];
    }
    if (project) orc_project_c(op, m, w);
    s = 0;
    for (size_t q = 0; q < dim; ++q) s += w[q] * w[q];
    b[k] = csqrt(s);
    if (cabs(b[k]) < 1.0e-12) break;
    for (size_t q = 0; q < dim; ++q) { up[q] = u[q]; u[q] = w[q] / b[k]; }
  }
  free(u); free(up); free(w); free(dg); free(re); free(im); free(sre); free(sim); free(wk); free(btab);
  return mm;
}

/* A batch of real chains, one chain per thread, dynamic schedule (the CPU baseline of the chain engine). */
int orc_lanczos_batch(const orc_op *op, const orc_member *ms, const int *iv, const int *nlev, int nchain, int ldab, double *a,
                      double *b, int *mused) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < nchain; ++c) mused[c] = orc_lanczos(op, &ms[c], iv[c], nlev[c], 1.0e-12, a + (size_t)c * ldab,
                                                          b + (size_t)c * ldab);
  return 0;
}

int orc_edc_batch(const orc_op *op, const orc_member *ms, const cplx *ep, const int *iv, const int *nlev, int nchain, int ldab,
                  int project, cplx *a, cplx *b, int *mused) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < nchain; ++c)
    mused[c] = orc_edc_chain(op, &ms[c], ep + (size_t)c * op->nsp * 9, iv[c], nlev[c], project, a + (size_t)c * ldab,
                             b + (size_t)c * ldab);
  return 0;
}

/* ------------------------------------------------------------------ supercell realisation and KPM */
typedef struct {
  int nsite, nslot, nsp;
  const int *nbr;               /* nsite x nslot */
  const unsigned char *species; /* nsite */
  const double *c, *d, *o, *e;  /* nsp x 9 */
  const double *block, *block0; /* nslot x 81, 81: [l][l2] */
} orc_real;

/* hx = c x + d (S (d x)) with the on-site block (h_apply) */
static void real_h_apply(const orc_real *R, const double *x, double *hx, double *t) {
  const int nl = 9;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < R->nsite; ++i) {
    const double *dp = R->d + R->species[i] * 9;
    for (int l = 0; l < nl; ++l) t[nl * i + l] = dp[l] * x[nl * i + l];
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < R->nsite; ++i) {
    const double *dp = R->d + R->species[i] * 9, *cp = R->c + R->species[i] * 9;
    for (int l = 0; l < nl; ++l) {
      double acc = 0.0;
      for (int l2 = 0; l2 < nl; ++l2) acc = acc + R->block0[l * 9 + l2] * t[nl * i + l2];
      for (int k = 0; k < R->nslot; ++k) {
        int j = R->nbr[(size_t)i * R->nslot + k];
        const double *bk = R->block + (size_t)k * 81;
        for (int l2 = 0; l2 < nl; ++l2) acc = acc + bk[l * 9 + l2] * t[nl * j + l2];
      }
      hx[nl * i + l] = cp[l] * x[nl * i + l] + dp[l] * acc;
    }
  }
}

/* y = e x + h x - h o h x (real_apply, order 2); wk: 3 work vectors.  With o = 0 for every species the second h apply
 * acts on a zero vector and gives exactly zero, so it is skipped (y = e x + h x, the same rounding). */
void orc_real_apply(const orc_real *R, const double *x, double *y, double *wk) {
  const size_t n = (size_t)9 * R->nsite;
  double *hx = wk, *w = wk + n, *hw = wk + 2 * n, *t = y;
  real_h_apply(R, x, hx, t);
  int zero_o = 1;
  for (int q = 0; q < R->nsp * 9; ++q) zero_o = zero_o && R->o[q] == 0.0;
  if (zero_o) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < R->nsite; ++i)
      for (int l = 0; l < 9; ++l) y[9 * i + l] = R->e[R->species[i] * 9 + l] * x[9 * i + l] + hx[9 * i + l];
    return;
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < R->nsite; ++i)
    for (int l = 0; l < 9; ++l) w[9 * i + l] = R->o[R->species[i] * 9 + l] * hx[9 * i + l];
  double *t2 = (double *)malloc(sizeof(double) * n);
  real_h_apply(R, w, hw, t2);
  free(t2);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < R->nsite; ++i)
    for (int l = 0; l < 9; ++l)
      y[9 * i + l] = R->e[R->species[i] * 9 + l] * x[9 * i + l] + hx[9 * i + l] - hw[9 * i + l];
}

/* The number of OpenMP threads the parallel loops above use (the timed CPU baseline reports it). */
int orc_threads(void) {
  int n = 1;
#pragma omp parallel
  {
#pragma omp single
    n = omp_get_num_threads();
  }
  return n;
}

/* Counter-based Rademacher entry r(seed, vector, row) = +-1 (splitmix64 of a packed counter; the GPU uses the same). */
static inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double orc_rademacher(uint64_t seed, uint64_t v, uint64_t row) {
  uint64_t z = splitmix64((seed << 40) ^ (v << 28) ^ row);
  return (z >> 63) ? -1.0 : 1.0;
}

/* mu_n = (1/R) sum_r <r|T_n(Hs)|r> / dim, Hs = (H - c)/s, n = 0..nmom-1 (kpm_moments); mu_per (optional, nvec x nmom)
 * receives the per-vector terms <r|t_n>/dim. */
int orc_kpm_moments(const orc_real *R, double c, double s, int nmom, int nvec, uint64_t seed, double *mu, double *mu_per) {
  const size_t dim = (size_t)9 * R->nsite;
  double *r = (double *)malloc(sizeof(double) * dim), *t0 = (double *)malloc(sizeof(double) * dim);
  double *t1 = (double *)malloc(sizeof(double) * dim), *t2 = (double *)malloc(sizeof(double) * dim);
  double *w = (double *)malloc(sizeof(double) * dim), *wk = (double *)malloc(sizeof(double) * 3 * dim);
  for (int n = 0; n < nmom; ++n) mu[n] = 0.0;
  for (int iv = 0; iv < nvec; ++iv) {
    double *mv = mu_per ? mu_per + (size_t)iv * nmom : NULL;
    for (size_t q = 0; q < dim; ++q) r[q] = orc_rademacher(seed, (uint64_t)iv, q);
    memcpy(t0, r, sizeof(double) * dim);
    orc_real_apply(R, r, w, wk);
    for (size_t q = 0; q < dim; ++q) t1[q] = (w[q] - c * r[q]) / s;
    double m0 = dot(r, t0, dim) / (double)dim;
    mu[0] += m0;
    if (mv) mv[0] = m0;
    if (nmom > 1) {
      double m1 = dot(r, t1, dim) / (double)dim;
      mu[1] += m1;
      if (mv) mv[1] = m1;
    }
    for (int n = 2; n < nmom; ++n) {
      orc_real_apply(R, t1, w, wk);
#pragma omp parallel for schedule(static)
      for (size_t q = 0; q < dim; ++q) t2[q] = 2.0 * (w[q] - c * t1[q]) / s - t0[q];
      double mn = dot(r, t2, dim) / (double)dim;
      mu[n] += mn;
      if (mv) mv[n] = mn;
      double *tmp = t0; t0 = t1; t1 = t2; t2 = tmp;
    }
  }
  for (int n = 0; n < nmom; ++n) mu[n] /= (double)nvec;
  free(r); free(t0); free(t1); free(t2); free(w); free(wk);
  return 0;
}

/* Lanczos coefficients of nstep steps from a seeded random start (the recursion part of spectral_bounds). */
int orc_real_lanczos(const orc_real *R, int nstep, uint64_t seed, double *a, double *b) {
  const size_t dim = (size_t)9 * R->nsite;
  double *u = (double *)malloc(sizeof(double) * dim), *up = (double *)calloc(dim, sizeof(double));
  double *w = (double *)malloc(sizeof(double) * dim), *wk = (double *)malloc(sizeof(double) * 3 * dim);
  for (size_t q = 0; q < dim; ++q) u[q] = (double)(splitmix64((seed << 40) ^ q) >> 11) * 0x1.0p-53 - 0.5;
  double nrm = sqrt(dot(u, u, dim));
  for (size_t q = 0; q < dim; ++q) u[q] /= nrm;
  double bn = 0.0;
  for (int i = 0; i < nstep; ++i) {
    orc_real_apply(R, u, w, wk);
    a[i] = dot(u, w, dim);
    for (size_t q = 0; q < dim; ++q) w[q] = w[q] - a[i] * u[q] - bn * up[q];
    bn = sqrt(dot(w, w, dim));
    b[i] = bn;
    double inv = 1.0 / (bn > 1.0e-300 ? bn : 1.0e-300);
    for (size_t q = 0; q < dim; ++q) { up[q] = u[q]; u[q] = w[q] * inv; }
  }
  free(u); free(up); free(w); free(wk);
  return 0;
}
