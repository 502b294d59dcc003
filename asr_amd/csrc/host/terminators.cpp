// Host side of the chain boundary: terminated continued fractions, square-root tails, Beer-Pettifor asymptote, the Jackson
// KPM density of states and the occupation contours.  These stay on the host: the GPU returns the chain coefficients
// (a_n, b_n) or the Chebyshev moments, and the routines below turn them into G(z), DOS and contour sums.
//
//   cf_species      src/asr_chain.f90:67-167          (hull / bp / avg2 / none tails of a complex-symmetric chain)
//   bp_asymptote    src/reference/asr_terminator.f90:39-79
//   sqrt_tail       src/reference/asr_greens.f90:139-149
//   cf_green        src/reference/asr_greens.f90:207-217
//   kpm_dos         src/reference/asr_kpm.f90:68-87
//   contour_nodes_closed, gauss_legendre   src/asr_contour.f90:39-58, 132-151
//
// LAPACK is not used: the reference calls zgeev (complex general eigenvalues) on the m x m chain matrix and dsterf on a real
// symmetric tridiagonal.  Both are replaced by textbook routines for these tridiagonal matrices (m <= a few hundred): a
// shifted QR iteration on the complex upper-Hessenberg (here tridiagonal) matrix and the implicit QL iteration.
#include <cmath>
#include <complex>
#include <vector>
#include <algorithm>
#include <cstring>

#include "asr_host.h"

using cplx = std::complex<double>;

namespace {

// Eigenvalues of a real symmetric tridiagonal matrix (diagonal d[0..n-1], off-diagonal e[0..n-2]); ascending in d.
// Implicit QL with Wilkinson shifts (the method of EISPACK tql1 / LAPACK dsterf).  Returns 0 on success.
int tridiag_eigvals(int n, double* d, const double* e_in) {
  if (n <= 0) return 0;
  std::vector<double> e(n, 0.0);
  for (int i = 0; i < n - 1; ++i) e[i] = e_in[i];
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; ++m) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-300 || std::fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
      }
      if (m != l) {
        if (iter++ == 200) return 1;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  std::sort(d, d + n);
  return 0;
}

// Eigenvalues of a complex upper-Hessenberg matrix H (n x n, row-major, overwritten): single-shift QR with Givens
// rotations and Wilkinson shifts, deflating from the bottom (the complex QR algorithm of EISPACK comqr).
int hessenberg_eigvals(int n, std::vector<cplx>& H, cplx* ev) {
  auto A = [&](int i, int j) -> cplx& { return H[(size_t)i * n + j]; };
  int hi = n - 1;
  int iter = 0;
  const double eps = 2.220446049250313e-16;
  std::vector<double> cs(n);
  std::vector<cplx> sn(n);
  while (hi >= 0) {
    if (hi == 0) { ev[0] = A(0, 0); break; }
    int l = hi;
    while (l > 0) {
      double s = std::abs(A(l - 1, l - 1)) + std::abs(A(l, l));
      if (s == 0.0) s = 1.0;
      if (std::abs(A(l, l - 1)) <= eps * s) { A(l, l - 1) = 0.0; break; }
      --l;
    }
    if (l == hi) { ev[hi] = A(hi, hi); --hi; iter = 0; continue; }
    if (++iter > 60 * n + 300) return 1;
    // Wilkinson shift from the trailing 2x2 block; exceptional shifts every 10 iterations
    cplx a = A(hi - 1, hi - 1), b = A(hi - 1, hi), c = A(hi, hi - 1), d = A(hi, hi);
    cplx tr = a + d, det = a * d - b * c;
    cplx disc = std::sqrt(tr * tr - 4.0 * det);
    cplx l1 = 0.5 * (tr + disc), l2 = 0.5 * (tr - disc);
    cplx mu = (std::abs(l1 - d) < std::abs(l2 - d)) ? l1 : l2;
    if (iter % 11 == 10) mu = d + cplx(std::abs(A(hi, hi - 1).real()) + std::abs(A(hi - 1, hi - 2 >= 0 ? hi - 2 : 0).real()), 0.0);
    for (int i = l; i <= hi; ++i) A(i, i) -= mu;
    // QR by Givens on rows l..hi
    for (int k = l; k < hi; ++k) {
      cplx x = A(k, k), y = A(k + 1, k);
      double nrm = std::sqrt(std::norm(x) + std::norm(y));
      double c0;
      cplx s0;
      if (nrm == 0.0) { c0 = 1.0; s0 = 0.0; }
      else {
        double ax = std::abs(x);
        if (ax == 0.0) { c0 = 0.0; s0 = std::conj(y) / nrm; s0 = cplx(1.0, 0.0) * (std::conj(y) / std::abs(y)); }
        else { c0 = ax / nrm; cplx ph = x / ax; s0 = ph * std::conj(y) / nrm; }
      }
      cs[k] = c0; sn[k] = s0;
      // apply G = [c, s; -conj(s), c] to rows k, k+1 (columns k..n-1 of the active part)
      for (int j = k; j <= hi; ++j) {
        cplx t1 = A(k, j), t2 = A(k + 1, j);
        A(k, j) = c0 * t1 + s0 * t2;
        A(k + 1, j) = -std::conj(s0) * t1 + c0 * t2;
      }
    }
    // RQ: apply G^H on the right to columns k, k+1
    for (int k = l; k < hi; ++k) {
      double c0 = cs[k];
      cplx s0 = sn[k];
      int rmax = std::min(hi, k + 2);
      for (int i = l; i <= rmax; ++i) {
        cplx t1 = A(i, k), t2 = A(i, k + 1);
        A(i, k) = t1 * c0 + t2 * std::conj(s0);
        A(i, k + 1) = -t1 * s0 + t2 * c0;
      }
    }
    for (int i = l; i <= hi; ++i) A(i, i) += mu;
  }
  return 0;
}

inline void chain_eigs(int m, const cplx* a, const cplx* b, cplx am, std::vector<cplx>& ev, int* info) {
  std::vector<cplx> H((size_t)m * m, 0.0);
  for (int j = 0; j < m; ++j) {
    H[(size_t)j * m + j] = a[j];
    if (j < m - 1) { H[(size_t)j * m + j + 1] = b[j]; H[(size_t)(j + 1) * m + j] = b[j]; }
  }
  H[(size_t)(m - 1) * m + m - 1] = am;
  ev.assign(m, 0.0);
  *info = hessenberg_eigvals(m, H, ev.data());
}

inline int argmin_re(const std::vector<cplx>& v) {
  int k = 0;
  for (int i = 1; i < (int)v.size(); ++i) if (v[i].real() < v[k].real()) k = i;
  return k;
}
inline int argmax_re(const std::vector<cplx>& v) {
  int k = 0;
  for (int i = 1; i < (int)v.size(); ++i) if (v[i].real() > v[k].real()) k = i;
  return k;
}

}  // namespace

// g11 and f = sum_k c_k [g]_k1 of a complex-symmetric chain that closed (Krylov space exhausted): the finite fraction, no
// tail (cf_closed, src/asr_ed3.f90:214-230).
extern "C" int asr_host_cf_closed(const double* a_, const double* b_, const double* c_, int m, const double* z_, double* g11_,
                                  double* f_) {
  if (m < 1) return 1;
  const cplx* a = reinterpret_cast<const cplx*>(a_);
  const cplx* b = reinterpret_cast<const cplx*>(b_);
  const cplx* c = reinterpret_cast<const cplx*>(c_);
  const cplx z(z_[0], z_[1]);
  std::vector<cplx> gt(m);
  gt[m - 1] = 1.0 / (z - a[m - 1]);
  for (int j = m - 2; j >= 0; --j) gt[j] = 1.0 / (z - a[j] - b[j] * b[j] * gt[j + 1]);
  cplx f(0.0, 0.0), prod = gt[0];
  for (int k = 0; k < m; ++k) {
    f += c[k] * prod;
    if (k < m - 1) prod = prod * b[k] * gt[k + 1];
  }
  g11_[0] = gt[0].real(); g11_[1] = gt[0].imag();
  f_[0] = f.real(); f_[1] = f.imag();
  return 0;
}

// g11 and f = sum_k c_k [g]_k1 of the complex-symmetric chain (a, b) at z with a square-root tail (src/asr_chain.f90:67-121).
// tail: 0 hull, 1 bp, 2 avg2, 3 none.  a, b, c, z, g11, f, edges, pv are interleaved (re, im) pairs; edges and pv may be NULL.
// Returns the number of fallbacks (bp fixed point not converged or eigen-solver failure), 0 normally.
extern "C" int asr_host_cf_species(const double* a_, const double* b_, const double* c_, int m, const double* z_, int tail,
                                   double* g11_, double* f_, double* edges_, double* pv_) {
  const cplx* a = reinterpret_cast<const cplx*>(a_);
  const cplx* b = reinterpret_cast<const cplx*>(b_);
  const cplx* c = reinterpret_cast<const cplx*>(c_);
  cplx z(z_[0], z_[1]);
  int nfall = 0;
  cplx aa, bb;
  std::vector<cplx> ev;
  int info = 1;
  if (m >= 2 && tail != 3) chain_eigs(m, a, b, a[m - 1], ev, &info);
  if (info == 0) {
    int ilo = argmin_re(ev), ihi = argmax_re(ev);
    aa = 0.5 * (ev[ilo] + ev[ihi]);
    bb = 0.25 * (ev[ihi] - ev[ilo]);
    if (tail == 1) {  // Beer-Pettifor fixed point (src/asr_chain.f90:126-156)
      bool done = false;
      for (int it = 0; it < 200 && !done; ++it) {
        int inf2;
        chain_eigs(m, a, b, a[m - 1] + bb, ev, &inf2);
        if (inf2 != 0) { ++nfall; done = true; break; }
        cplx ep = ev[argmax_re(ev)];
        chain_eigs(m, a, b, a[m - 1] - bb, ev, &inf2);
        if (inf2 != 0) { ++nfall; done = true; break; }
        cplx em = ev[argmin_re(ev)];
        cplx an = 0.5 * (ep + em), bn = 0.25 * (ep - em);
        if (std::abs(an - aa) + std::abs(bn - bb) < 1e-12 * (1.0 + std::abs(bn))) { aa = an; bb = bn; done = true; break; }
        aa = an; bb = bn;
        if (it == 199) {  // not converged: back to the hull
          ++nfall;
          chain_eigs(m, a, b, a[m - 1], ev, &inf2);
          int lo = argmin_re(ev), hi2 = argmax_re(ev);
          aa = 0.5 * (ev[lo] + ev[hi2]);
          bb = 0.25 * (ev[hi2] - ev[lo]);
        }
      }
    }
  } else {
    aa = a[m - 1];
    bb = (m > 1) ? b[m - 2] : cplx(0.0);
  }
  if (tail == 2 && m >= 3) {
    aa = 0.5 * (a[m - 1] + a[m - 2]);
    bb = std::sqrt(0.5 * (b[m - 2] * b[m - 2] + b[m - 3] * b[m - 3]));
  }
  if (tail == 3) { aa = 0.0; bb = 0.0; }
  if (edges_) { edges_[0] = aa.real(); edges_[1] = aa.imag(); edges_[2] = bb.real(); edges_[3] = bb.imag(); }
  cplx d = z - aa, tl;
  if (std::abs(bb) < 1e-14) {
    tl = 0.0;
  } else {
    cplx s = std::sqrt(d * d - 4.0 * bb * bb);
    cplx g1 = (d - s) / (2.0 * bb * bb), g2 = (d + s) / (2.0 * bb * bb);
    tl = (std::abs(bb * g1) <= std::abs(bb * g2)) ? bb * bb * g1 : bb * bb * g2;
  }
  std::vector<cplx> gt(m + 1, 0.0);
  gt[m - 1] = 1.0 / (z - a[m - 1] - tl);
  for (int j = m - 2; j >= 0; --j) gt[j] = 1.0 / (z - a[j] - b[j] * b[j] * gt[j + 1]);
  cplx g11 = gt[0], f = 0.0, prod = gt[0];
  for (int k = 0; k < m; ++k) {
    f += c[k] * prod;
    if (pv_) { pv_[2 * k] = prod.real(); pv_[2 * k + 1] = prod.imag(); }
    if (k < m - 1) prod = prod * b[k] * gt[k + 1];
  }
  g11_[0] = g11.real(); g11_[1] = g11.imag();
  f_[0] = f.real(); f_[1] = f.imag();
  return nfall;
}

// Eigenvalues of a real symmetric tridiagonal matrix (the role of dsterf).  d (n, in: diagonal, out: ascending eigenvalues).
extern "C" int asr_host_tridiag_eigvals(int n, double* d, const double* e) { return tridiag_eigvals(n, d, e); }

// Eigenvalues of the complex-symmetric tridiagonal chain matrix (the role of zgeev in cf_species); ev: n (re, im) pairs.
extern "C" int asr_host_chain_eigvals(int n, const double* a, const double* b, double* ev) {
  std::vector<cplx> e;
  int info;
  const cplx* ac = reinterpret_cast<const cplx*>(a);
  chain_eigs(n, ac, reinterpret_cast<const cplx*>(b), ac[n - 1], e, &info);
  for (int i = 0; i < n; ++i) { ev[2 * i] = e[i].real(); ev[2 * i + 1] = e[i].imag(); }
  return info;
}

// Beer-Pettifor asymptote (src/reference/asr_terminator.f90:39-67): fixed point from the avg2 estimate, 500 iterations at most.
extern "C" int asr_host_bp_asymptote(const double* a, const double* b, int n, double* ainf, double* binf, int* niter) {
  double A = 0.5 * (a[n - 1] + a[n - 2]);
  double B = std::sqrt(0.5 * (b[n - 1] * b[n - 1] + b[n - 2] * b[n - 2]));
  std::vector<double> d(n), e(std::max(n - 1, 1));
  int it, ok = 0;
  for (it = 1; it <= 500; ++it) {
    for (int i = 0; i < n; ++i) d[i] = a[i];
    for (int i = 0; i < n - 1; ++i) e[i] = b[i];
    d[n - 1] = a[n - 1] + B;
    if (tridiag_eigvals(n, d.data(), e.data())) return -1;
    double ep = d[n - 1];
    for (int i = 0; i < n; ++i) d[i] = a[i];
    d[n - 1] = a[n - 1] - B;
    if (tridiag_eigvals(n, d.data(), e.data())) return -1;
    double em = d[0];
    double an = 0.5 * (ep + em), bn = 0.25 * (ep - em);
    if (std::fabs(an - A) + std::fabs(bn - B) < 1e-13) { A = an; B = bn; ok = 1; break; }
    A = an; B = bn;
  }
  *ainf = A; *binf = B;
  if (niter) *niter = std::min(it, 500);
  return ok ? 0 : 1;
}

// Square-root tail t(z) = binf^2 g(z) on the decaying sheet (src/reference/asr_greens.f90:139-149).
static cplx sqrt_tail(cplx z, double ainf, double binf) {
  cplx w = z - ainf;
  cplx s = std::sqrt(w * w - 4.0 * binf * binf);
  cplx g1 = (w - s) / (2.0 * binf * binf), g2 = (w + s) / (2.0 * binf * binf);
  cplx g = (std::abs(g1) <= std::abs(g2)) ? g1 : g2;
  return binf * binf * g;
}

// G(z) of the real chain terminated by the square-root tail of (ainf, binf) (src/reference/asr_terminator.f90:117-126 with
// src/reference/asr_greens.f90:207-217), at ne energies z (re, im pairs) -> g (re, im pairs).  binf = 0: no tail.
extern "C" int asr_host_terminated_green(const double* a, const double* b, int n, double ainf, double binf, int ne,
                                         const double* z_, double* g_) {
  for (int ie = 0; ie < ne; ++ie) {
    cplx z(z_[2 * ie], z_[2 * ie + 1]);
    cplx tl = (binf != 0.0) ? sqrt_tail(z, ainf, binf) : cplx(0.0);
    cplx g = 1.0 / (z - a[n - 1] - tl);
    for (int i = n - 2; i >= 0; --i) g = 1.0 / (z - a[i] - b[i] * b[i] * g);
    g_[2 * ie] = g.real(); g_[2 * ie + 1] = g.imag();
  }
  return 0;
}

// Jackson-kernel DOS from Chebyshev moments (src/reference/asr_kpm.f90:68-87).
extern "C" int asr_host_kpm_dos(const double* mu, int nmom, double c, double s, int ne, const double* e, double* rho) {
  const double pi = 3.14159265358979323846;
  std::vector<double> g(nmom);
  for (int n = 0; n < nmom; ++n)
    g[n] = ((nmom - n + 1) * std::cos(pi * n / (nmom + 1)) + std::sin(pi * n / (nmom + 1)) / std::tan(pi / (nmom + 1))) /
           (nmom + 1);
  for (int i = 0; i < ne; ++i) {
    double x = (e[i] - c) / s;
    if (std::fabs(x) >= 1.0) { rho[i] = 0.0; continue; }
    double th = std::acos(x);
    double sum = g[0] * mu[0];
    for (int n = 1; n < nmom; ++n) sum += 2.0 * g[n] * mu[n] * std::cos(n * th);
    rho[i] = sum / (pi * std::sqrt(1.0 - x * x) * s);
  }
  return 0;
}

// Gauss-Legendre nodes and weights on [-1, 1] (src/asr_contour.f90:132-151).
extern "C" int asr_host_gauss_legendre(int n, double* x, double* w) {
  const double pi = 3.14159265358979323846;
  for (int i = 1; i <= (n + 1) / 2; ++i) {
    double z = std::cos(pi * (i - 0.25) / (n + 0.5)), z1, pp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1.0, p2 = 0.0, p3;
      for (int k = 1; k <= n; ++k) {
        p3 = p2; p2 = p1;
        p1 = ((2 * k - 1) * z * p2 - (k - 1) * p3) / k;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs(z - z1) < 1e-15) break;
    }
    x[i - 1] = -z; x[n - i] = z;
    w[i - 1] = 2.0 / ((1.0 - z * z) * pp * pp); w[n - i] = w[i - 1];
  }
  return 0;
}

// Arc from E_bot to E_F + i eta closed by the vertical segment to E_F (src/asr_contour.f90:39-58).  z, w: (re, im) pairs.
extern "C" int asr_host_contour_closed(double ebot, double ef, double eta, int n_arc, int n_vert, double* z_, double* w_) {
  const double pi = 3.14159265358979323846;
  std::vector<double> xa(n_arc), wa(n_arc), xv(n_vert), wv(n_vert);
  asr_host_gauss_legendre(n_arc, xa.data(), wa.data());
  asr_host_gauss_legendre(n_vert, xv.data(), wv.data());
  double c = (ef * ef + eta * eta - ebot * ebot) / (2.0 * (ef - ebot)), r = c - ebot;
  double th0 = std::atan2(eta, ef - c);
  for (int j = 0; j < n_arc; ++j) {
    double th = th0 + 0.5 * (pi - th0) * (1.0 - xa[j]);
    cplx z(c + r * std::cos(th), r * std::sin(th));
    cplx w = wa[j] * (-0.5 * (pi - th0)) * cplx(0.0, r) * std::exp(cplx(0.0, th));
    z_[2 * j] = z.real(); z_[2 * j + 1] = z.imag(); w_[2 * j] = w.real(); w_[2 * j + 1] = w.imag();
  }
  for (int j = 0; j < n_vert; ++j) {
    int k = n_arc + j;
    z_[2 * k] = ef; z_[2 * k + 1] = 0.5 * eta * (1.0 + xv[j]);
    w_[2 * k] = 0.0; w_[2 * k + 1] = -0.5 * eta * wv[j];
  }
  return 0;
}

// G(z) of n complex-symmetric chains at once (cf_species per chain, the tail of every chain the same; the chains are
// rows of a, b of leading dimension ld with m[j] levels): the host end of a contour pass.  Chains in parallel.
extern "C" int asr_host_cf_species_batch(int n, const double* a, const double* b, int ld, const int* m, const double* z,
                                         int tail, double* g11) {
  if (n < 0 || ld < 1) return -1;
  std::vector<double> zero((size_t)2 * ld, 0.0);
#pragma omp parallel for schedule(dynamic, 16)
  for (int j = 0; j < n; ++j) {
    const int mj = std::max(1, std::min(m[j], ld));
    double f[2];
    (void)asr_host_cf_species(a + (size_t)2 * ld * j, b + (size_t)2 * ld * j, zero.data(), mj, z + 2 * (size_t)j, tail,
                              g11 + 2 * (size_t)j, f, nullptr, nullptr);
  }
  return 0;
}

// Semicircle from E_bot to E_F through the upper half plane, optionally stopped at Im z = eta_end (eta_end <= 0: down to the
// real axis) (contour_nodes, src/asr_contour.f90:17-34).
extern "C" int asr_host_contour_nodes(double ebot, double ef, int n, double eta_end, double* z_, double* w_) {
  const double pi = 3.14159265358979323846;
  if (n < 1 || !(ef > ebot)) return 1;
  std::vector<double> x(n), wg(n);
  asr_host_gauss_legendre(n, x.data(), wg.data());
  const double c = 0.5 * (ebot + ef), r = 0.5 * (ef - ebot);
  const double th0 = eta_end > 0.0 ? std::asin(std::min(eta_end / r, 0.5)) : 0.0;
  for (int j = 0; j < n; ++j) {
    const double th = th0 + 0.5 * (pi - th0) * (1.0 - x[j]);  // pi at E_bot down to th0 near E_F
    const cplx z(c + r * std::cos(th), r * std::sin(th));
    const cplx w = wg[j] * (-0.5 * (pi - th0)) * cplx(0.0, r) * std::exp(cplx(0.0, th));
    z_[2 * j] = z.real(); z_[2 * j + 1] = z.imag(); w_[2 * j] = w.real(); w_[2 * j + 1] = w.imag();
  }
  return 0;
}

// Matsubara poles of the Fermi contour, Gamma = 2 npole pi kT ~ 0.1 Ry, 2 <= npole <= 40 (fermi_npole,
// src/asr_contour.f90:61-64), and the node count of the contour (fermi_nodes_count :66-69).
extern "C" int asr_host_fermi_npole(double kt) {
  const double pi = 3.14159265358979323846;
  const long n = std::lround(0.1 / (2.0 * pi * kt));
  return (int)std::max(2L, std::min(40L, n));
}
extern "C" int asr_host_fermi_nodes_count(int npole, int n_main, int n_step) { return 4 + n_main + n_step + 8 + npole; }

// Fermi-smeared contour (contour_nodes_fermi, src/asr_contour.f90:77-121; Wildberger et al., PRB 52, 11502 (1995)):
//   int f(E) [-(1/pi) Im h(E + i0)] dE = -(1/pi) Im sum_j w_j h(z_j)   for h analytic in the upper half plane.
// Legs: the vertical E_bot -> E_bot + i Gamma (4 nodes, weight times the complex Fermi function), the line Im z = Gamma from
// E_bot to mu - 10 kT (n_main), across the step to mu + 10 kT (n_step), to mu + 40 kT (8), where f(x + i Gamma) = f(x) is
// real; then the npole Matsubara poles mu + i (2n + 1) pi kT with w = -2 pi i kT.  z, w: fermi_nodes_count (re, im) pairs.
extern "C" int asr_host_contour_fermi(double ebot, double mu, double kt, int npole, int n_main, int n_step, double* z_,
                                      double* w_) {
  const double pi = 3.14159265358979323846;
  if (!(kt > 0.0) || npole < 0 || n_main < 1 || n_step < 1) return 1;
  auto fermi_r = [&](double e) {
    const double t = (e - mu) / kt;
    return t > 0.0 ? std::exp(-t) / (1.0 + std::exp(-t)) : 1.0 / (1.0 + std::exp(t));
  };
  auto fermi_c = [&](cplx zz) {
    const cplx t = (zz - mu) / kt;
    return t.real() < -600.0 ? cplx(1.0, 0.0) : 1.0 / (std::exp(t) + 1.0);
  };
  const double gam = 2.0 * npole * pi * kt;
  int k = 0;
  auto put = [&](cplx z, cplx w) {
    z_[2 * k] = z.real(); z_[2 * k + 1] = z.imag(); w_[2 * k] = w.real(); w_[2 * k + 1] = w.imag();
    ++k;
  };
  std::vector<double> xv(4), wv(4);
  asr_host_gauss_legendre(4, xv.data(), wv.data());
  for (int j = 0; j < 4; ++j) {
    const cplx z(ebot, 0.5 * gam * (1.0 + xv[j]));
    put(z, cplx(0.0, 0.5 * gam * wv[j]) * fermi_c(z));
  }
  auto line = [&](int n, double lo, double hi) {
    std::vector<double> x(n), w(n);
    asr_host_gauss_legendre(n, x.data(), w.data());
    for (int j = 0; j < n; ++j) {
      const double e = 0.5 * (lo + hi) + 0.5 * (hi - lo) * x[j];
      put(cplx(e, gam), cplx(0.5 * (hi - lo) * w[j] * fermi_r(e), 0.0));
    }
  };
  line(n_main, ebot, mu - 10.0 * kt);
  line(n_step, mu - 10.0 * kt, mu + 10.0 * kt);
  line(8, mu + 10.0 * kt, mu + 40.0 * kt);
  for (int j = 0; j < npole; ++j) put(cplx(mu, (2 * j + 1) * pi * kt), cplx(0.0, -2.0 * pi * kt));
  return 0;
}
