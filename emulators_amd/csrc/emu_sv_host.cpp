// emu_sv_host.cpp -- host-side numerics of the Krylov step.
//
// expm of the small (j+3)x(j+3) extended Krylov matrix, computed in complex
// double by scaling and squaring with the [13/13] Pade approximant
// (N. J. Higham, "The scaling and squaring method for the matrix exponential
// revisited", SIAM J. Matrix Anal. Appl. 26(4), 2005, Algorithm 2.3 with the
// degree-13 branch).  The matrices are at most a few hundred wide, so a plain
// O(m^3) implementation with partial-pivoting LU is enough.

#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "emu_sv_hip.h"

namespace {

using cd = std::complex<double>;

struct Mat {
  int n;
  std::vector<cd> a;
  explicit Mat(int n_) : n(n_), a((size_t)n_ * n_, cd(0.0, 0.0)) {}
  cd &operator()(int i, int j) { return a[(size_t)i * n + j]; }
  const cd &operator()(int i, int j) const { return a[(size_t)i * n + j]; }
};

Mat mul(const Mat &x, const Mat &y) {
  const int n = x.n;
  Mat z(n);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      const cd xik = x(i, k);
      if (xik == cd(0.0, 0.0)) continue;
      for (int j = 0; j < n; ++j) z(i, j) += xik * y(k, j);
    }
  return z;
}

double norm1(const Mat &x) {
  double best = 0.0;
  for (int j = 0; j < x.n; ++j) {
    double s = 0.0;
    for (int i = 0; i < x.n; ++i) s += std::abs(x(i, j));
    if (s > best) best = s;
  }
  return best;
}

// Solve P X = Q in place (Q <- X) with partial pivoting.  Returns false on a
// singular pivot.
bool solve(Mat p, Mat &q) {
  const int n = p.n;
  for (int col = 0; col < n; ++col) {
    int piv = col;
    double best = std::abs(p(col, col));
    for (int r = col + 1; r < n; ++r)
      if (std::abs(p(r, col)) > best) {
        best = std::abs(p(r, col));
        piv = r;
      }
    if (best == 0.0) return false;
    if (piv != col)
      for (int j = 0; j < n; ++j) {
        std::swap(p(col, j), p(piv, j));
        std::swap(q(col, j), q(piv, j));
      }
    const cd inv = 1.0 / p(col, col);
    for (int r = col + 1; r < n; ++r) {
      const cd f = p(r, col) * inv;
      if (f == cd(0.0, 0.0)) continue;
      for (int j = col; j < n; ++j) p(r, j) -= f * p(col, j);
      for (int j = 0; j < n; ++j) q(r, j) -= f * q(col, j);
    }
  }
  for (int col = n - 1; col >= 0; --col) {
    const cd inv = 1.0 / p(col, col);
    for (int j = 0; j < n; ++j) q(col, j) *= inv;
    for (int r = 0; r < col; ++r) {
      const cd f = p(r, col);
      if (f == cd(0.0, 0.0)) continue;
      for (int j = 0; j < n; ++j) q(r, j) -= f * q(col, j);
    }
  }
  return true;
}

// Pade [13/13] coefficients b_0..b_13 (Higham 2005, Table 2.3 context).
const double kB13[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0,
                         1187353796428800.0,  129060195264000.0,   10559470521600.0,
                         670442572800.0,      33522128640.0,       1323241920.0,
                         40840800.0,          960960.0,            16380.0,
                         182.0,               1.0};
const double kTheta13 = 5.371920351148152;

}  // namespace

extern "C" int emu_sv_expm(int m, const double *a_in, double *e_out) {
  if (m < 1 || m > 4096 || !a_in || !e_out) return EMU_SV_EINVAL;
  Mat a(m);
  for (int i = 0; i < m * m; ++i) a.a[i] = cd(a_in[2 * i], a_in[2 * i + 1]);
  const double nrm = norm1(a);
  int s = 0;
  if (nrm > kTheta13) s = (int)std::ceil(std::log2(nrm / kTheta13));
  if (s > 0) {
    const double f = std::ldexp(1.0, -s);
    for (auto &x : a.a) x *= f;
  }
  Mat a2 = mul(a, a), a4 = mul(a2, a2), a6 = mul(a4, a2);
  Mat u(m), v(m), t(m);
  for (int i = 0; i < m * m; ++i) {
    t.a[i] = kB13[13] * a6.a[i] + kB13[11] * a4.a[i] + kB13[9] * a2.a[i];
  }
  t = mul(a6, t);
  for (int i = 0; i < m * m; ++i) t.a[i] += kB13[7] * a6.a[i] + kB13[5] * a4.a[i] + kB13[3] * a2.a[i];
  for (int i = 0; i < m; ++i) t(i, i) += kB13[1];
  u = mul(a, t);
  Mat t2(m);
  for (int i = 0; i < m * m; ++i) t2.a[i] = kB13[12] * a6.a[i] + kB13[10] * a4.a[i] + kB13[8] * a2.a[i];
  v = mul(a6, t2);
  for (int i = 0; i < m * m; ++i) v.a[i] += kB13[6] * a6.a[i] + kB13[4] * a4.a[i] + kB13[2] * a2.a[i];
  for (int i = 0; i < m; ++i) v(i, i) += kB13[0];
  Mat p(m), q(m);
  for (int i = 0; i < m * m; ++i) {
    p.a[i] = v.a[i] - u.a[i];
    q.a[i] = v.a[i] + u.a[i];
  }
  if (!solve(p, q)) return EMU_SV_EINVAL;
  for (int k = 0; k < s; ++k) q = mul(q, q);
  for (int i = 0; i < m * m; ++i) {
    e_out[2 * i] = q.a[i].real();
    e_out[2 * i + 1] = q.a[i].imag();
  }
  return EMU_SV_OK;
}

// First column of expm(A): exp(A) e_0 = e^c exp(A - cI) e_0 with c the mean
// of the diagonal, advanced in s = ceil(||A - cI||_inf / 4) substeps of a
// 32-term Taylor series (truncation 4^33/33! < 1e-17 per substep, largest
// term 4^4/4! ~ 11).  O(s m^2) per call instead of the O(m^3) products of the
// Pade form; beyond ||A - cI||_inf = 256 the Pade form is used.
extern "C" int emu_sv_expm_column(int m, const double *a_in, double *col_out) {
  if (m < 1 || m > 4096 || !a_in || !col_out) return EMU_SV_EINVAL;
  std::vector<cd> a((size_t)m * m);
  cd c(0.0, 0.0);
  for (int i = 0; i < m * m; ++i) a[i] = cd(a_in[2 * i], a_in[2 * i + 1]);
  for (int i = 0; i < m; ++i) c += a[(size_t)i * m + i];
  c /= (double)m;
  for (int i = 0; i < m; ++i) a[(size_t)i * m + i] -= c;
  double rowmax = 0.0;
  std::vector<int> first(m), last(m);  // nonzero span of each row
  for (int i = 0; i < m; ++i) {
    double rs = 0.0;
    int f = m, l = -1;
    for (int j = 0; j < m; ++j) {
      const cd x = a[(size_t)i * m + j];
      if (x != cd(0.0, 0.0)) {
        rs += std::abs(x);
        if (f == m) f = j;
        l = j;
      }
    }
    first[i] = f;
    last[i] = l;
    rowmax = rs > rowmax ? rs : rowmax;
  }
  if (rowmax > 256.0) {
    std::vector<double> e((size_t)2 * m * m);
    if (int rc = emu_sv_expm(m, a_in, e.data())) return rc;
    for (int i = 0; i < m; ++i) {
      col_out[2 * i] = e[(size_t)2 * i * m];
      col_out[2 * i + 1] = e[(size_t)2 * i * m + 1];
    }
    return EMU_SV_OK;
  }
  const int steps = rowmax > 0.0 ? (int)std::ceil(rowmax / 4.0) : 1;
  const double h = 1.0 / steps;
  // real arithmetic (std::complex products go through the library's
  // Inf/NaN-aware multiply)
  std::vector<double> ar((size_t)m * m), ai((size_t)m * m);
  for (size_t i = 0; i < (size_t)m * m; ++i) {
    ar[i] = a[i].real();
    ai[i] = a[i].imag();
  }
  std::vector<double> ur(m, 0.0), ui(m, 0.0), tr(m), ti(m), yr(m), yi(m);
  ur[0] = 1.0;
  for (int st = 0; st < steps; ++st) {
    tr = ur;
    ti = ui;
    for (int k = 1; k <= 32; ++k) {
      const double f = h / k;
      for (int i = 0; i < m; ++i) {
        double sr = 0.0, si = 0.0;
        const double *rr = &ar[(size_t)i * m], *ri = &ai[(size_t)i * m];
        for (int j = first[i]; j <= last[i]; ++j) {
          sr += rr[j] * tr[j] - ri[j] * ti[j];
          si += rr[j] * ti[j] + ri[j] * tr[j];
        }
        yr[i] = sr * f;
        yi[i] = si * f;
      }
      for (int i = 0; i < m; ++i) {
        tr[i] = yr[i];
        ti[i] = yi[i];
        ur[i] += yr[i];
        ui[i] += yi[i];
      }
    }
  }
  const double em = std::exp(c.real()), er = em * std::cos(c.imag()), ei = em * std::sin(c.imag());
  for (int i = 0; i < m; ++i) {
    col_out[2 * i] = er * ur[i] - ei * ui[i];
    col_out[2 * i + 1] = er * ui[i] + ei * ur[i];
  }
  return EMU_SV_OK;
}

// ---------------------------------------------------------------------------
// Lanczos exponential from the symmetric tridiagonal form.
//
// For op = sigma*H with H Hermitian, the Lanczos matrix T (T[k,k] = sigma
// alpha_k, T[k+1,k] = |sigma| beta_k) satisfies T = W^-1 (sigma J) W with J
// the real symmetric tridiagonal (alpha, beta) and W = diag(w^k), w =
// sigma/|sigma|.  With J = Z diag(lambda) Z^T:
//   (exp(T) e_0)_k            = w^-k     sum_i Z[k,i] exp(sigma lambda_i) Z[0,i]
//   e_{m-1}^T phi_p(T) e_0    = w^-(m-1) sum_i Z[m-1,i] phi_p(sigma lambda_i) Z[0,i]
// where phi_1(z) = (e^z - 1)/z and phi_2(z) = (e^z - 1 - z)/z^2 give the two
// extra entries of exp of the Expokit extended matrix (Sidje 1998).  The
// eigen-decomposition is the implicit symmetric QR iteration with Wilkinson
// shifts (Golub & Van Loan, Matrix Computations, section 8.3); only the rows
// of Z that are needed are accumulated: rows 0 and m-1 per iteration, all
// rows when the combination coefficients are requested.
// ---------------------------------------------------------------------------
namespace {

bool tridiag_eig(int n, double *d, double *e, int nrow, const int *rows, double *zr) {
  for (int i = 0; i < nrow; ++i)
    for (int j = 0; j < n; ++j) zr[(size_t)i * n + j] = (rows[i] == j) ? 1.0 : 0.0;
  const double eps = 2.220446049250313e-16;
  int hi = n - 1, iters = 0;
  while (hi > 0) {
    if (std::fabs(e[hi - 1]) <= eps * (std::fabs(d[hi - 1]) + std::fabs(d[hi]))) {
      e[hi - 1] = 0.0;
      --hi;
      continue;
    }
    int lo = hi - 1;
    while (lo > 0 && std::fabs(e[lo - 1]) > eps * (std::fabs(d[lo - 1]) + std::fabs(d[lo]))) --lo;
    if (lo > 0) e[lo - 1] = 0.0;
    if (++iters > 100 * n) return false;
    // Wilkinson shift: eigenvalue of the trailing 2x2 closer to d[hi].
    const double dd = 0.5 * (d[hi - 1] - d[hi]);
    const double eh = e[hi - 1];
    const double mu = d[hi] - eh * eh / (dd + std::copysign(std::hypot(dd, eh), dd));
    double x = d[lo] - mu, z = e[lo], bulge = 0.0;
    for (int k = lo; k < hi; ++k) {
      // |x|, |z| are bounded by the matrix entries (|J| << 1e150), so the
      // plain root cannot overflow; it is several times cheaper than hypot
      const double r = std::sqrt(x * x + z * z);
      const double c = r == 0.0 ? 1.0 : x / r, s = r == 0.0 ? 0.0 : z / r;
      if (k > lo) e[k - 1] = r;
      const double a = d[k], b = e[k], d1 = d[k + 1];
      d[k] = c * c * a + 2.0 * c * s * b + s * s * d1;
      d[k + 1] = s * s * a - 2.0 * c * s * b + c * c * d1;
      e[k] = c * s * (d1 - a) + (c * c - s * s) * b;
      if (k + 1 < hi) {
        bulge = s * e[k + 1];
        e[k + 1] = c * e[k + 1];
      }
      for (int i = 0; i < nrow; ++i) {
        double *row = zr + (size_t)i * n;
        const double zk = row[k], zk1 = row[k + 1];
        row[k] = c * zk + s * zk1;
        row[k + 1] = -s * zk + c * zk1;
      }
      x = e[k];
      z = bulge;
    }
  }
  return true;
}

// phi_p(z) for p = 1, 2: series near 0, closed form elsewhere.  Written
// in real arithmetic: std::complex products and quotients go through the
// library's NaN-aware helpers, which dominated the cost of this routine.
cd phi(int p, cd zc) {
  const double zr = zc.real(), zi = zc.imag();
  if (zr * zr + zi * zi < 0.25) {
    // sum_k z^k / (k + p)!
    double tr = 1.0, ti = 0.0, sr = 0.0, si = 0.0;
    double fact = p == 1 ? 1.0 : 2.0;  // p!
    for (int k = 0; k < 30; ++k) {
      sr += tr / fact;
      si += ti / fact;
      const double nr = tr * zr - ti * zi;
      ti = tr * zi + ti * zr;
      tr = nr;
      fact *= (k + p + 1);
    }
    return cd(sr, si);
  }
  const double ea = std::exp(zr);
  double nr = ea * std::cos(zi) - 1.0, ni = ea * std::sin(zi);  // e^z - 1
  double dr = zr, di = zi;                                      // z
  if (p == 2) {
    nr -= zr;  // e^z - 1 - z
    ni -= zi;
    dr = zr * zr - zi * zi;  // z^2
    di = 2.0 * zr * zi;
  }
  const double den = dr * dr + di * di;
  return cd((nr * dr + ni * di) / den, (ni * dr - nr * di) / den);
}

}  // namespace

extern "C" int emu_sv_lanczos_exp(int m, const double *tdiag, const double *tsub, double sigma_re,
                                  double sigma_im, double n2, int want_coeffs, double *out) {
  if (m < 1 || m > 4096 || !tdiag || (m > 1 && !tsub) || !out) return EMU_SV_EINVAL;
  const cd sigma(sigma_re, sigma_im);
  const double asig = std::abs(sigma);
  if (asig == 0.0) return EMU_SV_EINVAL;
  const cd w = sigma / asig;
  std::vector<double> d(m), e(m > 1 ? m - 1 : 1, 0.0);
  for (int k = 0; k < m; ++k) d[k] = (cd(tdiag[2 * k], tdiag[2 * k + 1]) / sigma).real();
  for (int k = 0; k + 1 < m; ++k) e[k] = tsub[k] / asig;
  std::vector<int> rows;
  if (want_coeffs) {
    for (int k = 0; k < m; ++k) rows.push_back(k);
  } else {
    rows.push_back(0);
    if (m > 1) rows.push_back(m - 1);
  }
  const int nrow = (int)rows.size();
  std::vector<double> zr((size_t)nrow * m);
  if (!tridiag_eig(m, d.data(), e.data(), nrow, rows.data(), zr.data())) return EMU_SV_EINVAL;
  const double *z0 = zr.data();
  const double *zl = zr.data() + (size_t)(nrow - 1) * m;  // row m-1 (or row 0 when m == 1)
  cd s1 = 0.0, s2 = 0.0;
  for (int i = 0; i < m; ++i) {
    const cd x = sigma * d[i];
    s1 += zl[i] * phi(1, x) * z0[i];
    s2 += zl[i] * phi(2, x) * z0[i];
  }
  const cd wl = std::pow(std::conj(w), m - 1);  // w^-(m-1), |w| = 1
  const cd c_ext = n2 * wl * s1;
  out[0] = std::abs(c_ext);
  out[1] = std::abs(n2 * wl * s2);
  if (want_coeffs) {
    std::vector<cd> ex(m);
    for (int i = 0; i < m; ++i) ex[i] = std::exp(sigma * d[i]) * z0[i];
    cd wk = 1.0;
    for (int k = 0; k < m; ++k) {
      cd c = 0.0;
      const double *zk = zr.data() + (size_t)k * m;
      for (int i = 0; i < m; ++i) c += zk[i] * ex[i];
      c *= wk;
      out[2 + 2 * k] = c.real();
      out[3 + 2 * k] = c.imag();
      wk *= std::conj(w);
    }
    out[2 + 2 * m] = c_ext.real();
    out[3 + 2 * m] = c_ext.imag();
  }
  return EMU_SV_OK;
}

extern "C" int emu_sv_tridiag_eig(int m, const double *d_in, const double *e_in, double *lambda_out,
                                  double *z_out) {
  if (m < 1 || m > 4096 || !d_in || (m > 1 && !e_in) || !lambda_out || !z_out) return EMU_SV_EINVAL;
  std::vector<double> d(d_in, d_in + m), e(m > 1 ? m - 1 : 1, 0.0);
  for (int k = 0; k + 1 < m; ++k) e[k] = e_in[k];
  std::vector<int> rows(m);
  for (int k = 0; k < m; ++k) rows[k] = k;
  if (!tridiag_eig(m, d.data(), e.data(), m, rows.data(), z_out)) return EMU_SV_EINVAL;
  for (int k = 0; k < m; ++k) lambda_out[k] = d[k];
  return EMU_SV_OK;
}
