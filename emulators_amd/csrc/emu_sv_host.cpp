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
