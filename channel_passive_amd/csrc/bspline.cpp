// bspline.cpp -- host-side generation of the B-spline collocation
// operators (replaces LABS_fill_full + LABS_initialize, which need GSL:
// SOURCE_PASSOPB/labs_fill_full.cpp:35-256, module_labs.f90:77-328).
//
// GSL's bspline module is a port of de Boor's PPPACK routines BSPLVB and
// BSPLVD (C. de Boor, "A Practical Guide to Splines", 1978, ch. X); those
// published recurrences are implemented here directly.  Greville abscissae
// use the extended-precision running mean of gsl_stats_mean; integration
// weights use the (k+1)/2-point Gauss-Legendre rule per breakpoint interval
// with nodes from a Newton iteration on P_n.
#include "bspline.h"

#include <algorithm>
#include <cmath>

namespace pb {
namespace {

struct Basis {
  int k, nbreak, n;
  std::vector<double> t;

  Basis(int order, const std::vector<double> &brk) : k(order), nbreak((int)brk.size()) {
    const int l = nbreak - 1;
    n = l + k - 1;
    t.resize(n + k);
    for (int i = 0; i < k; ++i) t[i] = brk.front();
    for (int i = 1; i < l; ++i) t[k - 1 + i] = brk[i];
    for (int i = n; i < n + k; ++i) t[i] = brk.back();
  }

  // left such that t[left] <= x < t[left+1]; the right end belongs to the last interval
  int locate(double x) const {
    if (x < t[0]) return -1;
    const int l = nbreak - 1;
    for (int i = k - 1; i < k + l - 1; ++i) {
      if (t[i] <= x && x < t[i + 1]) return i;
      if (t[i] < x && x == t[i + 1] && t[i + 1] == t[k + l - 1]) return i;
    }
    return -1;
  }

  // BSPLVB: raise the order of the nonzero B-spline values in b from
  // jcur+1 to jhigh at x.
  void raise(int jhigh, double x, int left, int &jcur, double *dl, double *dr, double *b) const {
    for (; jcur < jhigh - 1; ++jcur) {
      dr[jcur] = t[left + jcur + 1] - x;
      dl[jcur] = x - t[left - jcur];
      double saved = 0.0;
      for (int i = 0; i <= jcur; ++i) {
        const double term = b[i] / (dr[i] + dl[jcur - i]);
        b[i] = saved + dr[i] * term;
        saved = dl[jcur - i] * term;
      }
      b[jcur + 1] = saved;
    }
  }

  void values(double x, int left, double *b) const {
    std::vector<double> dl(k + 1), dr(k + 1);
    int j = 0;
    b[0] = 1.0;
    raise(k, x, left, j, dl.data(), dr.data(), b);
  }

  // BSPLVD: d[i][m] = m-th derivative of the i-th nonzero B-spline at x.
  void derivatives(double x, int left, int nd, std::vector<std::vector<double>> &d) const {
    d.assign(k, std::vector<double>(nd + 1, 0.0));
    std::vector<double> dl(k + 1), dr(k + 1), col(k, 0.0);
    std::vector<std::vector<double>> a(k, std::vector<double>(k, 0.0));
    const int mhigh = std::min(nd, k - 1);
    int j = 0;
    col[0] = 1.0;
    raise(k - mhigh, x, left, j, dl.data(), dr.data(), col.data());
    for (int i = 0; i < k; ++i) d[i][0] = col[i];
    if (nd == 0) return;
    int ideriv = mhigh;
    for (int m = 1; m <= mhigh; ++m) {
      for (int jj = ideriv, src = 0; jj < k; ++jj, ++src) d[jj][ideriv] = d[src][0];
      --ideriv;
      for (int i = 0; i < k; ++i) col[i] = d[i][0];
      raise(k - ideriv, x, left, j, dl.data(), dr.data(), col.data());
      for (int i = 0; i < k; ++i) d[i][0] = col[i];
    }
    int jlow = 0;
    for (int i = 0; i < k; ++i) {
      for (int jj = jlow; jj < k; ++jj) a[jj][i] = 0.0;
      jlow = i;
      a[i][i] = 1.0;
    }
    for (int m = 1; m <= mhigh; ++m) {
      const int kmm = k - m;
      const double fkmm = (double)(float)kmm;
      int il = left, i = k - 1;
      for (int rep = 0; rep < kmm; ++rep) {
        const double factor = fkmm / (t[il + kmm] - t[il]);
        for (int jj = 0; jj <= i; ++jj) a[i][jj] = factor * (a[i][jj] - a[i - 1][jj]);
        --il;
        --i;
      }
      for (i = 0; i < k; ++i) {
        double sum = 0;
        for (int jj = std::max(i, m); jj < k; ++jj) sum += a[jj][i] * d[jj][m];
        d[i][m] = sum;
      }
    }
  }

  double greville(int i) const {
    int km1 = k - 1;
    const double *p = t.data() + i + 1;
    if (km1 == 0) { km1 = 2; --p; }
    long double mean = 0;
    for (int m = 0; m < km1; ++m) mean += (p[m] - mean) / (m + 1);
    return (double)mean;
  }
};

// positive Gauss-Legendre nodes (ascending; 0 first for odd n) and weights
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
  const int h = (n + 1) / 2;
  std::vector<long double> xs(h), ws(h);
  const long double pi = 3.14159265358979323846264338327950288L;
  for (int i = 0; i < h; ++i) {
    long double z = cosl(pi * (i + 0.75L) / (n + 0.5L)), dp = 0;
    auto eval = [&](long double zz, long double &p1, long double &p2) {
      p1 = 1; p2 = 0;
      for (int j = 1; j <= n; ++j) {
        const long double p3 = p2;
        p2 = p1;
        p1 = ((2 * j - 1) * zz * p2 - (j - 1) * p3) / j;
      }
    };
    for (int it = 0; it < 100; ++it) {
      long double p1, p2;
      eval(z, p1, p2);
      dp = n * (z * p1 - p2) / (z * z - 1);
      const long double z1 = z;
      z = z1 - p1 / dp;
      if (fabsl(z - z1) < 1e-19L) {
        eval(z, p1, p2);
        dp = n * (z * p1 - p2) / (z * z - 1);
        break;
      }
    }
    xs[i] = z;
    ws[i] = 2 / ((1 - z * z) * dp * dp);
  }
  x.clear(); w.clear();
  for (int i = h - 1; i >= 0; --i) {
    x.push_back((n % 2 == 1 && i == h - 1) ? 0.0 : (double)xs[i]);
    w.push_back((double)ws[i]);
  }
}

}  // namespace

std::vector<double> breakpoints(int nbreak, int knot_map, double stretch, double ly) {
  std::vector<double> b(nbreak);
  const double pi = 3.141592653589793238462643;
  for (int i = 0; i < nbreak; ++i) {
    if (knot_map == 1)
      b[i] = 0.5 * ly * std::tanh(stretch * (-1.0 + (2.0 * i) / (nbreak - 1))) / std::tanh(stretch);
    else if (std::fabs(stretch) <= 0.0000000000000001)
      b[i] = -1.0 + 2.0 / (nbreak - 1) * i;
    else
      b[i] = std::sin(stretch * pi / 2.0 * (-1.0 + (2.0 * i) / (nbreak - 1))) / std::sin(stretch * pi / 2.0);
  }
  return b;
}

int build_collocation(int order, int n, int knot_map, double stretch, double ly, Collocation &out) {
  const int nbreak = n - order + 2;
  if (nbreak < 2 || order < 2) return -1;
  out.brk = breakpoints(nbreak, knot_map, stretch, ly);
  for (int i = 1; i < nbreak; ++i)
    if (!(out.brk[i] > out.brk[i - 1])) return -2;
  Basis B(order, out.brk);
  if (B.n != n) return -3;
  out.n = n;
  out.order = order;
  for (int m = 0; m < 3; ++m) out.D[m].assign((size_t)n * n, 0.0);
  out.y.resize(n);
  std::vector<std::vector<double>> d;
  std::vector<double> v(order);
  for (int r = 0; r < n; ++r) {
    const double y = B.greville(r);
    out.y[r] = y;
    const int left = B.locate(y);
    if (left < 0) return -4;
    B.derivatives(y, left, 4, d);
    B.values(y, left, v.data());
    const int c0 = left - order + 1;
    for (int j = 0; j < order; ++j) {
      out.D[0][(size_t)r * n + c0 + j] = v[j];
      out.D[1][(size_t)r * n + c0 + j] = (1 <= order - 1) ? d[j][1] : 0.0;
      out.D[2][(size_t)r * n + c0 + j] = (2 <= order - 1) ? d[j][2] : 0.0;
    }
  }
  // integration weights: ngl-point Gauss-Legendre per breakpoint interval
  const int ngl = (order + 1) / 2;
  std::vector<double> gx, gw;
  gauss_legendre(ngl, gx, gw);
  out.integ.assign(n, 0.0);
  for (int i = 0; i + 1 < nbreak; ++i) {
    const double a = B.t[order - 1 + i], b = B.t[order + i];
    const double A = (b - a) / 2, M = (a + b) / 2;
    for (int q = 0; q < ngl; ++q) {
      double xq, wq;
      if (ngl % 2 == 1) {
        const int kk = q - ngl / 2;
        const int sg = kk < 0 ? -1 : 1;
        xq = M + sg * A * gx[sg * kk];
        wq = A * gw[sg * kk];
      } else if (q < ngl / 2) {
        const int idx = ngl / 2 - 1 - q;
        xq = M - A * gx[idx];
        wq = A * gw[idx];
      } else {
        const int idx = q - ngl / 2;
        xq = M + A * gx[idx];
        wq = A * gw[idx];
      }
      const int left = B.locate(xq);
      if (left < 0) return -5;
      B.values(xq, left, v.data());
      for (int j = 0; j < order; ++j) out.integ[left - order + 1 + j] += wq * v[j];
    }
  }
  return 0;
}

int basis_row(const Collocation &c, double y, int &first, std::vector<double> &vals) {
  Basis B(c.order, c.brk);
  const int left = B.locate(y);
  if (left < 0) return -1;
  vals.assign(c.order, 0.0);
  B.values(y, left, vals.data());
  first = left - c.order + 1;
  return 0;
}

int detection_row(const Collocation &c, const std::vector<double> &d0lu, int bw, double y, std::vector<double> &r) {
  int first = 0;
  std::vector<double> vals;
  if (basis_row(c, y, first, vals)) return -1;
  const int n = c.n, W = 2 * bw + 1;
  auto fst = [&](int row) { return row < bw ? 0 : (row >= n - bw ? n - 1 - 2 * bw : row - bw); };
  auto lst = [&](int row) { return row < bw ? 2 * bw : (row >= n - bw ? n - 1 : row + bw); };
  auto at = [&](int row, int col) {
    int slot;
    if (row < bw && col > row + bw) slot = 2 * bw - col;
    else if (row >= n - bw && col < row - bw) slot = n - 1 - col;
    else slot = bw + col - row;
    return d0lu[(size_t)row * W + slot];
  };
  std::vector<double> b(n, 0.0);
  for (size_t j = 0; j < vals.size(); ++j) b[first + j] = vals[j];
  // U^T w = b (unit lower), rows of U that reach column col, ascending
  std::vector<double> w(n);
  for (int col = 0; col < n; ++col) {
    double v = b[col];
    for (int row = std::max(0, col - 2 * bw); row < col; ++row)
      if (col <= lst(row)) v = v - at(row, col) * w[row];
    w[col] = v;
  }
  // L^T r = w (upper with the diagonal), descending
  r.assign(n, 0.0);
  for (int p = n - 1; p >= 0; --p) {
    double v = w[p];
    for (int q = p + 1; q <= std::min(n - 1, p + 2 * bw); ++q)
      if (fst(q) <= p) v = v - at(q, p) * r[q];
    r[p] = v / at(p, p);
  }
  return 0;
}

// LABS_eval_bw (module_labs.f90:245-293): smallest k for which every entry
// of D0, D1, D2 outside the packed band pattern is exactly zero.
int eval_bandwidth(const Collocation &c) {
  const int n = c.n;
  for (int k = 1; 2 * k + 1 < n; ++k) {
    bool ok = true;
    auto zero = [&](int r, int col) {
      for (int m = 0; m < 3; ++m)
        if (c.D[m][(size_t)r * n + col] != 0.0) ok = false;
    };
    for (int r = 0; r <= k && ok; ++r)
      for (int col = 2 * k + 1; col < n; ++col) zero(r, col);
    for (int r = k + 1; r < n - k - 1 && ok; ++r) {
      for (int col = 0; col < r - k; ++col) zero(r, col);
      for (int col = r + k + 1; col < n; ++col) zero(r, col);
    }
    for (int r = n - k - 1; r < n && ok; ++r)
      for (int col = 0; col < n - 2 * k - 1; ++col) zero(r, col);
    if (ok) return k;
  }
  return -1;
}

void to_band(const std::vector<double> &full, int n, int bw, std::vector<double> &band) {
  const int W = 2 * bw + 1;
  band.assign((size_t)n * W, 0.0);
  for (int r = 0; r < n; ++r) {
    const int f = r < bw ? 0 : (r >= n - bw ? n - 1 - 2 * bw : r - bw);
    const int l = r < bw ? 2 * bw : (r >= n - bw ? n - 1 : r + bw);
    for (int col = f; col <= l; ++col) {
      int s;
      if (r < bw && col > r + bw) s = 2 * bw - col;
      else if (r >= n - bw && col < r - bw) s = n - 1 - col;
      else s = bw + (col - r);
      band[(size_t)r * W + s] = full[(size_t)r * n + col];
    }
  }
}

}  // namespace pb
