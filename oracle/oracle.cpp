// oracle/oracle.cpp -- CPU restatement of the reference RK3-CN substep.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the CPU
// baseline: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it.  The product library (channel_passive_amd/csrc) never
// links it.
//
// It restates, in plain C++ with strict IEEE FP64 (built with
// -ffp-contract=off, no FMA), the algorithm of yaojieuh/CHANNEL_PASSIVE
// (files cited relative to SOURCE_PASSOPB/):
//   * B-spline operators          labs_fill_full.cpp:35-256  (GSL 1.16 / 2.x
//     bspline + glfixed are not available; de Boor's BSPLVB/BSPLVD, the
//     published algorithms GSL ports, are restated instead)
//   * LABS packed band algebra    module_labs.f90:245-328, 401-413, 3487-3636,
//                                 4252-4318, 4395-4430
//   * 3/2-dealiased transforms    module_ppdfftw.f90:794-861, dft_x_realspace.f90:68-213
//     (FFTW 3.3.x is not available; an own mixed-radix FFT computes the same
//     unnormalised DFTs)
//   * rescale, wavespace          rescale_data.f90, wavespace.f90:46-107
//   * per-line solvers            hvhg.f90, solve_g.f90, solve_t.f90,
//                                 solve_v_f.f90, solve_u_w.f90, wave00.f90
//   * RK3 coefficients            module_runge_kutta.f90:34-48
//   * stores from u,v,w           initial_step.f90:139-179
//
// Complex arithmetic is written out component-wise so that the order of
// every rounding is explicit: complex*real and complex/real act on each
// component (what gfortran emits for a real operand), complex*complex is
// (ar*br - ai*bi, ar*bi + ai*br), complex/complex is Smith's algorithm
// (gfortran's -fcx-fortran-rules division).
//
// Parity pins: see tests/test_oracle_*.py (scipy BSpline, numpy.fft, dense
// LU, closed-form laminar states, divergence/wall invariants).

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <omp.h>

#include "../include/pb_capi.h"

namespace {

// Planted mistakes for the tests that check the oracle's own pins (0 = the
// restatement as written).  Each nonzero value changes one term the way a
// plausible transcription slip would; a pin that does not fail under it does
// not cover that term.
//   1 hvhg: sign of (kz^2 - kx^2) P5 in h_g
//   2 hvhg: i (kx^3 + kx kz^2) P2 dropped from h_v
//   3 solve_g/solve_t store: (1 + a k^2) in place of (1 - a k^2)
//   4 solve_g/solve_t store: zeta of substep s in place of s + 1
//   5 solve_v_f: C2 v_h(i) in place of C2 v_h(ny + 1 - i) in v
//   6 solve_v_f: Poisson operator D2 + k^2 D0
//   7 Helmholtz operator (1 - b k^2) D0 - b D2
//   8 wave00: + c d(uv)/dy in place of - c d(uv)/dy in the mean-U right-hand side
//   9 mixed-type scalar source with the wrong sign, -(nu/pr) u / bulk_u
//  10 statistics: <2p du/dx> with the wrong sign on kx > 0
static int g_mutant = 0;

// ----------------------------------------------------------------- complex
struct cd {
  double re, im;
};
static inline cd C(double r, double i) { cd z; z.re = r; z.im = i; return z; }
static inline cd add(cd a, cd b) { return C(a.re + b.re, a.im + b.im); }
static inline cd sub(cd a, cd b) { return C(a.re - b.re, a.im - b.im); }
static inline cd mulr(cd a, double r) { return C(a.re * r, a.im * r); }
static inline cd divr(cd a, double r) { return C(a.re / r, a.im / r); }
static inline cd mul(cd a, cd b) {
  return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// Smith's complex division (GCC expand_complex_div_wide, Fortran rules).
static inline cd cdiv(cd a, cd b) {
  double tr, ti;
  if (std::fabs(b.re) < std::fabs(b.im)) {
    double ratio = b.re / b.im;
    double div = (b.re * ratio) + b.im;
    tr = ((a.re * ratio) + a.im) / div;
    ti = ((a.im * ratio) - a.re) / div;
  } else {
    double ratio = b.im / b.re;
    double div = (b.im * ratio) + b.re;
    tr = ((a.im * ratio) + a.re) / div;
    ti = (a.im - (a.re * ratio)) / div;
  }
  return C(tr, ti);
}
// i*k as the reference forms it: ci*k with ci = (0,1) -> (0*k - 1*0, 0*0 + 1*k)
static inline cd imag_unit_times(double k) { return C(0.0 * k - 1.0 * 0.0, 0.0 * 0.0 + 1.0 * k); }

const double PI_REF = 3.141592653589793238462643;  // labs_fill_full.cpp:24

// ---------------------------------------------------------------- B-splines
// GSL workspace equivalent: order k, ncoeffs n = nbreak + k - 2, knots n + k.
struct Bspl {
  int k = 0, n = 0, nbreak = 0;
  std::vector<double> t;  // knots
};

// gsl_bspline_knots: k-fold end knots, interior breakpoints.
static void bspl_knots(Bspl &w, const std::vector<double> &brk) {
  w.nbreak = (int)brk.size();
  const int l = w.nbreak - 1;
  w.n = l + w.k - 1;
  w.t.assign(w.n + w.k, 0.0);
  for (int i = 0; i < w.k; ++i) w.t[i] = brk[0];
  for (int i = 1; i < l; ++i) w.t[w.k - 1 + i] = brk[i];
  for (int i = w.n; i < w.n + w.k; ++i) w.t[i] = brk[l];
}

// Interval search: t[i] <= x < t[i+1], right end point belongs to the last
// interval (GSL bspline_find_interval).  Returns -1 when outside.
static int bspl_interval(const Bspl &w, double x) {
  if (x < w.t[0]) return -1;
  const int l = w.nbreak - 1;
  int i;
  for (i = w.k - 1; i < w.k + l - 1; ++i) {
    const double ti = w.t[i], tip1 = w.t[i + 1];
    if (ti <= x && x < tip1) break;
    if (ti < x && x == tip1 && tip1 == w.t[w.k + l - 1]) break;
  }
  if (i == w.k + l - 1) return -1;
  return i;
}

// de Boor BSPLVB (A Practical Guide to Splines, 1978, ch. X): values of the
// jhigh nonzero B-splines of order jhigh at x, left = interval index.
static void bsplvb(const std::vector<double> &t, int jhigh, int index, double x, int left,
                   int &j, double *deltal, double *deltar, double *biatx) {
  if (index == 1) {
    j = 0;
    biatx[0] = 1.0;
  }
  for (; j < jhigh - 1; ++j) {
    deltar[j] = t[left + j + 1] - x;
    deltal[j] = x - t[left - j];
    double saved = 0.0;
    for (int i = 0; i <= j; ++i) {
      const double term = biatx[i] / (deltar[i] + deltal[j - i]);
      biatx[i] = saved + deltar[i] * term;
      saved = deltal[j - i] * term;
    }
    biatx[j + 1] = saved;
  }
}

// de Boor BSPLVD: dbiatx[i*(nd+1) + m] = m-th derivative of the i-th nonzero
// B-spline (from the left) at x, m = 0..nderiv.
static void bsplvd(const std::vector<double> &t, int k, double x, int left, int nderiv,
                   std::vector<double> &dbiatx) {
  const int nd1 = nderiv + 1;
  dbiatx.assign((size_t)k * nd1, 0.0);
  std::vector<double> a((size_t)k * k, 0.0), deltal(k + 1), deltar(k + 1), col(k, 0.0);
  const int mhigh = std::min(nderiv, k - 1);
  int jj = 0;
  auto DB = [&](int i, int m) -> double & { return dbiatx[(size_t)i * nd1 + m]; };
  auto A = [&](int i, int j) -> double & { return a[(size_t)i * k + j]; };
  bsplvb(t, k - mhigh, 1, x, left, jj, deltal.data(), deltar.data(), col.data());
  for (int i = 0; i < k; ++i) DB(i, 0) = col[i];
  if (nderiv > 0) {
    int ideriv = mhigh;
    for (int m = 1; m <= mhigh; ++m) {
      for (int j = ideriv, jp1mid = 0; j < k; ++j, ++jp1mid) DB(j, ideriv) = DB(jp1mid, 0);
      --ideriv;
      for (int i = 0; i < k; ++i) col[i] = DB(i, 0);
      bsplvb(t, k - ideriv, 2, x, left, jj, deltal.data(), deltar.data(), col.data());
      for (int i = 0; i < k; ++i) DB(i, 0) = col[i];
    }
    int jlow = 0;
    for (int i = 0; i < k; ++i) {
      for (int j = jlow; j < k; ++j) A(j, i) = 0.0;
      jlow = i;
      A(i, i) = 1.0;
    }
    for (int m = 1; m <= mhigh; ++m) {
      const int kmm = k - m;
      const double fkmm = (double)(float)kmm;
      int il = left;
      int i = k - 1;
      for (int ld = 0; ld < kmm; ++ld) {
        const double factor = fkmm / (t[il + kmm] - t[il]);
        for (int j = 0; j <= i; ++j) A(i, j) = factor * (A(i, j) - A(i - 1, j));
        --il;
        --i;
      }
      for (i = 0; i < k; ++i) {
        double sum = 0;
        const int jl = std::max(i, m);
        for (int j = jl; j < k; ++j) sum += A(j, i) * DB(j, m);
        DB(i, m) = sum;
      }
    }
  }
}

// Greville abscissa: mean of t[i+1 .. i+k-1] with the long-double running
// mean of gsl_stats_mean.
static double bspl_greville(const Bspl &w, int i) {
  int km1 = w.k - 1;
  const double *data = w.t.data() + (i + 1);
  if (km1 == 0) { km1 = 2; data -= 1; }
  long double mean = 0;
  for (int m = 0; m < km1; ++m) mean += (data[m] - mean) / (m + 1);
  return (double)mean;
}

// Gauss-Legendre nodes (positive half, ascending) and weights for n points
// by Newton iteration on P_n in long double.
static void gauss_legendre_half(int n, std::vector<double> &x, std::vector<double> &wt) {
  const int h = (n + 1) / 2;
  std::vector<long double> xs, ws;
  for (int i = 0; i < h; ++i) {
    long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
    long double pp = 0;
    for (int it = 0; it < 100; ++it) {
      long double p1 = 1, p2 = 0;
      for (int j = 1; j <= n; ++j) {
        long double p3 = p2;
        p2 = p1;
        p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1);
      long double z1 = z;
      z = z1 - p1 / pp;
      if (fabsl(z - z1) < 1e-19L) {
        // one more evaluation at the converged node for the weight
        p1 = 1; p2 = 0;
        for (int j = 1; j <= n; ++j) {
          long double p3 = p2;
          p2 = p1;
          p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
        }
        pp = n * (z * p1 - p2) / (z * z - 1);
        break;
      }
    }
    xs.push_back(z);
    ws.push_back(2 / ((1 - z * z) * pp * pp));
  }
  // ascending positive nodes; for odd n the first is 0
  x.clear(); wt.clear();
  for (int i = h - 1; i >= 0; --i) {
    long double xv = xs[i];
    if (n % 2 == 1 && i == h - 1) xv = 0;
    x.push_back((double)xv);
    wt.push_back((double)ws[i]);
  }
}

// gsl_integration_glfixed_point: i-th point (ascending) of the n-point rule on [a,b].
static void gl_point(double a, double b, int i, int n, const std::vector<double> &x,
                     const std::vector<double> &wt, double &xi, double &wi) {
  const double A = (b - a) / 2;
  const double B = (a + b) / 2;
  if (n % 2 == 1) {
    const int k = i - n / 2;
    const int s = k < 0 ? -1 : 1;
    xi = B + s * A * x[s * k];
    wi = A * wt[s * k];
  } else if (i < n / 2) {
    const int idx = (n / 2) - 1 - i;
    xi = B - A * x[idx];
    wi = A * wt[idx];
  } else {
    const int idx = i - n / 2;
    xi = B + A * x[idx];
    wi = A * wt[idx];
  }
}

// LABS_fill_full (labs_fill_full.cpp:35-195): k = GSL order, n = ny.
// M0..M4 row-major: M[r*n + c] = d^m B_c / dy^m at collocation point r.
static int labs_fill_full(int k, int n, int knot_map, double stretch, double ly,
                          std::vector<double> M[5], std::vector<double> &yc,
                          std::vector<double> &brk, std::vector<double> &integ,
                          std::vector<double> &center) {
  const int nbreak = n - k + 2;
  if (nbreak < 2) return -1;
  brk.assign(nbreak, 0.0);
  if (knot_map == PB_KNOTS_TANH) {
    for (int i = 0; i < nbreak; ++i)
      brk[i] = 0.5 * ly * std::tanh(stretch * (-1.0 + (2.0 * i) / (nbreak - 1))) / std::tanh(stretch);
  } else if (std::fabs(stretch) <= 0.0000000000000001) {
    for (int i = 0; i < nbreak; ++i) brk[i] = -1.0 + 2.0 / (nbreak - 1) * i;
  } else {
    for (int i = 0; i < nbreak; ++i)
      brk[i] = std::sin(stretch * PI_REF / 2.0 * (-1.0 + (2.0 * i) / (nbreak - 1))) /
               std::sin(stretch * PI_REF / 2.0);
  }
  Bspl w;
  w.k = k;
  bspl_knots(w, brk);
  if (w.n != n) return -2;
  for (int m = 0; m < 5; ++m) M[m].assign((size_t)n * n, 0.0);
  yc.assign(n, 0.0);
  std::vector<double> db;
  for (int i = 0; i < n; ++i) {
    const double y = bspl_greville(w, i);
    yc[i] = y;
    const int left = bspl_interval(w, y);
    if (left < 0) return -3;
    bsplvd(w.t, k, y, left, 4, db);
    const int mmax = std::min(4, k - 1);
    for (int j = 0; j < k; ++j) {
      const int c = left - k + 1 + j;
      for (int m = 1; m <= 4; ++m) M[m][(size_t)i * n + c] = (m <= mmax) ? db[(size_t)j * 5 + m] : 0.0;
    }
  }
  std::vector<double> dl(k + 1), dr(k + 1), bia(k);
  for (int i = 0; i < n; ++i) {
    const double y = yc[i];
    const int left = bspl_interval(w, y);
    int jj = 0;
    bsplvb(w.t, k, 1, y, left, jj, dl.data(), dr.data(), bia.data());
    for (int j = 0; j < k; ++j) M[0][(size_t)i * n + (left - k + 1 + j)] = bia[j];
  }
  // integration weights (labs_fill_full.cpp:206-256), nderiv = 0
  const int ngl = (k - 0 + 1) / 2;
  std::vector<double> glx, glw;
  gauss_legendre_half(ngl, glx, glw);
  integ.assign(n, 0.0);
  for (int i = 0; i < nbreak - 1; ++i) {
    const double a = w.t[k - 1 + i], b = w.t[k + i];
    for (int j = 0; j < ngl; ++j) {
      double xj, wj;
      gl_point(a, b, j, ngl, glx, glw, xj, wj);
      const int left = bspl_interval(w, xj);
      int jj = 0;
      bsplvb(w.t, k, 1, xj, left, jj, dl.data(), dr.data(), bia.data());
      const int ks = left - k + 1;
      for (int m = 0; m < k; ++m) integ[ks + m] += wj * bia[m];
    }
  }
  // centre-line values (labs_fill_full.cpp:146-149)
  center.assign(n, 0.0);
  {
    const int left = bspl_interval(w, 0.0);
    int jj = 0;
    bsplvb(w.t, k, 1, 0.0, left, jj, dl.data(), dr.data(), bia.data());
    for (int j = 0; j < k; ++j) center[left - k + 1 + j] = bia[j];
  }
  return 0;
}

// B-spline values at an arbitrary y (used for the opposition-control
// detection plane).  Returns the first nonzero index; vals gets k values.
static int bspl_row_at(int k, int n, const std::vector<double> &brk, double y, std::vector<double> &vals) {
  Bspl w;
  w.k = k;
  bspl_knots(w, brk);
  (void)n;
  const int left = bspl_interval(w, y);
  if (left < 0) return -1;
  std::vector<double> dl(k + 1), dr(k + 1);
  vals.assign(k, 0.0);
  int jj = 0;
  bsplvb(w.t, k, 1, y, left, jj, dl.data(), dr.data(), vals.data());
  return left - k + 1;
}

// ------------------------------------------------------------------- LABS
// LABS_eval_bw (module_labs.f90:245-293) on row-major matrices (row = collocation point).
static int labs_eval_bw(const std::vector<double> M[5], int n) {
  for (int k = 1;; ++k) {
    double err = 0.0;
    auto chk = [&](int r, int c) {  // 1-based matrix row r, column c
      for (int m = 0; m < 3; ++m) err = std::max(err, std::fabs(M[m][(size_t)(r - 1) * n + (c - 1)]));
    };
    for (int r = 1; r <= k + 1; ++r)
      for (int c = 2 * k + 2; c <= n; ++c) chk(r, c);
    for (int r = k + 2; r <= n - k - 1; ++r) {
      for (int c = 1; c <= r - k - 1; ++c) chk(r, c);
      for (int c = r + k + 1; c <= n; ++c) chk(r, c);
    }
    for (int r = n - k; r <= n; ++r)
      for (int c = 1; c <= n - 2 * k - 1; ++c) chk(r, c);
    if (err <= 0.0) return k;
    if (2 * k + 1 >= n) return -1;
  }
}

// Packed band geometry (LABS_full_to_band, module_labs.f90:301-328), 0-based.
struct Band {
  int n, bw, w;  // w = 2bw+1
  int first(int r) const {
    if (r < bw) return 0;
    if (r >= n - bw) return n - 1 - 2 * bw;
    return r - bw;
  }
  int last(int r) const {
    if (r < bw) return 2 * bw;
    if (r >= n - bw) return n - 1;
    return r + bw;
  }
  int slot(int r, int c) const {
    if (r < bw && c > r + bw) return 2 * bw - c;                 // extra right entries, top rows
    if (r >= n - bw && c < r - bw) return n - 1 - c;             // extra left entries, bottom rows
    return bw + (c - r);
  }
};

static void labs_full_to_band(const double *full, double *band, int bw, int n) {
  Band B{n, bw, 2 * bw + 1};
  std::fill(band, band + (size_t)n * B.w, 0.0);
  for (int r = 0; r < n; ++r)
    for (int c = B.first(r); c <= B.last(r); ++c) band[(size_t)r * B.w + B.slot(r, c)] = full[(size_t)r * n + c];
}

// No-pivot Crout LU in place (fixed_bandec*, module_labs.f90:943-1198, 3487-3636):
// L (with diagonal) below, unit U above; every entry's update subtracts the
// products of the nearest previous row first; U rows are scaled by the
// reciprocal of the pivot.
static void labs_lu(double *A, int n, int bw) {
  Band B{n, bw, 2 * bw + 1};
  auto at = [&](int r, int c) -> double & { return A[(size_t)r * B.w + B.slot(r, c)]; };
  for (int r = 0; r < n; ++r) {
    const int f = B.first(r);
    for (int c = f; c <= r; ++c) {
      double v = at(r, c);
      for (int p = c - 1; p >= f; --p) {
        if (c > B.last(p)) break;
        v = v - at(r, p) * at(p, c);
      }
      at(r, c) = v;
    }
    const double inv = 1.0 / at(r, r);
    for (int c = r + 1; c <= B.last(r); ++c) {
      double v = at(r, c);
      for (int p = r - 1; p >= f; --p) {
        if (c > B.last(p)) break;
        v = v - at(r, p) * at(p, c);
      }
      at(r, c) = inv * v;
    }
  }
}

// Forward/back substitution with complex RHS (fixed_banbks*, module_labs.f90:3769-3823, 4252-4318).
static void labs_bs(const double *A, int n, int bw, cd *b) {
  Band B{n, bw, 2 * bw + 1};
  auto at = [&](int r, int c) -> double { return A[(size_t)r * B.w + B.slot(r, c)]; };
  for (int r = 0; r < n; ++r) {
    cd v = b[r];
    for (int p = B.first(r); p < r; ++p) v = sub(v, mulr(b[p], at(r, p)));
    b[r] = divr(v, at(r, r));
  }
  for (int r = n - 1; r >= 0; --r) {
    cd v = b[r];
    for (int c = r + 1; c <= B.last(r); ++c) v = sub(v, mulr(b[c], at(r, c)));
    b[r] = v;
  }
}

// Band matvec b = A x, ascending column order (LABS_mat_mul_4, module_labs.f90:4395-4430).
static void labs_mv(const double *A, int n, int bw, const cd *x, cd *b) {
  Band B{n, bw, 2 * bw + 1};
  for (int r = 0; r < n; ++r) {
    const int f = B.first(r), l = B.last(r);
    cd v = mulr(x[f], A[(size_t)r * B.w + B.slot(r, f)]);
    for (int c = f + 1; c <= l; ++c) v = add(v, mulr(x[c], A[(size_t)r * B.w + B.slot(r, c)]));
    b[r] = v;
  }
}

// LABS_integration (module_labs.f90:401-413)
static cd labs_integrate(const std::vector<double> &w, const cd *src, int n) {
  cd t = C(0.0, 0.0);
  for (int i = 0; i < n; ++i) t = add(t, mulr(src[i], w[i]));
  return t;
}

// -------------------------------------------------------------------- FFT
// Unnormalised DFT  X[k] = sum_j x[j] exp(sign*2*pi*i*j*k/N)  by recursive
// mixed-radix decimation in time.
struct FFTPlan {
  int n = 0;
  std::vector<int> radix;       // n = prod radix
  std::vector<cd> tw_f, tw_b;   // exp(-2 pi i m/n), exp(+2 pi i m/n)
};

static FFTPlan make_plan(int n) {
  FFTPlan p;
  p.n = n;
  int m = n;
  for (int r : {4, 2, 3, 5, 7}) {
    while (m % r == 0) { p.radix.push_back(r); m /= r; }
  }
  for (int r = 11; m > 1; r += 2)
    while (m % r == 0) { p.radix.push_back(r); m /= r; }
  p.tw_f.resize(n);
  p.tw_b.resize(n);
  for (int j = 0; j < n; ++j) {
    const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)n;
    const double c = (double)cosl(a), s = (double)sinl(a);
    p.tw_f[j] = C(c, -s);
    p.tw_b[j] = C(c, s);
  }
  return p;
}

// out[0..n) = DFT of in[0, stride, 2*stride, ...]; level selects the radix.
static void fft_rec(const FFTPlan &p, const cd *in, cd *out, int n, int stride, int level, int sign, cd *scratch) {
  if (n == 1) { out[0] = in[0]; return; }
  const int r = p.radix[level];
  const int m = n / r;
  const std::vector<cd> &tw = sign < 0 ? p.tw_f : p.tw_b;
  const int tstep = p.n / n;  // twiddle index scale for this level
  // sub-transforms of the decimated sequences into out[q*m .. q*m+m)
  for (int q = 0; q < r; ++q) fft_rec(p, in + (size_t)q * stride, out + (size_t)q * m, m, stride * r, level + 1, sign, scratch);
  // combine: X[k + s*m] = sum_q W_n^{q(k+s m)} Y_q[k]
  for (int k = 0; k < m; ++k) {
    for (int q = 0; q < r; ++q) {
      const cd y = out[(size_t)q * m + k];
      scratch[q] = (q == 0) ? y : mul(y, tw[((size_t)q * k * tstep) % p.n]);
    }
    for (int s = 0; s < r; ++s) {
      cd acc = scratch[0];
      for (int q = 1; q < r; ++q) {
        const int e = (int)(((long)q * s * m * tstep) % p.n);
        acc = add(acc, mul(scratch[q], tw[e]));
      }
      scratch[r + s] = acc;
    }
    for (int s = 0; s < r; ++s) out[(size_t)s * m + k] = scratch[r + s];
  }
}

static void fft(const FFTPlan &p, const cd *in, cd *out, int sign) {
  std::vector<cd> scratch(64);
  fft_rec(p, in, out, p.n, 1, 0, sign, scratch.data());
}

// ------------------------------------------------------------ the context
struct Ctx {
  pb_config cfg;
  int nx, ny, nz, nxh, nx32, nz32, nh, bw, W, kgsl;
  double Lx, Ly, Lz, lmfxr, lmfzr, dxr, dzr, cflx, rescale;
  std::vector<double> y, brk, integ, center, dyr;
  std::vector<double> full[5];
  std::vector<double> d0, d1, d2, d0lu;   // packed bands (n x W)
  std::vector<int> kx, kz;
  double rk[4][3];
  // opposition control: B-spline rows at the two detection planes
  std::vector<double> opp_lo_row, opp_hi_row;  // detection rows r: v(y_d) = r . v
  // wall actuation (one complex value per (kx,kz), index j*nxh + k), added to
  // the wall v of both walls; empty: none
  std::vector<cd> act;
  FFTPlan px, pz;
};

static void rk3_coef(double c[4][3]) {  // module_runge_kutta.f90:34-48, c[row][substep]
  c[0][0] = 29.0 / 96.0;  c[0][1] = -3.0 / 40.0;  c[0][2] = 1.0 / 6.0;
  c[1][0] = 37.0 / 160.0; c[1][1] = 5.0 / 24.0;   c[1][2] = 1.0 / 6.0;
  c[2][0] = 8.0 / 15.0;   c[2][1] = 5.0 / 12.0;   c[2][2] = 3.0 / 4.0;
  c[3][0] = 0.0;          c[3][1] = -17.0 / 60.0; c[3][2] = -5.0 / 12.0;
}

// Index into a reference-layout array A(ny, nxh, nz, slot).
static inline size_t ix(const Ctx &c, int slot, int j, int k, int y) {
  return (((size_t)slot * c.nz + j) * c.nxh + k) * c.ny + y;
}

// --------------------------------------------------------- per-line solvers
struct Line {
  const Ctx *c;
  int n, bw;
  const double *D0lu, *D0, *D1, *D2;
  std::vector<double> Dbuf;
  double dt;
  int s;  // 1-based substep
  double coef(int row, int sub) const { return c->rk[row][sub - 1]; }
};

// hvhg.f90:141-276
static void hvhg(Line &L, cd *u1, cd *u2, cd *u3, cd *u4, cd *u5, cd *u7, cd *u8, cd *u9, double kx, double kz) {
  const int n = L.n, bw = L.bw;
  const double kz2 = kz * kz, kx2 = kx * kx, kxkz = kx * kz;
  const cd ikx = imag_unit_times(kx), ikz = imag_unit_times(kz);
  const cd ikx3 = imag_unit_times(kx * kx * kx);
  const cd ikxkz2 = imag_unit_times(kx * (kz * kz));
  const cd ikx2kz = imag_unit_times((kx * kx) * kz);
  const cd ikz3 = imag_unit_times(kz * kz * kz);
  for (int i = 0; i < n; ++i) {
    const cd dummy = u1[i];
    u1[i] = add(mulr(sub(u1[i], u3[i]), kxkz), mulr(u5[i], g_mutant == 1 ? kx2 - kz2 : kz2 - kx2));
    u3[i] = sub(sub(mulr(dummy, -kx2), mulr(u5[i], 2.0 * kxkz)), mulr(u3[i], kz2));
  }
  labs_bs(L.D0lu, n, bw, u3);
  labs_mv(L.D1, n, bw, u3, u5);
  const cd a = add(ikx3, ikxkz2), b = add(ikx2kz, ikz3);
  for (int i = 0; i < n; ++i) u5[i] = add(add(u5[i], g_mutant == 2 ? C(0.0, 0.0) : mul(a, u2[i])), mul(b, u4[i]));
  for (int i = 0; i < n; ++i) {
    const cd t3 = sub(mul(ikx, u4[i]), mul(ikz, u2[i]));
    const cd t4 = add(mul(ikz, u4[i]), mul(ikx, u2[i]));
    u3[i] = t3;
    u4[i] = t4;
  }
  labs_bs(L.D0lu, n, bw, u3);
  labs_mv(L.D1, n, bw, u3, u2);
  labs_bs(L.D0lu, n, bw, u4);
  labs_mv(L.D2, n, bw, u4, u3);
  for (int i = 0; i < n; ++i) {
    u5[i] = add(u5[i], u3[i]);
    u1[i] = add(u1[i], u2[i]);
  }
  for (int i = 0; i < n; ++i) u7[i] = add(mul(ikx, u7[i]), mul(ikz, u9[i]));
  labs_bs(L.D0lu, n, bw, u8);
  labs_mv(L.D1, n, bw, u8, u9);
  for (int i = 0; i < n; ++i) {
    const cd t = add(u7[i], u9[i]);
    u7[i] = C(-t.re, -t.im);
  }
}

// LHS of the Helmholtz problem with diagonal-only wall rows (solve_g.f90:42-71).
static void helmholtz_lhs(Line &L, double k2, double beta, bool mean) {
  const int n = L.n, W = 2 * L.bw + 1, m = L.bw;
  double *Db = L.Dbuf.data();
  const double dummy = g_mutant == 7 ? 1.0 - beta * k2 : 1.0 + beta * k2;
  for (int i = 0; i < W; ++i) Db[i] = 0.0;
  Db[m] = mean ? (L.D0[m] - beta * L.D2[m]) : (dummy * L.D0[m] - beta * L.D2[m]);
  for (int j = 1; j < n - 1; ++j)
    for (int i = 0; i < W; ++i) {
      const size_t q = (size_t)j * W + i;
      Db[q] = mean ? (L.D0[q] - beta * L.D2[q]) : (dummy * L.D0[q] - beta * L.D2[q]);
    }
  const size_t e = (size_t)(n - 1) * W;
  for (int i = 0; i < W; ++i) Db[e + i] = 0.0;
  Db[e + m] = mean ? (L.D0[e + m] - beta * L.D2[e + m]) : (dummy * L.D0[e + m] - beta * L.D2[e + m]);
  labs_lu(Db, n, L.bw);
}

// solve_g.f90:28-201 (nuv = nu) and solve_t.f90 (nuv = nu/pr); buf1, buf2 scratch.
static void solve_gt(Line &L, cd *u1, cd *store, cd *buf1, cd *buf2, double kx, double kz, double nu, double pr, bool scalar) {
  const int n = L.n, bw = L.bw, s = L.s, S = L.c->cfg.rk_step;
  const double k2 = kx * kx + kz * kz;
  const double dt = L.dt;
  auto visc = [&](double coefv) { return scalar ? dt * coefv * nu / pr : dt * coefv * nu; };
  const double beta = visc(L.coef(1, s));
  helmholtz_lhs(L, k2, beta, false);
  if (s == 1) {
    const double alpha = visc(L.coef(0, s));
    const double gamma = dt * L.coef(2, s);
    const double dummy = 1.0 - alpha * k2;
    for (int i = 0; i < n; ++i) buf1[i] = add(mulr(store[i], dummy), mulr(u1[i], gamma));
    labs_bs(L.D0lu, n, bw, store);
    labs_mv(L.D2, n, bw, store, buf2);
    for (int i = 0; i < n; ++i) buf1[i] = add(buf1[i], mulr(buf2[i], alpha));
  } else {
    const double gamma = dt * L.coef(2, s);
    for (int i = 0; i < n; ++i) buf1[i] = add(store[i], mulr(u1[i], gamma));
  }
  buf1[0] = C(0.0, 0.0);
  buf1[n - 1] = C(0.0, 0.0);
  if (s != S) {
    const double zeta = dt * L.coef(3, g_mutant == 4 ? s : s + 1);
    for (int i = 0; i < n; ++i) store[i] = mulr(u1[i], zeta);
  }
  labs_bs(L.Dbuf.data(), n, bw, buf1);
  labs_mv(L.D0, n, bw, buf1, u1);
  labs_mv(L.D2, n, bw, buf1, buf2);
  if (s != S) {
    const double alpha = visc(L.coef(0, s + 1));
    const double dummy = g_mutant == 3 ? 1.0 + alpha * k2 : 1.0 - alpha * k2;
    for (int i = 0; i < n; ++i) store[i] = add(add(store[i], mulr(u1[i], dummy)), mulr(buf2[i], alpha));
  } else {
    for (int i = 0; i < n; ++i) store[i] = u1[i];
  }
}

// solve_v_f.f90:38-375.  vwall_lo / vwall_hi are the wall values of v
// (0 in the reference, -alpha*v(y_d) with opposition control).
static void solve_v_f(Line &L, cd *v, cd *dv, cd *hv, cd *store_phi, cd *buf1, cd *buf2, cd *buf3,
                      double kx, double kz, double nu, cd vwall_lo, cd vwall_hi) {
  const int n = L.n, bw = L.bw, s = L.s, S = L.c->cfg.rk_step, W = 2 * bw + 1, m = bw;
  const double k2 = kx * kx + kz * kz;
  const double dt = L.dt;
  double *Db = L.Dbuf.data();
  if (s == 1) {
    const double alpha = dt * L.coef(0, s) * nu;
    const double gamma = dt * L.coef(2, s);
    const double dummy = 1.0 - alpha * k2;
    for (int i = 0; i < n; ++i) buf1[i] = add(mulr(store_phi[i], dummy), mulr(hv[i], gamma));
    labs_bs(L.D0lu, n, bw, store_phi);
    labs_mv(L.D2, n, bw, store_phi, buf2);
    for (int i = 0; i < n; ++i) buf1[i] = add(buf1[i], mulr(buf2[i], alpha));
  } else {
    const double gamma = dt * L.coef(2, s);
    for (int i = 0; i < n; ++i) buf1[i] = add(store_phi[i], mulr(hv[i], gamma));
  }
  buf1[0] = C(0.0, 0.0);
  buf1[n - 1] = C(0.0, 0.0);
  if (s != S) {
    const double zeta = dt * L.coef(3, s + 1);
    for (int i = 0; i < n; ++i) store_phi[i] = mulr(hv[i], zeta);
  }
  buf2[0] = C(Db[m], Db[m]);
  for (int i = 1; i < n; ++i) buf2[i] = C(0.0, 0.0);
  labs_bs(Db, n, bw, buf1);
  labs_bs(Db, n, bw, buf2);
  labs_mv(L.D0, n, bw, buf1, hv);
  labs_mv(L.D0, n, bw, buf2, buf3);
  labs_mv(L.D2, n, bw, buf1, dv);
  if (s != S) {
    const double alpha = dt * nu * L.coef(0, s + 1);
    const double dummy = 1.0 - alpha * k2;
    for (int i = 0; i < n; ++i) store_phi[i] = add(add(store_phi[i], mulr(hv[i], dummy)), mulr(dv[i], alpha));
  } else {
    for (int i = 0; i < n; ++i) store_phi[i] = hv[i];
  }
  // Poisson LHS [D2 - k2 D0] with D0-diagonal wall rows (solve_v_f.f90:205-230)
  for (int i = 0; i < W; ++i) Db[i] = 0.0;
  Db[m] = L.D0[m];
  for (int j = 1; j < n - 1; ++j)
    for (int i = 0; i < W; ++i) {
      const size_t q = (size_t)j * W + i;
      Db[q] = (g_mutant == 6 ? k2 : -k2) * L.D0[q] + L.D2[q];
    }
  const size_t e = (size_t)(n - 1) * W;
  for (int i = 0; i < W; ++i) Db[e + i] = 0.0;
  Db[e + m] = L.D0[e + m];
  labs_lu(Db, n, bw);
  hv[0] = vwall_lo;
  hv[n - 1] = vwall_hi;
  labs_bs(Db, n, bw, hv);
  for (int i = 0; i < n; ++i) buf1[i] = buf3[i];
  buf1[0] = C(0.0, 0.0);
  buf1[n - 1] = C(0.0, 0.0);
  labs_bs(Db, n, bw, buf1);
  labs_mv(L.D0, n, bw, hv, v);
  labs_mv(L.D1, n, bw, hv, dv);
  labs_mv(L.D1, n, bw, buf1, hv);
  const cd den = sub(mul(hv[0], hv[0]), mul(hv[n - 1], hv[n - 1]));
  const cd c1 = cdiv(sub(mul(hv[n - 1], dv[n - 1]), mul(hv[0], dv[0])), den);
  const cd c2 = cdiv(sub(mul(hv[0], dv[n - 1]), mul(hv[n - 1], dv[0])), den);
  for (int i = 0; i < n; ++i) dv[i] = sub(add(dv[i], mul(c1, hv[i])), mul(c2, hv[n - 1 - i]));
  labs_mv(L.D2, n, bw, buf2, hv);
  if (s != S) {
    const double alpha = dt * nu * L.coef(0, s + 1);
    const double dummy = 1.0 - alpha * k2;
    for (int i = 0; i < n; ++i) {
      // store_phi + C1*dummy*buf3(i) + C2*dummy*buf3(ny+1-i) + C1*alpha*hv(i) + C2*alpha*hv(ny+1-i)
      const cd c1d = mulr(c1, dummy), c2d = mulr(c2, dummy), c1a = mulr(c1, alpha), c2a = mulr(c2, alpha);
      cd acc = add(store_phi[i], mul(c1d, buf3[i]));
      acc = add(acc, mul(c2d, buf3[n - 1 - i]));
      acc = add(acc, mul(c1a, hv[i]));
      acc = add(acc, mul(c2a, hv[n - 1 - i]));
      store_phi[i] = acc;
    }
  } else {
    for (int i = 0; i < n; ++i) store_phi[i] = add(add(store_phi[i], mul(c1, buf3[i])), mul(c2, buf3[n - 1 - i]));
  }
  labs_mv(L.D0, n, bw, buf1, hv);
  for (int i = 0; i < n; ++i) v[i] = add(add(v[i], mul(c1, hv[i])), mul(c2, hv[g_mutant == 5 ? i : n - 1 - i]));
}

// solve_u_w.f90:18-42 (g -> u, f -> w in place)
static void solve_u_w(int n, cd *g, cd *f, double kx, double kz) {
  const cd ci = C(0.0, 1.0);
  const double den = kx * kx + kz * kz;
  // kxx = ci*kx/(kx*kx+kz*kz): complex*real then complex/real
  const cd kxx = divr(mul(ci, C(kx, 0.0)), den);
  const cd kzz = divr(mul(ci, C(kz, 0.0)), den);
  for (int i = 0; i < n; ++i) {
    const cd dxg = mul(kxx, g[i]), dzg = mul(kzz, g[i]);
    const cd dxf = mul(kxx, f[i]), dzf = mul(kzz, f[i]);
    g[i] = sub(dxf, dzg);
    f[i] = add(dxg, dzf);
  }
}

// Mixed-type scalar condition, mean line: -h_Theta = d(v theta)/dy - ks U
// (u9 holds d(v theta)/dy), ks = (nu/pr) / bulk_u.
static void mixed_source(cd *u9, const cd *usrc, int n, double ks) {
  for (int i = 0; i < n; ++i) u9[i] = sub(u9[i], mulr(usrc[i], ks));
}

struct MeanOut {
  double phi_u = 0, phi_t = 0, tau_w = 0, q_w = 0;
  bool wall = false;
};

// wave00.f90:41-685 and calculate_phi (704-714)
// usrc: with the mixed-type scalar condition, the mean u of the substep's
// input state (source (nu/pr) U / bulk_u in the Theta equation, no held
// bulk); nullptr for the reference's held-bulk uniform source.
static void wave00(Line &L, cd *u1, cd *u2, cd *u3, cd *u4, cd *u5, cd *u7, cd *u8, cd *u9,
                   cd *st1, cd *st2, cd *st3, cd *buf1, double nu, double pr, double bulk_u,
                   double bulk_w, double bulk_t, double u_gal, const cd *usrc, MeanOut &mo) {
  const int n = L.n, bw = L.bw, s = L.s, S = L.c->cfg.rk_step, W = 2 * bw + 1, m = bw;
  const double dt = L.dt;
  const std::vector<double> &integ = L.c->integ;
  const double *D1 = L.D1;
  auto calc_phi = [&](double bulk, cd I) { return (bulk - 0.5 * I.re) / (0.5 * I.im); };
  auto wall_d1 = [&](const cd *c) {
    return (((D1[m] * c[0].re + D1[m + 1] * c[1].re) - D1[(size_t)(n - 1) * W + m - 1] * c[n - 2].re) -
            D1[(size_t)(n - 1) * W + m] * c[n - 1].re);
  };
  // ---- U
  double beta = dt * L.coef(1, s) * nu;
  helmholtz_lhs(L, 0.0, beta, true);
  if (s == 1) {
    for (int i = 0; i < n; ++i) u1[i] = st1[i];
    labs_bs(L.D0lu, n, bw, u1);
    labs_bs(L.D0lu, n, bw, u2);
    labs_mv(L.D2, n, bw, u1, buf1);
    labs_mv(L.D1, n, bw, u2, u3);
    const double alpha = dt * L.coef(0, s) * nu;
    const double gamma = dt * L.coef(2, s);
    for (int i = 1; i < n; ++i) buf1[i] = sub(add(st1[i], mulr(buf1[i], alpha)), mulr(u3[i], g_mutant == 8 ? -gamma : gamma));
  } else {
    labs_bs(L.D0lu, n, bw, u2);
    labs_mv(L.D1, n, bw, u2, u3);
    const double gamma = dt * L.coef(2, s);
    for (int i = 1; i < n; ++i) buf1[i] = sub(st1[i], mulr(u3[i], g_mutant == 8 ? -gamma : gamma));
  }
  buf1[0] = C(0.0, 0.0);
  for (int i = 1; i < n - 1; ++i) buf1[i] = C(buf1[i].re + u_gal, 1.0);
  buf1[n - 1] = C(0.0, 0.0);
  labs_bs(L.Dbuf.data(), n, bw, buf1);
  double phi = calc_phi(bulk_u, labs_integrate(integ, buf1, n));
  for (int i = 0; i < n; ++i) buf1[i] = C(buf1[i].re + phi * buf1[i].im, 0.0);
  mo.phi_u = phi;
  if (s == S) {
    mo.tau_w = 0.5 * nu * wall_d1(buf1);
    mo.wall = true;
  }
  labs_mv(L.D0, n, bw, buf1, u1);
  for (int i = 0; i < n; ++i) u1[i] = C(u1[i].re - u_gal, u1[i].im);
  if (s != S) {
    labs_mv(L.D2, n, bw, buf1, st1);
    const double alpha = dt * L.coef(0, s + 1) * nu;
    const double zeta = dt * L.coef(3, s + 1);
    for (int i = 0; i < n; ++i) st1[i] = sub(add(mulr(st1[i], alpha), u1[i]), mulr(u3[i], zeta));
  } else {
    for (int i = 0; i < n; ++i) st1[i] = u1[i];
  }
  // ---- W
  if (s == 1) {
    for (int i = 0; i < n; ++i) u2[i] = st2[i];
    labs_bs(L.D0lu, n, bw, u2);
    labs_mv(L.D2, n, bw, u2, buf1);
    labs_bs(L.D0lu, n, bw, u4);
    labs_mv(L.D1, n, bw, u4, u5);
    const double alpha = dt * L.coef(0, s) * nu;
    const double gamma = dt * L.coef(2, s);
    for (int i = 1; i < n; ++i) buf1[i] = sub(add(st2[i], mulr(buf1[i], alpha)), mulr(u5[i], gamma));
  } else {
    labs_bs(L.D0lu, n, bw, u4);
    labs_mv(L.D1, n, bw, u4, u5);
    const double gamma = dt * L.coef(2, s);
    for (int i = 1; i < n; ++i) buf1[i] = sub(st2[i], mulr(u5[i], gamma));
  }
  buf1[0] = C(0.0, 0.0);
  for (int i = 1; i < n - 1; ++i) buf1[i] = C(buf1[i].re, 1.0);
  buf1[n - 1] = C(0.0, 0.0);
  labs_bs(L.Dbuf.data(), n, bw, buf1);
  phi = calc_phi(bulk_w, labs_integrate(integ, buf1, n));
  phi = 0.0;
  for (int i = 0; i < n; ++i) buf1[i] = C(buf1[i].re + phi * buf1[i].im, 0.0);
  labs_mv(L.D0, n, bw, buf1, u3);
  if (s != S) {
    labs_mv(L.D2, n, bw, buf1, st2);
    const double alpha = dt * L.coef(0, s + 1) * nu;
    const double zeta = dt * L.coef(3, s + 1);
    for (int i = 0; i < n; ++i) st2[i] = sub(add(mulr(st2[i], alpha), u3[i]), mulr(u5[i], zeta));
  } else {
    for (int i = 0; i < n; ++i) st2[i] = u3[i];
  }
  for (int i = 0; i < n; ++i) u2[i] = C(0.0, 0.0);
  // ---- Theta
  beta = dt * L.coef(1, s) * nu / pr;
  helmholtz_lhs(L, 0.0, beta, true);
  if (s == 1) {
    for (int i = 0; i < n; ++i) u7[i] = st3[i];
    labs_bs(L.D0lu, n, bw, u7);
    labs_mv(L.D2, n, bw, u7, buf1);
    labs_bs(L.D0lu, n, bw, u8);
    labs_mv(L.D1, n, bw, u8, u9);
    if (usrc) mixed_source(u9, usrc, n, (g_mutant == 9 ? -1.0 : 1.0) * nu / pr / bulk_u);
    const double alpha = dt * L.coef(0, s) * nu / pr;
    const double gamma = dt * L.coef(2, s);
    for (int i = 1; i < n; ++i) buf1[i] = sub(add(st3[i], mulr(buf1[i], alpha)), mulr(u9[i], gamma));
  } else {
    labs_bs(L.D0lu, n, bw, u8);
    labs_mv(L.D1, n, bw, u8, u9);
    if (usrc) mixed_source(u9, usrc, n, (g_mutant == 9 ? -1.0 : 1.0) * nu / pr / bulk_u);
    const double gamma = dt * L.coef(2, s);
    for (int i = 1; i < n; ++i) buf1[i] = sub(st3[i], mulr(u9[i], gamma));
  }
  buf1[0] = C(0.0, 0.0);
  for (int i = 1; i < n - 1; ++i) buf1[i] = C(buf1[i].re, usrc ? 0.0 : 1.0);
  buf1[n - 1] = C(0.0, 0.0);
  labs_bs(L.Dbuf.data(), n, bw, buf1);
  phi = usrc ? 0.0 : calc_phi(bulk_t, labs_integrate(integ, buf1, n));
  for (int i = 0; i < n; ++i) buf1[i] = C(buf1[i].re + phi * buf1[i].im, 0.0);
  mo.phi_t = phi;
  if (s == S) mo.q_w = 0.5 * nu / pr * wall_d1(buf1);
  labs_mv(L.D0, n, bw, buf1, u7);
  if (s != S) {
    labs_mv(L.D2, n, bw, buf1, st3);
    const double alpha = dt * L.coef(0, s + 1) * nu / pr;
    const double zeta = dt * L.coef(3, s + 1);
    for (int i = 0; i < n; ++i) st3[i] = sub(add(mulr(st3[i], alpha), u7[i]), mulr(u9[i], zeta));
  } else {
    for (int i = 0; i < n; ++i) st3[i] = u7[i];
  }
}

// B-spline interpolation of a collocation-value column at a detection plane:
// c = LABS_bs(D0_LU, v); v(y_d) = sum_j B_j(y_d) c_j (ascending j).
// Opposition-control detection row (the project's definition; the
// reference's hooks are commented out, solve_v_f.f90:253-257):
// v(y_d) = B(y_d) . D0^-1 v = r . v with r = (LU(D0))^-T b, b = the basis
// values at y_d in their columns.  D0 = L U (Crout: L with its diagonal,
// unit U), so U^T w = b forward (rows of U that reach the column, ascending),
// then L^T r = w backward (rows whose band starts at or before the column,
// ascending), each sum subtracted term by term.
static int detection_row(const Ctx &c, double yd, std::vector<double> &r) {
  std::vector<double> vals;
  const int first = bspl_row_at(c.kgsl, c.ny, c.brk, yd, vals);
  if (first < 0) return -1;
  const int n = c.ny, bw = c.bw;
  Band B{n, bw, 2 * bw + 1};
  auto at = [&](int row, int col) { return c.d0lu[(size_t)row * B.w + B.slot(row, col)]; };
  std::vector<double> b(n, 0.0), w(n);
  for (size_t j = 0; j < vals.size(); ++j) b[first + j] = vals[j];
  for (int col = 0; col < n; ++col) {
    double v = b[col];
    for (int row = std::max(0, col - 2 * bw); row < col; ++row)
      if (col <= B.last(row)) v = v - at(row, col) * w[row];
    w[col] = v;
  }
  r.assign(n, 0.0);
  for (int p = n - 1; p >= 0; --p) {
    double v = w[p];
    for (int q = p + 1; q <= std::min(n - 1, p + 2 * bw); ++q)
      if (B.first(q) <= p) v = v - at(q, p) * r[q];
    r[p] = v / at(p, p);
  }
  return 0;
}
static cd detect_dot(const Ctx &c, const cd *v, const std::vector<double> &r) {
  cd acc = mulr(v[0], r[0]);
  for (int i = 1; i < c.ny; ++i) acc = add(acc, mulr(v[i], r[i]));
  return acc;
}

// solve_ppe (get_pressure.f90:55-133) for one (kx,kz) != (0,0) line:
// u1..u5 = u^2-v^2, uv, w^2-v^2, vw, uw are read; u6 = v^2 on entry, p on
// exit; bc1, bc2 = the Neumann values of the two walls.
static void solve_ppe(Line &L, const cd *u1, const cd *u2, const cd *u3, const cd *u4, const cd *u5, cd *u6,
                      cd bc1, cd bc2, cd *buf1, cd *buf2, double kx, double kz) {
  const int n = L.n, bw = L.bw, W = 2 * bw + 1;
  const cd ikx = C(0.0, kx), ikz = C(0.0, kz);  // dcmplx(0d0, k)
  const double k2 = kx * kx + kz * kz, kxkz = kx * kz;
  for (int i = 1; i < n - 1; ++i)
    buf1[i] = add(add(add(mulr(u1[i], kx * kx), mulr(u3[i], kz * kz)), mulr(u6[i], k2)), mulr(u5[i], 2.0 * kxkz));
  labs_bs(L.D0lu, n, bw, u6);
  labs_mv(L.D2, n, bw, u6, buf2);
  for (int i = 1; i < n - 1; ++i) {
    buf1[i] = sub(buf1[i], buf2[i]);
    u6[i] = add(mul(ikx, u2[i]), mul(ikz, u4[i]));  // rows 0 and n-1 keep the coefficients above
  }
  labs_bs(L.D0lu, n, bw, u6);
  labs_mv(L.D1, n, bw, u6, buf2);
  buf1[0] = bc1;
  for (int i = 1; i < n - 1; ++i) buf1[i] = sub(buf1[i], mulr(buf2[i], 2.0));
  buf1[n - 1] = bc2;
  double *Db = L.Dbuf.data();
  for (int i = 0; i < W; ++i) Db[i] = L.D1[i];
  for (int j = 1; j < n - 1; ++j)
    for (int i = 0; i < W; ++i) {
      const size_t q = (size_t)j * W + i;
      Db[q] = -(k2 * L.D0[q]) + L.D2[q];
    }
  const size_t e = (size_t)(n - 1) * W;
  for (int i = 0; i < W; ++i) Db[e + i] = L.D1[e + i];
  labs_lu(Db, n, bw);
  labs_bs(Db, n, bw, buf1);
  labs_mv(L.D0, n, bw, buf1, u6);
}

// Wall values of the pressure Poisson equation from v
// (get_statistics.f90:689-716): c = LU(D0)^-1 v,
// bc1 = nu*(D2 row 0 . c_0..2), bc2 = nu*(D2 row n-1 . c_n-3..n-1).
static void ppe_wall_values(const Ctx &c, const cd *v, cd *bc1, cd *bc2) {
  const int n = c.ny, bw = c.bw, W = 2 * bw + 1;
  std::vector<cd> t(v, v + n);
  labs_bs(c.d0lu.data(), n, bw, t.data());
  const double *D2 = c.d2.data();
  const double nu = c.cfg.nu;
  *bc1 = mulr(add(add(mulr(t[0], D2[bw]), mulr(t[1], D2[bw + 1])), mulr(t[2], D2[bw + 2])), nu);
  const size_t e = (size_t)(n - 1) * W;
  *bc2 = mulr(add(add(mulr(t[n - 3], D2[e + bw - 2]), mulr(t[n - 2], D2[e + bw - 1])), mulr(t[n - 1], D2[e + bw])),
              nu);
}

}  // namespace

// ======================================================================= ABI
extern "C" {

// LABS setup as a standalone call (for tests).  full: 5 row-major ny x ny.
int orc_labs_setup(int ny, int labs_k, int knot_map, double stretch, double ly, double *y,
                   double *brk, double *integ, double *center, double *full5, int *bw_out) {
  std::vector<double> M[5], yc, b, in, ce;
  const int r = labs_fill_full(labs_k + 1, ny, knot_map, stretch, ly, M, yc, b, in, ce);
  if (r) return r;
  if (y) std::copy(yc.begin(), yc.end(), y);
  if (brk) std::copy(b.begin(), b.end(), brk);
  if (integ) std::copy(in.begin(), in.end(), integ);
  if (center) std::copy(ce.begin(), ce.end(), center);
  if (full5)
    for (int m = 0; m < 5; ++m) std::copy(M[m].begin(), M[m].end(), full5 + (size_t)m * ny * ny);
  if (bw_out) *bw_out = labs_eval_bw(M, ny);
  return 0;
}

int orc_bspline_row(int ny, int labs_k, int knot_map, double stretch, double ly, double yv, double *vals,
                    int *first) {
  std::vector<double> M[5], yc, b, in, ce;
  const int r = labs_fill_full(labs_k + 1, ny, knot_map, stretch, ly, M, yc, b, in, ce);
  if (r) return r;
  std::vector<double> v;
  const int f = bspl_row_at(labs_k + 1, ny, b, yv, v);
  if (f < 0) return -4;
  std::copy(v.begin(), v.end(), vals);
  *first = f;
  return 0;
}

void orc_full_to_band(const double *full, double *band, int bw, int n) { labs_full_to_band(full, band, bw, n); }
void orc_band_lu(double *A, int n, int bw) { labs_lu(A, n, bw); }
void orc_band_solve(const double *A, int n, int bw, double *b) { labs_bs(A, n, bw, reinterpret_cast<cd *>(b)); }
void orc_band_matvec(const double *A, int n, int bw, const double *x, double *b) {
  labs_mv(A, n, bw, reinterpret_cast<const cd *>(x), reinterpret_cast<cd *>(b));
}

// Unnormalised complex DFT; sign -1 forward (FFTW_FORWARD), +1 backward.
void orc_fft(const double *in, double *out, int n, int sign) {
  FFTPlan p = make_plan(n);
  fft(p, reinterpret_cast<const cd *>(in), reinterpret_cast<cd *>(out), sign);
}

void *orc_create(const pb_config *cfg) {
  Ctx *c = new Ctx();
  c->cfg = *cfg;
  c->nx = (int)cfg->nx; c->ny = (int)cfg->ny; c->nz = (int)cfg->nz;
  c->nxh = c->nx / 2; c->nx32 = c->nx * 3 / 2; c->nz32 = c->nz * 3 / 2; c->nh = c->nz / 2;
  c->kgsl = cfg->labs_k + 1;
  const double pi = 3.14159265358979323;  // module_global_variable.f90:18
  c->Lx = cfg->lmfx * 2 * pi;
  c->Ly = cfg->lmfy * 2;
  c->Lz = cfg->lmfz * 2 * pi;
  c->rescale = 1.0 / (double)(c->nx32 * c->nz32);
  c->dxr = (double)c->nx / c->Lx;
  c->dzr = (double)c->nz / c->Lz;
  c->cflx = cfg->cfl / pi;
  c->lmfxr = 1.0 / cfg->lmfx;
  c->lmfzr = 1.0 / cfg->lmfz;
  if (labs_fill_full(c->kgsl, c->ny, cfg->knot_map, cfg->stretch, c->Ly, c->full, c->y, c->brk, c->integ,
                     c->center)) {
    delete c;
    return nullptr;
  }
  c->bw = labs_eval_bw(c->full, c->ny);
  if (c->bw < 1) { delete c; return nullptr; }
  c->W = 2 * c->bw + 1;
  const size_t bs = (size_t)c->ny * c->W;
  c->d0.resize(bs); c->d1.resize(bs); c->d2.resize(bs);
  labs_full_to_band(c->full[0].data(), c->d0.data(), c->bw, c->ny);
  labs_full_to_band(c->full[1].data(), c->d1.data(), c->bw, c->ny);
  labs_full_to_band(c->full[2].data(), c->d2.data(), c->bw, c->ny);
  c->d0lu = c->d0;
  labs_lu(c->d0lu.data(), c->ny, c->bw);
  rk3_coef(c->rk);
  // wavenumber tables (module_ppdfftw.f90:347-384)
  c->kx.resize(c->nxh);
  for (int i = 0; i < c->nxh; ++i) c->kx[i] = i;
  c->kz.resize(c->nz);
  for (int i = 0; i <= c->nh; ++i) c->kz[i] = i;
  for (int i = c->nh + 1; i < c->nz; ++i) c->kz[i] = i + 1 - 2 * c->nh - 1;
  // dyr (module_ppdfftw.f90:572-583)
  c->dyr.resize(c->ny);
  for (int i = 0; i < c->ny; ++i) {
    double d;
    if (i == 0) d = std::fabs(c->y[1] - c->y[0]);
    else if (i == c->ny - 1) d = std::fabs(c->y[c->ny - 1] - c->y[c->ny - 2]);
    else d = std::min(std::fabs(c->y[i] - c->y[i - 1]), std::fabs(c->y[i] - c->y[i + 1]));
    c->dyr[i] = 1.0 / (cfg->dy_factor * d);
  }
  // detection rows whenever the plane lies inside the channel, so that a gain
  // set later (orc_set_control) has them; a nonzero gain requires them
  if (cfg->opp_alpha != 0.0 || (cfg->opp_ydist > 0.0 && cfg->opp_ydist < 1.0)) {
    if (detection_row(*c, -1.0 + cfg->opp_ydist, c->opp_lo_row) ||
        detection_row(*c, 1.0 - cfg->opp_ydist, c->opp_hi_row)) {
      if (cfg->opp_alpha != 0.0) { delete c; return nullptr; }
      c->opp_lo_row.clear();
      c->opp_hi_row.clear();
    }
  }
  c->px = make_plan(c->nx32);
  c->pz = make_plan(c->nz32);
  return c;
}

void orc_destroy(void *h) { delete static_cast<Ctx *>(h); }

int orc_info(void *h, int *bw, double *y, double *d0, double *d1, double *d2, double *d0lu, double *integ,
             double *dyr) {
  Ctx *c = static_cast<Ctx *>(h);
  if (bw) *bw = c->bw;
  const size_t bs = (size_t)c->ny * c->W;
  if (y) std::copy(c->y.begin(), c->y.end(), y);
  if (d0) std::copy(c->d0.begin(), c->d0.begin() + bs, d0);
  if (d1) std::copy(c->d1.begin(), c->d1.begin() + bs, d1);
  if (d2) std::copy(c->d2.begin(), c->d2.begin() + bs, d2);
  if (d0lu) std::copy(c->d0lu.begin(), c->d0lu.begin() + bs, d0lu);
  if (integ) std::copy(c->integ.begin(), c->integ.end(), integ);
  if (dyr) std::copy(c->dyr.begin(), c->dyr.end(), dyr);
  return 0;
}

// set_initial_data_uvw (initial_step.f90:139-179): state U(ny,nxh,nz,4)
// slots u,v,w,theta -> RK(ny,nxh,nz,3).
int orc_init_stores(void *h, const double *Ud, double *RKd) {
  Ctx *c = static_cast<Ctx *>(h);
  const cd *U = reinterpret_cast<const cd *>(Ud);
  cd *RK = reinterpret_cast<cd *>(RKd);
  const int n = c->ny;
#pragma omp parallel for schedule(static)
  for (int l = 0; l < c->nxh * c->nz; ++l) {
    const int k = l % c->nxh, j = l / c->nxh;
    std::vector<cd> d(n);
    const cd *u = U + ix(*c, 0, j, k, 0), *v = U + ix(*c, 1, j, k, 0), *w = U + ix(*c, 2, j, k, 0),
             *th = U + ix(*c, 3, j, k, 0);
    cd *r1 = RK + ix(*c, 0, j, k, 0), *r2 = RK + ix(*c, 1, j, k, 0), *r3 = RK + ix(*c, 2, j, k, 0);
    if (c->kx[k] == 0 && c->kz[j] == 0) {
      for (int i = 0; i < n; ++i) { r1[i] = u[i]; r2[i] = w[i]; r3[i] = th[i]; }
    } else {
      for (int i = 0; i < n; ++i) d[i] = v[i];
      labs_bs(c->d0lu.data(), n, c->bw, d.data());
      labs_mv(c->d2.data(), n, c->bw, d.data(), r2);
      const double kxp = c->kx[k] * c->lmfxr, kzp = c->kz[j] * c->lmfzr;
      const double k2 = kxp * kxp + kzp * kzp;
      const cd ikz = C(0.0, kzp), ikx = C(0.0, kxp);
      for (int i = 0; i < n; ++i) {
        r2[i] = sub(r2[i], mulr(v[i], k2));
        r1[i] = sub(mul(ikz, u[i]), mul(ikx, w[i]));
      }
      for (int i = 0; i < n; ++i) r3[i] = th[i];
    }
  }
  return 0;
}

// Restart from v, g, theta and the mean lines (set_initial_data_vg,
// initial_step.f90:301-353).  In: S(ny,nxh,nz,4) slot 1 = v; RK slot 0 = g
// (mean U - u_gal on kx = kz = 0), slot 1 = W mean on kx = kz = 0, slot 2 =
// theta (mean Theta on kx = kz = 0).  Out: S slots 0, 2, 3 = u, w, theta,
// RK slot 1 = phi = (D2 LU(D0)^-1 - k^2) v off the mean line.
int orc_init_from_vg(void *h, double *Sd, double *RKd) {
  Ctx *c = static_cast<Ctx *>(h);
  cd *S = reinterpret_cast<cd *>(Sd);
  cd *RK = reinterpret_cast<cd *>(RKd);
  const int n = c->ny;
#pragma omp parallel for schedule(static)
  for (int l = 0; l < c->nxh * c->nz; ++l) {
    const int k = l % c->nxh, j = l / c->nxh;
    cd *u = S + ix(*c, 0, j, k, 0), *v = S + ix(*c, 1, j, k, 0), *w = S + ix(*c, 2, j, k, 0),
       *th = S + ix(*c, 3, j, k, 0);
    cd *r1 = RK + ix(*c, 0, j, k, 0), *r2 = RK + ix(*c, 1, j, k, 0), *r3 = RK + ix(*c, 2, j, k, 0);
    if (c->kx[k] == 0 && c->kz[j] == 0) {
      for (int i = 0; i < n; ++i) { u[i] = r1[i]; v[i] = C(0.0, 0.0); w[i] = r2[i]; th[i] = r3[i]; }
      continue;
    }
    std::vector<cd> d(v, v + n), f(n);
    labs_bs(c->d0lu.data(), n, c->bw, d.data());
    labs_mv(c->d1.data(), n, c->bw, d.data(), f.data());
    labs_mv(c->d2.data(), n, c->bw, d.data(), r2);
    const double kxp = c->kx[k] * c->lmfxr, kzp = c->kz[j] * c->lmfzr;
    const double k2 = kxp * kxp + kzp * kzp;
    for (int i = 0; i < n; ++i) r2[i] = sub(r2[i], mulr(v[i], k2));
    for (int i = 0; i < n; ++i) { u[i] = r1[i]; w[i] = f[i]; }
    solve_u_w(n, u, w, kxp, kzp);  // g -> u, f -> w
    for (int i = 0; i < n; ++i) th[i] = r3[i];
  }
  return 0;
}

// Restart from g, phi, theta and the mean lines (set_initial_data_gphi,
// initial_step.f90:430-513).  In: RK slots 0, 1, 2 = g, phi, theta (the mean
// lines on kx = kz = 0 as in orc_init_from_vg).  Out: S = u, v, w, theta:
// v from [D2 - k^2 D0] c = phi with the operator's own diagonal on the wall
// rows and zero wall values, v = D0 c, f = D1 c, u and w from g and f.
int orc_init_from_gphi(void *h, double *Sd, double *RKd) {
  Ctx *c = static_cast<Ctx *>(h);
  cd *S = reinterpret_cast<cd *>(Sd);
  const cd *RK = reinterpret_cast<const cd *>(RKd);
  const int n = c->ny, W = c->W, m = c->bw;
#pragma omp parallel for schedule(static)
  for (int l = 0; l < c->nxh * c->nz; ++l) {
    const int k = l % c->nxh, j = l / c->nxh;
    cd *u = S + ix(*c, 0, j, k, 0), *v = S + ix(*c, 1, j, k, 0), *w = S + ix(*c, 2, j, k, 0),
       *th = S + ix(*c, 3, j, k, 0);
    const cd *r1 = RK + ix(*c, 0, j, k, 0), *r2 = RK + ix(*c, 1, j, k, 0), *r3 = RK + ix(*c, 2, j, k, 0);
    if (c->kx[k] == 0 && c->kz[j] == 0) {
      for (int i = 0; i < n; ++i) { u[i] = r1[i]; v[i] = C(0.0, 0.0); w[i] = r2[i]; th[i] = r3[i]; }
      continue;
    }
    const double kxp = c->kx[k] * c->lmfxr, kzp = c->kz[j] * c->lmfzr;
    const double k2 = kzp * kzp + kxp * kxp;
    std::vector<double> Db((size_t)n * W, 0.0);
    Db[m] = -k2 * c->d0[m] + c->d2[m];
    for (int r = 1; r < n - 1; ++r)
      for (int q = 0; q < W; ++q) Db[(size_t)r * W + q] = -k2 * c->d0[(size_t)r * W + q] + c->d2[(size_t)r * W + q];
    const size_t e = (size_t)(n - 1) * W;
    Db[e + m] = -k2 * c->d0[e + m] + c->d2[e + m];
    labs_lu(Db.data(), n, c->bw);
    std::vector<cd> cc(r2, r2 + n), f(n);
    cc[0] = C(0.0, 0.0);
    cc[n - 1] = C(0.0, 0.0);
    labs_bs(Db.data(), n, c->bw, cc.data());
    labs_mv(c->d0.data(), n, c->bw, cc.data(), v);
    labs_mv(c->d1.data(), n, c->bw, cc.data(), f.data());
    for (int i = 0; i < n; ++i) { u[i] = r1[i]; w[i] = f[i]; }
    solve_u_w(n, u, w, kxp, kzp);
    for (int i = 0; i < n; ++i) th[i] = r3[i];
  }
  return 0;
}

// Strided view of a field set: element (slot, kz index j, kx index k, y).
struct View {
  cd *p;
  size_t sl, j, k, y;
  cd &operator()(int f, int jj, int kk, int yy) const { return p[f * sl + jj * j + kk * k + yy * y]; }
};

// Nonlinear part of the substep for one y plane: y_to_x (x4),
// dft_x_realspace and x_to_y (9 or 10 fields) and rescale_data.
// Statistics-time moments of the physical velocities at one x point
// (dft_x_realspace.f90:128-152), in the order of its accumulators 3..21:
// u^3 v^3 w^3, u^4 v^4 w^4, -u^2 v, -v w^2, -v^3, -u v^2, u^6 v^6 w^6,
// u^8 v^8 w^8, u^10 v^10 w^10 (raw moments, the means included).
constexpr int kNMom = 19;
static inline void add_moments(double *acc, double u, double v, double w) {
  const double u2 = u * u, v2 = v * v, w2 = w * w;
  acc[0] += u * u * u; acc[1] += v * v * v; acc[2] += w * w * w;
  acc[3] += u * u * u * u; acc[4] += v * v * v * v; acc[5] += w * w * w * w;
  acc[6] -= u * u * v; acc[7] -= v * w * w; acc[8] -= v * v * v; acc[9] -= u * v * v;
  const double u4 = u2 * u2, v4 = v2 * v2, w4 = w2 * w2;
  acc[10] += u4 * u2; acc[11] += v4 * v2; acc[12] += w4 * w2;
  acc[13] += u4 * u4; acc[14] += v4 * v4; acc[15] += w4 * w4;
  acc[16] += u4 * u4 * u2; acc[17] += v4 * v4 * v2; acc[18] += w4 * w4 * w2;
}

static double nonlinear_plane(const Ctx *c, const View &S, const View &P, int want_vv, int y,
                              std::vector<cd> &G, std::vector<cd> &H, double *mom = nullptr) {
  const int nxh = c->nxh, nz32 = c->nz32, nx32 = c->nx32, nh = c->nh;
  std::vector<cd> col(nz32), colo(nz32), xs(nx32), xo(nx32);
  std::vector<double> phys((size_t)4 * nx32), out((size_t)10 * nx32);
  double vmax = 0.0;
  // z inverse with 3/2 padding (module_ppdfftw.f90:830-861)
  for (int f = 0; f < 4; ++f)
    for (int k = 0; k < nxh; ++k) {
      for (int jj = 0; jj < nh; ++jj) col[jj] = S(f, jj, k, y);
      for (int jj = nh; jj <= 2 * nh; ++jj) col[jj] = C(0.0, 0.0);
      for (int jj = 2 * nh + 1; jj < nz32; ++jj) col[jj] = S(f, jj - nh, k, y);
      fft(c->pz, col.data(), colo.data(), +1);
      for (int z = 0; z < nz32; ++z) G[((size_t)f * nz32 + z) * nxh + k] = colo[z];
    }
  for (int z = 0; z < nz32; ++z) {
    // c2r with zero padding nxh -> nx32/2+1 (dft_x_realspace.f90:72-105)
    for (int f = 0; f < 4; ++f) {
      for (int m = 0; m < nx32; ++m) xs[m] = C(0.0, 0.0);
      const cd *g = &G[((size_t)f * nz32 + z) * nxh];
      xs[0] = C(g[0].re, 0.0);  // FFTW c2r ignores Im of the DC bin
      for (int m = 1; m < nxh; ++m) {
        xs[m] = g[m];
        xs[nx32 - m] = C(g[m].re, -g[m].im);
      }
      fft(c->px, xs.data(), xo.data(), +1);
      for (int m = 0; m < nx32; ++m) phys[(size_t)f * nx32 + m] = xo[m].re;
    }
    // products (dft_x_realspace.f90:109-127), 0-based slots
    const double *p = phys.data();
    for (int m = 0; m < nx32; ++m) {
      const double t1 = p[m], t2 = p[nx32 + m], t3 = p[2 * nx32 + m], t4 = p[3 * nx32 + m];
      const double bv2 = t2 * t2;
      out[(size_t)8 * nx32 + m] = t3 * t4;
      out[(size_t)7 * nx32 + m] = t2 * t4;
      out[(size_t)6 * nx32 + m] = t1 * t4;
      out[(size_t)4 * nx32 + m] = t1 * t3;
      out[(size_t)3 * nx32 + m] = t2 * t3;
      out[(size_t)2 * nx32 + m] = t3 * t3 - bv2;
      out[(size_t)1 * nx32 + m] = t1 * t2;
      out[(size_t)0 * nx32 + m] = t1 * t1 - bv2;
      out[(size_t)9 * nx32 + m] = t2;
      out[(size_t)5 * nx32 + m] = bv2;
      const double vel = (std::fabs(t1) * c->dxr) + (std::fabs(t2) * c->dyr[y]) + (std::fabs(t3) * c->dzr);
      vmax = std::max(vel, vmax);
      if (mom) add_moments(mom, t1, t2, t3);
    }
    // r2c, keep 0..nxh-1
    for (int f = 0; f < 10; ++f) {
      if (f == 5 && !want_vv) continue;
      for (int m = 0; m < nx32; ++m) xs[m] = C(out[(size_t)f * nx32 + m], 0.0);
      fft(c->px, xs.data(), xo.data(), -1);
      for (int m = 0; m < nxh; ++m) H[((size_t)f * nz32 + z) * nxh + m] = xo[m];
    }
  }
  // z forward with truncation, then rescale (module_ppdfftw.f90:794-823; rescale_data.f90)
  for (int f = 0; f < 10; ++f) {
    if (f == 5 && !want_vv) continue;
    for (int k = 0; k < nxh; ++k) {
      for (int z = 0; z < nz32; ++z) col[z] = H[((size_t)f * nz32 + z) * nxh + k];
      fft(c->pz, col.data(), colo.data(), -1);
      for (int jj = 0; jj < nh; ++jj) P(f, jj, k, y) = mulr(colo[jj], c->rescale);
      P(f, nh, k, y) = mulr(C(0.0, 0.0), c->rescale);
      for (int jj = 2 * nh + 1; jj < nz32; ++jj) P(f, jj - nh, k, y) = mulr(colo[jj], c->rescale);
    }
  }
  return vmax;
}

// Nonlinear part for the y planes [y0, y1) of reference-layout arrays
// S(ny,nxh,nz,4) -> P(ny,nxh,nz,10); velmax = max over the range.
int orc_nonlinear(void *h, const double *Sd, double *Pd, int want_vv, int y0, int y1, double *velmax) {
  Ctx *c = static_cast<Ctx *>(h);
  const size_t ny = c->ny, nxh = c->nxh, nz = c->nz;
  View S{reinterpret_cast<cd *>(const_cast<double *>(Sd)), nz * nxh * ny, nxh * ny, ny, 1};
  View P{reinterpret_cast<cd *>(Pd), nz * nxh * ny, nxh * ny, ny, 1};
  double vmax = 0.0;
#pragma omp parallel reduction(max : vmax)
  {
    std::vector<cd> G((size_t)4 * c->nz32 * nxh), H((size_t)10 * c->nz32 * nxh);
#pragma omp for schedule(dynamic, 1)
    for (int y = y0; y < y1; ++y) vmax = std::max(vmax, nonlinear_plane(c, S, P, want_vv, y, G, H));
  }
  if (velmax) *velmax = vmax;
  return 0;
}

// orc_nonlinear with the statistics-time moments: mom(19, ny) receives the
// per-plane sums of add_moments over the padded physical grid (unscaled).
int orc_nonlinear_mom(void *h, const double *Sd, double *Pd, int want_vv, double *velmax, double *mom) {
  Ctx *c = static_cast<Ctx *>(h);
  const size_t ny = c->ny, nxh = c->nxh, nz = c->nz;
  View S{reinterpret_cast<cd *>(const_cast<double *>(Sd)), nz * nxh * ny, nxh * ny, ny, 1};
  View P{reinterpret_cast<cd *>(Pd), nz * nxh * ny, nxh * ny, ny, 1};
  double vmax = 0.0;
#pragma omp parallel reduction(max : vmax)
  {
    std::vector<cd> G((size_t)4 * c->nz32 * nxh), H((size_t)10 * c->nz32 * nxh);
    std::vector<double> acc(kNMom);
#pragma omp for schedule(dynamic, 1)
    for (int y = 0; y < (int)ny; ++y) {
      std::fill(acc.begin(), acc.end(), 0.0);
      vmax = std::max(vmax, nonlinear_plane(c, S, P, want_vv, y, G, H, acc.data()));
      for (int q = 0; q < kNMom; ++q) mom[(size_t)q * ny + y] = acc[q];
    }
  }
  if (velmax) *velmax = vmax;
  return 0;
}

// get_statistics_2 (get_statistics.f90:376-539 with the moments of
// dft_x_realspace.f90:128-484).  S(ny,nxh,nz,4): the substep's input state;
// P(ny,nxh,nz,10): its products with p in slot 5 (after orc_pressure); mom:
// orc_nonlinear_mom's sums.  out(29, ny), the datasets in the reference's
// order: uuu vvv www, uuuu vvvv wwww, u6 v6 w6, u8 v8 w8, u10 v10 w10
// (rescale * sum, zero on the wall rows), -duuvdy -dvvvdy -dvwwdy -duvvdy
// (D1 LU(D0)^-1 of rescale * sum of -uuv, -vvv, -vww, -uvv), up, -dpudy,
// vp, -2dpvdy, wp, pp, 2pdudx, 2pdvdy, 2pdwdz, pdudy+pdvdx.
int orc_statistics2(void *h, const double *Sd, const double *Pd, const double *mom, double *out) {
  Ctx *c = static_cast<Ctx *>(h);
  const cd *S = reinterpret_cast<const cd *>(Sd);
  const cd *P = reinterpret_cast<const cd *>(Pd);
  const int n = c->ny;
  auto line = [&](int q) { return out + (size_t)q * n; };
  // dataset order of the first 15 lines -> accumulator index (3..21 - 3)
  const int mom_of[15] = {0, 1, 2, 3, 4, 5, 10, 11, 12, 13, 14, 15, 16, 17, 18};
  for (int q = 0; q < 15; ++q) {
    double *o = line(q);
    o[0] = 0.0;
    for (int i = 1; i < n - 1; ++i) o[i] = c->rescale * mom[(size_t)mom_of[q] * n + i];
    o[n - 1] = 0.0;
  }
  // D1 LU(D0)^-1 of a real line (dcmplx(x, 0), real part)
  auto ddy = [&](const double *x, double scale_in, double *o, double scale_out) {
    std::vector<cd> a(n), b(n);
    for (int i = 0; i < n; ++i) a[i] = C(x[i] * scale_in, 0.0);
    labs_bs(c->d0lu.data(), n, c->bw, a.data());
    labs_mv(c->d1.data(), n, c->bw, a.data(), b.data());
    for (int i = 0; i < n; ++i) o[i] = scale_out * b[i].re;
  };
  // -uuv, -vvv, -vww, -uvv are accumulators 6, 8, 7, 9 (the reference's
  // dataset order -duuvdy, -dvvvdy, -dvwwdy, -duvvdy)
  const int tri[4] = {6, 8, 7, 9};
  for (int q = 0; q < 4; ++q) ddy(mom + (size_t)tri[q] * n, c->rescale, line(15 + q), 1.0);
  // Parseval sums over all modes in the reference's line order (kx fastest):
  // get_uv (982-1017) and get_turb_transport_PI1_ij (817-902)
  std::vector<double> up(n, 0.0), vp(n, 0.0), wp(n, 0.0), pp(n, 0.0), puu(n, 0.0), pvv(n, 0.0), pww(n, 0.0),
      puv(n, 0.0);
  const int L = c->nxh * c->nz;
  auto uv_sum = [&](std::vector<double> &acc, int fa, int fb, bool fb_is_p) {
    for (int l = 0; l < L; ++l) {
      const int k = l % c->nxh, j = l / c->nxh;
      const cd *a = S + ix(*c, fa, j, k, 0);
      const cd *b = fb_is_p ? P + ix(*c, 5, j, k, 0) : S + ix(*c, fb, j, k, 0);
      for (int i = 0; i < n; ++i) {
        if (k == 0) acc[i] = acc[i] + a[i].re * b[i].re + a[i].im * b[i].im;
        else acc[i] = acc[i] + 2.0 * (a[i].re * b[i].re + a[i].im * b[i].im);
      }
    }
  };
  uv_sum(up, 0, 0, true);
  uv_sum(vp, 1, 0, true);
  uv_sum(wp, 2, 0, true);
  for (int l = 0; l < L; ++l) {  // <pp>
    const int k = l % c->nxh, j = l / c->nxh;
    const cd *a = P + ix(*c, 5, j, k, 0);
    for (int i = 0; i < n; ++i) {
      if (k == 0) pp[i] = pp[i] + a[i].re * a[i].re + a[i].im * a[i].im;
      else pp[i] = pp[i] + 2.0 * (a[i].re * a[i].re + a[i].im * a[i].im);
    }
  }
  for (int l = 0; l < L; ++l) {  // PI1, first loop: x and z derivatives
    const int k = l % c->nxh, j = l / c->nxh;
    const double kxw = c->lmfxr * c->kx[k], kzw = c->lmfzr * c->kz[j];
    const cd *u = S + ix(*c, 0, j, k, 0), *v = S + ix(*c, 1, j, k, 0), *w = S + ix(*c, 2, j, k, 0);
    const cd *p = P + ix(*c, 5, j, k, 0);
    for (int i = 0; i < n; ++i) {
      if (k == 0) {
        puu[i] = puu[i] + 2.0 * kxw * (u[i].re * p[i].im - u[i].im * p[i].re);
        puv[i] = puv[i] + kxw * (v[i].re * p[i].im - v[i].im * p[i].re);
        pww[i] = pww[i] + 2.0 * kzw * (w[i].re * p[i].im - w[i].im * p[i].re);
      } else {
        puu[i] = puu[i] + (g_mutant == 10 ? -4.0 : 4.0) * kxw * (u[i].re * p[i].im - u[i].im * p[i].re);
        puv[i] = puv[i] + 2.0 * kxw * (v[i].re * p[i].im - v[i].im * p[i].re);
        pww[i] = pww[i] + 4.0 * kzw * (w[i].re * p[i].im - w[i].im * p[i].re);
      }
    }
  }
  std::vector<cd> du(n), dv(n), t(n);
  for (int l = 0; l < L; ++l) {  // PI1, second loop: y derivatives
    const int k = l % c->nxh, j = l / c->nxh;
    const cd *u = S + ix(*c, 0, j, k, 0), *v = S + ix(*c, 1, j, k, 0), *p = P + ix(*c, 5, j, k, 0);
    for (int i = 0; i < n; ++i) t[i] = u[i];
    labs_bs(c->d0lu.data(), n, c->bw, t.data());
    labs_mv(c->d1.data(), n, c->bw, t.data(), du.data());
    for (int i = 0; i < n; ++i) t[i] = v[i];
    labs_bs(c->d0lu.data(), n, c->bw, t.data());
    labs_mv(c->d1.data(), n, c->bw, t.data(), dv.data());
    for (int i = 0; i < n; ++i) {
      if (k == 0) {
        pvv[i] = pvv[i] + 2.0 * (dv[i].re * p[i].re + dv[i].im * p[i].im);
        puv[i] = puv[i] + (du[i].re * p[i].re + du[i].im * p[i].im);
      } else {
        pvv[i] = pvv[i] + 4.0 * (dv[i].re * p[i].re + dv[i].im * p[i].im);
        puv[i] = puv[i] + 2.0 * (du[i].re * p[i].re + du[i].im * p[i].im);
      }
    }
  }
  std::copy(up.begin(), up.end(), line(19));
  ddy(up.data(), 1.0, line(20), -1.0);
  std::copy(vp.begin(), vp.end(), line(21));
  ddy(vp.data(), 1.0, line(22), -2.0);
  std::copy(wp.begin(), wp.end(), line(23));
  std::copy(pp.begin(), pp.end(), line(24));
  std::copy(puu.begin(), puu.end(), line(25));
  std::copy(pvv.begin(), pvv.end(), line(26));
  std::copy(pww.begin(), pww.end(), line(27));
  std::copy(puv.begin(), puv.end(), line(28));
  return 0;
}

// Detection-plane values of v for opposition control: out(nxh, nz, 2)
// complex, [0] at y = -1 + ydist, [1] at y = 1 - ydist.
int orc_opp_planes(void *h, const double *Sd, double *out) {
  Ctx *c = static_cast<Ctx *>(h);
  const cd *S = reinterpret_cast<const cd *>(Sd);
  cd *o = reinterpret_cast<cd *>(out);
  const size_t nl = (size_t)c->nxh * c->nz;
#pragma omp parallel for schedule(static)
  for (long l = 0; l < (long)nl; ++l) {
    const int k = (int)(l % c->nxh), j = (int)(l / c->nxh);
    const cd *v = S + ix(*c, 1, j, k, 0);
    o[l] = detect_dot(*c, v, c->opp_lo_row);
    o[nl + l] = detect_dot(*c, v, c->opp_hi_row);
  }
  return 0;
}

// wavespace (wavespace.f90:46-107) over lines [l0, l1) of the (kx,kz) grid
// (l = kz_index*nxh + kx).  U(ny,nxh,nz,10) holds the products on entry and
// u,v,w,theta in slots 0,1,2,6 on exit; RK(ny,nxh,nz,3) is updated.
// vd (nxh,nz,2) = detection-plane v for opposition control (may be NULL).
// mean[4] = phi_u, phi_t, tau_w, q_w (written by the line that holds (0,0)).
int orc_wavespace(void *h, int s, double dt, double *Ud, double *RKd, const double *vdd, const double *usrcd,
                  int l0, int l1, double *mean) {
  Ctx *c = static_cast<Ctx *>(h);
  cd *U = reinterpret_cast<cd *>(Ud);
  cd *RK = reinterpret_cast<cd *>(RKd);
  const cd *vd = reinterpret_cast<const cd *>(vdd);
  // mixed-type scalar condition: u of the substep's input state, reference
  // layout (ny, nxh, nz)
  const bool mixed = c->cfg.scalar_bc == PB_SCALAR_MIXED;
  if (mixed && !usrcd) return PB_E_ARG;
  const cd *usrc = mixed ? reinterpret_cast<const cd *>(usrcd) : nullptr;
  const double ks = (g_mutant == 9 ? -1.0 : 1.0) * c->cfg.nu / c->cfg.pr / c->cfg.bulk_u;
  const int n = c->ny;
  const size_t nl = (size_t)c->nxh * c->nz;
  const pb_config &g = c->cfg;
#pragma omp parallel
  {
    Line L;
    L.c = c; L.n = n; L.bw = c->bw;
    L.D0lu = c->d0lu.data(); L.D0 = c->d0.data(); L.D1 = c->d1.data(); L.D2 = c->d2.data();
    L.Dbuf.assign((size_t)n * c->W, 0.0);
    L.dt = dt; L.s = s;
    std::vector<cd> b1(n), b2(n);
#pragma omp for schedule(dynamic, 16)
    for (int l = l0; l < l1; ++l) {
      const int k = l % c->nxh, j = l / c->nxh;
      auto slot = [&](int f) { return U + ix(*c, f, j, k, 0); };
      auto rk = [&](int f) { return RK + ix(*c, f, j, k, 0); };
      if (c->kx[k] == 0 && c->kz[j] == 0) {
        MeanOut mo;
        wave00(L, slot(0), slot(1), slot(2), slot(3), slot(4), slot(6), slot(7), slot(8), rk(0), rk(1), rk(2),
               b1.data(), g.nu, g.pr, g.bulk_u, g.bulk_w, g.bulk_t, g.u_gal, usrc ? usrc + ix(*c, 0, j, k, 0) : nullptr,
               mo);
        if (mean) { mean[0] = mo.phi_u; mean[1] = mo.phi_t; mean[2] = mo.tau_w; mean[3] = mo.q_w; }
      } else {
        const double kxp = c->kx[k] * c->lmfxr, kzp = c->kz[j] * c->lmfzr;
        hvhg(L, slot(0), slot(1), slot(2), slot(3), slot(4), slot(6), slot(7), slot(8), kxp, kzp);
        solve_gt(L, slot(0), rk(0), b1.data(), b2.data(), kxp, kzp, g.nu, g.pr, false);
        cd wlo = C(0.0, 0.0), whi = C(0.0, 0.0);
        if (g.opp_alpha != 0.0 && vd) {
          wlo = mulr(vd[(size_t)j * c->nxh + k], -g.opp_alpha);
          whi = mulr(vd[nl + (size_t)j * c->nxh + k], -g.opp_alpha);
        }
        if (!c->act.empty()) {
          wlo = add(wlo, c->act[(size_t)j * c->nxh + k]);
          whi = add(whi, c->act[(size_t)j * c->nxh + k]);
        }
        solve_v_f(L, slot(1), slot(2), slot(4), rk(1), slot(3), b1.data(), b2.data(), kxp, kzp, g.nu, wlo, whi);
        solve_u_w(n, slot(0), slot(2), kxp, kzp);
        if (usrc) {  // h_theta += (nu/pr) u / bulk_u
          cd *ht = slot(6);
          const cd *us = usrc + ix(*c, 0, j, k, 0);
          for (int i = 0; i < n; ++i) ht[i] = add(ht[i], mulr(us[i], ks));
        }
        solve_gt(L, slot(6), rk(2), b1.data(), b2.data(), kxp, kzp, g.nu, g.pr, true);
      }
    }
  }
  return 0;
}

// hvhg alone (hvhg.f90:141-276) on every line except kx = kz = 0, in place on
// reference-layout products U(ny,nxh,nz,10): afterwards slot 0 holds h_g,
// slot 4 h_v and slot 6 h_theta (slots 1-3, 7, 8 are scratch).
int orc_hvhg(void *h, double *Ud) {
  Ctx *c = static_cast<Ctx *>(h);
  cd *U = reinterpret_cast<cd *>(Ud);
  Line L;
  L.c = c; L.n = c->ny; L.bw = c->bw;
  L.D0lu = c->d0lu.data(); L.D0 = c->d0.data(); L.D1 = c->d1.data(); L.D2 = c->d2.data();
  L.dt = 0.0; L.s = 1;
  for (int j = 0; j < c->nz; ++j)
    for (int k = 0; k < c->nxh; ++k) {
      if (c->kx[k] == 0 && c->kz[j] == 0) continue;
      auto slot = [&](int f) { return U + ix(*c, f, j, k, 0); };
      hvhg(L, slot(0), slot(1), slot(2), slot(3), slot(4), slot(6), slot(7), slot(8), c->kx[k] * c->lmfxr,
           c->kz[j] * c->lmfzr);
    }
  return 0;
}

void orc_set_mutant(int m) { g_mutant = m; }

// Bounded CPU-baseline sample at full grid size without full-size arrays:
// the nonlinear stage on nplanes y planes (plane-local buffers) and the
// wavespace on nlines (kx,kz) systems spread over the grid, on synthetic
// data.  Returns wall seconds of each part (OpenMP over planes / lines).
int orc_sample(void *h, int nplanes, int nlines, int s, double dt, double *t_nonlin, double *t_wave) {
  Ctx *c = static_cast<Ctx *>(h);
  const size_t nxh = c->nxh, nz = c->nz, ny = c->ny;
  uint64_t st = 0x9E3779B97F4A7C15ull;
  auto rnd = [&]() {
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    return (double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  };
  // --- nonlinear: each thread owns one plane of input and output
  const int nt = omp_get_max_threads();
  std::vector<std::vector<cd>> Sp(nt), Pp(nt);
  for (int t = 0; t < nt; ++t) {
    Sp[t].resize(4 * nz * nxh);
    Pp[t].resize(10 * nz * nxh);
    for (auto &z : Sp[t]) z = C(rnd(), rnd());
  }
  double t0 = omp_get_wtime();
#pragma omp parallel
  {
    const int t = omp_get_thread_num();
    std::vector<cd> G((size_t)4 * c->nz32 * nxh), H((size_t)10 * c->nz32 * nxh);
    View S{Sp[t].data(), nz * nxh, nxh, 1, 0}, P{Pp[t].data(), nz * nxh, nxh, 1, 0};
#pragma omp for schedule(dynamic, 1)
    for (int q = 0; q < nplanes; ++q) nonlinear_plane(c, S, P, 0, (int)((size_t)q * ny / nplanes), G, H);
  }
  *t_nonlin = omp_get_wtime() - t0;
  // --- wavespace lines: (kx,kz) spread over the grid, skipping (0,0)
  std::vector<cd> U((size_t)10 * nlines * ny), RK((size_t)3 * nlines * ny);
  for (auto &z : U) z = C(1e-3 * rnd(), 1e-3 * rnd());
  for (auto &z : RK) z = C(1e-3 * rnd(), 1e-3 * rnd());
  t0 = omp_get_wtime();
#pragma omp parallel
  {
    Line L;
    L.c = c; L.n = (int)ny; L.bw = c->bw;
    L.D0lu = c->d0lu.data(); L.D0 = c->d0.data(); L.D1 = c->d1.data(); L.D2 = c->d2.data();
    L.Dbuf.assign(ny * c->W, 0.0);
    L.dt = dt; L.s = s;
    std::vector<cd> b1(ny), b2(ny);
    const pb_config &g = c->cfg;
#pragma omp for schedule(dynamic, 4)
    for (int l = 0; l < nlines; ++l) {
      const size_t sys = 1 + (size_t)l * (nxh * nz - 1) / nlines;
      const int k = (int)(sys % nxh), j = (int)(sys / nxh);
      const double kxp = c->kx[k] * c->lmfxr, kzp = c->kz[j] * c->lmfzr;
      auto sl = [&](int f) { return U.data() + ((size_t)f * nlines + l) * ny; };
      auto rk = [&](int f) { return RK.data() + ((size_t)f * nlines + l) * ny; };
      hvhg(L, sl(0), sl(1), sl(2), sl(3), sl(4), sl(6), sl(7), sl(8), kxp, kzp);
      solve_gt(L, sl(0), rk(0), b1.data(), b2.data(), kxp, kzp, g.nu, g.pr, false);
      solve_v_f(L, sl(1), sl(2), sl(4), rk(1), sl(3), b1.data(), b2.data(), kxp, kzp, g.nu, C(0, 0), C(0, 0));
      solve_u_w((int)ny, sl(0), sl(2), kxp, kzp);
      solve_gt(L, sl(6), rk(2), b1.data(), b2.data(), kxp, kzp, g.nu, g.pr, true);
    }
  }
  *t_wave = omp_get_wtime() - t0;
  return nt;
}

// One full substep on reference-layout arrays.  S(ny,nxh,nz,4) state in/out,
// RK(ny,nxh,nz,3) in/out.  out6 = dt, velmax, phi_u, phi_t, tau_w, q_w.
// The opposition gain and the wall actuation (nxh*nz complex, [kz][kx], or
// NULL for none) of later substeps: v(wall) = -alpha v(y_d) + act.
int orc_set_control(void *h, double alpha, const double *act) {
  Ctx *c = static_cast<Ctx *>(h);
  if (alpha != 0.0 && c->opp_lo_row.empty()) return PB_E_ARG;
  c->cfg.opp_alpha = alpha;
  c->act.clear();
  if (act) c->act.assign(reinterpret_cast<const cd *>(act), reinterpret_cast<const cd *>(act) + (size_t)c->nxh * c->nz);
  return 0;
}

int orc_substep(void *h, double *Sd, double *RKd, int s, double *dt_inout, double *out6) {
  Ctx *c = static_cast<Ctx *>(h);
  const size_t N = (size_t)c->ny * c->nxh * c->nz;
  std::vector<double> P(N * 10 * 2);
  std::vector<double> vd;
  if (c->cfg.opp_alpha != 0.0) {
    vd.resize((size_t)c->nxh * c->nz * 2 * 2);
    orc_opp_planes(h, Sd, vd.data());
  }
  double velmax = 0.0;
  orc_nonlinear(h, Sd, P.data(), 0, 0, c->ny, &velmax);
  double dt = *dt_inout;
  if (s == 1) {
    dt = c->cfg.dt_factor * dt + (1.0 - c->cfg.dt_factor) * c->cflx / velmax;
    if (dt <= 1e-14 || dt > 1e5) return PB_E_NUMERIC;
  }
  double mean[4] = {0, 0, 0, 0};
  const int rw = orc_wavespace(h, s, dt, P.data(), RKd, vd.empty() ? nullptr : vd.data(), Sd, 0,
                               (int)(c->nxh * c->nz), mean);
  if (rw) return rw;
  // copy u,v,w,theta back into the state
  const int map[4] = {0, 1, 2, 6};
  for (int f = 0; f < 4; ++f) std::memcpy(Sd + (size_t)f * N * 2, P.data() + (size_t)map[f] * N * 2, N * 16);
  *dt_inout = dt;
  if (out6) {
    out6[0] = dt; out6[1] = velmax; out6[2] = mean[0]; out6[3] = mean[1];
    out6[4] = mean[2]; out6[5] = mean[3];
  }
  return 0;
}

// Profiles of get_statistics_1 (get_statistics.f90:54-155) from the state
// S(ny,nxh,nz,4) = u, v, w, theta.  out: 16 lines of ny doubles
//   0 U (get_u, 1024-1048, + u_gal)  1 W  2 Theta
//   3 dU/dy  4 dW/dy  5 dTheta/dy (D1 LU(D0)^-1 of the mean column, 546-652)
//   6 uu 7 vv 8 ww 9 tt 10 uv 11 uw 12 vw 13 ut 14 wt 15 vt (get_uv, 982-1017)
// then u_center, t_center (centre-line coefficients, 566-570, 604-607).
static const int kUvPairs[10][2] = {{0, 0}, {1, 1}, {2, 2}, {3, 3}, {0, 1}, {0, 2}, {1, 2}, {0, 3}, {3, 2}, {1, 3}};
int orc_statistics(void *h, const double *Sd, double *out) {
  Ctx *c = static_cast<Ctx *>(h);
  const int n = c->ny, bw = c->bw;
  const size_t nl = (size_t)c->nxh * c->nz;
  const cd *S = reinterpret_cast<const cd *>(Sd);
  auto line = [&](int q) { return out + (size_t)q * n; };
  const double ugal = c->cfg.u_gal;
  const int fld[3] = {0, 2, 3};
  for (int q = 0; q < 3; ++q) {
    const cd *m = S + ix(*c, fld[q], 0, 0, 0);
    for (int y = 0; y < n; ++y) line(q)[y] = m[y].re + (q == 0 ? ugal : 0.0);
  }
  double centre[2] = {0.0, 0.0};
  for (int q = 0; q < 3; ++q) {
    std::vector<cd> c1(S + ix(*c, fld[q], 0, 0, 0), S + ix(*c, fld[q], 0, 0, 0) + n), c2(n);
    labs_bs(c->d0lu.data(), n, bw, c1.data());
    labs_mv(c->d1.data(), n, bw, c1.data(), c2.data());
    for (int y = 0; y < n; ++y) line(3 + q)[y] = c2[y].re;
    if (q != 1) {
      double a = (q == 0) ? ugal : 0.0;
      for (int k = 0; k < n; ++k) a = a + c1[k].re * c->center[k];
      centre[q == 0 ? 0 : 1] = a;
    }
  }
  for (int q = 0; q < 10; ++q) {
    double *st = line(6 + q);
    for (int y = 0; y < n; ++y) st[y] = 0.0;
    for (size_t l = 0; l < nl; ++l) {
      const int k = (int)(l % c->nxh), j = (int)(l / c->nxh);
      const cd *a = S + ix(*c, kUvPairs[q][0], j, k, 0), *b = S + ix(*c, kUvPairs[q][1], j, k, 0);
      for (int y = 0; y < n; ++y) {
        if (k == 0)
          st[y] = st[y] + a[y].re * b[y].re + a[y].im * b[y].im;
        else
          st[y] = st[y] + 2.0 * (a[y].re * b[y].re + a[y].im * b[y].im);
      }
    }
  }
  out[(size_t)16 * n] = centre[0];
  out[(size_t)16 * n + 1] = centre[1];
  return 0;
}

// 1-D energy (co)spectra of a state S(ny,nxh,nz,4) (write_energy_spectra.f90:9-300 without
// the HDF5 output).
//   du/dy, dw/dy: labs_bs(D0_LU) then labs_mat_mul(D1) on every line, kx = kz = 0 included
//     (get_y_dir_deriv_of_field, get_statistics.f90:658-681)
//   vorticity (get_vorticity_field, get_statistics.f90:912-948), with (0,k) = dcmplx(0d0,k):
//     ox = dw/dy - (0,kz) v,  oy = (0,kz) u - (0,kx) w,  oz = (0,kx) v - du/dy
//   E_ab_x(y, kx) = sum over the kz index in storage order of w (Re a Re b + Im a Im b),
//     w = 1 at kx = 0, else 2 (get_E_uu_x, write_energy_spectra.f90:326-359);
//   E_ab_z(y, kz) = the same sum over kx in ascending order (get_E_uu_z, 366-399).
// Fields 0 u, 1 v, 2 w, 3 theta, 4 ox, 5 oy, 6 oz; pairs in the reference's dataset order
// (write_energy_spectra.f90:86-239).  Its kz list computes "Eozoz_kz" from (ox, oz)
// (write_energy_spectra.f90:218-219), and so does this list.
//   ex: 18 x nxh x ny, ez: 18 x nz x ny (y fastest); either may be null
static const int kSpecX[18][2] = {{0, 0}, {1, 1}, {2, 2}, {3, 3}, {0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3},
                                  {3, 2}, {4, 4}, {5, 5}, {6, 6}, {4, 5}, {4, 6}, {5, 6}, {1, 6}, {2, 5}};
static const int kSpecZ[18][2] = {{0, 0}, {1, 1}, {2, 2}, {3, 3}, {0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3},
                                  {2, 3}, {4, 4}, {5, 5}, {4, 6}, {4, 5}, {4, 6}, {5, 6}, {1, 6}, {2, 5}};
int orc_spectra(void *h, const double *Sd, double *ex, double *ez) {
  Ctx *c = static_cast<Ctx *>(h);
  const int n = c->ny, bw = c->bw, nxh = c->nxh, nz = c->nz;
  const size_t line = (size_t)n, nl = (size_t)nxh * nz, fsz = nl * line;
  const cd *S = reinterpret_cast<const cd *>(Sd);
  std::vector<cd> F((size_t)7 * fsz);
  for (int f = 0; f < 4; ++f) std::copy(S + (size_t)f * fsz, S + (size_t)(f + 1) * fsz, F.begin() + (size_t)f * fsz);
  std::vector<cd> buf(n), du(n), dw(n);
  for (int j = 0; j < nz; ++j)
    for (int k = 0; k < nxh; ++k) {
      const size_t o = ((size_t)j * nxh + k) * line;
      const cd *u = &F[o], *v = &F[fsz + o], *w = &F[2 * fsz + o];
      std::copy(u, u + n, buf.begin());
      labs_bs(c->d0lu.data(), n, bw, buf.data());
      labs_mv(c->d1.data(), n, bw, buf.data(), du.data());
      std::copy(w, w + n, buf.begin());
      labs_bs(c->d0lu.data(), n, bw, buf.data());
      labs_mv(c->d1.data(), n, bw, buf.data(), dw.data());
      const cd ikz = C(0.0, c->kz[j] * c->lmfzr), ikx = C(0.0, c->kx[k] * c->lmfxr);
      cd *ox = &F[4 * fsz + o], *oy = &F[5 * fsz + o], *oz = &F[6 * fsz + o];
      for (int y = 0; y < n; ++y) {
        ox[y] = sub(dw[y], mul(ikz, v[y]));
        oy[y] = sub(mul(ikz, u[y]), mul(ikx, w[y]));
        oz[y] = sub(mul(ikx, v[y]), du[y]);
      }
    }
  auto term = [](cd a, cd b) { return a.re * b.re + a.im * b.im; };
  for (int q = 0; q < 18; ++q) {
    if (ex) {
      const cd *A = &F[(size_t)kSpecX[q][0] * fsz], *B = &F[(size_t)kSpecX[q][1] * fsz];
      double *o = ex + (size_t)q * nxh * n;
      for (size_t i = 0; i < (size_t)nxh * n; ++i) o[i] = 0.0;
      for (size_t l = 0; l < nl; ++l) {
        const int k = (int)(l % nxh);
        const cd *a = A + l * line, *b = B + l * line;
        double *st = o + (size_t)k * n;
        for (int y = 0; y < n; ++y)
          st[y] = (k == 0) ? st[y] + term(a[y], b[y]) : st[y] + 2.0 * term(a[y], b[y]);
      }
    }
    if (ez) {
      const cd *A = &F[(size_t)kSpecZ[q][0] * fsz], *B = &F[(size_t)kSpecZ[q][1] * fsz];
      double *o = ez + (size_t)q * nz * n;
      for (size_t i = 0; i < (size_t)nz * n; ++i) o[i] = 0.0;
      for (size_t l = 0; l < nl; ++l) {
        const int k = (int)(l % nxh), j = (int)(l / nxh);
        const cd *a = A + l * line, *b = B + l * line;
        double *st = o + (size_t)j * n;
        for (int y = 0; y < n; ++y)
          st[y] = (k == 0) ? st[y] + term(a[y], b[y]) : st[y] + 2.0 * term(a[y], b[y]);
      }
    }
  }
  return 0;
}

// Statistics-time pressure (get_statistics_1 -> get_uv(v, v) and the PPE
// wall values, then get_pressure.f90 after the products).
//   Sd: state S(ny,nxh,nz,4) at the start of the substep (v = slot 1)
//   Pd: products P(ny,nxh,nz,10) of orc_nonlinear(want_vv = 1); slot 5
//       (v^2) is replaced by p
//   vv: <vv>(y), ny doubles (get_statistics.f90:982-1017: modes in
//       kz-major, kx-minor order, kx = 0 once, kx > 0 twice)
int orc_pressure(void *h, const double *Sd, double *Pd, double *vv) {
  Ctx *c = static_cast<Ctx *>(h);
  const int n = c->ny;
  const size_t nl = (size_t)c->nxh * c->nz;
  const cd *S = reinterpret_cast<const cd *>(Sd);
  cd *P = reinterpret_cast<cd *>(Pd);
  std::vector<double> acc(n, 0.0);
  for (size_t l = 0; l < nl; ++l) {
    const int k = (int)(l % c->nxh), j = (int)(l / c->nxh);
    const cd *v = S + ix(*c, 1, j, k, 0);
    for (int y = 0; y < n; ++y) {
      if (k == 0)
        acc[y] = acc[y] + v[y].re * v[y].re + v[y].im * v[y].im;
      else
        acc[y] = acc[y] + 2.0 * (v[y].re * v[y].re + v[y].im * v[y].im);
    }
  }
  if (vv) std::memcpy(vv, acc.data(), n * sizeof(double));
#pragma omp parallel
  {
    Line L{c, n, c->bw, c->d0lu.data(), c->d0.data(), c->d1.data(), c->d2.data(),
           std::vector<double>((size_t)n * (2 * c->bw + 1)), 0.0, 1};
    std::vector<cd> b1(n), b2(n);
#pragma omp for schedule(static)
    for (long l = 0; l < (long)nl; ++l) {
      const int k = (int)(l % c->nxh), j = (int)(l / c->nxh);
      cd *u6 = P + ix(*c, 5, j, k, 0);
      if (c->kx[k] == 0 && c->kz[j] == 0) {
        for (int y = 0; y < n; ++y) u6[y] = C(-acc[y], 0.0);
        continue;
      }
      cd bc1, bc2;
      ppe_wall_values(*c, S + ix(*c, 1, j, k, 0), &bc1, &bc2);
      solve_ppe(L, P + ix(*c, 0, j, k, 0), P + ix(*c, 1, j, k, 0), P + ix(*c, 2, j, k, 0), P + ix(*c, 3, j, k, 0),
                P + ix(*c, 4, j, k, 0), u6, bc1, bc2, b1.data(), b2.data(), c->kx[k] * c->lmfxr,
                c->kz[j] * c->lmfzr);
    }
  }
  return 0;
}

}  // extern "C"
