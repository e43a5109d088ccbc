// solver_adj.h -- discrete adjoint of the per-(kx,kz) wall-normal solve of
// one RK3-CN substep (the transpose of wavespace_line in solver.h).
//
// For a fixed dt the wall-normal stage is affine in its inputs: the product
// spectra, the three RK stores, the detection-plane v of the opposition
// control, the wall actuation and (mixed-type scalar condition) the source
// u.  Its matrices (D0, D1, D2, LU(D0), the Helmholtz and Poisson operators
// and their homogeneous solutions) depend on k^2, dt and the substep only,
// so the adjoint needs no forward intermediates: every function below runs
// the statements of its forward counterpart in reverse order, with
//   band solve  x = A^-1 b  ->  b_bar = A^-T x_bar  (adj_bs),
//   band product b = A x    ->  x_bar += A^T b_bar  (adj_mv),
//   w = c z (c complex)     ->  z_bar += conj(c) w_bar,
// under the real inner product <a, b> = sum Re(conj(a) b) over every stored
// complex number (each stored (kx, kz) column is an independent unknown of
// the code, as it is of the forward).  Values overwritten by a forward
// statement get a zero adjoint; wall rows and imaginary parts set to
// constants kill theirs.  Host + device (one GPU thread per system); a CPU
// build is checked against the forward by dot-product tests.
#pragma once

#include "solver_fused.h"  // solver.h and the register windows CWin

namespace pb {

PB_HD cplx conjc(cplx a) { return mk(a.re, -a.im); }
// conj(c) * w
PB_HD cplx cmulc(cplx c, cplx w) { return cmul(conjc(c), w); }

// strided real column (reads and writes)
struct RVc {
  double *p;
  size_t s;
  PB_HD double &operator[](int i) const { return p[(size_t)i * s]; }
};

// Every loop below reads a batch of ADJ_BATCH rows (all its inputs) before
// it writes any of them: a write through one column might alias a later read
// for all the compiler knows, so row-at-a-time loops would wait for memory
// once per row; batched, a thread waits once per batch.
constexpr int ADJ_BATCH = 8;

// rows 0..n-1: ld(i) reads row i's inputs, st(i, v) computes and writes
template <class Ld, class St>
PB_HD void rows_batched(int n, Ld &&ld, St &&st) {
  using T = decltype(ld(0));
  int i = 0;
  for (; i + ADJ_BATCH <= n; i += ADJ_BATCH) {
    T v[ADJ_BATCH];
#pragma unroll
    for (int q = 0; q < ADJ_BATCH; ++q) v[q] = ld(i + q);
#pragma unroll
    for (int q = 0; q < ADJ_BATCH; ++q) st(i + q, v[q]);
  }
  for (; i < n; ++i) st(i, ld(i));
}
struct C2 {
  cplx a, b;
};
// y = y + a x;  y = a x;  y = x1 + a x2
PB_HD void col_axpy(const SV &y, const SV &x, double a, int n) {
  rows_batched(n, [&](int i) { return C2{y[i], x[i]}; }, [&](int i, C2 v) { y[i] = cadd(v.a, cscale(v.b, a)); });
}
PB_HD void col_scale(const SV &y, const SV &x, double a, int n) {
  rows_batched(n, [&](int i) { return x[i]; }, [&](int i, cplx v) { y[i] = cscale(v, a); });
}
PB_HD void col_add_scaled(const SV &y, const SV &x1, const SV &x2, double a, int n) {
  rows_batched(n, [&](int i) { return C2{x1[i], x2[i]}; }, [&](int i, C2 v) { y[i] = cadd(v.a, cscale(v.b, a)); });
}

// b_bar <- A^-T b_bar for the packed LU of band_bs (L with its diagonal,
// unit U): y = U^-T b by rows ascending, then L^-T y by rows descending.
// Both are recurrences over a window of the last 2bw results held in
// registers; a batch's right-hand sides and band entries are read before
// the batch runs its chain.
template <int BW, class M>
PB_HD void adj_bs(const BandGeom &g, const M &A, const SV &b) {
  constexpr int B = ADJ_BATCH;
  const int n = g.n;
  CWin<2 * BW> w;
  w.clear();
  // y[c] = b[c] - sum U(r, c) y[r] over r < c with c <= last(r):
  // r = 0..c-1 for c <= 2bw; r = c-1-t, t < bw (slot bw+1+t) after
  auto first_rows = [&](int c) {
    cplx v = b[c];
#pragma unroll
    for (int q = 0; q < 2 * BW; ++q) {
      const int r = c - 1 - q;
      if (r >= 0) v = csub(v, cscale(w.w[q], A(r, g.slot(r, c))));
    }
    b[c] = v;
    w.push(v);
  };
  int c = 0;
  for (; c <= 2 * BW && c < n; ++c) first_rows(c);
  auto rows1 = [&](int c0, int cnt) {
    cplx bb[B];
    double uu[B][BW];
#pragma unroll
    for (int q = 0; q < B; ++q)
      if (q < cnt) {
        bb[q] = b[c0 + q];
#pragma unroll
        for (int t = 0; t < BW; ++t) uu[q][t] = A(c0 + q - 1 - t, BW + 1 + t);
      }
#pragma unroll
    for (int q = 0; q < B; ++q)
      if (q < cnt) {
        cplx v = bb[q];
#pragma unroll
        for (int t = 0; t < BW; ++t) v = csub(v, cscale(w.w[t], uu[q][t]));
        b[c0 + q] = v;
        w.push(v);
      }
  };
  for (; c + B <= n; c += B) rows1(c, B);
  if (c < n) rows1(c, n - c);
  // z[r] = (y[r] - sum L(p, r) z[p] over p > r with first(p) <= r) / L(r, r):
  // p = r+1..n-1 for r >= n-1-2bw; p = r+1+t, t < bw (slot bw-1-t) before
  w.clear();
  auto last_rows = [&](int r) {
    cplx v = b[r];
#pragma unroll
    for (int q = 0; q < 2 * BW; ++q) {
      const int p = r + 1 + q;
      if (p < n) v = csub(v, cscale(w.w[q], A(p, g.slot(p, r))));
    }
    v = cdivr(v, A(r, BW));
    b[r] = v;
    w.push(v);
  };
  int r = n - 1;
  for (; r >= n - 1 - 2 * BW && r >= 0; --r) last_rows(r);
  auto rows2 = [&](int r0, int cnt) {  // rows r0, r0-1, ..., r0-cnt+1
    cplx bb[B];
    double ll[B][BW], dg[B];
#pragma unroll
    for (int q = 0; q < B; ++q)
      if (q < cnt) {
        const int rr = r0 - q;
        bb[q] = b[rr];
        dg[q] = A(rr, BW);
#pragma unroll
        for (int t = 0; t < BW; ++t) ll[q][t] = A(rr + 1 + t, BW - 1 - t);
      }
#pragma unroll
    for (int q = 0; q < B; ++q)
      if (q < cnt) {
        cplx v = bb[q];
#pragma unroll
        for (int t = 0; t < BW; ++t) v = csub(v, cscale(w.w[t], ll[q][t]));
        v = cdivr(v, dg[q]);
        b[r0 - q] = v;
        w.push(v);
      }
  };
  for (; r - B + 1 >= 0; r -= B) rows2(r, B);
  if (r >= 0) rows2(r, r + 1);
}

// x_bar += scale * A^T b_bar, gathered per column c over the rows r that
// hold c (first(r) <= c <= last(r), all within 2bw of c).  Interior columns
// (2bw < c < n-1-2bw) take rows c-bw..c+bw, all interior (slot 2bw-q for
// row c-bw+q), from a register window of b_bar that advances one row per
// column, in batches.
template <int BW>
PB_HD void adj_mv(const BandGeom &g, const SharedBand &A, const SV &b, double scale, const SV &x) {
  constexpr int B = ADJ_BATCH;
  const int n = g.n;
  auto gather = [&](int c) {
    cplx t = mk(0.0, 0.0);
#pragma unroll
    for (int q = -2 * BW; q <= 2 * BW; ++q) {
      const int r = c + q;
      if (r >= 0 && r < n && g.first(r) <= c && c <= g.last(r)) t = cadd(t, cscale(b[r], A(r, g.slot(r, c))));
    }
    x[c] = cadd(x[c], cscale(t, scale));
  };
  const int lo = 2 * BW + 1, hi = n - 1 - 2 * BW;
  if (hi <= lo) {
    for (int c = 0; c < n; ++c) gather(c);
    return;
  }
  for (int c = 0; c < lo; ++c) gather(c);
  cplx w[2 * BW + 1];
#pragma unroll
  for (int q = 0; q < 2 * BW; ++q) w[q] = b[lo - BW + q];
  auto cols = [&](int c0, int cnt) {
    cplx nb[B], xo[B];
#pragma unroll
    for (int q = 0; q < B; ++q)
      if (q < cnt) {
        nb[q] = b[c0 + q + BW];
        xo[q] = x[c0 + q];
      }
#pragma unroll
    for (int q = 0; q < B; ++q)
      if (q < cnt) {
        const int c = c0 + q;
        w[2 * BW] = nb[q];
        cplx t = mk(0.0, 0.0);
#pragma unroll
        for (int k = 0; k <= 2 * BW; ++k) t = cadd(t, cscale(w[k], A(c - BW + k, 2 * BW - k)));
        x[c] = cadd(xo[q], cscale(t, scale));
#pragma unroll
        for (int k = 0; k < 2 * BW; ++k) w[k] = w[k + 1];
      }
  };
  int c = lo;
  for (; c + B <= hi; c += B) cols(c, B);
  if (c < hi) cols(c, hi - c);
  for (c = hi; c < n; ++c) gather(c);
}
PB_HD void zero_col(const SV &x, int n) {
  for (int i = 0; i < n; ++i) x[i] = mk(0.0, 0.0);
}
// sum_i conj(a[i or n-1-i]) * w[i]
PB_HD cplx dotc(const SV &a, const SV &w, int n, bool rev) {
  cplx t = mk(0.0, 0.0);
  rows_batched(n, [&](int i) { return C2{a[rev ? n - 1 - i : i], w[i]}; },
               [&](int, C2 v) { t = cadd(t, cmulc(v.a, v.b)); });
  return t;
}
// the same for a real column a standing for the complex column (a, a)
struct DC {
  double e;
  cplx w;
};
PB_HD cplx dotc_rr(const RVc &a, const SV &w, int n, bool rev) {
  cplx t = mk(0.0, 0.0);
  rows_batched(n, [&](int i) { return DC{a[rev ? n - 1 - i : i], w[i]}; },
               [&](int, DC v) { t = cadd(t, cmulc(mk(v.e, v.e), v.w)); });
  return t;
}

// The matrices of the transposed substep for one substep index at a fixed
// dt, per system, formed once (adj_factor_line) and read by every transposed
// substep: the LU factors of the Helmholtz operators of nu and nu/Pr (k^2 of
// the system; for (0,0) they equal wave00's mean-line operators) and of the
// Poisson operator, and four real columns (x stands for the complex (x, x)
// of solve_v_f's homogeneous solutions X2 = H^-1 (d00, d00) e0 and Y2 =
// Poisson^-1 (D0 X2 with zero walls)): 0 B3 = D0 X2, 1 hvD = D2 X2, 2 hvC =
// D1 Y2, 3 hvE = D0 Y2.  For the (0,0) system columns 0 and 1 hold the mean
// lines' homogeneous responses H^-1 (0, 1 .. 1, 0) of nu and of nu/Pr.
struct AdjFactors {
  double *luh, *lut, *lup, *hom;
  size_t plane, field;
  int W;
  PB_HD OwnBand band(double *base, size_t sys) const { return OwnBand{base + sys, plane, W}; }
  PB_HD RVc col(int q, size_t sys) const { return RVc{hom + (size_t)q * field + sys, plane}; }
};
// Adjoint of solve_gt (solver.h) with the LU of its Helmholtz operator:
// on entry u1 / store hold the adjoints of the new field and the new store;
// on return the adjoints of the right-hand side h and of the old store.
template <int BW>
PB_HD void adj_solve_gt(const LineParams &P, const OwnBand &LU, SV u1, SV store, SV buf1, SV buf2, double kx, double kz,
                        bool scalar) {
  const int n = P.g.n, s = P.s, S = P.S;
  const double k2 = kx * kx + kz * kz, dt = P.dt;
  auto visc = [&](double cf) { return scalar ? dt * cf * P.nu / P.pr : dt * cf * P.nu; };
  // store_out = store_mid + (1 - a' k2) u1_new + a' D2 x  (s < S), = u1_new (s = S)
  zero_col(buf1, n);
  if (s != S) {
    const double alpha = visc(P.coef(0, s + 1)), dummy = 1.0 - alpha * k2;
    col_axpy(u1, store, dummy, n);
    adj_mv<BW>(P.g, P.D2, store, alpha, buf1);
  } else {
    col_axpy(u1, store, 1.0, n);
  }
  adj_mv<BW>(P.g, P.D0, u1, 1.0, buf1);  // x_bar
  adj_bs<BW>(P.g, LU, buf1);             // buf1 (rhs) bar
  // store_mid = zeta h (s < S): h_bar starts from zeta store_bar
  if (s != S) {
    const double zeta = dt * P.coef(3, s + 1);
    col_scale(u1, store, zeta, n);
  } else {
    zero_col(u1, n);
  }
  buf1[0] = mk(0.0, 0.0);
  buf1[n - 1] = mk(0.0, 0.0);
  const double gamma = dt * P.coef(2, s);
  if (s == 1) {
    // rhs = (1 - a k2) store_in + gamma h + a D2 LU(D0)^-1 store_in
    const double alpha = visc(P.coef(0, s)), dummy = 1.0 - alpha * k2;
    zero_col(buf2, n);
    adj_mv<BW>(P.g, P.D2, buf1, alpha, buf2);
    adj_bs<BW>(P.g, P.D0LU, buf2);
    col_add_scaled(store, buf2, buf1, dummy, n);
  } else {
    col_scale(store, buf1, 1.0, n);
  }
  col_axpy(u1, buf1, gamma, n);
}

// Adjoint of solve_v_f (solver.h) with the LU factors of its Helmholtz (luh)
// and Poisson (lup) operators and its homogeneous columns (hom 0..3): on
// entry v, dv, sphi hold the adjoints of v, dv/dy (f) and the new store; on
// return hv holds the adjoint of h_v, sphi that of the old store, *vlo /
// *vhi those of the wall values; v and dv are left as scratch.  c0, c3, c4
// scratch.
template <int BW>
PB_HD void adj_solve_v_f(const LineParams &P, const OwnBand &luh, const OwnBand &lup, const RVc &B3, const RVc &hvD,
                         const RVc &hvC, const RVc &hvE, SV v, SV dv, SV hv, SV sphi, SV c0, SV c3, SV c4, double kx,
                         double kz, cplx *vlo, cplx *vhi) {
  const int n = P.g.n, s = P.s, S = P.S;
  const double k2 = kx * kx + kz * kz, dt = P.dt, nu = P.nu;
  const cplx h0 = mk(hvC[0], hvC[0]), hn = mk(hvC[n - 1], hvC[n - 1]);
  const cplx den = csub(cmul(h0, h0), cmul(hn, hn));
  // c1, c2 adjoints from v = vA + c1 hvE + c2 hvE~, sphi, dv = dvB + c1 hvC - c2 hvC~
  cplx cb1 = cadd(dotc_rr(hvE, v, n, false), dotc_rr(hvC, dv, n, false));
  cplx cb2 = csub(dotc_rr(hvE, v, n, true), dotc_rr(hvC, dv, n, true));
  if (s != S) {
    const double alpha = dt * nu * P.coef(0, s + 1), dummy = 1.0 - alpha * k2;
    struct R5 {
      cplx w;
      double b, d, br, dr;
    };
    rows_batched(n, [&](int i) { return R5{sphi[i], B3[i], hvD[i], B3[n - 1 - i], hvD[n - 1 - i]}; },
                 [&](int, R5 v) {
                   const double f = v.b * dummy + v.d * alpha, r = v.br * dummy + v.dr * alpha;
                   cb1 = cadd(cb1, cmulc(mk(f, f), v.w));
                   cb2 = cadd(cb2, cmulc(mk(r, r), v.w));
                 });
  } else {
    cb1 = cadd(cb1, dotc_rr(B3, sphi, n, false));
    cb2 = cadd(cb2, dotc_rr(B3, sphi, n, true));
  }
  // c1 = (hn dn - h0 d0)/den, c2 = (h0 dn - hn d0)/den
  const cplx ph0 = cdiv(h0, den), phn = cdiv(hn, den);
  const cplx dnb = cadd(cmulc(phn, cb1), cmulc(ph0, cb2));
  const cplx d0b = cneg(cadd(cmulc(ph0, cb1), cmulc(phn, cb2)));
  dv[0] = cadd(dv[0], d0b);
  dv[n - 1] = cadd(dv[n - 1], dnb);
  // v = D0 Y1, dv = D1 Y1, Y1 = Poisson^-1 (hvA with wall values)
  zero_col(c0, n);
  adj_mv<BW>(P.g, P.D0, v, 1.0, c0);
  adj_mv<BW>(P.g, P.D1, dv, 1.0, c0);
  adj_bs<BW>(P.g, lup, c0);
  *vlo = c0[0];
  *vhi = c0[n - 1];
  c0[0] = mk(0.0, 0.0);
  c0[n - 1] = mk(0.0, 0.0);  // hvA_bar (wall rows replaced)
  // sphi_m2 = sphi_m + (1 - a' k2) hvA + a' dvA  (s < S), = hvA (s = S)
  zero_col(c3, n);
  if (s != S) {
    const double alpha = dt * nu * P.coef(0, s + 1), dummy = 1.0 - alpha * k2;
    col_axpy(c0, sphi, dummy, n);
    adj_mv<BW>(P.g, P.D2, sphi, alpha, c3);
  } else {
    col_axpy(c0, sphi, 1.0, n);
  }
  adj_mv<BW>(P.g, P.D0, c0, 1.0, c3);  // X1_bar (hvA = D0 X1, dvA = D2 X1)
  adj_bs<BW>(P.g, luh, c3);            // B1_bar
  if (s != S) {
    const double zeta = dt * P.coef(3, s + 1);
    col_scale(hv, sphi, zeta, n);
  } else {
    zero_col(hv, n);
  }
  c3[0] = mk(0.0, 0.0);
  c3[n - 1] = mk(0.0, 0.0);
  const double gamma = dt * P.coef(2, s);
  if (s == 1) {
    const double alpha = dt * P.coef(0, s) * nu, dummy = 1.0 - alpha * k2;
    zero_col(c4, n);
    adj_mv<BW>(P.g, P.D2, c3, alpha, c4);
    adj_bs<BW>(P.g, P.D0LU, c4);
    col_add_scaled(sphi, c4, c3, dummy, n);
  } else {
    col_scale(sphi, c3, 1.0, n);
  }
  col_axpy(hv, c3, gamma, n);
}

// Adjoint of one mean line (mean_u / mean_w / mean_t of solver.h), in one
// function: kind 0 = U (held bulk, tau_w), 1 = W (phi forced to 0), 2 =
// Theta (held bulk, or the mixed-type source with phi = 0; q_w).  LU: the
// line's Helmholtz factor; xh: its homogeneous response to the unit interior
// right-hand side (the imaginary part the forward carries).  On entry out
// holds the adjoint of the new mean profile and st that of the new store;
// wall_bar is the adjoint of tau_w / q_w (last substep).  On return st holds
// the adjoint of the old store, prod that of the product column whose
// y-derivative drives the line (uv, vw, v theta), and, for the mixed source,
// *src that of the source u (src may be null otherwise).  t1..t3 scratch.
template <int BW>
PB_HD void adj_mean_line(const LineParams &P, const OwnBand &LU, const RVc &xh, int kind, bool mixed, double ks, SV out,
                         SV st, SV prod, SV t1, SV t2, SV t3, double wall_bar, const SV *src) {
  const int n = P.g.n, s = P.s, S = P.S, m = P.g.bw;
  const double dt = P.dt, nu = P.nu;
  const double kap = kind == 2 ? nu / P.pr : nu;
  const bool held = kind == 0 || (kind == 2 && !mixed);
  double iim = 0.0;
  if (held)
    for (int i = 0; i < n; ++i) iim = iim + xh[i] * P.integ[i];
  // out = D0 Y (- u_gal); st_out = a' D2 Y + out - zeta d  (s < S), = out (s = S)
  zero_col(t1, n);  // Y_bar
  zero_col(t2, n);  // d_bar (d = D1 LU(D0)^-1 prod [- ks src])
  col_axpy(out, st, 1.0, n);
  if (s != S) {
    const double alpha = dt * P.coef(0, s + 1) * kap, zeta = dt * P.coef(3, s + 1);
    adj_mv<BW>(P.g, P.D2, st, alpha, t1);
    col_scale(t2, st, -zeta, n);
  }
  adj_mv<BW>(P.g, P.D0, out, 1.0, t1);
  if (s == S && kind != 1) {
    // wall = 0.5 kap (D1(0,m) Y0 + D1(0,m+1) Y1 - D1(n-1,m-1) Y_{n-2} - D1(n-1,m) Y_{n-1}) (real parts)
    const double w = 0.5 * kap * wall_bar;
    t1[0].re += w * P.D1(0, m);
    t1[1].re += w * P.D1(0, m + 1);
    t1[n - 2].re -= w * P.D1(n - 1, m - 1);
    t1[n - 1].re -= w * P.D1(n - 1, m);
  }
  // Y = (X.re + phi X.im, 0): X.re_bar = Y.re_bar, phi_bar = sum X.im Y.re_bar
  double phib = 0.0;
  if (held)
    for (int i = 0; i < n; ++i) phib = phib + xh[i] * t1[i].re;
  // phi = (bulk - 0.5 I.re) / (0.5 I.im), I = sum integ X: X.re_bar += -phi_bar / I.im integ
  const double ire = held ? -phib / iim : 0.0;
  rows_batched(n, [&](int i) { return t1[i].re; },
               [&](int i, double v) { t1[i] = mk(held ? v + ire * P.integ[i] : v, 0.0); });
  adj_bs<BW>(P.g, LU, t1);  // rhs_bar; rows 0, n-1 and the interior imaginary parts are constants
  t1[0] = mk(0.0, 0.0);
  t1[n - 1] = mk(0.0, 0.0);
  for (int i = 1; i < n - 1; ++i) t1[i].im = 0.0;
  // rhs[i] = st_in[i] (+ a D2 LU(D0)^-1 st_in [i], s = 1) - gamma d[i], i >= 1
  const double gamma = dt * P.coef(2, s);
  col_axpy(t2, t1, -gamma, n);
  if (s == 1) {
    const double alpha = dt * P.coef(0, s) * kap;
    zero_col(t3, n);
    adj_mv<BW>(P.g, P.D2, t1, alpha, t3);
    adj_bs<BW>(P.g, P.D0LU, t3);
    col_add_scaled(st, t3, t1, 1.0, n);
  } else {
    col_scale(st, t1, 1.0, n);
  }
  if (mixed && kind == 2 && src) col_scale(*src, t2, -ks, n);
  zero_col(prod, n);
  adj_mv<BW>(P.g, P.D1, t2, 1.0, prod);
  adj_bs<BW>(P.g, P.D0LU, prod);
}

// The factors of one system for substep P.s (dt = P.dt), formed as the
// forward forms its matrices; c0, c1 complex scratch columns.
PB_HD void adj_factor_line(const LineParams &P, const AdjFactors &F, size_t sys, int nxh, int nh, double lmfxr,
                           double lmfzr, SV c0, SV c1) {
  const int n = P.g.n, W = P.g.W(), m = P.g.bw, s = P.s;
  const int k = (int)(sys % nxh), j = (int)(sys / nxh);
  const int kzi = (j <= nh) ? j : j + 1 - 2 * nh - 1;
  const double kx = k * lmfxr, kz = kzi * lmfzr, k2 = kx * kx + kz * kz;
  const OwnBand H = F.band(F.luh, sys), T = F.band(F.lut, sys), Q = F.band(F.lup, sys);
  helmholtz_lu(P, H, k2, P.dt * P.coef(1, s) * P.nu, false);
  helmholtz_lu(P, T, k2, P.dt * P.coef(1, s) * P.nu / P.pr, false);
  if (k == 0 && kzi == 0) {
    const OwnBand *lu[2] = {&H, &T};
    for (int q = 0; q < 2; ++q) {
      c0[0] = mk(0.0, 0.0);
      for (int i = 1; i < n - 1; ++i) c0[i] = mk(0.0, 1.0);
      c0[n - 1] = mk(0.0, 0.0);
      band_bs(P.g, *lu[q], c0);
      const RVc xh = F.col(q, sys);
      for (int i = 0; i < n; ++i) xh[i] = c0[i].im;
    }
    return;
  }
  // X2 = H^-1 (d00, d00) e0 (d00 = the factor's (0,0) entry), B3 = D0 X2, hvD = D2 X2
  const double d00 = H(0, m);
  c0[0] = mk(d00, d00);
  for (int i = 1; i < n; ++i) c0[i] = mk(0.0, 0.0);
  band_bs(P.g, H, c0);
  band_mv(P.g, P.D0, c0, c1);
  const RVc B3 = F.col(0, sys), hvD = F.col(1, sys), hvC = F.col(2, sys), hvE = F.col(3, sys);
  for (int i = 0; i < n; ++i) B3[i] = c1[i].re;
  band_mv(P.g, P.D2, c0, c1);
  for (int i = 0; i < n; ++i) hvD[i] = c1[i].re;
  // Poisson [D2 - k2 D0] with D0-diagonal wall rows (solve_v_f.f90:205-230)
  for (int i = 0; i < W; ++i) Q.at(0, i) = 0.0;
  Q.at(0, m) = P.D0(0, m);
  for (int jj = 1; jj < n - 1; ++jj)
    for (int i = 0; i < W; ++i) Q.at(jj, i) = -k2 * P.D0(jj, i) + P.D2(jj, i);
  for (int i = 0; i < W; ++i) Q.at(n - 1, i) = 0.0;
  Q.at(n - 1, m) = P.D0(n - 1, m);
  band_lu(P.g, Q);
  // Y2 = Poisson^-1 (B3 with zero walls), hvC = D1 Y2, hvE = D0 Y2
  for (int i = 0; i < n; ++i) c0[i] = mk(B3[i], B3[i]);
  c0[0] = mk(0.0, 0.0);
  c0[n - 1] = mk(0.0, 0.0);
  band_bs(P.g, Q, c0);
  band_mv(P.g, P.D1, c0, c1);
  for (int i = 0; i < n; ++i) hvC[i] = c1[i].re;
  band_mv(P.g, P.D0, c0, c1);
  for (int i = 0; i < n; ++i) hvE[i] = c1[i].re;
}

// Everything the adjoint of one system needs ([slot][y][kz][kx] layout).
struct AdjArgs {
  cplx *AU;          // 9 slots: in 0,1,2,6 = adjoints of u, v, w, theta out;
                     // out 0-4, 6-8 = adjoints of the products, 5 = of the source u (mixed)
  cplx *ARK;         // 3 slots: adjoints of the stores, out -> in (in place)
  cplx *abuf;        // ADJ_SCRATCH scratch columns per system
  AdjFactors F;      // the matrices of this substep index (adj_factor_line)
  const cplx *vd;    // detection-plane v of the forward (2 planes) or nullptr
  cplx *avd;         // out: adjoint of vd (2 planes) or nullptr
  cplx *aact;        // accumulated: adjoint of the wall actuation (one plane) or nullptr
  double *aalpha;    // accumulated: per-system term of d/d alpha (one plane) or nullptr
  double ks;         // mixed-type source factor (nu/pr)/bulk_u; 0 for the held-bulk source
  int mixed;
  double tau_bar, q_bar;  // adjoints of tau_w and q_w (last substep)
  size_t plane, field;
  int nxh, nz, nh;
  double lmfxr, lmfzr, opp_alpha;
};

// The transpose of one system's wall-normal stage runs as three
// independent chains (phase 1), then the transpose of hvhg (phase 2):
//   chain 0: theta (solve_t; the mixed-type source);
//   chain 1: g (solve_g) from u_bar, w_bar through solve_u_w;
//   chain 2: v (solve_v_f, the wall values: opposition control and
//            actuation) from v_bar and u_bar, w_bar through solve_u_w;
// for (0,0) the chains take the mean lines U, W, Theta.  Scratch columns
// (abuf): chain 0 uses 0-2, chain 1 3 (g_bar, then h_g_bar) and 8-9, chain 2
// 4 (f_bar) and 5-7; phase 2 uses 0-2 and reads column 3.
constexpr int ADJ_SCRATCH = 10;

template <int BW>
PB_HD void adj_chain(const LineParams &P, const AdjArgs &A, size_t sys, int chain) {
  const int n = P.g.n;
  const int k = (int)(sys % A.nxh), j = (int)(sys / A.nxh);
  const int kzi = (j <= A.nh) ? j : j + 1 - 2 * A.nh - 1;
  auto col = [&](cplx *base, int slot) { return SV{base + (size_t)slot * A.field + sys, A.plane}; };
  const OwnBand luh = A.F.band(A.F.luh, sys), lut = A.F.band(A.F.lut, sys);
  if (k == 0 && kzi == 0) {
    // U: out slot 0, store 0, product uv (slot 1); W: slot 2, store 1, vw
    // (slot 3), the V column (slot 1 after U's use of it) set to 0 -- slot 4
    // cleared; Theta: slot 6, store 2, v theta (slot 7), source (slot 5) --
    // slot 8 cleared.
    if (chain == 0) {
      adj_mean_line<BW>(P, luh, A.F.col(0, sys), 0, false, 0.0, col(A.AU, 0), col(A.ARK, 0), col(A.AU, 1),
                        col(A.abuf, 0), col(A.abuf, 1), col(A.abuf, 2), A.tau_bar, nullptr);
      zero_col(col(A.AU, 0), n);
    } else if (chain == 1) {
      adj_mean_line<BW>(P, luh, A.F.col(0, sys), 1, false, 0.0, col(A.AU, 2), col(A.ARK, 1), col(A.AU, 3),
                        col(A.abuf, 3), col(A.abuf, 8), col(A.abuf, 9), 0.0, nullptr);
      zero_col(col(A.AU, 2), n);
      zero_col(col(A.AU, 4), n);
    } else {
      const SV src = col(A.AU, 5);
      zero_col(src, n);
      adj_mean_line<BW>(P, lut, A.F.col(1, sys), 2, A.mixed != 0, A.ks, col(A.AU, 6), col(A.ARK, 2), col(A.AU, 7),
                        col(A.abuf, 4), col(A.abuf, 5), col(A.abuf, 6), A.q_bar, &src);
      zero_col(col(A.AU, 6), n);
      zero_col(col(A.AU, 8), n);
      if (A.avd) A.avd[sys] = A.avd[A.plane + sys] = mk(0.0, 0.0);
    }
    return;
  }
  const double kx = k * A.lmfxr, kz = kzi * A.lmfzr;
  if (chain == 0) {
    adj_solve_gt<BW>(P, lut, col(A.AU, 6), col(A.ARK, 2), col(A.abuf, 0), col(A.abuf, 1), kx, kz, true);
    const SV src = col(A.AU, 5), ht = col(A.AU, 6);
    if (A.mixed) col_scale(src, ht, A.ks, n);
    else zero_col(src, n);
    return;
  }
  // u = kxx f - kzz g, w = kxx g + kzz f (solve_u_w)
  const cplx ci = mk(0.0, 1.0);
  const double den = kx * kx + kz * kz;
  const cplx kxx = cdivr(cmul(ci, mk(kx, 0.0)), den), kzz = cdivr(cmul(ci, mk(kz, 0.0)), den);
  const SV ub = col(A.AU, 0), wb = col(A.AU, 2);
  if (chain == 1) {
    const SV gb = col(A.abuf, 3);
    rows_batched(n, [&](int i) { return C2{ub[i], wb[i]}; },
                 [&](int i, C2 v) { gb[i] = csub(cmulc(kxx, v.b), cmulc(kzz, v.a)); });
    adj_solve_gt<BW>(P, luh, gb, col(A.ARK, 0), col(A.abuf, 8), col(A.abuf, 9), kx, kz, false);
    return;
  }
  const SV fb = col(A.abuf, 4);
  rows_batched(n, [&](int i) { return C2{ub[i], wb[i]}; },
               [&](int i, C2 v) { fb[i] = cadd(cmulc(kxx, v.a), cmulc(kzz, v.b)); });
  cplx vlo, vhi;
  adj_solve_v_f<BW>(P, luh, A.F.band(A.F.lup, sys), A.F.col(0, sys), A.F.col(1, sys), A.F.col(2, sys),
                    A.F.col(3, sys), col(A.AU, 1), fb, col(A.AU, 4), col(A.ARK, 1), col(A.abuf, 5), col(A.abuf, 6),
                    col(A.abuf, 7), kx, kz, &vlo, &vhi);
  // wall values: -alpha vd + actuation
  const bool opp = A.opp_alpha != 0.0 && A.vd;
  if (A.avd) {
    A.avd[sys] = opp ? cscale(vlo, -A.opp_alpha) : mk(0.0, 0.0);
    A.avd[A.plane + sys] = opp ? cscale(vhi, -A.opp_alpha) : mk(0.0, 0.0);
  }
  if (A.aalpha) {
    double t = 0.0;
    if (A.vd) {
      const cplx a = A.vd[sys], b = A.vd[A.plane + sys];
      t = -((vlo.re * a.re + vlo.im * a.im) + (vhi.re * b.re + vhi.im * b.im));
    }
    A.aalpha[sys] += t;
  }
  if (A.aact) A.aact[sys] = cadd(A.aact[sys], cadd(vlo, vhi));
}

// Phase 2, the transpose of hvhg, in two steps.  Step a (part 0..3, four
// independent threads per system): part 0 the scalar products (P7, P8, P9
// into slots 6-8, from h_theta_bar in slot 6); parts 1-3 the three
// transposed solves of the velocity part, t4_bar = LU(D0)^-T D2^T h_v_bar,
// t3_bar = LU(D0)^-T D1^T h_g_bar, u3n_bar = LU(D0)^-T D1^T h_v_bar, into
// abuf columns 0, 1, 2.  Step b (elementwise): P1..P5 into slots 0-4.
template <int BW>
PB_HD void adj_hvhg_part(const LineParams &P, const AdjArgs &A, size_t sys, int part) {
  const int n = P.g.n;
  const int k = (int)(sys % A.nxh), j = (int)(sys / A.nxh);
  const int kzi = (j <= A.nh) ? j : j + 1 - 2 * A.nh - 1;
  if (k == 0 && kzi == 0) return;
  auto col = [&](cplx *base, int slot) { return SV{base + (size_t)slot * A.field + sys, A.plane}; };
  const double kx = k * A.lmfxr, kz = kzi * A.lmfzr;
  if (part == 0) {
    // u7 = -(ikx P7 + ikz P9 + D1 LU(D0)^-1 P8)
    const SV u7 = col(A.AU, 6), u8 = col(A.AU, 7), u9 = col(A.AU, 8);
    const cplx ikx = iunit(kx), ikz = iunit(kz);
    zero_col(u8, n);
    adj_mv<BW>(P.g, P.D1, u7, -1.0, u8);
    adj_bs<BW>(P.g, P.D0LU, u8);
    rows_batched(n, [&](int i) { return u7[i]; }, [&](int i, cplx v) {
      const cplx w = cneg(v);
      u9[i] = cmulc(ikz, w);
      u7[i] = cmulc(ikx, w);
    });
    return;
  }
  const SV hv = col(A.AU, 4), hg = col(A.abuf, 3), t = col(A.abuf, part - 1);
  zero_col(t, n);
  if (part == 1) adj_mv<BW>(P.g, P.D2, hv, 1.0, t);
  if (part == 2) adj_mv<BW>(P.g, P.D1, hg, 1.0, t);
  if (part == 3) adj_mv<BW>(P.g, P.D1, hv, 1.0, t);
  adj_bs<BW>(P.g, P.D0LU, t);
}
PB_HD void adj_hvhg_combine(const AdjArgs &A, size_t sys, int y) {
  const int k = (int)(sys % A.nxh), j = (int)(sys / A.nxh);
  const int kzi = (j <= A.nh) ? j : j + 1 - 2 * A.nh - 1;
  if (k == 0 && kzi == 0) return;
  const double kx = k * A.lmfxr, kz = kzi * A.lmfzr;
  const double kz2 = kz * kz, kx2 = kx * kx, kxkz = kx * kz;
  const cplx ikx = iunit(kx), ikz = iunit(kz);
  const cplx a = cadd(iunit(kx * kx * kx), iunit(kx * (kz * kz))), b = cadd(iunit((kx * kx) * kz), iunit(kz * kz * kz));
  const size_t e = (size_t)y * A.plane + sys, f = A.field;
  const cplx hv = A.AU[4 * f + e], hg = A.abuf[3 * f + e];
  const cplx t0 = A.abuf[e], t1 = A.abuf[f + e], t2 = A.abuf[2 * f + e];
  // t3 = ikx P4 - ikz P2, t4 = ikz P4 + ikx P2, u5 += a P2 + b P4
  A.AU[3 * f + e] = cadd(cadd(cmulc(ikx, t1), cmulc(ikz, t0)), cmulc(b, hv));
  A.AU[f + e] = cadd(csub(cmulc(ikx, t0), cmulc(ikz, t1)), cmulc(a, hv));
  // u1n = (P1 - P3) kxkz + P5 (kz2 - kx2); u3n = -kx2 P1 - 2 kxkz P5 - kz2 P3
  A.AU[e] = csub(cscale(hg, kxkz), cscale(t2, kx2));
  A.AU[2 * f + e] = csub(cscale(hg, -kxkz), cscale(t2, kz2));
  A.AU[4 * f + e] = csub(cscale(hg, kz2 - kx2), cscale(t2, 2.0 * kxkz));
}

// The whole transpose for one system (the two phases back to back).
template <int BW>
PB_HD void adj_wavespace_line(const LineParams &P, const AdjArgs &A, size_t sys) {
  for (int chain = 0; chain < 3; ++chain) adj_chain<BW>(P, A, sys, chain);
  for (int part = 0; part < 4; ++part) adj_hvhg_part<BW>(P, A, sys, part);
  for (int y = 0; y < P.g.n; ++y) adj_hvhg_combine(A, sys, y);
}

}  // namespace pb
