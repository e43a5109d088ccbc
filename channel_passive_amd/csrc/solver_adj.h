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

#include "solver.h"

namespace pb {

PB_HD cplx conjc(cplx a) { return mk(a.re, -a.im); }
// conj(c) * w
PB_HD cplx cmulc(cplx c, cplx w) { return cmul(conjc(c), w); }

// b_bar <- A^-T b_bar for the packed LU of band_bs: the reverse of its
// back substitution (rows ascending), then of its forward substitution
// (rows descending).
template <class M>
PB_HD void adj_bs(const BandGeom &g, const M &A, const SV &b) {
  for (int r = 0; r < g.n; ++r) {
    const cplx t = b[r];
    for (int c = r + 1; c <= g.last(r); ++c) b[c] = csub(b[c], cscale(t, A(r, g.slot(r, c))));
  }
  for (int r = g.n - 1; r >= 0; --r) {
    const cplx t = cdivr(b[r], A(r, g.bw));
    b[r] = t;
    for (int p = g.first(r); p < r; ++p) b[p] = csub(b[p], cscale(t, A(r, g.slot(r, p))));
  }
}

// x_bar += scale * A^T b_bar
PB_HD void adj_mv(const BandGeom &g, const SharedBand &A, const SV &b, double scale, const SV &x) {
  for (int r = 0; r < g.n; ++r) {
    const cplx t = cscale(b[r], scale);
    for (int c = g.first(r); c <= g.last(r); ++c) x[c] = cadd(x[c], cscale(t, A(r, g.slot(r, c))));
  }
}
PB_HD void zero_col(const SV &x, int n) {
  for (int i = 0; i < n; ++i) x[i] = mk(0.0, 0.0);
}
// sum_i conj(a[i or n-1-i]) * w[i]
PB_HD cplx dotc(const SV &a, const SV &w, int n, bool rev) {
  cplx t = mk(0.0, 0.0);
  for (int i = 0; i < n; ++i) t = cadd(t, cmulc(a[rev ? n - 1 - i : i], w[i]));
  return t;
}

// Adjoint of hvhg (solver.h): on entry hg = u1, hv = u5, ht = u7 hold the
// adjoints of h_g, h_v and h_theta; on return u1..u5, u7..u9 hold the
// adjoints of the products P1..P5, P7..P9.  t0..t2 scratch.
PB_HD void adj_hvhg(const LineParams &P, SV u1, SV u2, SV u3, SV u4, SV u5, SV u7, SV u8, SV u9, SV t0, SV t1, SV t2,
                    double kx, double kz) {
  const int n = P.g.n;
  const double kz2 = kz * kz, kx2 = kx * kx, kxkz = kx * kz;
  const cplx ikx = iunit(kx), ikz = iunit(kz);
  const cplx a = cadd(iunit(kx * kx * kx), iunit(kx * (kz * kz))), b = cadd(iunit((kx * kx) * kz), iunit(kz * kz * kz));
  // scalar: u7 = -(ikx P7 + ikz P9 + D1 LU(D0)^-1 P8)
  zero_col(u8, n);
  adj_mv(P.g, P.D1, u7, -1.0, u8);
  adj_bs(P.g, P.D0LU, u8);
  for (int i = 0; i < n; ++i) {
    const cplx w = cneg(u7[i]);
    u9[i] = cmulc(ikz, w);
    u7[i] = cmulc(ikx, w);
  }
  // velocity: u1 = u1n + D1 LU^-1 t3, u5 = D1 LU^-1 u3n + a P2 + b P4 + D2 LU^-1 t4
  zero_col(t0, n);
  adj_mv(P.g, P.D2, u5, 1.0, t0);  // t4_bar
  adj_bs(P.g, P.D0LU, t0);
  zero_col(t1, n);
  adj_mv(P.g, P.D1, u1, 1.0, t1);  // t3_bar
  adj_bs(P.g, P.D0LU, t1);
  for (int i = 0; i < n; ++i) {
    // t3 = ikx P4 - ikz P2, t4 = ikz P4 + ikx P2
    u4[i] = cadd(cadd(cmulc(ikx, t1[i]), cmulc(ikz, t0[i])), cmulc(b, u5[i]));
    u2[i] = cadd(csub(cmulc(ikx, t0[i]), cmulc(ikz, t1[i])), cmulc(a, u5[i]));
  }
  zero_col(t2, n);
  adj_mv(P.g, P.D1, u5, 1.0, t2);  // u3n_bar
  adj_bs(P.g, P.D0LU, t2);
  for (int i = 0; i < n; ++i) {
    // u1n = (P1 - P3) kxkz + P5 (kz2 - kx2); u3n = -kx2 P1 - 2 kxkz P5 - kz2 P3
    const cplx g1 = u1[i], g3 = t2[i];
    u1[i] = csub(cscale(g1, kxkz), cscale(g3, kx2));
    u3[i] = csub(cscale(g1, -kxkz), cscale(g3, kz2));
    u5[i] = csub(cscale(g1, kz2 - kx2), cscale(g3, 2.0 * kxkz));
  }
}

// Adjoint of solve_gt (solver.h): on entry u1 / store hold the adjoints of
// the new field and the new store; on return the adjoints of the right-hand
// side h and of the old store.
PB_HD void adj_solve_gt(const LineParams &P, const OwnBand &Db, SV u1, SV store, SV buf1, SV buf2, double kx, double kz,
                        bool scalar) {
  const int n = P.g.n, s = P.s, S = P.S;
  const double k2 = kx * kx + kz * kz, dt = P.dt;
  auto visc = [&](double cf) { return scalar ? dt * cf * P.nu / P.pr : dt * cf * P.nu; };
  helmholtz_lu(P, Db, k2, visc(P.coef(1, s)), false);
  // store_out = store_mid + (1 - a' k2) u1_new + a' D2 x  (s < S), = u1_new (s = S)
  zero_col(buf1, n);
  if (s != S) {
    const double alpha = visc(P.coef(0, s + 1)), dummy = 1.0 - alpha * k2;
    for (int i = 0; i < n; ++i) u1[i] = cadd(u1[i], cscale(store[i], dummy));
    adj_mv(P.g, P.D2, store, alpha, buf1);
  } else {
    for (int i = 0; i < n; ++i) u1[i] = cadd(u1[i], store[i]);
  }
  adj_mv(P.g, P.D0, u1, 1.0, buf1);  // x_bar
  adj_bs(P.g, Db, buf1);             // buf1 (rhs) bar
  // store_mid = zeta h (s < S): h_bar starts from zeta store_bar
  if (s != S) {
    const double zeta = dt * P.coef(3, s + 1);
    for (int i = 0; i < n; ++i) u1[i] = cscale(store[i], zeta);
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
    adj_mv(P.g, P.D2, buf1, alpha, buf2);
    adj_bs(P.g, P.D0LU, buf2);
    for (int i = 0; i < n; ++i) store[i] = cadd(buf2[i], cscale(buf1[i], dummy));
  } else {
    for (int i = 0; i < n; ++i) store[i] = buf1[i];
  }
  for (int i = 0; i < n; ++i) u1[i] = cadd(u1[i], cscale(buf1[i], gamma));
}

// Adjoint of solve_v_f (solver.h): on entry v, dv, sphi hold the adjoints of
// v, dv/dy (f) and the new store; on return hv holds the adjoint of h_v,
// sphi that of the old store, *vlo / *vhi those of the wall values; v and
// dv are left as scratch.  c0..c5 scratch.
PB_HD void adj_solve_v_f(const LineParams &P, const OwnBand &Db, SV v, SV dv, SV hv, SV sphi, SV c0, SV c1, SV c2,
                         SV c3, SV c4, SV c5, double kx, double kz, cplx *vlo, cplx *vhi) {
  const int n = P.g.n, s = P.s, S = P.S, W = P.g.W(), m = P.g.bw;
  const double k2 = kx * kx + kz * kz, dt = P.dt, nu = P.nu;
  // the homogeneous solutions (state-independent), as the forward forms them
  helmholtz_lu(P, Db, k2, dt * P.coef(1, s) * nu, false);
  const double d00 = Db(0, m);
  c0[0] = mk(d00, d00);
  for (int i = 1; i < n; ++i) c0[i] = mk(0.0, 0.0);
  band_bs(P.g, Db, c0);           // X2
  band_mv(P.g, P.D0, c0, c1);     // B3 = D0 X2
  band_mv(P.g, P.D2, c0, c2);     // hvD = D2 X2
  for (int i = 0; i < W; ++i) Db.at(0, i) = 0.0;
  Db.at(0, m) = P.D0(0, m);
  for (int j = 1; j < n - 1; ++j)
    for (int i = 0; i < W; ++i) Db.at(j, i) = -k2 * P.D0(j, i) + P.D2(j, i);
  for (int i = 0; i < W; ++i) Db.at(n - 1, i) = 0.0;
  Db.at(n - 1, m) = P.D0(n - 1, m);
  band_lu(P.g, Db);
  for (int i = 0; i < n; ++i) c3[i] = c1[i];
  c3[0] = mk(0.0, 0.0);
  c3[n - 1] = mk(0.0, 0.0);
  band_bs(P.g, Db, c3);           // Y2
  band_mv(P.g, P.D1, c3, c4);     // hvC = D1 Y2
  band_mv(P.g, P.D0, c3, c5);     // hvE = D0 Y2
  const cplx h0 = c4[0], hn = c4[n - 1];
  const cplx den = csub(cmul(h0, h0), cmul(hn, hn));
  // c1, c2 adjoints from v = vA + c1 hvE + c2 hvE~, sphi, dv = dvB + c1 hvC - c2 hvC~
  cplx cb1 = cadd(dotc(c5, v, n, false), dotc(c4, dv, n, false));
  cplx cb2 = csub(dotc(c5, v, n, true), dotc(c4, dv, n, true));
  if (s != S) {
    const double alpha = dt * nu * P.coef(0, s + 1), dummy = 1.0 - alpha * k2;
    for (int i = 0; i < n; ++i) {
      const cplx w = sphi[i];
      const cplx f = cadd(cscale(c1[i], dummy), cscale(c2[i], alpha));
      const cplx r = cadd(cscale(c1[n - 1 - i], dummy), cscale(c2[n - 1 - i], alpha));
      cb1 = cadd(cb1, cmulc(f, w));
      cb2 = cadd(cb2, cmulc(r, w));
    }
  } else {
    cb1 = cadd(cb1, dotc(c1, sphi, n, false));
    cb2 = cadd(cb2, dotc(c1, sphi, n, true));
  }
  // c1 = (hn dn - h0 d0)/den, c2 = (h0 dn - hn d0)/den
  const cplx ph0 = cdiv(h0, den), phn = cdiv(hn, den);
  const cplx dnb = cadd(cmulc(phn, cb1), cmulc(ph0, cb2));
  const cplx d0b = cneg(cadd(cmulc(ph0, cb1), cmulc(phn, cb2)));
  dv[0] = cadd(dv[0], d0b);
  dv[n - 1] = cadd(dv[n - 1], dnb);
  // v = D0 Y1, dv = D1 Y1, Y1 = Poisson^-1 (hvA with wall values)
  zero_col(c0, n);
  adj_mv(P.g, P.D0, v, 1.0, c0);
  adj_mv(P.g, P.D1, dv, 1.0, c0);
  adj_bs(P.g, Db, c0);
  *vlo = c0[0];
  *vhi = c0[n - 1];
  c0[0] = mk(0.0, 0.0);
  c0[n - 1] = mk(0.0, 0.0);       // hvA_bar (wall rows replaced)
  // sphi_m2 = sphi_m + (1 - a' k2) hvA + a' dvA  (s < S), = hvA (s = S)
  zero_col(c3, n);
  if (s != S) {
    const double alpha = dt * nu * P.coef(0, s + 1), dummy = 1.0 - alpha * k2;
    for (int i = 0; i < n; ++i) c0[i] = cadd(c0[i], cscale(sphi[i], dummy));
    adj_mv(P.g, P.D2, sphi, alpha, c3);
  } else {
    for (int i = 0; i < n; ++i) c0[i] = cadd(c0[i], sphi[i]);
  }
  adj_mv(P.g, P.D0, c0, 1.0, c3);  // X1_bar (hvA = D0 X1, dvA = D2 X1)
  helmholtz_lu(P, Db, k2, dt * P.coef(1, s) * nu, false);
  adj_bs(P.g, Db, c3);             // B1_bar
  if (s != S) {
    const double zeta = dt * P.coef(3, s + 1);
    for (int i = 0; i < n; ++i) hv[i] = cscale(sphi[i], zeta);
  } else {
    zero_col(hv, n);
  }
  c3[0] = mk(0.0, 0.0);
  c3[n - 1] = mk(0.0, 0.0);
  const double gamma = dt * P.coef(2, s);
  if (s == 1) {
    const double alpha = dt * P.coef(0, s) * nu, dummy = 1.0 - alpha * k2;
    zero_col(c4, n);
    adj_mv(P.g, P.D2, c3, alpha, c4);
    adj_bs(P.g, P.D0LU, c4);
    for (int i = 0; i < n; ++i) sphi[i] = cadd(c4[i], cscale(c3[i], dummy));
  } else {
    for (int i = 0; i < n; ++i) sphi[i] = c3[i];
  }
  for (int i = 0; i < n; ++i) hv[i] = cadd(hv[i], cscale(c3[i], gamma));
}

// Adjoint of solve_u_w: on entry g, f hold the adjoints of u, w; on return
// those of g and f (dv/dy).
PB_HD void adj_solve_u_w(int n, SV g, SV f, double kx, double kz) {
  const cplx ci = mk(0.0, 1.0);
  const double den = kx * kx + kz * kz;
  const cplx kxx = cdivr(cmul(ci, mk(kx, 0.0)), den), kzz = cdivr(cmul(ci, mk(kz, 0.0)), den);
  for (int i = 0; i < n; ++i) {
    const cplx ub = g[i], wb = f[i];  // u = kxx f - kzz g, w = kxx g + kzz f
    g[i] = csub(cmulc(kxx, wb), cmulc(kzz, ub));
    f[i] = cadd(cmulc(kxx, ub), cmulc(kzz, wb));
  }
}

// Adjoint of one mean line (mean_u / mean_w / mean_t of solver.h), in one
// function: kind 0 = U (held bulk, tau_w), 1 = W (phi forced to 0), 2 =
// Theta (held bulk, or the mixed-type source with phi = 0; q_w).  On entry
// out holds the adjoint of the new mean profile and st that of the new
// store; wall_bar is the adjoint of tau_w / q_w (last substep).  On return
// st holds the adjoint of the old store, prod that of the product column
// whose y-derivative drives the line (uv, vw, v theta), and, for the mixed
// source, *src that of the source u (src may be null otherwise).  t0..t3
// scratch.
PB_HD void adj_mean_line(const LineParams &P, const OwnBand &Db, int kind, bool mixed, double ks, SV out, SV st,
                         SV prod, SV t0, SV t1, SV t2, SV t3, double wall_bar, const SV *src) {
  const int n = P.g.n, s = P.s, S = P.S, m = P.g.bw;
  const double dt = P.dt, nu = P.nu;
  const double kap = kind == 2 ? nu / P.pr : nu;
  const bool held = kind == 0 || (kind == 2 && !mixed);
  helmholtz_lu(P, Db, 0.0, dt * P.coef(1, s) * kap, true);
  // homogeneous response X.im to the unit interior right-hand side
  double iim = 0.0;
  if (held) {
    t0[0] = mk(0.0, 0.0);
    for (int i = 1; i < n - 1; ++i) t0[i] = mk(0.0, 1.0);
    t0[n - 1] = mk(0.0, 0.0);
    band_bs(P.g, Db, t0);
    iim = integrate(P, t0).im;
  }
  // out = D0 Y (- u_gal); st_out = a' D2 Y + out - zeta d  (s < S), = out (s = S)
  zero_col(t1, n);  // Y_bar
  zero_col(t2, n);  // d_bar (d = D1 LU(D0)^-1 prod [- ks src])
  for (int i = 0; i < n; ++i) out[i] = cadd(out[i], st[i]);
  if (s != S) {
    const double alpha = dt * P.coef(0, s + 1) * kap, zeta = dt * P.coef(3, s + 1);
    adj_mv(P.g, P.D2, st, alpha, t1);
    for (int i = 0; i < n; ++i) t2[i] = cscale(st[i], -zeta);
  }
  adj_mv(P.g, P.D0, out, 1.0, t1);
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
    for (int i = 0; i < n; ++i) phib = phib + t0[i].im * t1[i].re;
  for (int i = 0; i < n; ++i) t1[i] = mk(t1[i].re, 0.0);
  if (held) {
    // phi = (bulk - 0.5 I.re) / (0.5 I.im), I = sum integ X
    const double ire = -phib / iim;
    for (int i = 0; i < n; ++i) t1[i].re += ire * P.integ[i];
  }
  adj_bs(P.g, Db, t1);  // rhs_bar; rows 0, n-1 and the interior imaginary parts are constants
  t1[0] = mk(0.0, 0.0);
  t1[n - 1] = mk(0.0, 0.0);
  for (int i = 1; i < n - 1; ++i) t1[i].im = 0.0;
  // rhs[i] = st_in[i] (+ a D2 LU(D0)^-1 st_in [i], s = 1) - gamma d[i], i >= 1
  const double gamma = dt * P.coef(2, s);
  for (int i = 0; i < n; ++i) t2[i] = csub(t2[i], cscale(t1[i], gamma));
  if (s == 1) {
    const double alpha = dt * P.coef(0, s) * kap;
    zero_col(t3, n);
    adj_mv(P.g, P.D2, t1, alpha, t3);
    adj_bs(P.g, P.D0LU, t3);
    for (int i = 0; i < n; ++i) st[i] = cadd(t3[i], t1[i]);
  } else {
    for (int i = 0; i < n; ++i) st[i] = t1[i];
  }
  if (mixed && kind == 2 && src)
    for (int i = 0; i < n; ++i) (*src)[i] = cscale(t2[i], -ks);
  zero_col(prod, n);
  adj_mv(P.g, P.D1, t2, 1.0, prod);
  adj_bs(P.g, P.D0LU, prod);
}

// Everything the adjoint of one system needs ([slot][y][kz][kx] layout).
struct AdjArgs {
  cplx *AU;          // 9 slots: in 0,1,2,6 = adjoints of u, v, w, theta out;
                     // out 0-4, 6-8 = adjoints of the products, 5 = of the source u (mixed)
  cplx *ARK;         // 3 slots: adjoints of the stores, out -> in (in place)
  cplx *abuf;        // 6 scratch columns per system
  double *dbuf;      // W*ny doubles per system
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

PB_HD void adj_wavespace_line(const LineParams &P, const AdjArgs &A, size_t sys) {
  const int n = P.g.n;
  const int k = (int)(sys % A.nxh), j = (int)(sys / A.nxh);
  const int kzi = (j <= A.nh) ? j : j + 1 - 2 * A.nh - 1;
  auto col = [&](cplx *base, int slot) { return SV{base + (size_t)slot * A.field + sys, A.plane}; };
  OwnBand Db{A.dbuf + sys, A.plane, P.g.W()};
  SV x0 = col(A.abuf, 0), x1 = col(A.abuf, 1), x2 = col(A.abuf, 2), x3 = col(A.abuf, 3), x4 = col(A.abuf, 4),
     x5 = col(A.abuf, 5);
  if (k == 0 && kzi == 0) {
    const SV src = col(A.AU, 5);
    zero_col(src, n);
    // U: out slot 0, store 0, product uv (slot 1); W: slot 2, store 1, vw (slot 3);
    // Theta: slot 6, store 2, v theta (slot 7).  The V column (slot 1) is set to 0.
    adj_mean_line(P, Db, 0, false, 0.0, col(A.AU, 0), col(A.ARK, 0), col(A.AU, 1), x0, x1, x2, x3, A.tau_bar, nullptr);
    adj_mean_line(P, Db, 1, false, 0.0, col(A.AU, 2), col(A.ARK, 1), col(A.AU, 3), x0, x1, x2, x3, 0.0, nullptr);
    adj_mean_line(P, Db, 2, A.mixed != 0, A.ks, col(A.AU, 6), col(A.ARK, 2), col(A.AU, 7), x0, x1, x2, x3, A.q_bar,
                  &src);
    zero_col(col(A.AU, 0), n);
    zero_col(col(A.AU, 2), n);
    zero_col(col(A.AU, 4), n);
    zero_col(col(A.AU, 6), n);
    zero_col(col(A.AU, 8), n);
    if (A.avd) A.avd[sys] = A.avd[A.plane + sys] = mk(0.0, 0.0);
    return;
  }
  const double kx = k * A.lmfxr, kz = kzi * A.lmfzr;
  adj_solve_gt(P, Db, col(A.AU, 6), col(A.ARK, 2), x0, x1, kx, kz, true);
  const SV src = col(A.AU, 5);
  for (int i = 0; i < n; ++i) src[i] = A.mixed ? cscale(col(A.AU, 6)[i], A.ks) : mk(0.0, 0.0);
  adj_solve_u_w(n, col(A.AU, 0), col(A.AU, 2), kx, kz);
  cplx vlo, vhi;
  adj_solve_v_f(P, Db, col(A.AU, 1), col(A.AU, 2), col(A.AU, 4), col(A.ARK, 1), x0, x1, x2, x3, x4, x5, kx, kz, &vlo,
                &vhi);
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
  adj_solve_gt(P, Db, col(A.AU, 0), col(A.ARK, 0), x0, x1, kx, kz, false);
  adj_hvhg(P, col(A.AU, 0), col(A.AU, 1), col(A.AU, 2), col(A.AU, 3), col(A.AU, 4), col(A.AU, 6), col(A.AU, 7),
           col(A.AU, 8), x0, x1, x2, kx, kz);
}

}  // namespace pb
