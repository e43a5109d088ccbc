// solver.h -- per-(kx,kz) wall-normal solvers of one RK3-CN substep.
//
// One call handles one (kx,kz) system (one GPU thread).  Every vector is a
// column of ny complex values in the device layout [y][kz][kx] (kx fastest),
// so consecutive threads touch consecutive addresses at every y.  The code
// is host+device so that a CPU build of the very same functions can be
// compared bit-for-bit with the oracle.  Arithmetic follows the reference
// statement by statement (SOURCE_PASSOPB/hvhg.f90, solve_g.f90, solve_t.f90,
// solve_v_f.f90, solve_u_w.f90, wave00.f90) and the packed-band kernels of
// module_labs.f90; compile with FP contraction off.
#pragma once

#include "pb_common.h"

namespace pb {

// strided complex column
struct SV {
  cplx *p;
  size_t s;
  PB_HD cplx &operator[](int i) const { return p[(size_t)i * s]; }
};

// Read of data that every lane of a wave reads at the same address and that
// no kernel writes (the shared band matrices): on the device the pointer is
// taken in the constant address space, so the compiler emits scalar loads
// into SGPRs instead of a vector load per lane.
PB_HD double ld_uniform(const double *p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return *(const __attribute__((address_space(4))) double *)p;
#else
  return *p;
#endif
}

// band read from a shared (k-independent) matrix: A[r*W + slot]
struct SharedBand {
  const double *a;
  int W;
  PB_HD double operator()(int r, int slot) const { return ld_uniform(a + r * W + slot); }
};

// per-system band: element (r, slot) at p[(r*W + slot) * s]
struct OwnBand {
  double *p;
  size_t s;
  int W;
  PB_HD double &at(int r, int slot) const { return p[(size_t)(r * W + slot) * s]; }
  PB_HD double operator()(int r, int slot) const { return p[(size_t)(r * W + slot) * s]; }
};

// ---------------------------------------------------------------- LABS ops
// No-pivot Crout (fixed_bandec*, module_labs.f90:943-1198, 3487-3636):
// L(r,c) = A(r,c) - sum_{p=c-1 down} L(r,p) U(p,c);
// U(r,c) = (1/L(r,r)) * (A(r,c) - sum_{p=r-1 down} L(r,p) U(p,c)).
PB_HD void band_lu(const BandGeom &g, const OwnBand &A) {
  for (int r = 0; r < g.n; ++r) {
    const int f = g.first(r);
    for (int c = f; c <= r; ++c) {
      double v = A(r, g.slot(r, c));
      for (int p = c - 1; p >= f; --p) {
        if (c > g.last(p)) break;
        v = v - A(r, g.slot(r, p)) * A(p, g.slot(p, c));
      }
      A.at(r, g.slot(r, c)) = v;
    }
    const double inv = 1.0 / A(r, g.bw);
    for (int c = r + 1; c <= g.last(r); ++c) {
      double v = A(r, g.slot(r, c));
      for (int p = r - 1; p >= f; --p) {
        if (c > g.last(p)) break;
        v = v - A(r, g.slot(r, p)) * A(p, g.slot(p, c));
      }
      A.at(r, g.slot(r, c)) = inv * v;
    }
  }
}

// forward (farthest first, divide by the pivot) and backward (nearest first)
// substitution (fixed_banbks*, module_labs.f90:3769-3823, 4252-4318)
template <class M>
PB_HD void band_bs(const BandGeom &g, const M &A, const SV &b) {
  for (int r = 0; r < g.n; ++r) {
    cplx v = b[r];
    for (int p = g.first(r); p < r; ++p) v = csub(v, cscale(b[p], A(r, g.slot(r, p))));
    b[r] = cdivr(v, A(r, g.bw));
  }
  for (int r = g.n - 1; r >= 0; --r) {
    cplx v = b[r];
    for (int c = r + 1; c <= g.last(r); ++c) v = csub(v, cscale(b[c], A(r, g.slot(r, c))));
    b[r] = v;
  }
}

// b = A x, ascending column order (LABS_mat_mul_4, module_labs.f90:4395-4430)
PB_HD void band_mv(const BandGeom &g, const SharedBand &A, const SV &x, const SV &b) {
  for (int r = 0; r < g.n; ++r) {
    const int f = g.first(r), l = g.last(r);
    cplx v = cscale(x[f], A(r, g.slot(r, f)));
    for (int c = f + 1; c <= l; ++c) v = cadd(v, cscale(x[c], A(r, g.slot(r, c))));
    b[r] = v;
  }
}

// -------------------------------------------------------------- line state
struct LineParams {
  BandGeom g;
  SharedBand D0, D1, D2, D0LU;
  const double *integ;  // LABS_INTEGRATE
  double rk[4][3];
  double dt, nu, pr, bulk_u, bulk_w, bulk_t, u_gal;
  int s, S;  // substep (1-based), steps per RK step
  // row is a literal at every call site; sub is selected without indexing so
  // that the parameter block stays in registers (no private-memory copy)
  PB_HD double coef(int row, int sub) const {
    const double *c = rk[row];
    return sub == 1 ? c[0] : (sub == 2 ? c[1] : c[2]);
  }
};

// Helmholtz LHS with diagonal-only wall rows (solve_g.f90:42-71, wave00.f90:58-82)
PB_HD void helmholtz_lu(const LineParams &P, const OwnBand &Db, double k2, double beta, bool mean) {
  const int n = P.g.n, W = P.g.W(), m = P.g.bw;
  const double dummy = 1.0 + beta * k2;
  for (int i = 0; i < W; ++i) Db.at(0, i) = 0.0;
  Db.at(0, m) = mean ? (P.D0(0, m) - beta * P.D2(0, m)) : (dummy * P.D0(0, m) - beta * P.D2(0, m));
  for (int j = 1; j < n - 1; ++j)
    for (int i = 0; i < W; ++i)
      Db.at(j, i) = mean ? (P.D0(j, i) - beta * P.D2(j, i)) : (dummy * P.D0(j, i) - beta * P.D2(j, i));
  for (int i = 0; i < W; ++i) Db.at(n - 1, i) = 0.0;
  Db.at(n - 1, m) = mean ? (P.D0(n - 1, m) - beta * P.D2(n - 1, m)) : (dummy * P.D0(n - 1, m) - beta * P.D2(n - 1, m));
  band_lu(P.g, Db);
}

// hvhg.f90:141-276
PB_HD void hvhg(const LineParams &P, SV u1, SV u2, SV u3, SV u4, SV u5, SV u7, SV u8, SV u9, double kx, double kz) {
  const int n = P.g.n;
  const double kz2 = kz * kz, kx2 = kx * kx, kxkz = kx * kz;
  const cplx ikx = iunit(kx), ikz = iunit(kz);
  const cplx ikx3 = iunit(kx * kx * kx), ikxkz2 = iunit(kx * (kz * kz));
  const cplx ikx2kz = iunit((kx * kx) * kz), ikz3 = iunit(kz * kz * kz);
  for (int i = 0; i < n; ++i) {
    const cplx d = u1[i];
    u1[i] = cadd(cscale(csub(u1[i], u3[i]), kxkz), cscale(u5[i], kz2 - kx2));
    u3[i] = csub(csub(cscale(d, -kx2), cscale(u5[i], 2.0 * kxkz)), cscale(u3[i], kz2));
  }
  band_bs(P.g, P.D0LU, u3);
  band_mv(P.g, P.D1, u3, u5);
  const cplx a = cadd(ikx3, ikxkz2), b = cadd(ikx2kz, ikz3);
  for (int i = 0; i < n; ++i) u5[i] = cadd(cadd(u5[i], cmul(a, u2[i])), cmul(b, u4[i]));
  for (int i = 0; i < n; ++i) {
    const cplx t3 = csub(cmul(ikx, u4[i]), cmul(ikz, u2[i]));
    const cplx t4 = cadd(cmul(ikz, u4[i]), cmul(ikx, u2[i]));
    u3[i] = t3;
    u4[i] = t4;
  }
  band_bs(P.g, P.D0LU, u3);
  band_mv(P.g, P.D1, u3, u2);
  band_bs(P.g, P.D0LU, u4);
  band_mv(P.g, P.D2, u4, u3);
  for (int i = 0; i < n; ++i) {
    u5[i] = cadd(u5[i], u3[i]);
    u1[i] = cadd(u1[i], u2[i]);
  }
  for (int i = 0; i < n; ++i) u7[i] = cadd(cmul(ikx, u7[i]), cmul(ikz, u9[i]));
  band_bs(P.g, P.D0LU, u8);
  band_mv(P.g, P.D1, u8, u9);
  for (int i = 0; i < n; ++i) u7[i] = cneg(cadd(u7[i], u9[i]));
}

// solve_g.f90:28-201 (scalar == false) and solve_t.f90:28-201 (scalar == true)
PB_HD void solve_gt(const LineParams &P, const OwnBand &Db, SV u1, SV store, SV buf1, SV buf2, double kx, double kz,
                    bool scalar) {
  const int n = P.g.n, s = P.s, S = P.S;
  const double k2 = kx * kx + kz * kz, dt = P.dt;
  auto visc = [&](double cf) { return scalar ? dt * cf * P.nu / P.pr : dt * cf * P.nu; };
  helmholtz_lu(P, Db, k2, visc(P.coef(1, s)), false);
  if (s == 1) {
    const double alpha = visc(P.coef(0, s)), gamma = dt * P.coef(2, s);
    const double dummy = 1.0 - alpha * k2;
    for (int i = 0; i < n; ++i) buf1[i] = cadd(cscale(store[i], dummy), cscale(u1[i], gamma));
    band_bs(P.g, P.D0LU, store);
    band_mv(P.g, P.D2, store, buf2);
    for (int i = 0; i < n; ++i) buf1[i] = cadd(buf1[i], cscale(buf2[i], alpha));
  } else {
    const double gamma = dt * P.coef(2, s);
    for (int i = 0; i < n; ++i) buf1[i] = cadd(store[i], cscale(u1[i], gamma));
  }
  buf1[0] = mk(0.0, 0.0);
  buf1[n - 1] = mk(0.0, 0.0);
  if (s != S) {
    const double zeta = dt * P.coef(3, s + 1);
    for (int i = 0; i < n; ++i) store[i] = cscale(u1[i], zeta);
  }
  band_bs(P.g, Db, buf1);
  band_mv(P.g, P.D0, buf1, u1);
  band_mv(P.g, P.D2, buf1, buf2);
  if (s != S) {
    const double alpha = visc(P.coef(0, s + 1));
    const double dummy = 1.0 - alpha * k2;
    for (int i = 0; i < n; ++i) store[i] = cadd(cadd(store[i], cscale(u1[i], dummy)), cscale(buf2[i], alpha));
  } else {
    for (int i = 0; i < n; ++i) store[i] = u1[i];
  }
}

// solve_v_f.f90:38-375; vlo/vhi = wall values of v (0 without opposition control)
PB_HD void solve_v_f(const LineParams &P, const OwnBand &Db, SV v, SV dv, SV hv, SV sphi, SV buf1, SV buf2, SV buf3,
                     double kx, double kz, cplx vlo, cplx vhi) {
  const int n = P.g.n, s = P.s, S = P.S, W = P.g.W(), m = P.g.bw;
  const double k2 = kx * kx + kz * kz, dt = P.dt, nu = P.nu;
  if (s == 1) {
    const double alpha = dt * P.coef(0, s) * nu, gamma = dt * P.coef(2, s);
    const double dummy = 1.0 - alpha * k2;
    for (int i = 0; i < n; ++i) buf1[i] = cadd(cscale(sphi[i], dummy), cscale(hv[i], gamma));
    band_bs(P.g, P.D0LU, sphi);
    band_mv(P.g, P.D2, sphi, buf2);
    for (int i = 0; i < n; ++i) buf1[i] = cadd(buf1[i], cscale(buf2[i], alpha));
  } else {
    const double gamma = dt * P.coef(2, s);
    for (int i = 0; i < n; ++i) buf1[i] = cadd(sphi[i], cscale(hv[i], gamma));
  }
  buf1[0] = mk(0.0, 0.0);
  buf1[n - 1] = mk(0.0, 0.0);
  if (s != S) {
    const double zeta = dt * P.coef(3, s + 1);
    for (int i = 0; i < n; ++i) sphi[i] = cscale(hv[i], zeta);
  }
  const double d00 = Db(0, m);
  buf2[0] = mk(d00, d00);
  for (int i = 1; i < n; ++i) buf2[i] = mk(0.0, 0.0);
  band_bs(P.g, Db, buf1);
  band_bs(P.g, Db, buf2);
  band_mv(P.g, P.D0, buf1, hv);
  band_mv(P.g, P.D0, buf2, buf3);
  band_mv(P.g, P.D2, buf1, dv);
  if (s != S) {
    const double alpha = dt * nu * P.coef(0, s + 1);
    const double dummy = 1.0 - alpha * k2;
    for (int i = 0; i < n; ++i) sphi[i] = cadd(cadd(sphi[i], cscale(hv[i], dummy)), cscale(dv[i], alpha));
  } else {
    for (int i = 0; i < n; ++i) sphi[i] = hv[i];
  }
  // Poisson LHS [D2 - k2 D0] with D0-diagonal wall rows (solve_v_f.f90:205-230)
  for (int i = 0; i < W; ++i) Db.at(0, i) = 0.0;
  Db.at(0, m) = P.D0(0, m);
  for (int j = 1; j < n - 1; ++j)
    for (int i = 0; i < W; ++i) Db.at(j, i) = -k2 * P.D0(j, i) + P.D2(j, i);
  for (int i = 0; i < W; ++i) Db.at(n - 1, i) = 0.0;
  Db.at(n - 1, m) = P.D0(n - 1, m);
  band_lu(P.g, Db);
  hv[0] = vlo;
  hv[n - 1] = vhi;
  band_bs(P.g, Db, hv);
  for (int i = 0; i < n; ++i) buf1[i] = buf3[i];
  buf1[0] = mk(0.0, 0.0);
  buf1[n - 1] = mk(0.0, 0.0);
  band_bs(P.g, Db, buf1);
  band_mv(P.g, P.D0, hv, v);
  band_mv(P.g, P.D1, hv, dv);
  band_mv(P.g, P.D1, buf1, hv);
  const cplx h0 = hv[0], hn = hv[n - 1], d0 = dv[0], dn = dv[n - 1];
  const cplx den = csub(cmul(h0, h0), cmul(hn, hn));
  const cplx c1 = cdiv(csub(cmul(hn, dn), cmul(h0, d0)), den);
  const cplx c2 = cdiv(csub(cmul(h0, dn), cmul(hn, d0)), den);
  for (int i = 0; i < n; ++i) dv[i] = csub(cadd(dv[i], cmul(c1, hv[i])), cmul(c2, hv[n - 1 - i]));
  band_mv(P.g, P.D2, buf2, hv);
  if (s != S) {
    const double alpha = dt * nu * P.coef(0, s + 1);
    const double dummy = 1.0 - alpha * k2;
    const cplx c1d = cscale(c1, dummy), c2d = cscale(c2, dummy), c1a = cscale(c1, alpha), c2a = cscale(c2, alpha);
    for (int i = 0; i < n; ++i) {
      cplx acc = cadd(sphi[i], cmul(c1d, buf3[i]));
      acc = cadd(acc, cmul(c2d, buf3[n - 1 - i]));
      acc = cadd(acc, cmul(c1a, hv[i]));
      acc = cadd(acc, cmul(c2a, hv[n - 1 - i]));
      sphi[i] = acc;
    }
  } else {
    for (int i = 0; i < n; ++i) sphi[i] = cadd(cadd(sphi[i], cmul(c1, buf3[i])), cmul(c2, buf3[n - 1 - i]));
  }
  band_mv(P.g, P.D0, buf1, hv);
  for (int i = 0; i < n; ++i) v[i] = cadd(cadd(v[i], cmul(c1, hv[i])), cmul(c2, hv[n - 1 - i]));
}

// solve_u_w.f90:18-42 (g -> u, f -> w)
PB_HD void solve_u_w(int n, SV g, SV f, double kx, double kz) {
  const cplx ci = mk(0.0, 1.0);
  const double den = kx * kx + kz * kz;
  const cplx kxx = cdivr(cmul(ci, mk(kx, 0.0)), den), kzz = cdivr(cmul(ci, mk(kz, 0.0)), den);
  for (int i = 0; i < n; ++i) {
    const cplx dxg = cmul(kxx, g[i]), dzg = cmul(kzz, g[i]);
    const cplx dxf = cmul(kxx, f[i]), dzf = cmul(kzz, f[i]);
    g[i] = csub(dxf, dzg);
    f[i] = cadd(dxg, dzf);
  }
}

// LABS_integration (module_labs.f90:401-413)
PB_HD cplx integrate(const LineParams &P, const SV &src) {
  cplx t = mk(0.0, 0.0);
  for (int i = 0; i < P.g.n; ++i) t = cadd(t, cscale(src[i], P.integ[i]));
  return t;
}

PB_HD double calc_phi(double bulk, cplx I) { return (bulk - 0.5 * I.re) / (0.5 * I.im); }  // wave00.f90:704-714
// two D1 coefficients at each wall (wave00.f90:210-211, 621-622)
PB_HD double wall_d1(const LineParams &P, const SV &c) {
  const int n = P.g.n, m = P.g.bw;
  return (((P.D1(0, m) * c[0].re + P.D1(0, m + 1) * c[1].re) - P.D1(n - 1, m - 1) * c[n - 2].re) -
          P.D1(n - 1, m) * c[n - 1].re);
}

// The (0,0) system of wave00.f90 is three independent solves; each of the
// three functions below touches only its own columns (and Db), so they may
// run one after another on one scratch Db or side by side on three.
//
// Mean streamwise velocity (wave00.f90:58-257): u1 <- U, st1 updated,
// u2/u3 left as scratch; out4[0] = phi_u, out4[2] = tau_w (last substep).
PB_HD void mean_u(const LineParams &P, const OwnBand &Db, SV u1, SV u2, SV u3, SV st1, SV buf1, double *out4) {
  const int n = P.g.n, s = P.s, S = P.S;
  const double dt = P.dt, nu = P.nu;
  helmholtz_lu(P, Db, 0.0, dt * P.coef(1, s) * nu, true);
  if (s == 1) {
    for (int i = 0; i < n; ++i) u1[i] = st1[i];
    band_bs(P.g, P.D0LU, u1);
    band_bs(P.g, P.D0LU, u2);
    band_mv(P.g, P.D2, u1, buf1);
    band_mv(P.g, P.D1, u2, u3);
    const double alpha = dt * P.coef(0, s) * nu, gamma = dt * P.coef(2, s);
    for (int i = 1; i < n; ++i) buf1[i] = csub(cadd(st1[i], cscale(buf1[i], alpha)), cscale(u3[i], gamma));
  } else {
    band_bs(P.g, P.D0LU, u2);
    band_mv(P.g, P.D1, u2, u3);
    const double gamma = dt * P.coef(2, s);
    for (int i = 1; i < n; ++i) buf1[i] = csub(st1[i], cscale(u3[i], gamma));
  }
  buf1[0] = mk(0.0, 0.0);
  for (int i = 1; i < n - 1; ++i) buf1[i] = mk(buf1[i].re + P.u_gal, 1.0);
  buf1[n - 1] = mk(0.0, 0.0);
  band_bs(P.g, Db, buf1);
  const double phi = calc_phi(P.bulk_u, integrate(P, buf1));
  for (int i = 0; i < n; ++i) buf1[i] = mk(buf1[i].re + phi * buf1[i].im, 0.0);
  out4[0] = phi;
  if (s == S) out4[2] = 0.5 * nu * wall_d1(P, buf1);
  band_mv(P.g, P.D0, buf1, u1);
  for (int i = 0; i < n; ++i) u1[i] = mk(u1[i].re - P.u_gal, u1[i].im);
  if (s != S) {
    band_mv(P.g, P.D2, buf1, st1);
    const double alpha = dt * P.coef(0, s + 1) * nu, zeta = dt * P.coef(3, s + 1);
    for (int i = 0; i < n; ++i) st1[i] = csub(cadd(cscale(st1[i], alpha), u1[i]), cscale(u3[i], zeta));
  } else {
    for (int i = 0; i < n; ++i) st1[i] = u1[i];
  }
}

// Mean spanwise velocity (wave00.f90:260-443, phi forced to 0 at 396): u3 <- W,
// u2 <- 0, st2 updated, u4/u5 left as scratch.  Its matrix is the one of
// mean_u; factor = true forms it again (same values) for a solve that does
// not follow mean_u on the same Db.
PB_HD void mean_w(const LineParams &P, const OwnBand &Db, bool factor, SV u2, SV u3, SV u4, SV u5, SV st2, SV buf1) {
  const int n = P.g.n, s = P.s, S = P.S;
  const double dt = P.dt, nu = P.nu;
  if (factor) helmholtz_lu(P, Db, 0.0, dt * P.coef(1, s) * nu, true);
  if (s == 1) {
    for (int i = 0; i < n; ++i) u2[i] = st2[i];
    band_bs(P.g, P.D0LU, u2);
    band_mv(P.g, P.D2, u2, buf1);
    band_bs(P.g, P.D0LU, u4);
    band_mv(P.g, P.D1, u4, u5);
    const double alpha = dt * P.coef(0, s) * nu, gamma = dt * P.coef(2, s);
    for (int i = 1; i < n; ++i) buf1[i] = csub(cadd(st2[i], cscale(buf1[i], alpha)), cscale(u5[i], gamma));
  } else {
    band_bs(P.g, P.D0LU, u4);
    band_mv(P.g, P.D1, u4, u5);
    const double gamma = dt * P.coef(2, s);
    for (int i = 1; i < n; ++i) buf1[i] = csub(st2[i], cscale(u5[i], gamma));
  }
  buf1[0] = mk(0.0, 0.0);
  for (int i = 1; i < n - 1; ++i) buf1[i] = mk(buf1[i].re, 1.0);
  buf1[n - 1] = mk(0.0, 0.0);
  band_bs(P.g, Db, buf1);
  double phi = calc_phi(P.bulk_w, integrate(P, buf1));
  phi = 0.0;  // wave00.f90:396
  for (int i = 0; i < n; ++i) buf1[i] = mk(buf1[i].re + phi * buf1[i].im, 0.0);
  band_mv(P.g, P.D0, buf1, u3);
  if (s != S) {
    band_mv(P.g, P.D2, buf1, st2);
    const double alpha = dt * P.coef(0, s + 1) * nu, zeta = dt * P.coef(3, s + 1);
    for (int i = 0; i < n; ++i) st2[i] = csub(cadd(cscale(st2[i], alpha), u3[i]), cscale(u5[i], zeta));
  } else {
    for (int i = 0; i < n; ++i) st2[i] = u3[i];
  }
  for (int i = 0; i < n; ++i) u2[i] = mk(0.0, 0.0);
}

// Mean scalar (wave00.f90:473-667): u7 <- Theta, st3 updated, u8/u9 left as
// scratch; out4[1] = phi_t, out4[3] = q_w (last substep).
// usrc (mixed-type scalar condition): the mean u of the substep's input
// state; the source ks U enters as -h, u9 = d(v theta)/dy - ks U, and no bulk
// is held (phi_t = 0).  Null: the reference's held-bulk uniform source.
PB_HD void mixed_source(const SV &u9, const SV &usrc, int n, double ks) {
  for (int i = 0; i < n; ++i) u9[i] = csub(u9[i], cscale(usrc[i], ks));
}
PB_HD void mean_t(const LineParams &P, const OwnBand &Db, SV u7, SV u8, SV u9, SV st3, SV buf1, double *out4,
                  const SV *usrc, double ks) {
  const int n = P.g.n, s = P.s, S = P.S;
  const double dt = P.dt, nu = P.nu, pr = P.pr;
  helmholtz_lu(P, Db, 0.0, dt * P.coef(1, s) * nu / pr, true);
  if (s == 1) {
    for (int i = 0; i < n; ++i) u7[i] = st3[i];
    band_bs(P.g, P.D0LU, u7);
    band_mv(P.g, P.D2, u7, buf1);
    band_bs(P.g, P.D0LU, u8);
    band_mv(P.g, P.D1, u8, u9);
    if (usrc) mixed_source(u9, *usrc, n, ks);
    const double alpha = dt * P.coef(0, s) * nu / pr, gamma = dt * P.coef(2, s);
    for (int i = 1; i < n; ++i) buf1[i] = csub(cadd(st3[i], cscale(buf1[i], alpha)), cscale(u9[i], gamma));
  } else {
    band_bs(P.g, P.D0LU, u8);
    band_mv(P.g, P.D1, u8, u9);
    if (usrc) mixed_source(u9, *usrc, n, ks);
    const double gamma = dt * P.coef(2, s);
    for (int i = 1; i < n; ++i) buf1[i] = csub(st3[i], cscale(u9[i], gamma));
  }
  buf1[0] = mk(0.0, 0.0);
  for (int i = 1; i < n - 1; ++i) buf1[i] = mk(buf1[i].re, usrc ? 0.0 : 1.0);
  buf1[n - 1] = mk(0.0, 0.0);
  band_bs(P.g, Db, buf1);
  const double phi = usrc ? 0.0 : calc_phi(P.bulk_t, integrate(P, buf1));
  for (int i = 0; i < n; ++i) buf1[i] = mk(buf1[i].re + phi * buf1[i].im, 0.0);
  out4[1] = phi;
  if (s == S) out4[3] = 0.5 * nu / pr * wall_d1(P, buf1);
  band_mv(P.g, P.D0, buf1, u7);
  if (s != S) {
    band_mv(P.g, P.D2, buf1, st3);
    const double alpha = dt * P.coef(0, s + 1) * nu / pr, zeta = dt * P.coef(3, s + 1);
    for (int i = 0; i < n; ++i) st3[i] = csub(cadd(cscale(st3[i], alpha), u7[i]), cscale(u9[i], zeta));
  } else {
    for (int i = 0; i < n; ++i) st3[i] = u7[i];
  }
}

// wave00.f90:41-685 with calculate_phi (704-714); out4 = phi_u, phi_t, tau_w, q_w
PB_HD void wave00(const LineParams &P, const OwnBand &Db, SV u1, SV u2, SV u3, SV u4, SV u5, SV u7, SV u8, SV u9,
                  SV st1, SV st2, SV st3, SV buf1, double *out4, const SV *usrc, double ks) {
  mean_u(P, Db, u1, u2, u3, st1, buf1, out4);
  mean_w(P, Db, false, u2, u3, u4, u5, st2, buf1);
  mean_t(P, Db, u7, u8, u9, st3, buf1, out4, usrc, ks);
}

// ------------------------------------------------- statistics-time pressure
// Wall values of the pressure Poisson equation (get_statistics.f90:689-716):
// c = LU(D0)^-1 v (in tmp), bc1 = nu*(D2 row 0 . c_0..2),
// bc2 = nu*(D2 row n-1 . c_n-3..n-1).
PB_HD void ppe_wall_values(const LineParams &P, const SV &v, const SV &tmp, double nu, cplx *bc1, cplx *bc2) {
  const int n = P.g.n, bw = P.g.bw;
  for (int i = 0; i < n; ++i) tmp[i] = v[i];
  band_bs(P.g, P.D0LU, tmp);
  *bc1 = cscale(cadd(cadd(cscale(tmp[0], P.D2(0, bw)), cscale(tmp[1], P.D2(0, bw + 1))), cscale(tmp[2], P.D2(0, bw + 2))),
                nu);
  *bc2 = cscale(cadd(cadd(cscale(tmp[n - 3], P.D2(n - 1, bw - 2)), cscale(tmp[n - 2], P.D2(n - 1, bw - 1))),
                     cscale(tmp[n - 1], P.D2(n - 1, bw))),
                nu);
}

// solve_ppe (get_pressure.f90:55-133) for one (kx,kz) != (0,0) system:
// u1..u5 = u^2-v^2, uv, w^2-v^2, vw, uw are read; u6 = v^2 in, p out.
PB_HD void solve_ppe(const LineParams &P, const OwnBand &Db, SV u1, SV u2, SV u3, SV u4, SV u5, SV u6, cplx bc1,
                     cplx bc2, SV buf1, SV buf2, double kx, double kz) {
  const int n = P.g.n, W = P.g.W();
  const cplx ikx = mk(0.0, kx), ikz = mk(0.0, kz);  // dcmplx(0d0, k)
  const double k2 = kx * kx + kz * kz, kxkz = kx * kz;
  for (int i = 1; i < n - 1; ++i)
    buf1[i] = cadd(cadd(cadd(cscale(u1[i], kx * kx), cscale(u3[i], kz * kz)), cscale(u6[i], k2)),
                   cscale(u5[i], 2.0 * kxkz));
  band_bs(P.g, P.D0LU, u6);
  band_mv(P.g, P.D2, u6, buf2);
  for (int i = 1; i < n - 1; ++i) {
    buf1[i] = csub(buf1[i], buf2[i]);
    u6[i] = cadd(cmul(ikx, u2[i]), cmul(ikz, u4[i]));  // rows 0 and n-1 keep the coefficients above
  }
  band_bs(P.g, P.D0LU, u6);
  band_mv(P.g, P.D1, u6, buf2);
  buf1[0] = bc1;
  for (int i = 1; i < n - 1; ++i) buf1[i] = csub(buf1[i], cscale(buf2[i], 2.0));
  buf1[n - 1] = bc2;
  for (int i = 0; i < W; ++i) Db.at(0, i) = P.D1(0, i);
  for (int j = 1; j < n - 1; ++j)
    for (int i = 0; i < W; ++i) Db.at(j, i) = -(k2 * P.D0(j, i)) + P.D2(j, i);
  for (int i = 0; i < W; ++i) Db.at(n - 1, i) = P.D1(n - 1, i);
  band_lu(P.g, Db);
  band_bs(P.g, Db, buf1);
  band_mv(P.g, P.D0, buf1, u6);
}

// <ab>(y) of get_uv (get_statistics.f90:982-1017) for one y: the modes of
// the plane rows in kz-major, kx-minor order, kx = 0 once, kx > 0 twice.
// Serial over the plane in memory order (kz rows, kx within a row); weight 1
// at kx = 0, 2 otherwise.  Within a row the loads of 8 modes are issued
// before their sums, which keep the serial order.
PB_HD double mean_uv_row(const cplx *a, const cplx *b, int nxh, size_t plane) {
  constexpr int NB = 8;
  double acc = 0.0;
  for (size_t r = 0; r < plane; r += nxh) {
    const cplx *ar = a + r, *br = b + r;
    acc = acc + ar[0].re * br[0].re + ar[0].im * br[0].im;
    int k = 1;
    for (; k + NB <= nxh; k += NB) {
      cplx u[NB], v[NB];
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        u[i] = ar[k + i];
        v[i] = br[k + i];
      }
#pragma unroll
      for (int i = 0; i < NB; ++i) acc = acc + 2.0 * (u[i].re * v[i].re + u[i].im * v[i].im);
    }
    for (; k < nxh; ++k) acc = acc + 2.0 * (ar[k].re * br[k].re + ar[k].im * br[k].im);
  }
  return acc;
}
PB_HD double mean_vv_row(const cplx *row, int nxh, size_t plane) { return mean_uv_row(row, row, nxh, plane); }

// Mean-line profiles of get_statistics_1 (get_statistics.f90:54-86 and the
// gradient routines 546-652): out lines (ny each) 0 U (+u_gal), 1 W,
// 2 Theta, 3 dU/dy, 4 dW/dy, 5 dTheta/dy; centre[0] = u_center (u_gal +
// sum Re(c_k) B_k(0)), centre[1] = t_center.  m[q] are the (0,0) columns of
// u, w, theta; t1, t2 scratch.
// One mean line q (0 U, 1 W, 2 Theta): its profile, its y derivative and
// (U, Theta) its centre value; the three lines are independent.
PB_HD void mean_line_stats_q(const LineParams &P, const SV &m, int q, const SV &t1, const SV &t2,
                             const double *center, double u_gal, double *out, double *centre) {
  const int n = P.g.n;
  for (int y = 0; y < n; ++y) out[(size_t)q * n + y] = m[y].re + (q == 0 ? u_gal : 0.0);
  for (int y = 0; y < n; ++y) t1[y] = m[y];
  band_bs(P.g, P.D0LU, t1);
  band_mv(P.g, P.D1, t1, t2);
  for (int y = 0; y < n; ++y) out[(size_t)(3 + q) * n + y] = t2[y].re;
  if (q != 1) {
    double a = (q == 0) ? u_gal : 0.0;
    for (int k = 0; k < n; ++k) a = a + t1[k].re * center[k];
    centre[q == 0 ? 0 : 1] = a;
  }
}
PB_HD void mean_line_stats(const LineParams &P, const SV m[3], const SV &t1, const SV &t2, const double *center,
                           double u_gal, double *out, double *centre) {
  for (int q = 0; q < 3; ++q) mean_line_stats_q(P, m[q], q, t1, t2, center, u_gal, out, centre);
}

// Detection-plane values of a collocation column, both planes: v(y_d) =
// r . v, ascending rows, with r = LU(D0)^-T B(y_d) formed on the host
// (rows = r_lo then r_hi, n each): the same value as B(y_d) . D0^-1 v, read
// in one pass over v.
PB_HD void opp_dots(const SV &v, const double *rows, int n, cplx *lo, cplx *hi) {
  const cplx x0 = v[0];
  cplx a = cscale(x0, ld_uniform(rows)), b = cscale(x0, ld_uniform(rows + n));
  for (int i = 1; i < n; ++i) {
    const cplx x = v[i];
    a = cadd(a, cscale(x, ld_uniform(rows + i)));
    b = cadd(b, cscale(x, ld_uniform(rows + n + i)));
  }
  *lo = a;
  *hi = b;
}

// Everything one system needs: field base pointers in the [slot][y][kz][kx]
// layout and per-system scratch.
struct WsArgs {
  cplx *U;        // 9 slots: products in, u,v,w,theta (slots 0,1,2,6) out
  cplx *RK;       // 3 slots
  cplx *buf;      // 2 scratch columns per system
  double *dbuf;   // W*ny doubles per system
  double *rbuf;   // one real scratch column per system (fused sweeps only)
  cplx *xcol;     // one complex scratch column per system (fused sweeps only): hvhg's u1a from
                  // sweep A to B_g, then v_p' from F to G
  const cplx *vd; // detection-plane v (2 planes) or nullptr
  // mixed-type scalar condition: u of the substep's input state (one
  // field, device layout) and the source factor (nu/pr)/bulk_u; nullptr for
  // the reference's held-bulk uniform source
  const cplx *usrc;
  double src_scale;
  // wall actuation: one complex value per system added to the wall v of
  // both walls ((kx,kz) != (0,0)), or nullptr
  const cplx *act;
  size_t plane;   // nz*nxh
  size_t field;   // ny*plane
  int nxh, nz, nh;
  double lmfxr, lmfzr, opp_alpha;
  double *mean_out;  // 4 doubles written by the (0,0) system
};

PB_HD void wavespace_line(const LineParams &P, const WsArgs &A, size_t sys) {
  const int k = (int)(sys % A.nxh), j = (int)(sys / A.nxh);
  const int kxi = k, kzi = (j <= A.nh) ? j : j + 1 - 2 * A.nh - 1;  // module_ppdfftw.f90:360-374
  auto col = [&](cplx *base, int slot) { return SV{base + (size_t)slot * A.field + sys, A.plane}; };
  OwnBand Db{A.dbuf + sys, A.plane, P.g.W()};
  SV b1 = col(A.buf, 0), b2 = col(A.buf, 1);
  const SV us{const_cast<cplx *>(A.usrc) + sys, A.plane};
  if (kxi == 0 && kzi == 0) {
    double o4[4] = {0, 0, 0, 0};
    wave00(P, Db, col(A.U, 0), col(A.U, 1), col(A.U, 2), col(A.U, 3), col(A.U, 4), col(A.U, 6), col(A.U, 7),
           col(A.U, 8), col(A.RK, 0), col(A.RK, 1), col(A.RK, 2), b1, o4, A.usrc ? &us : nullptr, A.src_scale);
    if (A.mean_out)
      for (int q = 0; q < 4; ++q) A.mean_out[q] = o4[q];
    return;
  }
  const double kx = kxi * A.lmfxr, kz = kzi * A.lmfzr;
  hvhg(P, col(A.U, 0), col(A.U, 1), col(A.U, 2), col(A.U, 3), col(A.U, 4), col(A.U, 6), col(A.U, 7), col(A.U, 8), kx,
       kz);
  solve_gt(P, Db, col(A.U, 0), col(A.RK, 0), b1, b2, kx, kz, false);
  cplx vlo = mk(0.0, 0.0), vhi = mk(0.0, 0.0);
  if (A.opp_alpha != 0.0 && A.vd) {
    vlo = cscale(A.vd[sys], -A.opp_alpha);
    vhi = cscale(A.vd[A.plane + sys], -A.opp_alpha);
  }
  if (A.act) {
    vlo = cadd(vlo, A.act[sys]);
    vhi = cadd(vhi, A.act[sys]);
  }
  solve_v_f(P, Db, col(A.U, 1), col(A.U, 2), col(A.U, 4), col(A.RK, 1), col(A.U, 3), b1, b2, kx, kz, vlo, vhi);
  solve_u_w(P.g.n, col(A.U, 0), col(A.U, 2), kx, kz);
  if (A.usrc) {  // mixed-type scalar condition: h_theta += ks u
    const SV ht = col(A.U, 6);
    for (int i = 0; i < P.g.n; ++i) ht[i] = cadd(ht[i], cscale(us[i], A.src_scale));
  }
  solve_gt(P, Db, col(A.U, 6), col(A.RK, 2), b1, b2, kx, kz, true);
}

}  // namespace pb
