// solver_fused.h -- the per-(kx,kz) wall-normal solve of one substep,
// reorganised into seven sweeps over y per system.
//
// Each value is produced by exactly the same floating-point expression as
// in the statement-by-statement port (solver.h, which follows hvhg.f90,
// solve_g.f90, solve_v_f.f90, solve_u_w.f90, solve_t.f90 and the packed band
// kernels of module_labs.f90), so the results are bit-identical; only the
// order in which independent values are produced changes:
//
//   A  (down) hvhg point-wise terms + forward substitution of 4 RHS with LU(D0)
//             (P8's in its own kernel A_t, the first of the theta chain)
//   B  (up)   back substitution + D1/D2 products -> h_g, h_v, h_theta, the
//             Helmholtz right-hand sides and the zeta*h part of the stores
//   C  (down) Crout LU of H_nu and H_theta row by row (matrix entries formed
//             on the fly from D0, D2 and k^2) fused with the forward
//             substitution of g, phi_p, phi_h and theta
//   D  (up)   back substitution + D0/D2 products -> g, theta, phi_p, phi_h,
//             D2 phi_h and the stores of g and theta
//   (B, C, D each run as one kernel per right-hand-side group)
//   E  (down) Crout LU of the Poisson operator + forward substitution of v_p, v_h
//   F  (up)   back substitution + D0/D1 products -> v_p, v_p', v_h', D0 c_h
//   G  (mirror pairs) influence coefficients C1, C2, f = v', v, u, w, store_phi
//
// Vectors live in the product slots of the system (layout [y][kz][kx]); a
// back substitution's last 2bw+1 results are held in a register shift
// window, which is also exactly the operand of the band products of the row
// that becomes complete (first(m) = r, last(m) = r + 2bw for every emitted
// row m), so every window index is a compile-time constant.
//
// Slot use per system (0-based product slots; s = substep):
//   after A, A_t: 0 yB, 1 P2, 2 yE, 3 P4, 4 yF, 6 P7, 7 y8, 8 P9; u1a in
//             the extra column X (P7 and P9 are read only by B_t, which forms u7a)
//   after B : 0 rhs_phi, 2 rhs_g, 7 rhs_theta, RK = zeta*h (s < S)
//   after C : 0 y_phip, 2 y_g, 7 y_theta, R0 y_phih; U factors in dbuf, dbuf2
//   after D : 0 phi_p, 2 g, 6 theta, R1 phi_h, R2 D2 phi_h; RK updated
//   after E : 0 y_vp, R3 y_vh; U factor in dbuf
//   after F : 0 v_p, X v_p', R4 v_h', R0 D0 c_h (E, F and G touch no slot of
//             the theta chain, which joins at the end of the stage)
//   after G : 0 u, 1 v, 2 w, 6 theta; RK2 = store_phi
// R0..R4 are real columns (the reference's (1+i)-scaled homogeneous
// solutions are (x, x): only x is stored); R0..R3 live in the memory of the
// two scratch columns, which hold A1's forward substitutions of the stores
// on s == 1 until B, R4 in rbuf.  The U factors and R0..R4 depend on k^2 only and are
// kept once per (kx, +-kz) pair, at the primary system's index; C and E run
// one thread per pair, the other sweeps one thread per system.
// B, C and D are split by right-hand-side group (theta | g | phi,v): the
// groups read disjoint data, so the split costs no traffic and keeps the
// register windows of each kernel small; the theta chain can run beside
// the others.
// s == 1 adds a down sweep (A1: forward substitution of the three stores
// with LU(D0) into buf0, buf1 and slot 5); the B sweeps of s == 1 finish
// D2 LU(D0)^-1 of the stores in their up sweeps (B_g1 for g and phi,
// B_t1 for theta).
#pragma once

#include "solver.h"

namespace pb {

template <int N>
struct CWin {
  cplx w[N];
  PB_HD void clear() {
#pragma unroll
    for (int q = 0; q < N; ++q) w[q] = mk(0.0, 0.0);
  }
  PB_HD void push(cplx x) {
#pragma unroll
    for (int q = N - 1; q > 0; --q) w[q] = w[q - 1];
    w[0] = x;
  }
};

template <int N>
struct RWin {
  double w[N];
  PB_HD void clear() {
#pragma unroll
    for (int q = 0; q < N; ++q) w[q] = 0.0;
  }
  PB_HD void push(double x) {
#pragma unroll
    for (int q = N - 1; q > 0; --q) w[q] = w[q - 1];
    w[0] = x;
  }
};

// strided real column
struct RV {
  double *p;
  size_t s;
  PB_HD double &operator[](int i) const { return p[(size_t)i * s]; }
};

// Per-system context shared by the sweeps.  The systems (kx, kz) and
// (kx, -kz) have the same k^2, hence bit-identical matrices: their LU
// factors and the homogeneous solutions (which depend on k^2 only) are
// formed once, by the pair's primary system prim (kz >= 0), and both
// systems read them there.
struct SysCtx {
  size_t sys, prim;
  int n;
  double kx, kz, k2;
  PB_HD bool is_prim() const { return sys == prim; }
  PB_HD SV col(const WsArgs &A, int slot) const { return SV{A.U + (size_t)slot * A.field + sys, A.plane}; }
  // shared real column k of the pair: 0..3 in the memory of the two scratch
  // columns (the complex scratch view is dead from sweep C on), 4 in rbuf
  PB_HD RV rcol(const WsArgs &A, int k) const {
    double *base = k < 4 ? reinterpret_cast<double *>(A.buf) + (size_t)k * A.field : A.rbuf;
    return RV{base + prim, A.plane};
  }
  PB_HD SV rk(const WsArgs &A, int slot) const { return SV{A.RK + (size_t)slot * A.field + sys, A.plane}; }
  PB_HD SV buf(const WsArgs &A, int slot) const { return SV{A.buf + (size_t)slot * A.field + sys, A.plane}; }
  PB_HD SV xcol(const WsArgs &A) const { return SV{A.xcol + sys, A.plane}; }
};

PB_HD bool make_sys(const WsArgs &A, int n, size_t sys, SysCtx &c) {
  const int k = (int)(sys % A.nxh), j = (int)(sys / A.nxh);
  const int kzi = (j <= A.nh) ? j : j + 1 - 2 * A.nh - 1;  // module_ppdfftw.f90:360-374
  c.sys = sys;
  c.prim = (j <= A.nh) ? sys : (size_t)(A.nz - j) * A.nxh + k;  // kz index nz-j holds -kz
  c.n = n;
  c.kx = k * A.lmfxr;
  c.kz = kzi * A.lmfzr;
  c.k2 = c.kx * c.kx + c.kz * c.kz;
  return !(k == 0 && kzi == 0);
}

// Forward windows hold the last bw results, w.w[q] = y(r-1-q), q < bw.
// Only the last bw rows reach further back (their packed entries start at
// column n-1-2bw); those older values are read back from the column the
// sweep stores y in.
template <int BW, class W, class Col>
PB_HD auto win_or(const W &w, int q, const Col &y, int p) -> decltype(w.w[0]) {
  return q < BW ? w.w[q < BW ? q : 0] : y[p];
}

// Row loops.  Interior rows bw .. n-bw-1 have first(r) = r-bw and
// last(r) = r+bw, so their packed slots are compile-time constants; the
// loops call row(r, RowKind<true>) for them and row(r, RowKind<false>) for
// the boundary rows.  A row's band entries are read into registers before
// they are used (interior rows unconditionally), so the loads of a row issue
// together instead of each waiting behind its own band condition.
template <bool B>
struct RowKind {
  static constexpr bool mid = B;
};
template <int BW, class F>
PB_HD void down_rows(int n, F &&row) {
  const int hi = n - BW, lo = BW < hi ? BW : hi;
  for (int r = 0; r < lo; ++r) row(r, RowKind<false>{});
  for (int r = lo; r < hi; ++r) row(r, RowKind<true>{});
  for (int r = hi > lo ? hi : lo; r < n; ++r) row(r, RowKind<false>{});
}
template <int BW, class F>
PB_HD void up_rows(int n, F &&row) {
  const int hi = n - BW, lo = BW < hi ? BW : hi;
  for (int r = n - 1; r >= hi && r >= lo; --r) row(r, RowKind<false>{});
  for (int r = hi - 1; r >= lo; --r) row(r, RowKind<true>{});
  for (int r = lo - 1; r >= 0; --r) row(r, RowKind<false>{});
}

// lr[q] = L(r, r-1-q) for the columns p >= first(r) of row r (others unused),
// lr[2bw] = L(r, r)
template <int BW, bool MID, class M>
PB_HD void load_lrow(const BandGeom &g, const M &L, int r, double (&lr)[2 * BW + 1]) {
  if (MID) {
#pragma unroll
    for (int q = 0; q < BW; ++q) lr[q] = L(r, BW - 1 - q);
#pragma unroll
    for (int q = BW; q < 2 * BW; ++q) lr[q] = 0.0;
  } else {
    const int f = g.first(r);
#pragma unroll
    for (int q = 0; q < 2 * BW; ++q) {
      const int p = r - 1 - q;
      lr[q] = p >= f ? L(r, g.slot(r, p)) : 0.0;
    }
  }
  lr[2 * BW] = L(r, BW);
}
// forward substitution of one row, farthest first
template <int BW, bool MID>
PB_HD cplx fwd_row(const BandGeom &g, int r, const double (&lr)[2 * BW + 1], const CWin<BW> &w, const SV &y, cplx b) {
  const int f = g.first(r);
  cplx v = b;
#pragma unroll
  for (int q = 2 * BW - 1; q >= 0; --q) {
    const int p = r - 1 - q;
    if (MID ? q < BW : p >= f) v = csub(v, cscale(win_or<BW>(w, q, y, p), lr[q]));
  }
  return cdivr(v, lr[2 * BW]);
}

// ur[q] = U(r, r+1+q) for the columns c <= last(r) of row r (others unused)
template <int BW, bool MID, class M>
PB_HD void load_urow(const BandGeom &g, const M &U, int r, double (&ur)[2 * BW]) {
  if (MID) {
#pragma unroll
    for (int q = 0; q < BW; ++q) ur[q] = U(r, BW + 1 + q);
#pragma unroll
    for (int q = BW; q < 2 * BW; ++q) ur[q] = 0.0;
  } else {
    const int l = g.last(r);
#pragma unroll
    for (int q = 0; q < 2 * BW; ++q) {
      const int c = r + 1 + q;
      ur[q] = c <= l ? U(r, g.slot(r, c)) : 0.0;
    }
  }
}
// back substitution of one row: w.w[q] = x(r+1+q); nearest first
template <int BW, bool MID>
PB_HD cplx bwd_row(const BandGeom &g, int r, const double (&ur)[2 * BW], const CWin<2 * BW + 1> &w, cplx y) {
  const int l = g.last(r);
  cplx v = y;
#pragma unroll
  for (int q = 0; q < 2 * BW; ++q) {
    if (MID ? q < BW : r + 1 + q <= l) v = csub(v, cscale(w.w[q], ur[q]));
  }
  return v;
}
template <int BW, bool MID>
PB_HD double bwd_row_r(const BandGeom &g, int r, const double (&ur)[2 * BW], const RWin<2 * BW + 1> &w, double y) {
  const int l = g.last(r);
  double v = y;
#pragma unroll
  for (int q = 0; q < 2 * BW; ++q) {
    if (MID ? q < BW : r + 1 + q <= l) v = v - w.w[q] * ur[q];
  }
  return v;
}

// band product of an emitted row m whose columns are r .. r+2bw, window w.w[q] = x(r+q)
template <int BW>
PB_HD cplx mv_row(const BandGeom &g, const SharedBand &A, int m, int r, const CWin<2 * BW + 1> &w) {
  cplx v = cscale(w.w[0], A(m, g.slot(m, r)));
#pragma unroll
  for (int q = 1; q <= 2 * BW; ++q) v = cadd(v, cscale(w.w[q], A(m, g.slot(m, r + q))));
  return v;
}
template <int BW>
PB_HD double mv_row_r(const BandGeom &g, const SharedBand &A, int m, int r, const RWin<2 * BW + 1> &w) {
  double v = w.w[0] * A(m, g.slot(m, r));
#pragma unroll
  for (int q = 1; q <= 2 * BW; ++q) v = v + w.w[q] * A(m, g.slot(m, r + q));
  return v;
}

// Rows whose band product becomes complete after x(r) is produced in an
// upward sweep: calls emit(m) for each (middle row r+bw; the last bw rows at
// r = n-1-2bw; the first bw rows at r = 0).
template <int BW, class F>
PB_HD void for_completed_rows(int n, int r, F &&emit) {
  if (r <= n - 1 - 2 * BW) emit(r + BW);
  // boundary rows: only m varies, the window indices stay compile-time constants
  if (r == n - 1 - 2 * BW) {
#pragma nounroll
    for (int t = 0; t < BW; ++t) emit(n - BW + t);
  }
  if (r == 0) {
#pragma nounroll
    for (int t = 0; t < BW; ++t) emit(t);
  }
}

// A system's own band factor: U entries of row p at p[(p*W + slot)*s]
struct OwnU {
  double *p;
  size_t s;
  int W;
  PB_HD double operator()(int r, int slot) const { return p[(size_t)(r * W + slot) * s]; }
  PB_HD void set(int r, int slot, double v) const { p[(size_t)(r * W + slot) * s] = v; }
};
// The sweeps' U factors in blocks of UBLK consecutive systems (one wave):
// [block][row][slot][UBLK].  A slot's values of a wave are one 512-byte
// segment, and all factors of a wave lie within n*W*512 bytes (a few memory
// pages) rather than one page per (row, slot).
constexpr int UBLK = 64;
PB_HD size_t ublk_elems(size_t plane, int n, int W) { return (plane + UBLK - 1) / UBLK * UBLK * (size_t)n * W; }
PB_HD OwnU own_u(double *base, size_t sys, int n, int W) {
  return OwnU{base + (sys / UBLK) * ((size_t)n * W * UBLK) + sys % UBLK, UBLK, W};
}

// Matrix entries formed on the fly.
struct HelmholtzA {  // (1 + beta k2) D0 - beta D2 with diagonal-only wall rows
  SharedBand D0, D2;
  double dummy, beta;
  int n, bw;
  PB_HD double operator()(int r, int slot) const {
    if (r == 0 || r == n - 1) return slot == bw ? dummy * D0(r, bw) - beta * D2(r, bw) : 0.0;
    return interior(r, slot);
  }
  PB_HD double interior(int r, int slot) const { return dummy * D0(r, slot) - beta * D2(r, slot); }
};
struct PoissonA {  // D2 - k2 D0 with D0-diagonal wall rows
  SharedBand D0, D2;
  double k2;
  int n, bw;
  PB_HD double operator()(int r, int slot) const {
    if (r == 0 || r == n - 1) return slot == bw ? D0(r, bw) : 0.0;
    return interior(r, slot);
  }
  PB_HD double interior(int r, int slot) const { return -k2 * D0(r, slot) + D2(r, slot); }
};

// One Crout row (fixed_bandec*): Lrow[q] = L(r, r-q); U(r, c) stored into Uo.
template <int BW, class M>
PB_HD void crout_row(const BandGeom &g, const M &A, const OwnU &Uo, int r, double (&Lrow)[2 * BW + 1]) {
  const int f = g.first(r);
#pragma unroll
  for (int qd = 2 * BW; qd >= 0; --qd) {
    const int c = r - qd;
    if (c >= f) {
      double v = A(r, g.slot(r, c));
#pragma unroll
      for (int qp = qd + 1; qp <= 2 * BW; ++qp) {
        const int p = r - qp;
        if (p >= f && c <= g.last(p)) v = v - Lrow[qp] * Uo(p, g.slot(p, c));
      }
      Lrow[qd] = v;
    } else {
      Lrow[qd] = 0.0;
    }
  }
  const double inv = 1.0 / Lrow[0];
#pragma unroll
  for (int qc = 1; qc <= 2 * BW; ++qc) {
    const int c = r + qc;
    if (c <= g.last(r)) {
      double v = A(r, g.slot(r, c));
#pragma unroll
      for (int qp = 1; qp <= 2 * BW; ++qp) {
        const int p = r - qp;
        if (p >= f && c <= g.last(p)) v = v - Lrow[qp] * Uo(p, g.slot(p, c));
      }
      Uo.set(r, g.slot(r, c), inv * v);
    }
  }
}

// Crout row of an interior row r in [2bw, n-bw): first(r) = r-bw,
// last(r) = r+bw and the bw rows above are interior as well, so every U(p,c)
// the row needs is in the register window uw[qp-1][d-1] = U(r-qp, r-qp+d),
// qp, d = 1..bw.  The terms and their order are those of crout_row (the
// terms left out are exactly the ones its conditions exclude).  The new U
// row is stored and shifted into the window.
template <int BW, class M>
PB_HD void crout_row_mid(const M &A, const OwnU &Uo, int r, double (&Lrow)[2 * BW + 1], double (&uw)[BW][BW]) {
#pragma unroll
  for (int qd = 2 * BW; qd > BW; --qd) Lrow[qd] = 0.0;
#pragma unroll
  for (int qd = BW; qd >= 0; --qd) {
    double v = A.interior(r, BW - qd);
#pragma unroll
    for (int qp = qd + 1; qp <= BW; ++qp) v = v - Lrow[qp] * uw[qp - 1][qp - qd - 1];
    Lrow[qd] = v;
  }
  const double inv = 1.0 / Lrow[0];
  double un[BW];
#pragma unroll
  for (int qc = 1; qc <= BW; ++qc) {
    double v = A.interior(r, BW + qc);
#pragma unroll
    for (int qp = 1; qp + qc <= BW; ++qp) v = v - Lrow[qp] * uw[qp - 1][qp + qc - 1];
    un[qc - 1] = inv * v;
    Uo.set(r, BW + qc, un[qc - 1]);
  }
#pragma unroll
  for (int qp = BW - 1; qp > 0; --qp)
#pragma unroll
    for (int d = 0; d < BW; ++d) uw[qp][d] = uw[qp - 1][d];
#pragma unroll
  for (int d = 0; d < BW; ++d) uw[0][d] = un[d];
}

// Crout LU of a whole system, row by row, calling row(r, Lrow, kind) after
// each row: the first 2bw and the last bw rows (whose packed entries reach
// further) read U back from Uo, the rows between (kind RowKind<true>: no wall
// row, first(r) = r-bw) keep it in registers.
template <int BW, class M, class F>
PB_HD void crout_sweep(const BandGeom &g, const M &A, const OwnU &Uo, F &&row) {
  const int n = g.n;
  const bool mid = n > 3 * BW;
  const int r1 = mid ? 2 * BW : n, r2 = mid ? n - BW : n;
  double L[2 * BW + 1];
  for (int r = 0; r < r1; ++r) {
    crout_row<BW>(g, A, Uo, r, L);
    row(r, L, RowKind<false>{});
  }
  if (mid) {
    double uw[BW][BW];
#pragma unroll
    for (int qp = 1; qp <= BW; ++qp)
#pragma unroll
      for (int d = 1; d <= BW; ++d) uw[qp - 1][d - 1] = Uo(r1 - qp, BW + d);
    for (int r = r1; r < r2; ++r) {
      crout_row_mid<BW>(A, Uo, r, L, uw);
      row(r, L, RowKind<true>{});
    }
  }
  for (int r = r2; r < n; ++r) {
    crout_row<BW>(g, A, Uo, r, L);
    row(r, L, RowKind<false>{});
  }
}

// forward substitution with the current row's L (Lrow[q] = L(r, r-q))
// (interior rows, MID: the terms q <= bw, all from the window)
template <int BW, bool MID>
PB_HD cplx fwd_lrow(const BandGeom &g, int r, const double (&Lrow)[2 * BW + 1], const CWin<BW> &w, const SV &y,
                    cplx b) {
  const int f = g.first(r);
  cplx v = b;
#pragma unroll
  for (int q = 2 * BW; q >= 1; --q) {
    if (MID ? q <= BW : r - q >= f) v = csub(v, cscale(win_or<BW>(w, q - 1, y, r - q), Lrow[q]));
  }
  return cdivr(v, Lrow[0]);
}
template <int BW, bool MID>
PB_HD double fwd_lrow_r(const BandGeom &g, int r, const double (&Lrow)[2 * BW + 1], const RWin<BW> &w, const RV &y,
                        double b) {
  const int f = g.first(r);
  double v = b;
#pragma unroll
  for (int q = 2 * BW; q >= 1; --q) {
    if (MID ? q <= BW : r - q >= f) v = v - win_or<BW>(w, q - 1, y, r - q) * Lrow[q];
  }
  return v / Lrow[0];
}

// ------------------------------------------------------------------ sweeps
// s == 1 only: scratch d2S_i = D2 LU(D0)^-1 RK_i  (solve_g.f90:97-98, solve_v_f.f90:72-73, solve_t.f90:97-98)
template <int BW>
PB_HD void sweep_A1(const LineParams &P, const WsArgs &A, const SysCtx &c) {
  const BandGeom g{c.n, BW};
  SV r1 = c.rk(A, 0), r2 = c.rk(A, 1), r3 = c.rk(A, 2);
  SV o1 = c.buf(A, 0), o2 = c.buf(A, 1), o3 = c.col(A, 5);
  CWin<BW> w1, w2, w3;
  w1.clear(); w2.clear(); w3.clear();
  down_rows<BW>(c.n, [&](int r, auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    double lr[2 * BW + 1];
    load_lrow<BW, MID>(g, P.D0LU, r, lr);
    const cplx y1 = fwd_row<BW, MID>(g, r, lr, w1, o1, r1[r]);
    const cplx y2 = fwd_row<BW, MID>(g, r, lr, w2, o2, r2[r]);
    const cplx y3 = fwd_row<BW, MID>(g, r, lr, w3, o3, r3[r]);
    o1[r] = y1; o2[r] = y2; o3[r] = y3;
    w1.push(y1); w2.push(y2); w3.push(y3);
  });
}

// A: hvhg point-wise terms (hvhg.f90:169-204, 256-258) and forward
// substitution of u3a, u3b, u4b with LU(D0).  u7a (from P7 and P9) is formed
// where it is used, in B_t, and P8's forward substitution is the theta
// chain's own first sweep (A_t): A reads only P1-P5.
template <int BW>
PB_HD void sweep_A(const LineParams &P, const WsArgs &A, const SysCtx &c) {
  const BandGeom g{c.n, BW};
  const double kx = c.kx, kz = c.kz;
  const double kz2 = kz * kz, kx2 = kx * kx, kxkz = kx * kz;
  const cplx ikx = iunit(kx), ikz = iunit(kz);
  SV s0 = c.col(A, 0), s1 = c.col(A, 1), s2 = c.col(A, 2), s3 = c.col(A, 3), s4 = c.col(A, 4), su = c.xcol(A);
  CWin<BW> wB, wE, wF;
  wB.clear(); wE.clear(); wF.clear();
  down_rows<BW>(c.n, [&](int r, auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    const cplx p1 = s0[r], p2 = s1[r], p3 = s2[r], p4 = s3[r], p5 = s4[r];
    double lr[2 * BW + 1];
    load_lrow<BW, MID>(g, P.D0LU, r, lr);
    const cplx u1a = cadd(cscale(csub(p1, p3), kxkz), cscale(p5, kz2 - kx2));
    const cplx u3a = csub(csub(cscale(p1, -kx2), cscale(p5, 2.0 * kxkz)), cscale(p3, kz2));
    const cplx u3b = csub(cmul(ikx, p4), cmul(ikz, p2));
    const cplx u4b = cadd(cmul(ikz, p4), cmul(ikx, p2));
    const cplx yB = fwd_row<BW, MID>(g, r, lr, wB, s0, u3a);
    const cplx yE = fwd_row<BW, MID>(g, r, lr, wE, s2, u3b);
    const cplx yF = fwd_row<BW, MID>(g, r, lr, wF, s4, u4b);
    s0[r] = yB; s2[r] = yE; s4[r] = yF; su[r] = u1a;
    wB.push(yB); wE.push(yE); wF.push(yF);
  });
}

// A_t: forward substitution of P8 with LU(D0) (hvhg.f90:256-258's operand),
// the first sweep of the theta chain
template <int BW>
PB_HD void sweep_At(const LineParams &P, const WsArgs &A, const SysCtx &c) {
  const BandGeom g{c.n, BW};
  SV s7 = c.col(A, 7);
  CWin<BW> w8;
  w8.clear();
  down_rows<BW>(c.n, [&](int r, auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    double lr[2 * BW + 1];
    load_lrow<BW, MID>(g, P.D0LU, r, lr);
    const cplx y8 = fwd_row<BW, MID>(g, r, lr, w8, s7, s7[r]);
    s7[r] = y8;
    w8.push(y8);
  });
}

// B, split by right-hand-side group (disjoint inputs, so the split moves no
// extra data): back substitution with LU(D0), hvhg's band products, the
// Helmholtz right-hand sides (solve_g.f90:84-123, solve_v_f.f90:56-95,
// solve_t.f90:84-123) and the zeta*h part of the stores (solve_g.f90:138-143,
// solve_v_f.f90:115-120).
//   B_g: h_g = u1a + D1 c_E          -> rhs_g in slot 2, store_g = zeta h_g
//   B_t: h_t = -(u7a + D1 c_8)       -> rhs_t in slot 7, store_t = zeta h_t
//   B_v: h_v = ((D1 c_B + a P2) + b P4) + D2 c_F -> rhs_phi in slot 0
struct RhsCoef {
  double gamma, al, du, zeta;
};
PB_HD RhsCoef rhs_coef(const LineParams &P, double k2, bool scalar) {
  RhsCoef r;
  const int s = P.s;
  r.gamma = P.dt * P.coef(2, s);
  r.al = scalar ? P.dt * P.coef(0, s) * P.nu / P.pr : P.dt * P.coef(0, s) * P.nu;
  r.du = 1.0 - r.al * k2;
  r.zeta = (s != P.S) ? P.dt * P.coef(3, s + 1) : 0.0;
  return r;
}
// right-hand side of one row and the zeta*h store; d2s = (D2 LU(D0)^-1 store)(m)
// on substep 1
PB_HD cplx rhs_row(const LineParams &P, const RhsCoef &q, int m, int n, SV &store, cplx d2s, cplx h) {
  cplx rhs;
  if (P.s == 1)
    rhs = cadd(cadd(cscale(store[m], q.du), cscale(h, q.gamma)), cscale(d2s, q.al));
  else
    rhs = cadd(store[m], cscale(h, q.gamma));
  if (m == 0 || m == n - 1) rhs = mk(0.0, 0.0);
  if (P.s != P.S) store[m] = cscale(h, q.zeta);
  return rhs;
}

// S1 (substep 1): the same up sweep also back-substitutes A1's forward
// substitutions of the g and phi stores (y1, y2) and forms D2 LU(D0)^-1 of
// them row by row: the g part is consumed at once, the phi part replaces y2
// in place for B_phi (a row m is emitted only after y2[m] has been read).
template <int BW, bool S1>
PB_HD void sweep_Bg(const LineParams &P, const WsArgs &A, const SysCtx &c) {
  const BandGeom g{c.n, BW};
  const int n = c.n;
  const RhsCoef q = rhs_coef(P, c.k2, false);
  SV s2 = c.col(A, 2), su = c.xcol(A), r1 = c.rk(A, 0), y1 = c.buf(A, 0), y2 = c.buf(A, 1);
  CWin<2 * BW + 1> cE, c1, c2;
  cE.clear(); c1.clear(); c2.clear();
  up_rows<BW>(n, [&](int r, auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    double ur[2 * BW];
    load_urow<BW, MID>(g, P.D0LU, r, ur);
    cE.push(bwd_row<BW, MID>(g, r, ur, cE, s2[r]));
    if (S1) {
      c1.push(bwd_row<BW, MID>(g, r, ur, c1, y1[r]));
      c2.push(bwd_row<BW, MID>(g, r, ur, c2, y2[r]));
    }
    for_completed_rows<BW>(n, r, [&](int m) {
      const cplx hg = cadd(su[m], mv_row<BW>(g, P.D1, m, r, cE));
      const cplx d2s = S1 ? mv_row<BW>(g, P.D2, m, r, c1) : mk(0.0, 0.0);
      if (S1) y2[m] = mv_row<BW>(g, P.D2, m, r, c2);
      s2[m] = rhs_row(P, q, m, n, r1, d2s, hg);
    });
  });
}

template <int BW, bool S1>
PB_HD void sweep_Bt(const LineParams &P, const WsArgs &A, const SysCtx &c) {
  const BandGeom g{c.n, BW};
  const int n = c.n;
  const RhsCoef q = rhs_coef(P, c.k2, true);
  SV s6 = c.col(A, 6), s7 = c.col(A, 7), s8 = c.col(A, 8), r3 = c.rk(A, 2), y3 = c.col(A, 5);
  const cplx ikx = iunit(c.kx), ikz = iunit(c.kz);
  CWin<2 * BW + 1> c8, c3;
  c8.clear(); c3.clear();
  up_rows<BW>(n, [&](int r, auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    double ur[2 * BW];
    load_urow<BW, MID>(g, P.D0LU, r, ur);
    c8.push(bwd_row<BW, MID>(g, r, ur, c8, s7[r]));
    if (S1) c3.push(bwd_row<BW, MID>(g, r, ur, c3, y3[r]));
    for_completed_rows<BW>(n, r, [&](int m) {
      const cplx u7a = cadd(cmul(ikx, s6[m]), cmul(ikz, s8[m]));
      cplx ht = cneg(cadd(u7a, mv_row<BW>(g, P.D1, m, r, c8)));
      if (A.usrc) ht = cadd(ht, cscale(A.usrc[c.sys + (size_t)m * A.plane], A.src_scale));  // mixed-type source
      const cplx d2s = S1 ? mv_row<BW>(g, P.D2, m, r, c3) : mk(0.0, 0.0);
      s7[m] = rhs_row(P, q, m, n, r3, d2s, ht);
    });
  });
}

template <int BW>
PB_HD void sweep_Bv(const LineParams &P, const WsArgs &A, const SysCtx &c) {
  const BandGeom g{c.n, BW};
  const int n = c.n;
  const double kx = c.kx, kz = c.kz;
  const cplx ikx3 = iunit(kx * kx * kx), ikxkz2 = iunit(kx * (kz * kz));
  const cplx ikx2kz = iunit((kx * kx) * kz), ikz3 = iunit(kz * kz * kz);
  const cplx a = cadd(ikx3, ikxkz2), b = cadd(ikx2kz, ikz3);
  const RhsCoef q = rhs_coef(P, c.k2, false);
  SV s0 = c.col(A, 0), s1 = c.col(A, 1), s3 = c.col(A, 3), s4 = c.col(A, 4), r2 = c.rk(A, 1), d2 = c.buf(A, 1);
  CWin<2 * BW + 1> cB, cF;
  cB.clear(); cF.clear();
  up_rows<BW>(n, [&](int r, auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    double ur[2 * BW];
    load_urow<BW, MID>(g, P.D0LU, r, ur);
    cB.push(bwd_row<BW, MID>(g, r, ur, cB, s0[r]));
    cF.push(bwd_row<BW, MID>(g, r, ur, cF, s4[r]));
    for_completed_rows<BW>(n, r, [&](int m) {
      const cplx u5b = cadd(cadd(mv_row<BW>(g, P.D1, m, r, cB), cmul(a, s1[m])), cmul(b, s3[m]));
      const cplx hv = cadd(u5b, mv_row<BW>(g, P.D2, m, r, cF));
      // s == 1: D2 LU(D0)^-1 of the phi store, formed by B_g1
      s0[m] = rhs_row(P, q, m, n, r2, P.s == 1 ? d2[m] : mk(0.0, 0.0), hv);
    });
  });
}

// C, split by matrix: Crout LU row by row fused with the forward
// substitutions of that matrix's right-hand sides (solve_g.f90:42-71, 158;
// solve_v_f.f90:142-148; solve_t.f90:42-71, 158).
//   C_nu: H_nu; rhs_g (slot 2), rhs_phi (slot 0), phi_h seed -> slot 1
//   C_t : H_theta; rhs_t (slot 7)
// Systems handled by one thread of a factorising sweep: the primary
// (kz >= 0) and, for 0 < kz index < nz/2, its partner at -kz.
struct PairCtx {
  SysCtx a, b;
  bool two;
};
PB_HD bool make_pair(const WsArgs &A, int n, size_t t, PairCtx &pc) {
  const int k = (int)(t % A.nxh), j = (int)(t / A.nxh);  // j = 0 .. nz/2
  if (!make_sys(A, n, (size_t)j * A.nxh + k, pc.a)) return false;
  pc.two = j >= 1 && j < A.nh;
  if (pc.two) make_sys(A, n, (size_t)(A.nz - j) * A.nxh + k, pc.b);
  else pc.b = pc.a;
  return true;
}

template <int BW, bool TWO>
PB_HD void sweep_Cnu_t(const LineParams &P, const WsArgs &A, const PairCtx &pc) {
  const SysCtx &c = pc.a;
  const BandGeom g{c.n, BW};
  const int n = c.n;
  const double be = P.dt * P.coef(1, P.s) * P.nu;
  const HelmholtzA H{P.D0, P.D2, 1.0 + be * c.k2, be, n, BW};
  const OwnU U = own_u(A.dbuf, c.prim, c.n, 2 * BW + 1);
  SV s0 = c.col(A, 0), s2 = c.col(A, 2), t0 = pc.b.col(A, 0), t2 = pc.b.col(A, 2);
  RV h1 = c.rcol(A, 0);
  CWin<BW> wg, wp, vg, vp;
  RWin<BW> wh;
  wg.clear(); wp.clear(); vg.clear(); vp.clear(); wh.clear();
  double d00 = 0.0;
  crout_sweep<BW>(g, H, U, [&](int r, const double(&L)[2 * BW + 1], auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    const bool r0 = !MID && r == 0;
    if (r0) d00 = L[0];
    const cplx yg = fwd_lrow<BW, MID>(g, r, L, wg, s2, s2[r]);
    const cplx yp = fwd_lrow<BW, MID>(g, r, L, wp, s0, s0[r]);
    const double yh = fwd_lrow_r<BW, MID>(g, r, L, wh, h1, r0 ? d00 : 0.0);
    s2[r] = yg; s0[r] = yp; h1[r] = yh;
    wg.push(yg); wp.push(yp); wh.push(yh);
    if (TWO) {
      const cplx zg = fwd_lrow<BW, MID>(g, r, L, vg, t2, t2[r]);
      const cplx zp = fwd_lrow<BW, MID>(g, r, L, vp, t0, t0[r]);
      t2[r] = zg; t0[r] = zp;
      vg.push(zg); vp.push(zp);
    }
  });
}
template <int BW>
PB_HD void sweep_Cnu(const LineParams &P, const WsArgs &A, const PairCtx &pc) {
  if (pc.two) sweep_Cnu_t<BW, true>(P, A, pc);
  else sweep_Cnu_t<BW, false>(P, A, pc);
}

template <int BW, bool TWO>
PB_HD void sweep_Ct_t(const LineParams &P, const WsArgs &A, const PairCtx &pc, double *dbuf2) {
  const SysCtx &c = pc.a;
  const BandGeom g{c.n, BW};
  const int n = c.n;
  const double be = P.dt * P.coef(1, P.s) * P.nu / P.pr;
  const HelmholtzA H{P.D0, P.D2, 1.0 + be * c.k2, be, n, BW};
  const OwnU U = own_u(dbuf2, c.prim, c.n, 2 * BW + 1);
  SV s7 = c.col(A, 7), t7 = pc.b.col(A, 7);
  CWin<BW> wt, vt;
  wt.clear(); vt.clear();
  crout_sweep<BW>(g, H, U, [&](int r, const double(&L)[2 * BW + 1], auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    const cplx yt = fwd_lrow<BW, MID>(g, r, L, wt, s7, s7[r]);
    s7[r] = yt;
    wt.push(yt);
    if (TWO) {
      const cplx zt = fwd_lrow<BW, MID>(g, r, L, vt, t7, t7[r]);
      t7[r] = zt;
      vt.push(zt);
    }
  });
}
template <int BW>
PB_HD void sweep_Ct(const LineParams &P, const WsArgs &A, const PairCtx &pc, double *dbuf2) {
  if (pc.two) sweep_Ct_t<BW, true>(P, A, pc, dbuf2);
  else sweep_Ct_t<BW, false>(P, A, pc, dbuf2);
}

// D, split by matrix: back substitution and the D0/D2 products
// (solve_g.f90:158-185, solve_v_f.f90:150-181, solve_t.f90:158-185).
//   D_nu: g -> slot 2, store_g; phi_p -> slot 0, phi_h -> slot 1,
//         D2 phi_h -> slot 3, the phi_p part of store_phi
//   D_t : theta -> slot 6, store_t
template <int BW>
PB_HD void sweep_Dnu(const LineParams &P, const WsArgs &A, const SysCtx &c) {
  const BandGeom g{c.n, BW};
  const int n = c.n, s = P.s, S = P.S;
  const double dt = P.dt, nu = P.nu, k2 = c.k2;
  const OwnU U = own_u(A.dbuf, c.prim, c.n, 2 * BW + 1);
  double al_g = 0, du_g = 0, al_p = 0, du_p = 0;
  if (s != S) {
    al_g = dt * P.coef(0, s + 1) * nu;
    du_g = 1.0 - al_g * k2;
    al_p = dt * nu * P.coef(0, s + 1);
    du_p = 1.0 - al_p * k2;
  }
  SV s0 = c.col(A, 0), s2 = c.col(A, 2);
  RV y1 = c.rcol(A, 0), h1 = c.rcol(A, 1), h3 = c.rcol(A, 2);
  SV r1 = c.rk(A, 0), r2 = c.rk(A, 1);
  CWin<2 * BW + 1> wg, wp;
  RWin<2 * BW + 1> wh;
  wg.clear(); wp.clear(); wh.clear();
  up_rows<BW>(n, [&](int r, auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    double ur[2 * BW];
    load_urow<BW, MID>(g, U, r, ur);
    wg.push(bwd_row<BW, MID>(g, r, ur, wg, s2[r]));
    wp.push(bwd_row<BW, MID>(g, r, ur, wp, s0[r]));
    wh.push(bwd_row_r<BW, MID>(g, r, ur, wh, y1[r]));
    for_completed_rows<BW>(n, r, [&](int m) {
      const cplx gv = mv_row<BW>(g, P.D0, m, r, wg);
      const cplx d2g = mv_row<BW>(g, P.D2, m, r, wg);
      const cplx pv = mv_row<BW>(g, P.D0, m, r, wp);
      const cplx d2p = mv_row<BW>(g, P.D2, m, r, wp);
      const double hv = mv_row_r<BW>(g, P.D0, m, r, wh);
      const double d2h = mv_row_r<BW>(g, P.D2, m, r, wh);
      if (s != S) {
        r1[m] = cadd(cadd(r1[m], cscale(gv, du_g)), cscale(d2g, al_g));
        r2[m] = cadd(cadd(r2[m], cscale(pv, du_p)), cscale(d2p, al_p));
      } else {
        r1[m] = gv;
        r2[m] = pv;
      }
      s2[m] = gv; s0[m] = pv;
      if (c.is_prim()) { h1[m] = hv; h3[m] = d2h; }
    });
  });
}

template <int BW>
PB_HD void sweep_Dt(const LineParams &P, const WsArgs &A, const SysCtx &c, double *dbuf2) {
  const BandGeom g{c.n, BW};
  const int n = c.n, s = P.s, S = P.S;
  const OwnU U = own_u(dbuf2, c.prim, c.n, 2 * BW + 1);
  double al = 0, du = 0;
  if (s != S) {
    al = P.dt * P.coef(0, s + 1) * P.nu / P.pr;
    du = 1.0 - al * c.k2;
  }
  SV s6 = c.col(A, 6), s7 = c.col(A, 7), r3 = c.rk(A, 2);
  CWin<2 * BW + 1> wt;
  wt.clear();
  up_rows<BW>(n, [&](int r, auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    double ur[2 * BW];
    load_urow<BW, MID>(g, U, r, ur);
    wt.push(bwd_row<BW, MID>(g, r, ur, wt, s7[r]));
    for_completed_rows<BW>(n, r, [&](int m) {
      const cplx tv = mv_row<BW>(g, P.D0, m, r, wt);
      if (s != S) {
        r3[m] = cadd(cadd(r3[m], cscale(tv, du)), cscale(mv_row<BW>(g, P.D2, m, r, wt), al));
      } else {
        r3[m] = tv;
      }
      s6[m] = tv;
    });
  });
}

// E: LU of the Poisson operator fused with the forward substitution of
// v_p (wall values -alpha v(y_d) with opposition control, else 0, plus the
// wall actuation) and v_h
// (solve_v_f.f90:205-262), for the pair of systems sharing the operator
template <int BW, bool TWO>
PB_HD void sweep_E_t(const LineParams &P, const WsArgs &A, const PairCtx &pc) {
  const SysCtx &c = pc.a, &d = pc.b;
  const BandGeom g{c.n, BW};
  const int n = c.n;
  const PoissonA Q{P.D0, P.D2, c.k2, n, BW};
  const OwnU Uq = own_u(A.dbuf, c.prim, c.n, 2 * BW + 1);
  const bool opp = A.opp_alpha != 0.0 && A.vd;
  SV s0 = c.col(A, 0), t0 = d.col(A, 0);
  RV h1 = c.rcol(A, 1), h5 = c.rcol(A, 3);
  CWin<BW> wp, vp;
  RWin<BW> wh;
  wp.clear(); vp.clear(); wh.clear();
  crout_sweep<BW>(g, Q, Uq, [&](int r, const double(&L)[2 * BW + 1], auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    cplx bp = s0[r], bq = TWO ? t0[r] : mk(0.0, 0.0);
    double bh = h1[r];
    if (!MID && (r == 0 || r == n - 1)) {
      // wall rows: the value is read here rather than held in a variable
      // chosen by r, which the compiler would turn into a private-memory
      // address select
      const size_t off = (r == 0 ? 0 : A.plane);
      bp = opp ? cscale(A.vd[off + c.sys], -A.opp_alpha) : mk(0.0, 0.0);
      if (TWO) bq = opp ? cscale(A.vd[off + d.sys], -A.opp_alpha) : mk(0.0, 0.0);
      if (A.act) {  // wall actuation, the same value at both walls
        bp = cadd(bp, A.act[c.sys]);
        if (TWO) bq = cadd(bq, A.act[d.sys]);
      }
      bh = 0.0;
    }
    const cplx yp = fwd_lrow<BW, MID>(g, r, L, wp, s0, bp);
    const double yh = fwd_lrow_r<BW, MID>(g, r, L, wh, h5, bh);
    s0[r] = yp; h5[r] = yh;
    wp.push(yp); wh.push(yh);
    if (TWO) {
      const cplx zp = fwd_lrow<BW, MID>(g, r, L, vp, t0, bq);
      t0[r] = zp;
      vp.push(zp);
    }
  });
}
template <int BW>
PB_HD void sweep_E(const LineParams &P, const WsArgs &A, const PairCtx &pc) {
  if (pc.two) sweep_E_t<BW, true>(P, A, pc);
  else sweep_E_t<BW, false>(P, A, pc);
}

// F: back substitution with U of the Poisson operator and the D0/D1
// products v_p -> slot 0, v_p' -> column X, v_h' -> R4, D0 c_h -> R0
// (solve_v_f.f90:258-271, 352)
template <int BW>
PB_HD void sweep_F(const LineParams &P, const WsArgs &A, const SysCtx &c) {
  const BandGeom g{c.n, BW};
  const int n = c.n;
  const OwnU Uq = own_u(A.dbuf, c.prim, c.n, 2 * BW + 1);
  SV s0 = c.col(A, 0), sx = c.xcol(A);
  RV h5 = c.rcol(A, 3), h4 = c.rcol(A, 4), h8 = c.rcol(A, 0);
  CWin<2 * BW + 1> wp;
  RWin<2 * BW + 1> wh;
  wp.clear(); wh.clear();
  up_rows<BW>(n, [&](int r, auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    double ur[2 * BW];
    load_urow<BW, MID>(g, Uq, r, ur);
    wp.push(bwd_row<BW, MID>(g, r, ur, wp, s0[r]));
    wh.push(bwd_row_r<BW, MID>(g, r, ur, wh, h5[r]));
    for_completed_rows<BW>(n, r, [&](int m) {
      const cplx vp = mv_row<BW>(g, P.D0, m, r, wp);
      const cplx dvp = mv_row<BW>(g, P.D1, m, r, wp);
      const double dvh = mv_row_r<BW>(g, P.D1, m, r, wh);
      const double d0h = mv_row_r<BW>(g, P.D0, m, r, wh);
      s0[m] = vp; sx[m] = dvp;
      if (c.is_prim()) { h4[m] = dvh; h8[m] = d0h; }
    });
  });
}

// G: influence coefficients and the final combination on mirror pairs
// (solve_v_f.f90:293-355, solve_u_w.f90:27-38): u -> slot 0, v -> 1, w -> 2
PB_HD void sweep_G(const LineParams &P, const WsArgs &A, const SysCtx &c) {
  const int n = c.n, s = P.s, S = P.S;
  const double dt = P.dt, nu = P.nu, k2 = c.k2;
  SV s0 = c.col(A, 0), s1 = c.col(A, 1), s2 = c.col(A, 2), sx = c.xcol(A);
  RV h1 = c.rcol(A, 1), h3 = c.rcol(A, 2), h5 = c.rcol(A, 4), h8 = c.rcol(A, 0);
  SV r2 = c.rk(A, 1);
  // the real columns stand for the complex (x, x) of the reference
  auto cx = [](double x) { return mk(x, x); };
  const cplx h0 = cx(h5[0]), hn = cx(h5[n - 1]), d0 = sx[0], dn = sx[n - 1];
  const cplx den = csub(cmul(h0, h0), cmul(hn, hn));
  const cplx c1 = cdiv(csub(cmul(hn, dn), cmul(h0, d0)), den);
  const cplx c2 = cdiv(csub(cmul(h0, dn), cmul(hn, d0)), den);
  cplx c1d = c1, c2d = c2, c1a = c1, c2a = c2;
  if (s != S) {
    const double alpha = dt * nu * P.coef(0, s + 1);
    const double dummy = 1.0 - alpha * k2;
    c1d = cscale(c1, dummy); c2d = cscale(c2, dummy); c1a = cscale(c1, alpha); c2a = cscale(c2, alpha);
  }
  const cplx ci = mk(0.0, 1.0);
  const cplx kxx = cdivr(cmul(ci, mk(c.kx, 0.0)), k2), kzz = cdivr(cmul(ci, mk(c.kz, 0.0)), k2);
  for (int i = 0; i <= (n - 1) / 2; ++i) {
    const int j = n - 1 - i;
    const cplx dvp_i = sx[i], dvp_j = sx[j], dvh_i = cx(h5[i]), dvh_j = cx(h5[j]);
    const cplx sp_i = r2[i], sp_j = r2[j], ph_i = cx(h1[i]), ph_j = cx(h1[j]), d2h_i = cx(h3[i]), d2h_j = cx(h3[j]);
    const cplx vp_i = s0[i], vp_j = s0[j], d0h_i = cx(h8[i]), d0h_j = cx(h8[j]), g_i = s2[i], g_j = s2[j];
    auto one = [&](int row, cplx dvp, cplx dvh, cplx dvh_m, cplx sp, cplx ph, cplx ph_m, cplx d2h, cplx d2h_m,
                   cplx vp, cplx d0h, cplx d0h_m, cplx gg) {
      const cplx f = csub(cadd(dvp, cmul(c1, dvh)), cmul(c2, dvh_m));
      cplx st;
      if (s != S) {
        st = cadd(sp, cmul(c1d, ph));
        st = cadd(st, cmul(c2d, ph_m));
        st = cadd(st, cmul(c1a, d2h));
        st = cadd(st, cmul(c2a, d2h_m));
      } else {
        st = cadd(cadd(sp, cmul(c1, ph)), cmul(c2, ph_m));
      }
      const cplx v = cadd(cadd(vp, cmul(c1, d0h)), cmul(c2, d0h_m));
      const cplx dxg = cmul(kxx, gg), dzg = cmul(kzz, gg), dxf = cmul(kxx, f), dzf = cmul(kzz, f);
      s0[row] = csub(dxf, dzg);
      s2[row] = cadd(dxg, dzf);
      s1[row] = v;
      r2[row] = st;
    };
    one(i, dvp_i, dvh_i, dvh_j, sp_i, ph_i, ph_j, d2h_i, d2h_j, vp_i, d0h_i, d0h_j, g_i);
    if (j != i) one(j, dvp_j, dvh_j, dvh_i, sp_j, ph_j, ph_i, d2h_j, d2h_i, vp_j, d0h_j, d0h_i, g_j);
  }
}

// y derivative of two columns of one system, d = D1 LU(D0)^-1 x, into the
// columns o1, o2 (which first hold LU(D0)'s forward substitution): the
// arithmetic of labs_bs then labs_mat_mul (get_y_dir_deriv_of_field,
// get_statistics.f90:658-681)
template <int BW>
PB_HD void deriv_line(const LineParams &P, const SV &x1, const SV &x2, const SV &o1, const SV &o2) {
  const BandGeom g{P.g.n, BW};
  const int n = g.n;
  CWin<BW> f1, f2;
  f1.clear(); f2.clear();
  down_rows<BW>(n, [&](int r, auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    double lr[2 * BW + 1];
    load_lrow<BW, MID>(g, P.D0LU, r, lr);
    const cplx y1 = fwd_row<BW, MID>(g, r, lr, f1, o1, x1[r]);
    const cplx y2 = fwd_row<BW, MID>(g, r, lr, f2, o2, x2[r]);
    o1[r] = y1; o2[r] = y2;
    f1.push(y1); f2.push(y2);
  });
  CWin<2 * BW + 1> w1, w2;
  w1.clear(); w2.clear();
  up_rows<BW>(n, [&](int r, auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    double ur[2 * BW];
    load_urow<BW, MID>(g, P.D0LU, r, ur);
    w1.push(bwd_row<BW, MID>(g, r, ur, w1, o1[r]));
    w2.push(bwd_row<BW, MID>(g, r, ur, w2, o2[r]));
    for_completed_rows<BW>(n, r, [&](int m) {
      o1[m] = mv_row<BW>(g, P.D1, m, r, w1);
      o2[m] = mv_row<BW>(g, P.D1, m, r, w2);
    });
  });
}

// One sweep: index is a system (kz-major, kx-minor) or, for the
// factorising sweeps C_nu, C_theta and E, a pair (kz index 0 .. nz/2).
// A substep runs, for all systems, A1 (s == 1 only), then the theta chain
// A_t, B_t, C_t, D_t and, in any order relative to it, A, B_g, B_phi, C_nu,
// D_nu; then E, F, G (on s == 1 B_g1 and B_t1 stand for B_g and B_t).
enum SweepId {
  SW_A1, SW_A, SW_BT, SW_CT, SW_DT, SW_BG, SW_BV, SW_CNU, SW_DNU, SW_E, SW_F, SW_G,
  SW_BT1, SW_BG1,  // substep 1: B with D2 LU(D0)^-1 of the stores
  SW_AT,           // theta chain: P8's forward substitution
  SW_COUNT
};

constexpr bool sweep_paired(int sw) { return sw == SW_CNU || sw == SW_CT || sw == SW_E; }
// threads of a sweep
PB_HD size_t sweep_threads(int sw, const WsArgs &A) {
  return sweep_paired(sw) ? (size_t)(A.nh + 1) * A.nxh : A.plane;
}

template <int BW, int SW>
PB_HD void run_sweep(const LineParams &P, const WsArgs &A, double *dbuf2, size_t idx) {
  if (sweep_paired(SW)) {
    PairCtx pc;
    if (!make_pair(A, P.g.n, idx, pc)) return;
    if (SW == SW_CT) sweep_Ct<BW>(P, A, pc, dbuf2);
    if (SW == SW_CNU) sweep_Cnu<BW>(P, A, pc);
    if (SW == SW_E) sweep_E<BW>(P, A, pc);
    return;
  }
  SysCtx c;
  if (!make_sys(A, P.g.n, idx, c)) return;
  if (SW == SW_A1) sweep_A1<BW>(P, A, c);
  if (SW == SW_A) sweep_A<BW>(P, A, c);
  if (SW == SW_AT) sweep_At<BW>(P, A, c);
  if (SW == SW_BT) sweep_Bt<BW, false>(P, A, c);
  if (SW == SW_BT1) sweep_Bt<BW, true>(P, A, c);
  if (SW == SW_DT) sweep_Dt<BW>(P, A, c, dbuf2);
  if (SW == SW_BG) sweep_Bg<BW, false>(P, A, c);
  if (SW == SW_BG1) sweep_Bg<BW, true>(P, A, c);
  if (SW == SW_BV) sweep_Bv<BW>(P, A, c);
  if (SW == SW_DNU) sweep_Dnu<BW>(P, A, c);
  if (SW == SW_F) sweep_F<BW>(P, A, c);
  if (SW == SW_G) sweep_G(P, A, c);
}

// the sweeps of substep s in a serial order valid for the stream schedule
inline int sweep_sequence(int s, int *ids) {
  int k = 0;
  const bool s1 = s == 1;
  if (s1) ids[k++] = SW_A1;
  const int rest[] = {SW_AT, s1 ? SW_BT1 : SW_BT, SW_CT, SW_DT, SW_A, s1 ? SW_BG1 : SW_BG, SW_BV,
                      SW_CNU, SW_DNU, SW_E, SW_F, SW_G};
  for (int q : rest) ids[k++] = q;
  return k;
}

}  // namespace pb
