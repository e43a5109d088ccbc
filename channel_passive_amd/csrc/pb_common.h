// pb_common.h -- scalar types, complex arithmetic and the packed band
// geometry shared by the host set-up code and the gfx950 kernels.
//
// Complex arithmetic is spelled out component-wise so that every rounding
// matches the reference's gfortran arithmetic:
//   complex*real, complex/real   -> per component
//   complex*complex              -> (ar*br - ai*bi, ar*bi + ai*br)
//   complex/complex              -> Smith's algorithm (-fcx-fortran-rules)
// Files that include this header are compiled with FP contraction off where
// bit-reproducibility matters (the banded solver).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PB_HD __host__ __device__ __forceinline__
#else
#define PB_HD inline
#endif

namespace pb {

struct cplx {
  double re, im;
};

PB_HD cplx mk(double r, double i) {
  cplx z;
  z.re = r;
  z.im = i;
  return z;
}
PB_HD cplx cadd(cplx a, cplx b) { return mk(a.re + b.re, a.im + b.im); }
PB_HD cplx csub(cplx a, cplx b) { return mk(a.re - b.re, a.im - b.im); }
PB_HD cplx cscale(cplx a, double r) { return mk(a.re * r, a.im * r); }
PB_HD cplx cdivr(cplx a, double r) { return mk(a.re / r, a.im / r); }
PB_HD cplx cmul(cplx a, cplx b) { return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
PB_HD cplx cneg(cplx a) { return mk(-a.re, -a.im); }
PB_HD cplx cdiv(cplx a, cplx b) {
  double tr, ti;
  if ((b.re < 0 ? -b.re : b.re) < (b.im < 0 ? -b.im : b.im)) {
    const double ratio = b.re / b.im;
    const double div = (b.re * ratio) + b.im;
    tr = ((a.re * ratio) + a.im) / div;
    ti = ((a.im * ratio) - a.re) / div;
  } else {
    const double ratio = b.im / b.re;
    const double div = (b.im * ratio) + b.re;
    tr = ((a.im * ratio) + a.re) / div;
    ti = (a.im - (a.re * ratio)) / div;
  }
  return mk(tr, ti);
}
// i*k formed as the reference forms ci*k with ci = (0,1).
PB_HD cplx iunit(double k) { return mk(0.0 * k - 1.0 * 0.0, 0.0 * 0.0 + 1.0 * k); }

// Geometry of the reference's near-boundary packed band
// (LABS_full_to_band, SOURCE_PASSOPB/module_labs.f90:301-328), 0-based.
// Row r stores columns first(r)..last(r) in W = 2bw+1 slots; the first bw
// rows keep their extra right entries (up to column 2bw) in the unused left
// slots, the last bw rows keep their extra left entries (from column
// n-1-2bw) in the unused right slots.
struct BandGeom {
  int n, bw;
  PB_HD int W() const { return 2 * bw + 1; }
  PB_HD int first(int r) const { return r < bw ? 0 : (r >= n - bw ? n - 1 - 2 * bw : r - bw); }
  PB_HD int last(int r) const { return r < bw ? 2 * bw : (r >= n - bw ? n - 1 : r + bw); }
  PB_HD int slot(int r, int c) const {
    if (r < bw && c > r + bw) return 2 * bw - c;
    if (r >= n - bw && c < r - bw) return n - 1 - c;
    return bw + (c - r);
  }
};

// RK3 coefficients alpha, beta, gamma, zeta (module_runge_kutta.f90:34-48),
// [row][substep-1].
struct RK3 {
  double c[4][3];
};
inline RK3 rk3_coefficients() {
  RK3 k;
  k.c[0][0] = 29.0 / 96.0;  k.c[0][1] = -3.0 / 40.0;  k.c[0][2] = 1.0 / 6.0;
  k.c[1][0] = 37.0 / 160.0; k.c[1][1] = 5.0 / 24.0;   k.c[1][2] = 1.0 / 6.0;
  k.c[2][0] = 8.0 / 15.0;   k.c[2][1] = 5.0 / 12.0;   k.c[2][2] = 3.0 / 4.0;
  k.c[3][0] = 0.0;          k.c[3][1] = -17.0 / 60.0; k.c[3][2] = -5.0 / 12.0;
  return k;
}

}  // namespace pb
