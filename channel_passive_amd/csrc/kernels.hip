// kernels.hip -- gfx950 kernels of one RK3-CN substep.
//
//   zinv_kernel     y_to_x: 3/2 zero-padded inverse c2c DFT along z
//                   (module_ppdfftw.f90:830-861)
//   xphys_kernel    dft_x_realspace: c2r along x, the nine products, CFL
//                   maximum, r2c along x (dft_x_realspace.f90:68-213)
//   zfwd_kernel     x_to_y + rescale_data: forward c2c along z, truncation,
//                   Nyquist zeroing, 1/(nx32*nz32) (module_ppdfftw.f90:794-823,
//                   rescale_data.f90)
//   wavespace_kernel  one thread per (kx,kz) system (wavespace.f90:46-107)
//
// Device layout of every spectral field: [y][kz][kx] (kx fastest); padded
// z-pencil fields: [y][z][kx].  All FFTs are unnormalised like FFTW's.
#include <hip/hip_runtime.h>

#include "pb_common.h"
#include "solver.h"
#include "solver_fused.h"
#include "solver_adj.h"

namespace pb {

struct FftPlan {
  int N;
  int np;
  int r[16];
};

__device__ __forceinline__ double2 c2(double r, double i) { return make_double2(r, i); }
__device__ __forceinline__ double2 cadd2(double2 a, double2 b) { return c2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub2(double2 a, double2 b) { return c2(a.x - b.x, a.y - b.y); }
// The transforms need not reproduce a reference's rounding (FFTW's order is
// unpinned; the transform stage is held to 1e-13 relative), so their complex
// products may contract to FMA; everything else in the library is built
// with contraction off.
__device__ __forceinline__ double2 cmul2(double2 a, double2 b) {
#pragma clang fp contract(fast)
  return c2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// CFL maximum that keeps a NaN: fmax (IEEE maxNum) would return the other
// operand, so a non-finite state could reduce to a finite velmax.
__device__ __forceinline__ double nan_max(double a, double b) { return (a > b || a != a) ? a : b; }
// multiply by sign*i  (sign = -1 forward, +1 backward)
__device__ __forceinline__ double2 mul_si(double2 a, int sign) { return sign < 0 ? c2(a.y, -a.x) : c2(-a.y, a.x); }

// in-register DFT of size R with exponent sign (W = exp(sign*2*pi*i/R))
template <int R>
__device__ __forceinline__ void dft(double2 *v, int sign);

template <>
__device__ __forceinline__ void dft<2>(double2 *v, int) {
  const double2 a = v[0], b = v[1];
  v[0] = cadd2(a, b);
  v[1] = csub2(a, b);
}
template <>
__device__ __forceinline__ void dft<4>(double2 *v, int sign) {
  const double2 a0 = cadd2(v[0], v[2]), a1 = csub2(v[0], v[2]);
  const double2 b0 = cadd2(v[1], v[3]), b1 = mul_si(csub2(v[1], v[3]), sign);
  v[0] = cadd2(a0, b0);
  v[2] = csub2(a0, b0);
  v[1] = cadd2(a1, b1);
  v[3] = csub2(a1, b1);
}
template <>
__device__ __forceinline__ void dft<3>(double2 *v, int sign) {
#pragma clang fp contract(fast)
  const double c = -0.5, s = sign * 0.86602540378443864676;
  const double2 t = cadd2(v[1], v[2]), d = csub2(v[1], v[2]);
  const double2 m = c2(v[0].x + c * t.x, v[0].y + c * t.y);
  const double2 rot = c2(-s * d.y, s * d.x);
  v[0] = cadd2(v[0], t);
  v[1] = cadd2(m, rot);
  v[2] = csub2(m, rot);
}
template <>
__device__ __forceinline__ void dft<5>(double2 *v, int sign) {
#pragma clang fp contract(fast)
  const double c1 = 0.30901699437494742410, c2v = -0.80901699437494742410;
  const double s1 = sign * 0.95105651629515357212, s2 = sign * 0.58778525229247312917;
  const double2 t1 = cadd2(v[1], v[4]), t2 = cadd2(v[2], v[3]);
  const double2 d1 = csub2(v[1], v[4]), d2 = csub2(v[2], v[3]);
  const double2 m1 = c2(v[0].x + c1 * t1.x + c2v * t2.x, v[0].y + c1 * t1.y + c2v * t2.y);
  const double2 m2 = c2(v[0].x + c2v * t1.x + c1 * t2.x, v[0].y + c2v * t1.y + c1 * t2.y);
  const double2 r1 = c2(-(s1 * d1.y + s2 * d2.y), s1 * d1.x + s2 * d2.x);
  const double2 r2 = c2(-(s2 * d1.y - s1 * d2.y), s2 * d1.x - s1 * d2.x);
  v[0] = cadd2(v[0], cadd2(t1, t2));
  v[1] = cadd2(m1, r1);
  v[4] = csub2(m1, r1);
  v[2] = cadd2(m2, r2);
  v[3] = csub2(m2, r2);
}
template <>
__device__ __forceinline__ void dft<8>(double2 *v, int sign) {
#pragma clang fp contract(fast)
  const double h = 0.70710678118654752440;
  double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
  dft<4>(e, sign);
  dft<4>(o, sign);
  // twiddle o[q] by W8^q
  o[1] = c2(h * (o[1].x - sign * o[1].y), h * (o[1].y + sign * o[1].x));
  o[2] = mul_si(o[2], sign);
  o[3] = c2(h * (-o[3].x - sign * o[3].y), h * (-o[3].y + sign * o[3].x));
  for (int q = 0; q < 4; ++q) {
    v[q] = cadd2(e[q], o[q]);
    v[q + 4] = csub2(e[q], o[q]);
  }
}

// cos(2 pi m / R), sin(2 pi m / R) for the composite radices (long-double
// values rounded to double; the quarter-turn points exact)
template <int R>
struct Wr;
template <>
struct Wr<6> {
  static constexpr double c[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
  static constexpr double s[6] = {0.0, 0.8660254037844386, 0.8660254037844386, 0.0, -0.8660254037844386,
                                  -0.8660254037844386};
};
template <>
struct Wr<9> {
  static constexpr double c[9] = {1.0, 0.766044443118978, 0.17364817766693036, -0.5, -0.9396926207859084,
                                  -0.9396926207859084, -0.5, 0.17364817766693036, 0.766044443118978};
  static constexpr double s[9] = {0.0, 0.6427876096865394, 0.984807753012208, 0.8660254037844386,
                                  0.3420201433256687, -0.3420201433256687, -0.8660254037844386, -0.984807753012208,
                                  -0.6427876096865394};
};
template <>
struct Wr<12> {
  static constexpr double c[12] = {1.0, 0.8660254037844386, 0.5, 0.0, -0.5, -0.8660254037844386,
                                   -1.0, -0.8660254037844386, -0.5, 0.0, 0.5, 0.8660254037844386};
  static constexpr double s[12] = {0.0, 0.5, 0.8660254037844386, 1.0, 0.8660254037844386, 0.5,
                                   0.0, -0.5, -0.8660254037844386, -1.0, -0.8660254037844386, -0.5};
};
template <>
struct Wr<16> {
  static constexpr double c[16] = {1.0, 0.9238795325112867, 0.7071067811865476, 0.3826834323650898,
                                   0.0, -0.3826834323650898, -0.7071067811865476, -0.9238795325112867,
                                   -1.0, -0.9238795325112867, -0.7071067811865476, -0.3826834323650898,
                                   0.0, 0.3826834323650898, 0.7071067811865476, 0.9238795325112867};
  static constexpr double s[16] = {0.0, 0.3826834323650898, 0.7071067811865476, 0.9238795325112867,
                                   1.0, 0.9238795325112867, 0.7071067811865476, 0.3826834323650898,
                                   0.0, -0.3826834323650898, -0.7071067811865476, -0.9238795325112867,
                                   -1.0, -0.9238795325112867, -0.7071067811865476, -0.3826834323650898};
};

template <>
struct Wr<18> {
  static constexpr double c[18] = {1.0, 0.9396926207859084, 0.766044443118978, 0.5, 0.17364817766693036,
                                   -0.17364817766693036, -0.5, -0.766044443118978, -0.9396926207859084, -1.0,
                                   -0.9396926207859084, -0.766044443118978, -0.5, -0.17364817766693036,
                                   0.17364817766693036, 0.5, 0.766044443118978, 0.9396926207859084};
  static constexpr double s[18] = {0.0, 0.3420201433256687, 0.6427876096865394, 0.8660254037844386,
                                   0.984807753012208, 0.984807753012208, 0.8660254037844386, 0.6427876096865394,
                                   0.3420201433256687, 0.0, -0.3420201433256687, -0.6427876096865394,
                                   -0.8660254037844386, -0.984807753012208, -0.984807753012208,
                                   -0.8660254037844386, -0.6427876096865394, -0.3420201433256687};
};

// Composite in-register DFT, R = P*Q (Cooley-Tukey): P-point DFTs over
// n1 of x[Q n1 + n2], twiddle W_R^(n2 k1), Q-point DFTs over n2, output
// X[k1 + P k2].
template <int P, int Q>
__device__ __forceinline__ void dft_pq(double2 *v, int sign) {
  constexpr int R = P * Q;
  double2 a[Q][P];
#pragma unroll
  for (int n2 = 0; n2 < Q; ++n2) {
#pragma unroll
    for (int n1 = 0; n1 < P; ++n1) a[n2][n1] = v[Q * n1 + n2];
    dft<P>(a[n2], sign);
  }
#pragma unroll
  for (int n2 = 1; n2 < Q; ++n2)
#pragma unroll
    for (int k1 = 1; k1 < P; ++k1) {
      const int m = (n2 * k1) % R;
      if ((4 * m) % R == 0) {  // quarter turns: exact, no arithmetic
        const int q = (4 * m / R) & 3;
        a[n2][k1] = q == 1 ? mul_si(a[n2][k1], sign) : (q == 2 ? c2(-a[n2][k1].x, -a[n2][k1].y)
                                                                 : mul_si(a[n2][k1], -sign));
      } else {
        a[n2][k1] = cmul2(a[n2][k1], c2(Wr<R>::c[m], sign * Wr<R>::s[m]));
      }
    }
#pragma unroll
  for (int k1 = 0; k1 < P; ++k1) {
    double2 t[Q];
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) t[n2] = a[n2][k1];
    dft<Q>(t, sign);
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) v[k1 + P * k2] = t[k2];
  }
}
template <>
__device__ __forceinline__ void dft<6>(double2 *v, int sign) { dft_pq<2, 3>(v, sign); }
template <>
__device__ __forceinline__ void dft<9>(double2 *v, int sign) { dft_pq<3, 3>(v, sign); }
template <>
__device__ __forceinline__ void dft<12>(double2 *v, int sign) { dft_pq<4, 3>(v, sign); }
template <>
__device__ __forceinline__ void dft<16>(double2 *v, int sign) { dft_pq<4, 4>(v, sign); }
template <>
__device__ __forceinline__ void dft<18>(double2 *v, int sign) { dft_pq<2, 9>(v, sign); }

// One Stockham pass of radix R on x[n*B + b] (n < N, b < B), in place via
// registers.  Ns = product of the radices already applied.
template <int R, int MAXB>
__device__ __forceinline__ void stockham_pass(double2 *x, int N, int B, int Ns, const double2 *__restrict__ tw,
                                              int sign) {
  const int NR = N / R;
  const int nbf = NR * B;
  const int T = blockDim.x;
  double2 v[MAXB][R];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int t = threadIdx.x + q * T;
    if (t < nbf) {
      const int b = t % B, j = t / B;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = x[(j + r * NR) * B + b];
    }
  }
  __syncthreads();
  const int tstep = N / (Ns * R);
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int t = threadIdx.x + q * T;
    if (t < nbf) {
      const int b = t % B, j = t / B;
      const int k = j % Ns;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        double2 w = tw[r * k * tstep];
        if (sign > 0) w.y = -w.y;
        v[q][r] = cmul2(v[q][r], w);
      }
      dft<R>(v[q], sign);
      const int base = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) x[(base + r * Ns) * B + b] = v[q][r];
    }
  }
  __syncthreads();
}

// Maximum elements (N*B) per block handled by lds_fft with 256 threads.
constexpr int FFT_MAX_ELEMS = 8192;
constexpr int FFT_THREADS = 256;

// EPT = upper bound of elements per thread (N*B/blockDim), a compile-time
// constant so that every register array below is statically sized.
template <int EPT>
__device__ void lds_fft(double2 *x, int B, const FftPlan &P, const double2 *__restrict__ tw, int sign) {
  int Ns = 1;
  for (int p = 0; p < P.np; ++p) {
    switch (P.r[p]) {
      case 2: stockham_pass<2, (EPT + 1) / 2>(x, P.N, B, Ns, tw, sign); break;
      case 3: stockham_pass<3, (EPT + 2) / 3>(x, P.N, B, Ns, tw, sign); break;
      case 4: stockham_pass<4, (EPT + 3) / 4>(x, P.N, B, Ns, tw, sign); break;
      case 5: stockham_pass<5, (EPT + 4) / 5>(x, P.N, B, Ns, tw, sign); break;
      case 8: stockham_pass<8, (EPT + 7) / 8>(x, P.N, B, Ns, tw, sign); break;
    }
    Ns *= P.r[p];
  }
}

// XCD-aware block numbering.  Workgroups are dealt to the 8 XCDs in turn
// and each XCD has its own L2; a z-kernel block touches B*16 = 64 bytes of
// every 128-byte line, its x-neighbour the other half.  Renumbering so that
// each XCD runs one contiguous range of logical blocks keeps both halves of
// a line in one L2.  Identity when the grid does not split evenly.
struct Blk {
  int x, y, z;
};
__device__ inline Blk xcd_block() {
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const size_t n = (size_t)gx * gy * gridDim.z;
  const size_t h = blockIdx.x + (size_t)gx * (blockIdx.y + (size_t)gy * blockIdx.z);
  const size_t L = (n % 8 == 0) ? (h % 8) * (n / 8) + h / 8 : h;
  Blk b;
  b.x = (int)(L % gx);
  b.y = (int)((L / gx) % gy);
  b.z = (int)(L / ((size_t)gx * gy));
  return b;
}

// ------------------------------------------------- compile-time plans
// The same Stockham passes with N, B, the radices and the direction known at
// compile time: index arithmetic folds to shifts and multiplies, the first
// pass needs no twiddles, and every register array has its exact size.
// Fixed-size passes form W^(rk) from one table value W^k by binary powering
// (at most log2 R complex products, a few ulp), instead of R-1 table loads.
#ifndef PB_FFT_TWPOW
#define PB_FFT_TWPOW 1
#endif

// LDS element index of the fixed-size transforms: an XOR swizzle of the
// low 3 bits of the 16-byte element index by bits 4..6, so that the strided
// stores of the first Stockham pass (and every other access pattern of the
// kernels, which are consecutive per lane) spread over all bank groups.
#ifndef PB_LDS_SWZ
#define PB_LDS_SWZ 1
#endif
__device__ __forceinline__ int lds_swz(int i) { return PB_LDS_SWZ ? i ^ ((i >> 4) & 7) : i; }

// Layout of B lines of N elements in LDS: interleaved (element n of line b
// at n*B + b, the z transforms, whose global rows hold B consecutive kx;
// XOR-swizzled, a permutation, as the streamed copies into LDS need) or
// separate (b*N + n, the x lines; one pad element after every 32).
// ds_read_b128 serves 32 consecutive lanes per clock and ds_write_b128 16:
// with one pad per 32, every aligned run of 32 consecutive elements (all
// reads of the passes) is conflict-free, and the element index of
// n0 + r*S stays linear in r for the strides the passes use, so they address
// their R elements with constant offsets from one base.  Only the
// stride-16 stores of a first radix-16 pass meet two lanes per bank group.
template <int N, int B, bool SEP>
struct Lines {
  static constexpr bool PAD = SEP;
  static constexpr int SIZE = PAD ? N * B + (N * B) / 32 : N * B;
  static_assert(!PAD || N % 32 == 0, "padded lines start on a 32-element boundary");
  __device__ static int at(int n, int b) {
    const int i = SEP ? b * N + n : n * B + b;
    return PAD ? i + (i >> 5) : lds_swz(i);
  }
  // butterfly t of a pass with N/R butterflies per line -> (line, index)
  __device__ static void split(int t, int NR, int &b, int &j) {
    if (SEP) { b = t / NR; j = t % NR; } else { b = t % B; j = t / B; }
  }
};

// Barrier between the passes.  LDS_ONLY waits for this wave's LDS operations
// only (lgkmcnt), not for its outstanding global loads: a kernel that copies
// its next tile global -> LDS while transforming the current one must not
// wait for that copy at every pass.  The empty asm statements keep the
// compiler from moving LDS accesses across the barrier.
template <bool LDS_ONLY>
__device__ __forceinline__ void pass_sync() {
  if (LDS_ONLY) {
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt, expcnt not waited for
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  } else {
    __syncthreads();
  }
}

// LDS index of element n0 + r*S of line b.  KIND 1: n0 % 16 == 0 and
// r*S < 16 (no pad crossed); KIND 2: S == 16 and n0 % 32 < 16 (one pad per
// two steps)
template <class L, int S, int KIND>
__device__ __forceinline__ int lds_at(int n0, int b, int r) {
  if constexpr (L::PAD && S % 32 == 0) return L::at(n0, b) + r * (S + S / 32);
  else if constexpr (L::PAD && KIND == 1) return L::at(n0, b) + r * S;
  else if constexpr (L::PAD && KIND == 2) return L::at(n0, b) + r * S + (r >> 1);
  else return L::at(n0 + r * S, b);
}

// Twiddles and DFT of one butterfly (index k within Ns) of a Stockham pass
template <int R, int SIGN, int Ns, int tstep>
__device__ __forceinline__ void bfly(double2 *v, const double2 *__restrict__ tw, int k) {
  if (Ns > 1) {
#if PB_FFT_TWPOW
    // one table load; W^(r k) = (W^k)^r by binary powering (depth <= log2 R)
    double2 w[R];
    w[1] = tw[k * tstep];
    if (SIGN > 0) w[1].y = -w[1].y;
#pragma unroll
    for (int r = 2; r < R; ++r) w[r] = (r & 1) ? cmul2(w[r - 1], w[1]) : cmul2(w[r / 2], w[r / 2]);
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul2(v[r], w[r]);
#else
#pragma unroll
    for (int r = 1; r < R; ++r) {
      double2 w = tw[r * k * tstep];
      if (SIGN > 0) w.y = -w.y;
      v[r] = cmul2(v[r], w);
    }
#endif
  }
  dft<R>(v, SIGN);
}

template <int T, class L, int N, int B, int SIGN, int Ns, int R, bool LDS_ONLY = false>
__device__ __forceinline__ void ct_pass(double2 *x, const double2 *__restrict__ tw) {
  constexpr int NR = N / R, NBF = NR * B, ROUNDS = (NBF + T - 1) / T;
  constexpr int tstep = N / (Ns * R);
  double2 v[ROUNDS][R];
#pragma unroll
  for (int q = 0; q < ROUNDS; ++q) {
    const int t = threadIdx.x + q * T;
    if (NBF % T == 0 || t < NBF) {
      int b, j;
      L::split(t, NR, b, j);
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = x[lds_at<L, NR, 0>(j, b, r)];
    }
  }
  // in the last pass (Ns * R == N: NR == Ns, base == j) every butterfly
  // writes back exactly the elements it read, so no thread can overwrite
  // another's inputs and the barrier between reads and writes is not needed
  if constexpr (Ns * R != N) pass_sync<LDS_ONLY>();
#pragma unroll
  for (int q = 0; q < ROUNDS; ++q) {
    const int t = threadIdx.x + q * T;
    if (NBF % T == 0 || t < NBF) {
      int b, j;
      L::split(t, NR, b, j);
      const int k = j % Ns;
      bfly<R, SIGN, Ns, tstep>(v[q], tw, k);
      const int base = (j / Ns) * Ns * R + k;
      // first pass of a radix-16 plan: base = 16 j; Ns = 16 with Ns*R a
      // multiple of 32: base % 32 = k < 16
      constexpr int KIND = (Ns == 1 && R == 16) ? 1 : ((Ns == 16 && (Ns * R) % 32 == 0) ? 2 : 0);
#pragma unroll
      for (int r = 0; r < R; ++r) x[lds_at<L, Ns, KIND>(base, b, r)] = v[q][r];
    }
  }
  pass_sync<LDS_ONLY>();
}
template <int T, class L, int N, int B, int SIGN, int Ns, int R, int... Rest>
__device__ __forceinline__ void ct_passes(double2 *x, const double2 *__restrict__ tw) {
  ct_pass<T, L, N, B, SIGN, Ns, R>(x, tw);
  if constexpr (sizeof...(Rest) > 0) ct_passes<T, L, N, B, SIGN, Ns * R, Rest...>(x, tw);
}
template <int T, class L, int N, int B, int SIGN, int Ns, int R, int... Rest>
__device__ __forceinline__ void ct_passes_lds(double2 *x, const double2 *__restrict__ tw) {
  ct_pass<T, L, N, B, SIGN, Ns, R, true>(x, tw);
  if constexpr (sizeof...(Rest) > 0) ct_passes_lds<T, L, N, B, SIGN, Ns * R, Rest...>(x, tw);
}

// Transform policies for the three kernels: N = B = 0 means sizes at run
// time (any 2,3,5-smooth length, generic passes, lines interleaved, no
// swizzle); otherwise fixed sizes.  at(n, b) is the LDS index of element n
// of line b.
template <int EPT>
struct RtFft {
  static constexpr int T = FFT_THREADS, N = 0, B = 0, ELEMS = EPT, MINW = 1;
  static size_t lds_elems(int n, int b) { return (size_t)n * b; }
  __device__ static int at(int n, int b, int B_) { return n * B_ + b; }
  __device__ static void run(double2 *x, int B_, const FftPlan &P, const double2 *__restrict__ tw, int sign) {
    lds_fft<EPT>(x, B_, P, tw, sign);
  }
};
template <int T_, int MINW_, int N_, int B_, bool SEP, int... R>
struct CtFft {
  // MINW: waves per SIMD the register budget must allow (the LDS of one
  // block limits how many blocks share a CU; more waves than that fit is
  // register space wasted)
  static constexpr int T = T_, N = N_, B = B_, ELEMS = (N_ * B_ + T_ - 1) / T_, MINW = MINW_;
  using L = Lines<N_, B_, SEP>;
  static size_t lds_elems(int, int) { return L::SIZE; }
  static_assert((R * ...) == N_, "radices must multiply to N");
  static_assert(T_ % 32 == 0 && T_ <= 1024, "block size");
  __device__ static int at(int n, int b, int) { return L::at(n, b); }
  __device__ static void run(double2 *x, int, const FftPlan &, const double2 *__restrict__ tw, int sign) {
    if (sign > 0)
      ct_passes<T_, L, N_, B_, +1, 1, R...>(x, tw);
    else
      ct_passes<T_, L, N_, B_, -1, 1, R...>(x, tw);
  }
};
// 3/2-padded production sizes.  nz32 = 1152 with 8 kx columns per block
// (128-byte row segments; 147 KB of LDS, one block per CU): 768 threads
// hold 12 elements each, radix 8, 12, 12.  nx32 = 1536, the two lines of a
// pair: 256 threads, radix 16, 16, 6.
#ifndef PB_FFTZ1152
#define PB_FFTZ1152 CtFft<768, 3, 1152, 8, false, 8, 12, 12>
#endif
#ifndef PB_FFTX1536
#define PB_FFTX1536 CtFft<256, 2, 1536, 2, true, 16, 16, 6>
#endif
using FftZ1152 = PB_FFTZ1152;
using FftX1536 = PB_FFTX1536;

// ------------------------------------------------------------------ z inverse
// Slot maps: field f of a pad <-> slot (map >> 4f) & 15 of U.
constexpr uint64_t MAP_STATE = 0x6210;          // u, v, w, theta in slots 0, 1, 2, 6
constexpr uint64_t MAP_PROD = 0x587643210ull;   // 8 products (+ v*v) in slots 0-4, 6-8 (5)
__device__ __forceinline__ int map_slot(uint64_t map, int f) { return (int)((map >> (4 * f)) & 15); }

// grid: (nxh/B, ny, nfield); field f comes from U slot map_slot(map, f)
// (MAP_STATE: the 4 state fields).  Output pad[f][y][z][kx], times scale.
template <class F>
__global__ void __launch_bounds__(F::T, F::MINW) zinv_kernel(const double2 *__restrict__ U, double2 *__restrict__ pad,
                                                           FftPlan P, const double2 *__restrict__ tw, int nxh, int nz,
                                                           int ny, int B_, size_t field, size_t pfield, uint64_t map,
                                                           double scale) {
  extern __shared__ double2 lds[];
  const int B = F::B ? F::B : B_;
  const Blk bk = xcd_block();
  const int kx0 = bk.x * B, y = bk.y, f = bk.z;
  const int slot = map_slot(map, f);
  const int N = F::N ? F::N : P.N, nh = nz / 2;
  const double2 *src = U + (size_t)slot * field + (size_t)y * nz * nxh + kx0;
  for (int e = threadIdx.x; e < N * B; e += F::T) {
    const int z = e / B, b = e % B;
    double2 val = c2(0.0, 0.0);
    if (z < nh) val = src[(size_t)z * nxh + b];
    else if (z > 2 * nh) val = src[(size_t)(z - nh) * nxh + b];
    lds[F::at(z, b, B)] = val;
  }
  __syncthreads();
  F::run(lds, B, P, tw, +1);
  double2 *dst = pad + (size_t)f * pfield + (size_t)y * N * nxh + kx0;
  for (int e = threadIdx.x; e < N * B; e += F::T) {
    const int z = e / B, b = e % B;
    const double2 v = lds[F::at(z, b, B)];
    dst[(size_t)z * nxh + b] = scale == 1.0 ? v : c2(v.x * scale, v.y * scale);
  }
}

// Statistics-time moments of the physical velocities at one point
// (dft_x_realspace.f90:128-152), the order of its accumulators 3..21:
// u^3 v^3 w^3, u^4 v^4 w^4, -u^2 v, -v w^2, -v^3, -u v^2, u^6 v^6 w^6,
// u^8 v^8 w^8, u^10 v^10 w^10.
constexpr int NMOM = 19;
__device__ __forceinline__ void add_moments(double (&acc)[NMOM], double u, double v, double w) {
  const double u2 = u * u, v2 = v * v, w2 = w * w;
  acc[0] += u * u * u; acc[1] += v * v * v; acc[2] += w * w * w;
  acc[3] += u * u * u * u; acc[4] += v * v * v * v; acc[5] += w * w * w * w;
  acc[6] -= u * u * v; acc[7] -= v * w * w; acc[8] -= v * v * v; acc[9] -= u * v * v;
  const double u4 = u2 * u2, v4 = v2 * v2, w4 = w2 * w2;
  acc[10] += u4 * u2; acc[11] += v4 * v2; acc[12] += w4 * w2;
  acc[13] += u4 * u4; acc[14] += v4 * v4; acc[15] += w4 * w4;
  acc[16] += u4 * u4 * u2; acc[17] += v4 * v4 * v2; acc[18] += w4 * w4 * w2;
}

// ---------------------------------------------------------------- x physical
// grid: (nz32, ny): one x line per block.  pad fields 0..3 = u,v,w,theta in,
// 0..7 = spectra of u^2-v^2, uv, w^2-v^2, vw, uw, u theta, v theta, w theta out.
template <class F>
__global__ void __launch_bounds__(F::T, F::MINW) xphys_kernel(double2 *__restrict__ pad, FftPlan P,
                                                            const double2 *__restrict__ tw, int nxh, int nz32,
                                                            size_t pfield, const double *__restrict__ dyr,
                                                            double dxr, double dzr, double *__restrict__ vpart,
                                                            int nround, double *__restrict__ mom) {
  extern __shared__ double2 lds[];  // 2 lines of N, layout F::at
  const int z = blockIdx.x, y = blockIdx.y, N = F::N ? F::N : P.N;
  double2 *line = pad + (size_t)y * nz32 * nxh + (size_t)z * nxh;
  // Hermitian completion of (u + i v) and (w + i theta); FFTW c2r ignores Im of the DC bin
  for (int m = threadIdx.x; m < N; m += F::T) {
    double2 z1 = c2(0.0, 0.0), z2 = c2(0.0, 0.0);
    if (m == 0) {
      z1 = c2(line[0].x, line[pfield].x);
      z2 = c2(line[2 * pfield].x, line[3 * pfield].x);
    } else if (m < nxh) {
      const double2 a = line[m], b = line[pfield + m], c = line[2 * pfield + m], d = line[3 * pfield + m];
      z1 = c2(a.x - b.y, a.y + b.x);
      z2 = c2(c.x - d.y, c.y + d.x);
    } else if (m > N - nxh) {
      const int q = N - m;
      const double2 a = line[q], b = line[pfield + q], c = line[2 * pfield + q], d = line[3 * pfield + q];
      z1 = c2(a.x + b.y, -a.y + b.x);
      z2 = c2(c.x + d.y, -c.y + d.x);
    }
    lds[F::at(m, 0, 2)] = z1;
    lds[F::at(m, 1, 2)] = z2;
  }
  __syncthreads();
  F::run(lds, 2, P, tw, +1);
  // products at this thread's points, kept in registers
  constexpr int MAXP = (F::ELEMS + 1) / 2;
  double pu[MAXP], pv[MAXP], pw[MAXP], pt[MAXP];
  const double dy = dyr[y];
  double vmax = 0.0;
#pragma unroll
  for (int q = 0; q < MAXP; ++q) {
    const int m = threadIdx.x + q * F::T;
    if (m < N) {
      const double2 a = lds[F::at(m, 0, 2)], c = lds[F::at(m, 1, 2)];
      pu[q] = a.x; pv[q] = a.y; pw[q] = c.x; pt[q] = c.y;
      const double vel = (fabs(pu[q]) * dxr) + (fabs(pv[q]) * dy) + (fabs(pw[q]) * dzr);
      vmax = nan_max(vel, vmax);
    }
  }
  // CFL maximum: across the wave by lane exchanges, one value per wave
  // (velmax_kernel takes the maximum of all of them; max is order-free)
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) vmax = nan_max(vmax, __shfl_xor(vmax, d, 64));
  constexpr int NW = (F::T + 63) / 64;
  if ((threadIdx.x & 63) == 0) vpart[((size_t)y * nz32 + z) * NW + threadIdx.x / 64] = vmax;
  // statistics time: the 19 physical-space moments of this line, one partial
  // sum per wave: mom[((q * ny + y) * nz32 + z) * NW + wave]
  if (mom) {
    double acc[NMOM];
#pragma unroll
    for (int k = 0; k < NMOM; ++k) acc[k] = 0.0;
#pragma unroll
    for (int q = 0; q < MAXP; ++q)
      if (threadIdx.x + q * F::T < N) add_moments(acc, pu[q], pv[q], pw[q]);
#pragma unroll
    for (int k = 0; k < NMOM; ++k) {
      double a = acc[k];
#pragma unroll
      for (int d = 32; d > 0; d >>= 1) a += __shfl_xor(a, d, 64);
      if ((threadIdx.x & 63) == 0)
        mom[(((size_t)k * gridDim.y + y) * nz32 + z) * NW + threadIdx.x / 64] = a;
    }
  }
  // two pairs of products per round: (P1 + i P2, P3 + i P4), (P5 + i P7, P8 + i P9);
  // a third round (statistics time) transforms v*v alone into pad field 8
  for (int round = 0; round < nround; ++round) {
#pragma unroll
    for (int q = 0; q < MAXP; ++q) {
      const int m = threadIdx.x + q * F::T;
      if (m < N) {
        const double t1 = pu[q], t2 = pv[q], t3 = pw[q], t4 = pt[q];
        const double bv2 = t2 * t2;
        if (round == 0) {
          lds[F::at(m, 0, 2)] = c2(t1 * t1 - bv2, t1 * t2);
          lds[F::at(m, 1, 2)] = c2(t3 * t3 - bv2, t2 * t3);
        } else if (round == 1) {
          lds[F::at(m, 0, 2)] = c2(t1 * t3, t1 * t4);
          lds[F::at(m, 1, 2)] = c2(t2 * t4, t3 * t4);
        } else {
          lds[F::at(m, 0, 2)] = c2(bv2, 0.0);
          lds[F::at(m, 1, 2)] = c2(0.0, 0.0);
        }
      }
    }
    __syncthreads();
    F::run(lds, 2, P, tw, -1);
    // unpack the two real spectra of each pair, keep modes 0..nxh-1
    for (int m = threadIdx.x; m < nxh; m += F::T) {
      const int q = (m == 0) ? 0 : N - m;
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const double2 Z = lds[F::at(m, b, 2)], Zc = lds[F::at(q, b, 2)];
        const double2 A = c2((Z.x + Zc.x) * 0.5, (Z.y - Zc.y) * 0.5);
        const double2 Bv = c2((Z.y + Zc.y) * 0.5, (Zc.x - Z.x) * 0.5);
        const int fa = round * 4 + 2 * b;
        if (round < 2) {
          line[(size_t)fa * pfield + m] = A;
          line[(size_t)(fa + 1) * pfield + m] = Bv;
        } else if (b == 0) {
          line[(size_t)8 * pfield + m] = A;
        }
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------- adjoint x pass
// The transpose of xphys_kernel's work on one x line (grid (nz32, ny)),
// for a fixed state: fwd = pad fields 0..3 of that state (u, v, w, theta,
// z-inverted), adj = pad fields 0..7 of the product adjoints in (the
// layout xphys_kernel writes its products in), fields 0..3 of the state
// adjoints out.  Each step of the forward is transposed: the unpack of two
// real spectra from one complex transform, the transform of sign -1 (its
// transpose is the transform of sign +1), the products (u^2-v^2, uv, w^2-v^2,
// vw, uw, u theta, v theta, w theta) at the point, the transform of sign +1
// and the Hermitian completion of u + i v and w + i theta.
template <class F>
__global__ void __launch_bounds__(F::T, F::MINW) xadj_kernel(const double2 *__restrict__ fwd, double2 *__restrict__ adj,
                                                           FftPlan P, const double2 *__restrict__ tw, int nxh,
                                                           int nz32, size_t pfield) {
  extern __shared__ double2 lds[];  // 2 lines of N, layout F::at
  const int z = blockIdx.x, y = blockIdx.y, N = F::N ? F::N : P.N;
  const size_t off = (size_t)y * nz32 * nxh + (size_t)z * nxh;
  const double2 *fl = fwd + off;
  double2 *al = adj + off;
  for (int m = threadIdx.x; m < N; m += F::T) {
    double2 z1 = c2(0.0, 0.0), z2 = c2(0.0, 0.0);
    if (m == 0) {
      z1 = c2(fl[0].x, fl[pfield].x);
      z2 = c2(fl[2 * pfield].x, fl[3 * pfield].x);
    } else if (m < nxh) {
      const double2 a = fl[m], b = fl[pfield + m], c = fl[2 * pfield + m], d = fl[3 * pfield + m];
      z1 = c2(a.x - b.y, a.y + b.x);
      z2 = c2(c.x - d.y, c.y + d.x);
    } else if (m > N - nxh) {
      const int q = N - m;
      const double2 a = fl[q], b = fl[pfield + q], c = fl[2 * pfield + q], d = fl[3 * pfield + q];
      z1 = c2(a.x + b.y, -a.y + b.x);
      z2 = c2(c.x + d.y, -c.y + d.x);
    }
    lds[F::at(m, 0, 2)] = z1;
    lds[F::at(m, 1, 2)] = z2;
  }
  __syncthreads();
  F::run(lds, 2, P, tw, +1);
  constexpr int MAXP = (F::ELEMS + 1) / 2;
  double pu[MAXP], pv[MAXP], pw[MAXP], pt[MAXP], au[MAXP], av[MAXP], aw[MAXP], at[MAXP];
#pragma unroll
  for (int q = 0; q < MAXP; ++q) {
    const int m = threadIdx.x + q * F::T;
    au[q] = av[q] = aw[q] = at[q] = 0.0;
    pu[q] = pv[q] = pw[q] = pt[q] = 0.0;
    if (m < N) {
      const double2 a = lds[F::at(m, 0, 2)], c = lds[F::at(m, 1, 2)];
      pu[q] = a.x; pv[q] = a.y; pw[q] = c.x; pt[q] = c.y;
    }
  }
  for (int round = 0; round < 2; ++round) {
    __syncthreads();
    // transpose of the unpack: Z[m] gets A_m/2 + (i/2) B_m (m < nxh) and
    // conj(A_q)/2 + (i/2) conj(B_q) for q = N-m < nxh (q = 0 for m = 0)
    for (int m = threadIdx.x; m < N; m += F::T) {
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int fa = round * 4 + 2 * b;
        double zr = 0.0, zi = 0.0;
        if (m < nxh) {
          const double2 A = al[(size_t)fa * pfield + m], B = al[(size_t)(fa + 1) * pfield + m];
          zr += 0.5 * A.x - 0.5 * B.y;
          zi += 0.5 * A.y + 0.5 * B.x;
        }
        const int q = (m == 0) ? 0 : N - m;
        if (q < nxh) {
          const double2 A = al[(size_t)fa * pfield + q], B = al[(size_t)(fa + 1) * pfield + q];
          zr += 0.5 * A.x + 0.5 * B.y;
          zi += -0.5 * A.y + 0.5 * B.x;
        }
        lds[F::at(m, b, 2)] = c2(zr, zi);
      }
    }
    __syncthreads();
    F::run(lds, 2, P, tw, +1);
#pragma unroll
    for (int q = 0; q < MAXP; ++q) {
      const int m = threadIdx.x + q * F::T;
      if (m < N) {
        const double2 Z0 = lds[F::at(m, 0, 2)], Z1 = lds[F::at(m, 1, 2)];
        const double u = pu[q], v = pv[q], w = pw[q], t = pt[q];
        if (round == 0) {  // P1 = u^2 - v^2, P2 = uv, P3 = w^2 - v^2, P4 = vw
          au[q] += 2.0 * u * Z0.x + v * Z0.y;
          av[q] += -2.0 * v * Z0.x + u * Z0.y - 2.0 * v * Z1.x + w * Z1.y;
          aw[q] += 2.0 * w * Z1.x + v * Z1.y;
        } else {  // P5 = uw, P7 = u theta, P8 = v theta, P9 = w theta
          au[q] += w * Z0.x + t * Z0.y;
          aw[q] += u * Z0.x + t * Z1.y;
          av[q] += t * Z1.x;
          at[q] += u * Z0.y + v * Z1.x + w * Z1.y;
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXP; ++q) {
    const int m = threadIdx.x + q * F::T;
    if (m < N) {
      lds[F::at(m, 0, 2)] = c2(au[q], av[q]);
      lds[F::at(m, 1, 2)] = c2(aw[q], at[q]);
    }
  }
  __syncthreads();
  F::run(lds, 2, P, tw, -1);
  __syncthreads();
  // transpose of the Hermitian completion: kx = q < nxh gathers Z[q] and
  // conj(Z[N-q]) (a + i b at q, conj(a) + i conj(b) at N-q; Re only at 0)
  for (int q = threadIdx.x; q < nxh; q += F::T) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const double2 Z = lds[F::at(q, b, 2)];
      double2 A, B;
      if (q == 0) {
        A = c2(Z.x, 0.0);
        B = c2(Z.y, 0.0);
      } else {
        const double2 W = lds[F::at(N - q, b, 2)];
        A = c2(Z.x + W.x, Z.y - W.y);
        B = c2(Z.y + W.y, W.x - Z.x);
      }
      al[(size_t)(2 * b) * pfield + q] = A;
      al[(size_t)(2 * b + 1) * pfield + q] = B;
    }
  }
}

// ------------------------------------------------------------------ z forward
// grid: (nxh/B, ny, nfield).  pad field g -> U slot map_slot(map, g)
// (MAP_PROD: (0,1,2,3,4,6,7,8,5)[g]), with truncation and rescale.
template <class F>
__global__ void __launch_bounds__(F::T, F::MINW) zfwd_kernel(const double2 *__restrict__ pad, double2 *__restrict__ U,
                                                           FftPlan P, const double2 *__restrict__ tw, int nxh, int nz,
                                                           int ny, int B_, size_t field, size_t pfield,
                                                           double rescale, uint64_t map) {
  extern __shared__ double2 lds[];
  const int B = F::B ? F::B : B_;
  const Blk bk = xcd_block();
  const int kx0 = bk.x * B, y = bk.y, g = bk.z;
  const int slot = map_slot(map, g);
  const int N = F::N ? F::N : P.N, nh = nz / 2;
  const double2 *src = pad + (size_t)g * pfield + (size_t)y * N * nxh + kx0;
  for (int e = threadIdx.x; e < N * B; e += F::T) {
    const int z = e / B, b = e % B;
    lds[F::at(z, b, B)] = src[(size_t)z * nxh + b];
  }
  __syncthreads();
  F::run(lds, B, P, tw, -1);
  double2 *dst = U + (size_t)slot * field + (size_t)y * nz * nxh + kx0;
  for (int e = threadIdx.x; e < nz * B; e += F::T) {
    const int j = e / B, b = e % B;
    double2 val;
    if (j < nh) val = lds[F::at(j, b, B)];
    else if (j == nh) val = c2(0.0, 0.0);
    else val = lds[F::at(j + nh, b, B)];
    dst[(size_t)j * nxh + b] = c2(val.x * rescale, val.y * rescale);
  }
}

// ---------------------------------------- streamed z transforms (N = 1152)
// One block per CU walks a strided list of tiles (4 kx columns of one y and
// one field) and keeps two tile buffers in LDS.  The next tile is copied
// global -> LDS directly (global_load_lds_dwordx4, no registers) while the
// current one is transformed, so HBM keeps streaming during the FFT; the
// passes wait for LDS operations only.  The buffers are separate arrays and
// the loop body is written out once per buffer, so the compiler can tell
// that the transform's LDS accesses do not touch the buffer being filled and
// does not hold them until the copy has landed.  Tiles are numbered kx tile
// fastest; each XCD takes a contiguous run of blocks.
namespace zs {
constexpr int N = 1152, B = 4, NE = N * B, T = 768, NTW = N / 8;  // twiddles W^k, k < N/8
using L = Lines<N, B, false>;
template <int SIGN>
__device__ __forceinline__ void fft(double2 *x, const double2 *tw) {
  ct_passes_lds<T, L, N, B, SIGN, 1, 8, 12, 12>(x, tw);
}
}  // namespace zs
__shared__ double2 zs_buf0[zs::NE];
__shared__ double2 zs_buf1[zs::NE];
__shared__ double2 zs_tw[zs::NTW];

struct ZTile {
  int x, y, f;
};
__device__ __forceinline__ ZTile z_tile(size_t t, int ntx, int ny) {
  ZTile z;
  z.x = (int)(t % ntx);
  const size_t r = t / ntx;
  z.y = (int)(r % ny);
  z.f = (int)(r / ny);
  return z;
}
__device__ __forceinline__ size_t xcd_rank() {
  const size_t G = gridDim.x;
  return (G % 8 == 0) ? (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8 : blockIdx.x;
}
// everything this wave has issued, LDS copies included, has landed
__device__ __forceinline__ void all_sync() {
  __builtin_amdgcn_s_waitcnt(0);
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}
// The wait at the top of the streamed x loop.  Vector memory operations count
// down in issue order (vmcnt), and each thread issues exactly STORES global
// stores (the previous line's output) after the copy of the line about to
// be transformed: vmcnt <= STORES means that copy has landed while those
// stores may still be in flight.  The barrier: every wave is done with the
// previous line.  Before the first line (no stores yet) the loop waits for
// everything.  The compiler must be able to count those stores (an unrolled
// store loop), or it adds waits of its own before the LDS accesses that
// could alias the older copy.
template <int STORES>
__device__ __forceinline__ void stream_sync() {
  static_assert(STORES >= 0 && STORES < 16, "low vmcnt field");
  __builtin_amdgcn_s_waitcnt(0x0070 | STORES);  // vmcnt(STORES) expcnt(7) lgkmcnt(0)
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}
// Copy of element e = swz(p) of a tile into LDS slot p (the swizzle is an
// involution, so slot p = L::at(e / B, e % B)); the wave's 64 lanes fill 64
// consecutive slots.  row(z) gives the global row of tile row z, or < 0 for
// a row of zeros (written by the threads; the copy skips those lanes).
// The zeros are written before any copy is issued: an LDS store issued
// after a copy into the same buffer would have to wait for that copy.
template <bool ZEROS, class Row>
__device__ __forceinline__ void zs_issue(const double2 *src, int nxh, double2 *buf, Row row) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (ZEROS) {
    for (int p = threadIdx.x; p < zs::NE; p += zs::T)
      if (row(lds_swz(p) / zs::B) < 0) buf[p] = c2(0.0, 0.0);
  }
  for (int i = wave; i < zs::NE / 64; i += zs::T / 64) {
    const int p = i * 64 + lane, e = lds_swz(p);
    const int gr = row(e / zs::B);
    if (!ZEROS || gr >= 0)
      __builtin_amdgcn_global_load_lds((const void *)(src + (size_t)gr * nxh + e % zs::B),
                                       (__attribute__((address_space(3))) void *)(buf + i * 64), 16, 0, 0);
  }
}

// y_to_x for the 4 state fields (slots 0,1,2,6), as zinv_kernel; grid = one
// block per CU, tiles ntx*ny*4
__device__ __forceinline__ void zinv_stream_tile(double2 *x, double2 *__restrict__ pad, size_t t, int ntx, int ny,
                                                 int nxh, size_t pfield) {
  zs::fft<+1>(x, zs_tw);
  const ZTile z = z_tile(t, ntx, ny);
  double2 *dst = pad + (size_t)z.f * pfield + (size_t)z.y * zs::N * nxh + (size_t)z.x * zs::B;
  for (int e = threadIdx.x; e < zs::NE; e += zs::T) {
    const int zz = e / zs::B, b = e % zs::B;
    dst[(size_t)zz * nxh + b] = x[zs::L::at(zz, b)];
  }
  pass_sync<true>();  // every wave has read x before it is refilled
}
__global__ void __launch_bounds__(zs::T, 3) zinv_stream_kernel(const double2 *__restrict__ U, double2 *__restrict__ pad,
                                                               const double2 *__restrict__ twg, int nxh, int nz,
                                                               int ny, size_t field, size_t pfield) {
  for (int i = threadIdx.x; i < zs::NTW; i += zs::T) zs_tw[i] = twg[i];
  const int ntx = nxh / zs::B, nh = nz / 2;
  const size_t ntiles = (size_t)ntx * ny * 4, G = gridDim.x;
  // padded row zz <- state row zz (zz < nz/2) or zz - nz/2 (zz > nz); zeros between
  auto issue = [&](size_t t, double2 *buf) {
    const ZTile z = z_tile(t, ntx, ny);
    const int slot = (z.f == 3) ? 6 : z.f;
    const double2 *src = U + (size_t)slot * field + (size_t)z.y * nz * nxh + (size_t)z.x * zs::B;
    zs_issue<true>(src, nxh, buf, [&](int zz) { return zz < nh ? zz : (zz > 2 * nh ? zz - nh : -1); });
  };
  size_t t = xcd_rank();
  if (t < ntiles) issue(t, zs_buf0);
  while (t < ntiles) {
    all_sync();  // tile t is in buffer 0
    if (t + G < ntiles) issue(t + G, zs_buf1);
    zinv_stream_tile(zs_buf0, pad, t, ntx, ny, nxh, pfield);
    t += G;
    if (t >= ntiles) break;
    all_sync();  // tile t is in buffer 1
    if (t + G < ntiles) issue(t + G, zs_buf0);
    zinv_stream_tile(zs_buf1, pad, t, ntx, ny, nxh, pfield);
    t += G;
  }
}

// x_to_y + rescale for nf = 8 (or 9) product fields, as zfwd_kernel
__device__ __forceinline__ void zfwd_stream_tile(double2 *x, double2 *__restrict__ U, size_t t, int ntx, int ny,
                                                 int nxh, int nz, size_t field, double rescale) {
  zs::fft<-1>(x, zs_tw);
  const ZTile z = z_tile(t, ntx, ny);
  const int nh = nz / 2;
  const int slot = z.f < 5 ? z.f : (z.f < 8 ? z.f + 1 : 5);
  double2 *dst = U + (size_t)slot * field + (size_t)z.y * nz * nxh + (size_t)z.x * zs::B;
  for (int e = threadIdx.x; e < nz * zs::B; e += zs::T) {
    const int j = e / zs::B, b = e % zs::B;
    double2 val;
    if (j < nh) val = x[zs::L::at(j, b)];
    else if (j == nh) val = c2(0.0, 0.0);
    else val = x[zs::L::at(j + nh, b)];
    dst[(size_t)j * nxh + b] = c2(val.x * rescale, val.y * rescale);
  }
  pass_sync<true>();
}
__global__ void __launch_bounds__(zs::T, 3) zfwd_stream_kernel(const double2 *__restrict__ pad, double2 *__restrict__ U,
                                                               const double2 *__restrict__ twg, int nxh, int nz,
                                                               int ny, size_t field, size_t pfield, double rescale,
                                                               int nf) {
  for (int i = threadIdx.x; i < zs::NTW; i += zs::T) zs_tw[i] = twg[i];
  const int ntx = nxh / zs::B;
  const size_t ntiles = (size_t)ntx * ny * nf, G = gridDim.x;
  auto issue = [&](size_t t, double2 *buf) {
    const ZTile z = z_tile(t, ntx, ny);
    const double2 *src = pad + (size_t)z.f * pfield + (size_t)z.y * zs::N * nxh + (size_t)z.x * zs::B;
    zs_issue<false>(src, nxh, buf, [](int zz) { return zz; });
  };
  size_t t = xcd_rank();
  if (t < ntiles) issue(t, zs_buf0);
  while (t < ntiles) {
    all_sync();
    if (t + G < ntiles) issue(t + G, zs_buf1);
    zfwd_stream_tile(zs_buf0, U, t, ntx, ny, nxh, nz, field, rescale);
    t += G;
    if (t >= ntiles) break;
    all_sync();
    if (t + G < ntiles) issue(t + G, zs_buf0);
    zfwd_stream_tile(zs_buf1, U, t, ntx, ny, nxh, nz, field, rescale);
    t += G;
  }
}

// ------------------------------------------- streamed x pass (nx32 = 1536)
// One block per CU walks a strided list of (y, z) lines (line l = y*nz32 + z).
// The four input rows of the next line (u, v, w, theta, modes 0..nxh-1) are
// copied global -> LDS directly while the current line is transformed.  All
// eight products go through one forward transform of four lines
// (P1 + i P2, P3 + i P4, P5 + i P7, P8 + i P9), so no product is held in
// registers across a transform; the completion is folded into the first
// inverse pass, the products into the last inverse pass and the unpack into
// the last forward pass, so a line takes 9 barriers instead of 24.
// Same passes, twiddles and arithmetic as xphys_kernel<FftX1536>; the
// complex products inside the transforms may pair into FMAs differently
// (they are contracted), so the two agree to a few ulp, not bit for bit.
// The v*v round of statistics time stays with
// xphys_kernel.  Each block writes one CFL partial maximum, over all its
// lines (vpart[blockIdx.x]).
// Vector memory operations count down in issue order (vmcnt), so no global
// load is issued between a copy and the wait for it (the twiddles live in
// LDS, dy is read before the copy): it would also wait for the copy.  The
// wait at the top of the loop lets the previous line's stores stay in
// flight (stream_sync).
namespace xs {
constexpr int N = 1536, NH = 512, NTW = N / 6;  // twiddles W^k, k < N/6 (radix 16, 16, 6)
using L4 = Lines<N, 4, true>;
using L2 = Lines<N, 2, true>;  // lines 0 and 1 of the L4 layout (index b*N + n, padded)
}  // namespace xs
__shared__ double2 xs_buf[xs::L4::SIZE];
__shared__ double2 xs_in[4 * xs::NH];
__shared__ double2 xs_tw[xs::NTW];

// Lanes 32..63 of a trade places with lanes 0..31 of b (v_permlane32_swap,
// a VALU exchange between the two halves of a wave), per 32-bit half.
__device__ __forceinline__ void swap_halves(double &a, double &b) {
  const unsigned long long ua = (unsigned long long)__double_as_longlong(a);
  const unsigned long long ub = (unsigned long long)__double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
  a = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
  b = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}

// Last inverse pass (radix 6 over Ns = 256; line 0 = u + i v, line 1 =
// w + i theta) fused with the products.  Butterflies (0, k) and (1, k) sit in
// lanes l and l + 32 of one wave (k = 32 w + l % 32).  After the DFT, a swap
// of the wave halves pairs outputs r and r + 3: lanes 0..31 then hold both
// lines at m = k + 256 r (r < 3), lanes 32..63 at m = k + 256 (r + 3).  Each
// lane forms the products (P1 + i P2, P3 + i P4, P5 + i P7, P8 + i P9) and
// CFL velocities of its three points and stores lines 0..3 there.  A wave
// writes only positions k + 256 r of its own k (read by its own lanes before)
// and of lines 2, 3, so no barrier is needed inside.  Same butterflies (bfly)
// and product arithmetic as ct_pass and xphys_kernel.
template <int T>
__device__ __forceinline__ void xs_last_inverse_products(double2 *x, const double2 *tw, double dxr, double dy,
                                                         double dzr, double &vmax) {
  using L4 = xs::L4;
  constexpr int N = xs::N, R = 6, Ns = 256, NR = N / R, tstep = N / (Ns * R);
  static_assert(T == 8 * 64 && NR == Ns && Ns == 8 * 32, "8 waves x 32 butterfly pairs");
  const int lane = threadIdx.x % 64, b = lane / 32, k = (threadIdx.x / 64) * 32 + lane % 32;
  double2 v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = x[lds_at<L4, NR, 0>(k, b, r)];
  bfly<R, +1, Ns, tstep>(v, tw, k);
  const int r0 = b ? 3 : 0;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double2 A = v[r], Bw = v[r + 3];  // lower half: line 0 at r, r + 3; upper half: line 1
    swap_halves(A.x, Bw.x);
    swap_halves(A.y, Bw.y);            // now A = line 0, Bw = line 1 at r + r0
    const int m = k + Ns * (r + r0);
    const double t1 = A.x, t2 = A.y, t3 = Bw.x, t4 = Bw.y;
    const double vel = (fabs(t1) * dxr) + (fabs(t2) * dy) + (fabs(t3) * dzr);
    vmax = nan_max(vel, vmax);
    const double bv2 = t2 * t2;
    x[L4::at(m, 0)] = c2(t1 * t1 - bv2, t1 * t2);
    x[L4::at(m, 1)] = c2(t3 * t3 - bv2, t2 * t3);
    x[L4::at(m, 2)] = c2(t1 * t3, t1 * t4);
    x[L4::at(m, 3)] = c2(t2 * t4, t3 * t4);
  }
}

template <int T>
__device__ __forceinline__ void xs_first_pass_completion(const double2 *in, double2 *x) {
  using L2 = xs::L2;
  constexpr int N = xs::N, NH = xs::NH, R = 16, NR = N / R, NBF = 2 * NR;
  static_assert(NBF <= T, "one butterfly per thread");
  const int t = threadIdx.x;
  if (t < NBF) {
    const int b = t / NR, j = t % NR;
    const double2 *p0 = in + (2 * b) * NH, *p1 = in + (2 * b + 1) * NH;
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int m = j + r * NR;
      const bool lo = m < NH, hi = m > N - NH;
      const int q = lo ? m : (hi ? N - m : 0);
      const double2 a = p0[q], c = p1[q];
      const double cy = lo ? c.y : -c.y, ay = lo ? a.y : -a.y;  // conjugate mirror: exact negations
      double2 z = c2(a.x - cy, ay + c.x);
      if (r == 0 && m == 0) z = c2(a.x, c.x);
      if (!lo && !hi) z = c2(0.0, 0.0);
      v[r] = z;
    }
    bfly<R, +1, 1, 1>(v, nullptr, 0);
#pragma unroll
    for (int r = 0; r < R; ++r) x[lds_at<L2, 1, 1>(16 * j, b, r)] = v[r];
  }
  pass_sync<true>();
}

// Last forward pass (radix 6 over Ns = 256: butterfly k of line b holds
// Z[k + 256 r], r < 6) fused with the unpack of the two real spectra of each
// line.  Thread (b, u) takes butterflies k1 = u and k2 = 256 - u (u = 0:
// k1 = 0, k2 = 128), which hold both halves of the pairs (m, N - m) for
// m = k1, k1 + 256, k2, k2 + 256: N - k1 - 256 r = k2 + 256 (5 - r).  The
// results stay in registers: no LDS store of the last pass, no reads of the
// unpack, one barrier less.  Same butterflies (bfly) and unpack arithmetic as
// ct_pass and xphys_kernel.  Every thread issues exactly 8 stores.
template <int T>
__device__ __forceinline__ void xs_last_pass_unpack(const double2 *x, const double2 *tw, double2 *line,
                                                    size_t pfield) {
  using L4 = xs::L4;
  constexpr int N = xs::N, R = 6, Ns = 256, NR = N / R, tstep = N / (Ns * R);
  static_assert(T == 4 * 128 && NR == Ns && xs::NH == 2 * Ns, "4 lines x 128 butterfly pairs");
  const int b = threadIdx.x / 128, u = threadIdx.x % 128;
  const int k1 = u, k2 = u ? Ns - u : Ns / 2;
  double2 v1[R], v2[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    v1[r] = x[lds_at<L4, NR, 0>(k1, b, r)];
    v2[r] = x[lds_at<L4, NR, 0>(k2, b, r)];
  }
  bfly<R, -1, Ns, tstep>(v1, tw, k1);
  bfly<R, -1, Ns, tstep>(v2, tw, k2);
  // Z[N - m]: u > 0 from the other butterfly; u = 0: Z[0] for m = 0 and the
  // same butterfly (k = 0, 128 pair with themselves)
  const double2 cA = u ? v2[5] : v1[0], cB = u ? v2[4] : v1[5];
  const double2 cC = u ? v1[5] : v2[5], cD = u ? v1[4] : v2[4];
  auto emit = [&](int m, double2 Z, double2 Zc) {
    line[(size_t)(2 * b) * pfield + m] = c2((Z.x + Zc.x) * 0.5, (Z.y - Zc.y) * 0.5);
    line[(size_t)(2 * b + 1) * pfield + m] = c2((Z.y + Zc.y) * 0.5, (Zc.x - Z.x) * 0.5);
  };
  emit(k1, v1[0], cA);
  emit(k1 + Ns, v1[1], cB);
  emit(k2, v2[0], cC);
  emit(k2 + Ns, v2[1], cD);
}

template <int T>
__global__ void __launch_bounds__(T, 1) xphys_stream_kernel(double2 *__restrict__ pad, const double2 *__restrict__ twg,
                                                            int nz32, int ny, size_t pfield,
                                                            const double *__restrict__ dyr, double dxr, double dzr,
                                                            double *__restrict__ vpart) {
  using namespace xs;
  constexpr int NW = T / 64, NCOPY = 4 * NH / 64, STORES = 8;  // stores per thread: xs_last_pass_unpack
  static_assert(T % 64 == 0 && NCOPY % NW == 0, "copy instructions split evenly over the waves");
  for (int i = threadIdx.x; i < NTW; i += T) xs_tw[i] = twg[i];
  const size_t nlines = (size_t)nz32 * ny, G = gridDim.x;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  // row f of line l, element m -> xs_in[f * NH + m]; instruction i fills 64 elements
  auto issue = [&](size_t l) {
    const double2 *src = pad + l * NH;
#pragma unroll
    for (int k = 0; k < NCOPY / NW; ++k) {
      const int i = wave + k * NW, f = i / (NH / 64), m = (i % (NH / 64)) * 64 + lane;
      __builtin_amdgcn_global_load_lds((const void *)(src + (size_t)f * pfield + m),
                                       (__attribute__((address_space(3))) void *)(xs_in + i * 64), 16, 0, 0);
    }
  };
  size_t l = xcd_rank();
  double vmax = 0.0;
  if (l < nlines) issue(l);
  __builtin_amdgcn_s_waitcnt(0);  // the first line's copy (no stores yet)
  while (l < nlines) {
    stream_sync<STORES>();  // line l is in xs_in; every wave is done with the previous line
    const double dy = dyr[l / nz32];
    xs_first_pass_completion<T>(xs_in, xs_buf);  // also: every wave is done with xs_in
    ct_passes_lds<T, L2, N, 2, +1, 16, 16>(xs_buf, xs_tw);  // the second pass; the third:
    xs_last_inverse_products<T>(xs_buf, xs_tw, dxr, dy, dzr, vmax);
    if (l + G < nlines) issue(l + G);
    pass_sync<true>();
    ct_passes_lds<T, L4, N, 4, -1, 1, 16, 16>(xs_buf, xs_tw);  // passes 1, 2; the third one:
    xs_last_pass_unpack<T>(xs_buf, xs_tw, pad + l * NH, pfield);
    l += G;
  }
  // the block's CFL maximum: across each wave by lane exchanges, then the waves
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) vmax = nan_max(vmax, __shfl_xor(vmax, d, 64));
  __shared__ double xs_vm[NW];
  if (lane == 0) xs_vm[wave] = vmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < NW; ++w) vmax = nan_max(vmax, xs_vm[w]);
    vpart[blockIdx.x] = vmax;
  }
}

// ------------------------------------------------- CFL maximum and dt update
// Device scalars of a substep: [0] dt, [1] sticky non-finite flag, [2]
// velmax (of the last substep 1), [3 + 4 (s-1) + q] the mean-line outputs of
// substep s (phi_u, phi_t, tau_w, q_w).
constexpr int SCAL_DT = 0, SCAL_NONFINITE = 1, SCAL_VELMAX = 2, SCAL_MEAN0 = 3;
constexpr int SCAL_N = SCAL_MEAN0 + 4 * 3;
// single block: velmax = max(vpart), stored when update has VM_STORE; with
// VM_DT (substep 1) dt = dt_factor*dt + (1-dt_factor)*cflx/velmax
// (dft_x_realspace.f90:489-498)
constexpr int VELMAX_THREADS = 1024;
constexpr int VM_STORE = 1, VM_DT = 2;  // velmax_kernel's update bits
__global__ void __launch_bounds__(VELMAX_THREADS) velmax_kernel(const double *__restrict__ vpart, size_t n,
                                                                double *__restrict__ scal, int update,
                                                                double dt_factor, double cflx) {
  __shared__ double red[VELMAX_THREADS / 64];
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  const size_t T = VELMAX_THREADS;
  size_t i = threadIdx.x;
  for (; i + 3 * T < n; i += 4 * T) {
    v0 = nan_max(v0, vpart[i]);
    v1 = nan_max(v1, vpart[i + T]);
    v2 = nan_max(v2, vpart[i + 2 * T]);
    v3 = nan_max(v3, vpart[i + 3 * T]);
  }
  for (; i < n; i += T) v0 = nan_max(v0, vpart[i]);
  double v = nan_max(nan_max(v0, v1), nan_max(v2, v3));
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v = nan_max(v, __shfl_xor(v, d, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < VELMAX_THREADS / 64; ++w) red[0] = nan_max(red[0], red[w]);
    const double vm = red[0];
    // velmax is a substep-1 quantity (dft_x_realspace.f90:125-127); the
    // other substeps only check that the state is finite
    if (update & VM_STORE) scal[SCAL_VELMAX] = vm;
    // a non-finite velocity anywhere: raise the sticky flag the host reads
    // after the substep (the state is left as it is)
    if (!(vm <= 1.79769313486231570e308)) scal[SCAL_NONFINITE] = 1.0;
    else if ((update & VM_DT) && dt_factor != 1.0)
      scal[SCAL_DT] = dt_factor * scal[SCAL_DT] + (1.0 - dt_factor) * cflx / vm;
  }
}

// ------------------------------------------------------------------ wavespace
__global__ void __launch_bounds__(64) wavespace_kernel(LineParams P, WsArgs A, size_t nsys, const double *scal) {
  const size_t sys = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sys >= nsys) return;
  LineParams Q = P;
  Q.dt = scal[0];
  wavespace_line(Q, A, sys);
}

// Transpose of the wall-normal stage, one thread per system (solver_adj.h);
// P.dt is the fixed dt of the substep, A.F its factors.
// Phase 1: grid.y = 3 chains; phase 2: the transpose of hvhg.
template <int BW>
__global__ void __launch_bounds__(64) adj_chain_kernel(LineParams P, AdjArgs A, size_t nsys) {
  const size_t sys = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sys >= nsys) return;
  adj_chain<BW>(P, A, sys, (int)blockIdx.y);
}
template <int BW>
__global__ void __launch_bounds__(64) adj_hvhg_part_kernel(LineParams P, AdjArgs A, size_t nsys) {
  const size_t sys = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sys >= nsys) return;
  adj_hvhg_part<BW>(P, A, sys, (int)blockIdx.y);
}
__global__ void adj_hvhg_combine_kernel(AdjArgs A, size_t field) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= field) return;
  adj_hvhg_combine(A, e % A.plane, (int)(e / A.plane));
}
// The factors of substep P.s at dt P.dt for every system; scr: 2 complex
// scratch columns per system.
__global__ void __launch_bounds__(64) adj_factor_kernel(LineParams P, AdjFactors F, size_t nsys, int nxh, int nh,
                                                        double lmfxr, double lmfzr, cplx *scr) {
  const size_t sys = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sys >= nsys) return;
  adj_factor_line(P, F, sys, nxh, nh, lmfxr, lmfzr, SV{scr + sys, F.plane}, SV{scr + F.field + sys, F.plane});
}

// The last transposes of a substep, per element of the [y][kz][kx] field:
// the detection planes (vd = r_lo . v, r_hi . v over y; rows: r_lo then r_hi)
// send avd back into the adjoint of v (slot 1), and the mixed-type source
// (u of the input state) sends slot 5 into the adjoint of u (slot 0).
__global__ void adj_finish_kernel(cplx *__restrict__ AU, const cplx *__restrict__ avd, const double *__restrict__ rows,
                                  size_t plane, size_t field, int ny, int mixed) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= field) return;
  const size_t sys = e % plane;
  const int y = (int)(e / plane);
  if (avd) {
    const cplx lo = avd[sys], hi = avd[plane + sys];
    AU[field + e] = cadd(AU[field + e], cadd(cscale(lo, rows[y]), cscale(hi, rows[ny + y])));
  }
  if (mixed) AU[e] = cadd(AU[e], AU[5 * field + e]);
}

// The (kx,kz) = (0,0) system: mean flow, mean spanwise velocity and mean
// scalar with the flux constraints (wave00.f90); one thread.
__global__ void wave00_kernel(LineParams P, WsArgs A, const double *scal) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  LineParams Q = P;
  Q.dt = scal[0];
  wavespace_line(Q, A, 0);
}

// The (0,0) system split by field, one block each: block 0 the mean
// streamwise velocity, 1 the spanwise velocity, 2 the scalar (the three
// solves of wave00.f90 are independent).  The block copies its columns of
// system 0 into LDS and runs mean_u / mean_w / mean_t of solver.h there with
// the same arithmetic: the band products and the elementwise updates row-
// parallel over the block's threads (each row's sum in band_mv's order), the
// recurrences on one thread each with register windows -- LU(D0)
// substitutions as deriv_line forms them, the Helmholtz LU fused with the
// forward substitution as sweep C forms it (crout_sweep with the mean
// operator: dummy = 1 + beta * 0 = 1, so 1 * D0 - beta D2 is helmholtz_lu's
// value), the two independent LU(D0) solves of substep 1 on two waves side
// by side -- and writes back what the one-thread path leaves in the state.
// U of the Helmholtz factor lives in LDS when db_scratch is null, otherwise
// in db_scratch + field * W * ny.
constexpr int MEAN_COLS = 6;
constexpr int MEAN_THREADS = 256;

// x <- LU(D0)^-1 x, one thread (band_bs with the shared factor)
template <int BW>
__device__ void mean_bs_d0(const BandGeom &g, const SharedBand &LU, const SV &x) {
  CWin<BW> f;
  f.clear();
  down_rows<BW>(g.n, [&](int r, auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    double lr[2 * BW + 1];
    load_lrow<BW, MID>(g, LU, r, lr);
    const cplx y = fwd_row<BW, MID>(g, r, lr, f, x, x[r]);
    x[r] = y;
    f.push(y);
  });
  CWin<2 * BW + 1> w;
  w.clear();
  up_rows<BW>(g.n, [&](int r, auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    double ur[2 * BW];
    load_urow<BW, MID>(g, LU, r, ur);
    const cplx v = bwd_row<BW, MID>(g, r, ur, w, x[r]);
    x[r] = v;
    w.push(v);
  });
}
// x <- H^-1 x for the mean-line Helmholtz operator D0 - beta D2 (diagonal
// wall rows), its LU formed row by row; U kept in Uo.  One thread.
template <int BW>
__device__ void mean_helmholtz(const LineParams &Q, double beta, const OwnU &Uo, const SV &x) {
  const BandGeom g{Q.g.n, BW};
  const HelmholtzA H{Q.D0, Q.D2, 1.0 + beta * 0.0, beta, g.n, BW};
  CWin<BW> f;
  f.clear();
  crout_sweep<BW>(g, H, Uo, [&](int r, const double(&L)[2 * BW + 1], auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    const cplx y = fwd_lrow<BW, MID>(g, r, L, f, x, x[r]);
    x[r] = y;
    f.push(y);
  });
  CWin<2 * BW + 1> w;
  w.clear();
  up_rows<BW>(g.n, [&](int r, auto kind) __attribute__((always_inline)) {
    constexpr bool MID = decltype(kind)::mid;
    double ur[2 * BW];
    load_urow<BW, MID>(g, Uo, r, ur);
    const cplx v = bwd_row<BW, MID>(g, r, ur, w, x[r]);
    x[r] = v;
    w.push(v);
  });
}
// b = A x row-parallel over the block (band_mv's order within a row)
__device__ void mean_mv(const BandGeom &g, const SharedBand &A, const SV &x, const SV &b) {
  for (int r = threadIdx.x; r < g.n; r += blockDim.x) {
    const int f = g.first(r), l = g.last(r);
    cplx v = cscale(x[f], A(r, g.slot(r, f)));
    for (int c = f + 1; c <= l; ++c) v = cadd(v, cscale(x[c], A(r, g.slot(r, c))));
    b[r] = v;
  }
}

template <int BW>
__global__ void __launch_bounds__(MEAN_THREADS) mean_kernel(LineParams P, WsArgs A, const double *scal,
                                                            double *db_scratch) {
  extern __shared__ double2 lds[];
  const int n = P.g.n, W = P.g.W(), f = blockIdx.x, tid = threadIdx.x, T = MEAN_THREADS;
  cplx *col = reinterpret_cast<cplx *>(lds);
  double *db = db_scratch ? db_scratch + (size_t)f * W * n : reinterpret_cast<double *>(lds + MEAN_COLS * n);
  auto gU = [&](int slot) { return A.U + (size_t)slot * A.field; };
  auto gR = [&](int slot) { return A.RK + (size_t)slot * A.field; };
  // global source of each LDS column (nullptr: scratch, not loaded)
  cplx *src[MEAN_COLS] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  if (f == 0) {         // u1 u2 u3 st1 buf1
    src[1] = gU(1); src[3] = gR(0);
  } else if (f == 1) {  // u2 u3 u4 u5 st2 buf1
    src[2] = gU(3); src[4] = gR(1);
  } else {              // u7 u8 u9 st3 buf1 (usrc)
    src[1] = gU(7); src[3] = gR(2);
    if (A.usrc) src[5] = const_cast<cplx *>(A.usrc);
  }
  for (int q = 0; q < MEAN_COLS; ++q)
    if (src[q])
      for (int i = tid; i < n; i += T) col[q * n + i] = src[q][(size_t)i * A.plane];
  __syncthreads();
  LineParams Q = P;
  Q.dt = scal[0];
  const BandGeom g{n, BW};
  const OwnU Uo{db, 1, W};
  auto L = [&](int q) { return SV{col + q * n, 1}; };
  const int s = Q.s, S = Q.S;
  const double dt = Q.dt, nu = Q.nu, pr = Q.pr;
  // columns of this line: out (the mean profile), prod (the y-derivative
  // source: uv, vw, v theta), d (D1 LU(D0)^-1 prod), st (the store), b (rhs)
  const bool isT = f == 2, isW = f == 1;
  const SV out = isW ? L(1) : L(0), prod = isW ? L(2) : L(1), d = isW ? L(3) : L(2), st = isW ? L(4) : L(3),
           b = isW ? L(5) : L(4), us = L(5);
  const bool mixed = isT && A.usrc;
  const double beta = isT ? dt * Q.coef(1, s) * nu / pr : dt * Q.coef(1, s) * nu;
  // substep 1: out <- LU(D0)^-1 st on wave 0 beside prod <- LU(D0)^-1 prod on wave 1
  if (s == 1) {
    for (int i = tid; i < n; i += T) out[i] = st[i];
    __syncthreads();
    if (tid == 0) mean_bs_d0<BW>(g, Q.D0LU, out);
  }
  if (tid == 64) mean_bs_d0<BW>(g, Q.D0LU, prod);
  __syncthreads();
  if (s == 1) mean_mv(g, Q.D2, out, b);
  mean_mv(g, Q.D1, prod, d);
  __syncthreads();
  const double gamma = dt * Q.coef(2, s);
  const double alpha = isT ? dt * Q.coef(0, s) * nu / pr : dt * Q.coef(0, s) * nu;
  const double ug = f == 0 ? Q.u_gal : 0.0, one_im = mixed ? 0.0 : 1.0;
  for (int i = tid; i < n; i += T) {
    if (mixed) d[i] = csub(d[i], cscale(us[i], A.src_scale));
    cplx v = mk(0.0, 0.0);
    if (i >= 1 && i < n - 1) {
      const cplx t = s == 1 ? csub(cadd(st[i], cscale(b[i], alpha)), cscale(d[i], gamma)) : csub(st[i], cscale(d[i], gamma));
      v = f == 0 ? mk(t.re + ug, one_im) : mk(t.re, one_im);
    }
    b[i] = v;
  }
  __syncthreads();
  __shared__ double phi_s;
  if (tid == 0) {
    mean_helmholtz<BW>(Q, beta, Uo, b);
    double phi = 0.0;
    if (f == 0) phi = calc_phi(Q.bulk_u, integrate(Q, b));
    if (f == 1) {
      phi = calc_phi(Q.bulk_w, integrate(Q, b));
      phi = 0.0;  // wave00.f90:396
    }
    if (f == 2) phi = mixed ? 0.0 : calc_phi(Q.bulk_t, integrate(Q, b));
    phi_s = phi;
    if (A.mean_out && f != 1) {
      A.mean_out[f == 0 ? 0 : 1] = phi;
      A.mean_out[f == 0 ? 2 : 3] = 0.0;
      if (s == S) {
        // wall_d1 of the corrected column (re + phi im on rows 0, 1, n-2, n-1)
        cplx c4[4];
        const int rows[4] = {0, 1, n - 2, n - 1};
        for (int q = 0; q < 4; ++q) c4[q] = mk(b[rows[q]].re + phi * b[rows[q]].im, 0.0);
        const int m = BW;
        const double wd = (((Q.D1(0, m) * c4[0].re + Q.D1(0, m + 1) * c4[1].re) - Q.D1(n - 1, m - 1) * c4[2].re) -
                           Q.D1(n - 1, m) * c4[3].re);
        A.mean_out[f == 0 ? 2 : 3] = f == 0 ? 0.5 * nu * wd : 0.5 * nu / pr * wd;
      }
    }
  }
  __syncthreads();
  const double phi = phi_s;
  for (int i = tid; i < n; i += T) b[i] = mk(b[i].re + phi * b[i].im, 0.0);
  __syncthreads();
  mean_mv(g, Q.D0, b, out);
  __syncthreads();
  if (f == 0)
    for (int i = tid; i < n; i += T) out[i] = mk(out[i].re - Q.u_gal, out[i].im);
  if (s != S) {
    const double al2 = isT ? dt * Q.coef(0, s + 1) * nu / pr : dt * Q.coef(0, s + 1) * nu;
    const double zeta = dt * Q.coef(3, s + 1);
    // st = D2 b row by row, then st = al2 st + out - zeta d (each row reads only its own values)
    for (int r = tid; r < n; r += T) {
      const int fr = g.first(r), lr = g.last(r);
      cplx v = cscale(b[fr], Q.D2(r, g.slot(r, fr)));
      for (int c = fr + 1; c <= lr; ++c) v = cadd(v, cscale(b[c], Q.D2(r, g.slot(r, c))));
      st[r] = csub(cadd(cscale(v, al2), out[r]), cscale(d[r], zeta));
    }
  } else {
    for (int i = tid; i < n; i += T) st[i] = out[i];
  }
  if (isW)
    for (int i = tid; i < n; i += T) L(0)[i] = mk(0.0, 0.0);  // mean_w leaves u2 = 0
  __syncthreads();
  // results: f 0 -> slot 0 (U), slot 1 (V = 0, written here because this
  // block is the one that reads slot 1), RK 0; f 1 -> slots 2-4, RK 1;
  // f 2 -> slots 6-8, RK 2
  cplx *dst[MEAN_COLS] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  if (f == 0) {
    dst[0] = gU(0); dst[3] = gR(0);
  } else if (f == 1) {
    dst[1] = gU(2); dst[2] = gU(3); dst[3] = gU(4); dst[4] = gR(1);
  } else {
    dst[0] = gU(6); dst[1] = gU(7); dst[2] = gU(8); dst[3] = gR(2);
  }
  for (int q = 0; q < MEAN_COLS; ++q)
    if (dst[q])
      for (int i = tid; i < n; i += T) dst[q][(size_t)i * A.plane] = col[q * n + i];
  if (f == 0)
    for (int i = tid; i < n; i += T) gU(1)[(size_t)i * A.plane] = mk(0.0, 0.0);
}

// Block order of the per-system sweeps that read what is stored once per
// (kx, +-kz) pair (U factors, homogeneous solutions): the blocks of kz rows
// j and nz-j with the same kx range are numbered back to back on one XCD
// (hardware blocks b and b+8 run on the same XCD), so the partner's reads
// of the shared rows mostly hit that XCD's L2.  Other shapes keep the
// plain order.
__device__ __forceinline__ size_t pair_block(const WsArgs &A) {
  const size_t G = gridDim.x, b = blockIdx.x;
  const int cpr = A.nxh / 64;  // blocks per kz row
  if (blockDim.x != 64 || A.nxh % 64 != 0 || G % 8 != 0 || G != (size_t)A.nz * cpr || A.nh < 1) return b;
  const size_t L = (b % 8) * (G / 8) + b / 8;  // each XCD a contiguous range
  const size_t npb = (size_t)(A.nh - 1) * cpr * 2;
  int row, c;
  if (L < npb) {
    const size_t q = L / 2;
    const int j = 1 + (int)(q / cpr);
    c = (int)(q % cpr);
    row = (L % 2) ? A.nz - j : j;
  } else {  // kz rows 0 and nz/2 have no partner
    const size_t l2 = L - npb;
    row = (l2 < (size_t)cpr) ? 0 : A.nh;
    c = (int)(l2 % cpr);
  }
  return (size_t)row * cpr + c;
}

// Fused sweeps over all (kx,kz) != (0,0) systems, one thread per system.
template <int BW, int SW>
__global__ void __launch_bounds__(64) sweep_kernel(LineParams P, WsArgs A, double *dbuf2, size_t nthreads,
                                                   const double *scal) {
  constexpr bool shared_reads = SW == SW_DNU || SW == SW_DT || SW == SW_F || SW == SW_G;
  const size_t blk = shared_reads ? pair_block(A) : (size_t)blockIdx.x;
  const size_t idx = blk * blockDim.x + threadIdx.x;
  if (idx >= nthreads) return;
  LineParams Q = P;
  Q.dt = scal[0];
  run_sweep<BW, SW>(Q, A, dbuf2, idx);
}

// opposition-control detection planes of the state's v (slot 1):
// vd[0][sys] = r_lo . v, vd[1][sys] = r_hi . v (rows = r_lo, r_hi)
__global__ void __launch_bounds__(64) opp_kernel(const cplx *U, cplx *vd, size_t plane, size_t nsys, int n,
                                                 const double *rows) {
  const size_t sys = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sys >= nsys) return;
  const SV v{const_cast<cplx *>(U) + (size_t)1 * plane * n + sys, plane};
  cplx lo, hi;
  opp_dots(v, rows, n, &lo, &hi);
  vd[sys] = lo;
  vd[plane + sys] = hi;
}

// --------------------------------------------- statistics-time pressure
// <vv>(y): one thread per y plane, serial over the plane's modes
__global__ void __launch_bounds__(64) vv_kernel(const cplx *U, size_t plane, int ny, int nxh, double *vv) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= ny) return;
  vv[y] = mean_vv_row(U + (size_t)1 * plane * ny + (size_t)y * plane, nxh, plane);
}
// statistics profiles (get_statistics_1): one thread per (y, product) for the
// ten Parseval sums, serial over the plane's modes; stat: 16 lines of ny,
// then u_center, t_center.  The ten products of one y are neighbouring
// threads, so a wave's loads touch the rows of about six y (each row a
// separate memory page) instead of 64.
__constant__ int kStatPairs[10][2] = {{0, 0}, {1, 1}, {2, 2}, {6, 6}, {0, 1}, {0, 2}, {1, 2}, {0, 6}, {6, 2}, {1, 6}};
__global__ void __launch_bounds__(64) stats_uv_kernel(const cplx *U, size_t plane, int ny, int nxh, double *stat) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 10 * ny) return;
  const int q = t % 10, y = t / 10;
  const size_t field = plane * ny;
  const cplx *a = U + (size_t)kStatPairs[q][0] * field + (size_t)y * plane;
  const cplx *b = U + (size_t)kStatPairs[q][1] * field + (size_t)y * plane;
  stat[(size_t)(6 + q) * ny + y] = mean_uv_row(a, b, nxh, plane);
}
// the mean-line profiles and centre values: thread q < 3 takes line q (U,
// W, Theta) with scratch columns q of the two scratch fields
__global__ void stats_mean_kernel(LineParams P, const cplx *U, cplx *tmp, size_t plane, const double *center,
                                  double u_gal, double *stat) {
  const int q = threadIdx.x;
  if (blockIdx.x != 0 || q >= 3) return;
  const size_t field = plane * P.g.n;
  const size_t slot = q == 0 ? 0 : (q == 1 ? 2 : 6);
  mean_line_stats_q(P, SV{const_cast<cplx *>(U) + slot * field, plane}, q, SV{tmp + q, plane},
                    SV{tmp + field + q, plane}, center, u_gal, stat, stat + (size_t)16 * P.g.n);
}

// ------------------------------------------ 1-D energy spectra (write_spectra)
// du/dy, dw/dy of every system, (0,0) included, into the scratch columns
template <int BW>
__global__ void __launch_bounds__(64) deriv_kernel(LineParams P, cplx *U, cplx *buf, size_t field, size_t plane) {
  const size_t sys = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sys >= plane) return;
  deriv_line<BW>(P, SV{U + sys, plane}, SV{U + 2 * field + sys, plane}, SV{buf + sys, plane},
                 SV{buf + field + sys, plane});
}
// ------------------------------------------------- get_statistics_2
// moment sums per (q, y): grid (ny, NMOM), the partials of all lines and waves
__global__ void __launch_bounds__(256) stats2_mom_kernel(const double *mom, size_t per_y, double *msum) {
  __shared__ double red[256];
  const int y = blockIdx.x, q = blockIdx.y;
  const double *src = mom + ((size_t)q * gridDim.x + y) * per_y;
  double a = 0.0;
  for (size_t i = threadIdx.x; i < per_y; i += 256) a += src[i];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) msum[(size_t)q * gridDim.x + y] = red[0];
}
// du/dy and dv/dy of the statistics copies of u and v (get_turb_transport_PI1_ij's
// labs_bs + labs_mat_mul) into the scratch columns
template <int BW>
__global__ void __launch_bounds__(64) stats2_deriv_kernel(LineParams P, const cplx *uvw, cplx *buf, size_t field,
                                                          size_t plane) {
  const size_t sys = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sys >= plane) return;
  deriv_line<BW>(P, SV{const_cast<cplx *>(uvw) + sys, plane}, SV{const_cast<cplx *>(uvw) + field + sys, plane},
                 SV{buf + sys, plane}, SV{buf + field + sys, plane});
}
// Parseval sums per y (get_uv, get_statistics.f90:982-1017;
// get_turb_transport_PI1_ij, 817-902): out[0..7][y] = up, vp, wp, pp,
// 2pdudx, 2pdvdy, 2pdwdz, pdudy+pdvdx.  One block per y: each thread sums
// the lines l = tid, tid + T, ... (weights 1 at kx = 0, 2 otherwise, the
// reference's factors), then a tree over the block.  The order of the sum
// differs from the reference's serial loop (a few ulp of the line).
constexpr int PV_THREADS = 256;
__global__ void __launch_bounds__(PV_THREADS) stats2_pv_kernel(const cplx *uvw, const cplx *p, const cplx *d,
                                                               size_t plane, int ny, int nxh, int nh, double lmfxr,
                                                               double lmfzr, double *out) {
  __shared__ double red[8][PV_THREADS];
  const int y = blockIdx.x, t = threadIdx.x;
  const size_t field = plane * ny, row = (size_t)y * plane;
  const cplx *u = uvw + row, *v = uvw + field + row, *w = uvw + 2 * field + row, *pr = p + row;
  const cplx *du = d + row, *dv = d + field + row;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t l = t; l < plane; l += PV_THREADS) {
    const int k = (int)(l % nxh), j = (int)(l / nxh);
    const int kzi = (j <= nh) ? j : j + 1 - 2 * nh - 1;  // module_ppdfftw.f90:360-374
    const double kxw = lmfxr * k, kzw = lmfzr * kzi, wt = (k == 0) ? 1.0 : 2.0;
    const cplx a = u[l], b = v[l], c = w[l], q = pr[l], da = du[l], db = dv[l];
    acc[0] += wt * (a.re * q.re + a.im * q.im);
    acc[1] += wt * (b.re * q.re + b.im * q.im);
    acc[2] += wt * (c.re * q.re + c.im * q.im);
    acc[3] += wt * (q.re * q.re + q.im * q.im);
    acc[4] += 2.0 * wt * kxw * (a.re * q.im - a.im * q.re);
    acc[5] += 2.0 * wt * (db.re * q.re + db.im * q.im);
    acc[6] += 2.0 * wt * kzw * (c.re * q.im - c.im * q.re);
    acc[7] += wt * (kxw * (b.re * q.im - b.im * q.re) + (da.re * q.re + da.im * q.im));
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) red[q][t] = acc[q];
  __syncthreads();
  for (int h = PV_THREADS / 2; h > 0; h >>= 1) {
    if (t < h)
#pragma unroll
      for (int q = 0; q < 8; ++q) red[q][t] += red[q][t + h];
    __syncthreads();
  }
  if (t < 8) out[(size_t)t * ny + y] = red[t][0];
}
// the 29 datasets of get_statistics_2 from the moment sums and the Parseval
// sums: the copies spread over the block, the six y derivatives (a band
// solve and a band product each, ny long) one per thread, each in its own
// two scratch columns tmp[2 q], tmp[2 q + 1] (ny complex each)
constexpr int STAT2_DDY = 6;
__global__ void __launch_bounds__(64) stats2_final_kernel(LineParams P, const double *msum, const double *pv,
                                                          double rescale, cplx *tmp, double *out) {
  if (blockIdx.x != 0) return;
  const int n = P.g.n, t = threadIdx.x;
  const int mom_of[15] = {0, 1, 2, 3, 4, 5, 10, 11, 12, 13, 14, 15, 16, 17, 18};
  for (int i = t; i < 15 * n; i += blockDim.x) {
    const int q = i / n, y = i % n;
    out[i] = (y == 0 || y == n - 1) ? 0.0 : rescale * msum[(size_t)mom_of[q] * n + y];
  }
  for (int i = t; i < n; i += blockDim.x) {
    out[(size_t)19 * n + i] = pv[i];
    out[(size_t)21 * n + i] = pv[(size_t)1 * n + i];
    out[(size_t)23 * n + i] = pv[(size_t)2 * n + i];
    out[(size_t)24 * n + i] = pv[(size_t)3 * n + i];
    for (int q = 0; q < 4; ++q) out[(size_t)(25 + q) * n + i] = pv[(size_t)(4 + q) * n + i];
  }
  if (t < STAT2_DDY) {
    // d/dy of the triple moments (-d<uuv>/dy ... into lines 15..18), of <up>
    // (line 20, times -1) and of <vp> (line 22, times -2)
    const int tri[4] = {6, 8, 7, 9};
    const double *x = t < 4 ? msum + (size_t)tri[t] * n : pv + (size_t)(t - 4) * n;
    const double scale_in = t < 4 ? rescale : 1.0;
    const double scale_out = t < 4 ? 1.0 : (t == 4 ? -1.0 : -2.0);
    double *o = out + (size_t)(t < 4 ? 15 + t : (t == 4 ? 20 : 22)) * n;
    const SV a{tmp + (size_t)(2 * t) * n, 1}, b{tmp + (size_t)(2 * t + 1) * n, 1};
    for (int i = 0; i < n; ++i) a[i] = mk(x[i] * scale_in, 0.0);
    band_bs(P.g, P.D0LU, a);
    band_mv(P.g, P.D1, a, b);
    for (int i = 0; i < n; ++i) o[i] = scale_out * b[i].re;
  }
}

// vorticity into slots 3, 4, 5 (get_vorticity_field, get_statistics.f90:912-948):
// ox = dw/dy - (0,kz) v, oy = (0,kz) u - (0,kx) w, oz = (0,kx) v - du/dy with (0,k) = dcmplx(0d0,k)
__global__ void vort_kernel(cplx *U, const cplx *buf, size_t field, size_t plane, int nxh, int nh, double lmfxr,
                            double lmfzr, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t sys = i % plane;
  const int k = (int)(sys % nxh), j = (int)(sys / nxh);
  const int kzi = (j <= nh) ? j : j + 1 - 2 * nh - 1;  // module_ppdfftw.f90:360-374
  const cplx ikx = mk(0.0, k * lmfxr), ikz = mk(0.0, kzi * lmfzr);
  const cplx u = U[i], v = U[field + i], w = U[2 * field + i], du = buf[i], dw = buf[field + i];
  U[3 * field + i] = csub(dw, cmul(ikz, v));
  U[4 * field + i] = csub(cmul(ikz, u), cmul(ikx, w));
  U[5 * field + i] = csub(cmul(ikx, v), du);
}
// field ids: 0 u, 1 v, 2 w, 3 theta, 4 ox, 5 oy, 6 oz -> device slots
__constant__ int kSpecSlot[7] = {0, 1, 2, 6, 3, 4, 5};
// pairs in write_energy_spectra.f90's dataset order (86-239); its kz list
// pairs "Eozoz_kz" with (ox, oz) (218-219)
__constant__ int kSpecX[18][2] = {{0, 0}, {1, 1}, {2, 2}, {3, 3}, {0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3},
                                  {3, 2}, {4, 4}, {5, 5}, {6, 6}, {4, 5}, {4, 6}, {5, 6}, {1, 6}, {2, 5}};
__constant__ int kSpecZ[18][2] = {{0, 0}, {1, 1}, {2, 2}, {3, 3}, {0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3},
                                  {2, 3}, {4, 4}, {5, 5}, {4, 6}, {4, 5}, {4, 6}, {5, 6}, {1, 6}, {2, 5}};
__device__ __forceinline__ double spec_term(cplx a, cplx b) { return a.re * b.re + a.im * b.im; }
// E_x(q, kx, y): one thread per output, serial sum over the kz index in
// storage order (get_E_uu_x, write_energy_spectra.f90:326-359)
__global__ void spectra_x_kernel(const cplx *U, size_t field, size_t plane, int ny, int nxh, int nz, double *ex) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)18 * ny * nxh) return;
  const int q = (int)(t / ((size_t)ny * nxh)), rem = (int)(t % ((size_t)ny * nxh));
  const int y = rem / nxh, kx = rem % nxh;
  const cplx *a = U + (size_t)kSpecSlot[kSpecX[q][0]] * field + (size_t)y * plane + kx;
  const cplx *b = U + (size_t)kSpecSlot[kSpecX[q][1]] * field + (size_t)y * plane + kx;
  double st = 0.0;
  for (int j = 0; j < nz; ++j) {
    const double tm = spec_term(a[(size_t)j * nxh], b[(size_t)j * nxh]);
    st = (kx == 0) ? st + tm : st + 2.0 * tm;
  }
  ex[((size_t)q * nxh + kx) * ny + y] = st;
}
// E_z(q, kz, y): serial sum over kx ascending (get_E_uu_z, 366-399)
__global__ void spectra_z_kernel(const cplx *U, size_t field, size_t plane, int ny, int nxh, int nz, double *ez) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)18 * ny * nz) return;
  const int q = (int)(t / ((size_t)ny * nz)), rem = (int)(t % ((size_t)ny * nz));
  const int y = rem / nz, j = rem % nz;
  const cplx *a = U + (size_t)kSpecSlot[kSpecZ[q][0]] * field + (size_t)y * plane + (size_t)j * nxh;
  const cplx *b = U + (size_t)kSpecSlot[kSpecZ[q][1]] * field + (size_t)y * plane + (size_t)j * nxh;
  double st = 0.0;
  for (int k = 0; k < nxh; ++k) {
    const double tm = spec_term(a[k], b[k]);
    st = (k == 0) ? st + tm : st + 2.0 * tm;
  }
  ez[((size_t)q * nz + j) * ny + y] = st;
}

// wall values of the pressure Poisson equation from v (slot 1), per system
__global__ void __launch_bounds__(64) ppe_bc_kernel(LineParams P, const cplx *U, cplx *tmp, cplx *bc, size_t plane,
                                                    size_t nsys, double nu) {
  const size_t sys = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sys >= nsys) return;
  const SV v{const_cast<cplx *>(U) + (size_t)1 * plane * P.g.n + sys, plane};
  const SV t{tmp + sys, plane};
  cplx b1, b2;
  ppe_wall_values(P, v, t, nu, &b1, &b2);
  bc[sys] = b1;
  bc[plane + sys] = b2;
}
// pressure of every system into product slot 5 (v*v in)
__global__ void __launch_bounds__(64) ppe_kernel(LineParams P, WsArgs A, const cplx *bc, const double *vv) {
  const size_t sys = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sys >= A.plane) return;
  const int k = (int)(sys % A.nxh), j = (int)(sys / A.nxh);
  const int kzi = (j <= A.nh) ? j : j + 1 - 2 * A.nh - 1;  // module_ppdfftw.f90:360-374
  auto col = [&](int slot) { return SV{A.U + (size_t)slot * A.field + sys, A.plane}; };
  if (k == 0 && kzi == 0) {  // get_pressure.f90:45-48
    const SV p = col(5);
    for (int y = 0; y < P.g.n; ++y) p[y] = mk(-vv[y], 0.0);
    return;
  }
  const OwnBand Db{A.dbuf + sys, A.plane, P.g.W()};
  solve_ppe(P, Db, col(0), col(1), col(2), col(3), col(4), col(5), bc[sys], bc[A.plane + sys],
            SV{A.buf + sys, A.plane}, SV{A.buf + A.field + sys, A.plane}, k * A.lmfxr, kzi * A.lmfzr);
}

// ------------------------------------------------------- layout conversions
// reference (slot, kz, kx, y) [y fastest]  <->  device (slot, y, kz, kx)
// one block per (slot, kz); 32x32 tiles through LDS.
__global__ void ref_to_dev_kernel(const double2 *__restrict__ in, double2 *__restrict__ out, int ny, int nxh, int nz,
                                  const int *slot_map) {
  __shared__ double2 tile[32][33];
  const int j = blockIdx.x, sl = blockIdx.y;
  const int dslot = slot_map[sl];
  const double2 *src = in + ((size_t)sl * nz + j) * nxh * ny;  // [kx][y]
  double2 *dst = out + (size_t)dslot * ny * nz * nxh + (size_t)j * nxh;  // [y][(plane)] + kx
  const size_t plane = (size_t)nz * nxh;
  for (int k0 = 0; k0 < nxh; k0 += 32)
    for (int y0 = 0; y0 < ny; y0 += 32) {
      for (int t = threadIdx.x; t < 32 * 32; t += blockDim.x) {
        const int kk = t / 32, yy = t % 32;
        if (k0 + kk < nxh && y0 + yy < ny) tile[kk][yy] = src[(size_t)(k0 + kk) * ny + y0 + yy];
      }
      __syncthreads();
      for (int t = threadIdx.x; t < 32 * 32; t += blockDim.x) {
        const int yy = t / 32, kk = t % 32;
        if (k0 + kk < nxh && y0 + yy < ny) dst[(size_t)(y0 + yy) * plane + k0 + kk] = tile[kk][yy];
      }
      __syncthreads();
    }
}

__global__ void dev_to_ref_kernel(const double2 *__restrict__ in, double2 *__restrict__ out, int ny, int nxh, int nz,
                                  const int *slot_map) {
  __shared__ double2 tile[32][33];
  const int j = blockIdx.x, sl = blockIdx.y;
  const int dslot = slot_map[sl];
  double2 *dst = out + ((size_t)sl * nz + j) * nxh * ny;
  const double2 *src = in + (size_t)dslot * ny * nz * nxh + (size_t)j * nxh;
  const size_t plane = (size_t)nz * nxh;
  for (int k0 = 0; k0 < nxh; k0 += 32)
    for (int y0 = 0; y0 < ny; y0 += 32) {
      for (int t = threadIdx.x; t < 32 * 32; t += blockDim.x) {
        const int yy = t / 32, kk = t % 32;
        if (k0 + kk < nxh && y0 + yy < ny) tile[kk][yy] = src[(size_t)(y0 + yy) * plane + k0 + kk];
      }
      __syncthreads();
      for (int t = threadIdx.x; t < 32 * 32; t += blockDim.x) {
        const int kk = t / 32, yy = t % 32;
        if (k0 + kk < nxh && y0 + yy < ny) dst[(size_t)(k0 + kk) * ny + y0 + yy] = tile[kk][yy];
      }
      __syncthreads();
    }
}

// stores from the state (set_initial_data_uvw, initial_step.f90:139-179)
__global__ void __launch_bounds__(64) init_stores_kernel(LineParams P, cplx *U, cplx *RK, cplx *buf, size_t plane,
                                                         size_t nsys, int nxh, int nh, double lmfxr, double lmfzr) {
  const size_t sys = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sys >= nsys) return;
  const int n = P.g.n;
  const size_t field = plane * n;
  const int k = (int)(sys % nxh), j = (int)(sys / nxh);
  const int kzi = (j <= nh) ? j : j + 1 - 2 * nh - 1;
  SV u{U + sys, plane}, v{U + field + sys, plane}, w{U + 2 * field + sys, plane}, th{U + 6 * field + sys, plane};
  SV r1{RK + sys, plane}, r2{RK + field + sys, plane}, r3{RK + 2 * field + sys, plane}, d{buf + sys, plane};
  if (k == 0 && kzi == 0) {
    for (int i = 0; i < n; ++i) { r1[i] = u[i]; r2[i] = w[i]; r3[i] = th[i]; }
    return;
  }
  for (int i = 0; i < n; ++i) d[i] = v[i];
  band_bs(P.g, P.D0LU, d);
  band_mv(P.g, P.D2, d, r2);
  const double kxp = k * lmfxr, kzp = kzi * lmfzr;
  const double k2 = kxp * kxp + kzp * kzp;
  const cplx ikz = mk(0.0, kzp), ikx = mk(0.0, kxp);
  for (int i = 0; i < n; ++i) {
    r2[i] = csub(r2[i], cscale(v[i], k2));
    r1[i] = csub(cmul(ikz, u[i]), cmul(ikx, w[i]));
  }
  for (int i = 0; i < n; ++i) r3[i] = th[i];
}

// Restart from v, g, theta and the mean lines (set_initial_data_vg,
// initial_step.f90:301-353): v in slot 1, g, theta (and the means on the
// (0,0) line) in RK 0 and 2, W mean in RK 1 at (0,0); u, w, theta -> slots
// 0, 2, 6, phi -> RK 1.
__global__ void __launch_bounds__(64) init_vg_kernel(LineParams P, cplx *U, cplx *RK, cplx *buf, size_t plane,
                                                     size_t nsys, int nxh, int nh, double lmfxr, double lmfzr) {
  const size_t sys = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sys >= nsys) return;
  const int n = P.g.n;
  const size_t field = plane * n;
  const int k = (int)(sys % nxh), j = (int)(sys / nxh);
  const int kzi = (j <= nh) ? j : j + 1 - 2 * nh - 1;
  SV u{U + sys, plane}, v{U + field + sys, plane}, w{U + 2 * field + sys, plane}, th{U + 6 * field + sys, plane};
  SV r1{RK + sys, plane}, r2{RK + field + sys, plane}, r3{RK + 2 * field + sys, plane};
  SV d{buf + sys, plane}, f{buf + field + sys, plane};
  if (k == 0 && kzi == 0) {
    for (int i = 0; i < n; ++i) { u[i] = r1[i]; v[i] = mk(0.0, 0.0); w[i] = r2[i]; th[i] = r3[i]; }
    return;
  }
  for (int i = 0; i < n; ++i) d[i] = v[i];
  band_bs(P.g, P.D0LU, d);
  band_mv(P.g, P.D1, d, f);
  band_mv(P.g, P.D2, d, r2);
  const double kxp = k * lmfxr, kzp = kzi * lmfzr;
  const double k2 = kxp * kxp + kzp * kzp;
  for (int i = 0; i < n; ++i) r2[i] = csub(r2[i], cscale(v[i], k2));
  for (int i = 0; i < n; ++i) { u[i] = r1[i]; w[i] = f[i]; }
  solve_u_w(n, u, w, kxp, kzp);
  for (int i = 0; i < n; ++i) th[i] = r3[i];
}

// Restart from g, phi, theta and the mean lines (set_initial_data_gphi,
// initial_step.f90:430-513): v from [D2 - k^2 D0] c = phi (the operator's own
// diagonal on the wall rows, zero wall values), v = D0 c, f = D1 c, u and w
// from g and f.  Db: per-system band scratch (W * ny doubles, stride plane).
__global__ void __launch_bounds__(64) init_gphi_kernel(LineParams P, cplx *U, const cplx *RK, cplx *buf, double *dbuf,
                                                       size_t plane, size_t nsys, int nxh, int nh, double lmfxr,
                                                       double lmfzr) {
  const size_t sys = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sys >= nsys) return;
  const int n = P.g.n, W = P.g.W(), m = P.g.bw;
  const size_t field = plane * n;
  const int k = (int)(sys % nxh), j = (int)(sys / nxh);
  const int kzi = (j <= nh) ? j : j + 1 - 2 * nh - 1;
  SV u{U + sys, plane}, v{U + field + sys, plane}, w{U + 2 * field + sys, plane}, th{U + 6 * field + sys, plane};
  const SV r1{const_cast<cplx *>(RK) + sys, plane}, r2{const_cast<cplx *>(RK) + field + sys, plane},
      r3{const_cast<cplx *>(RK) + 2 * field + sys, plane};
  SV cc{buf + sys, plane}, f{buf + field + sys, plane};
  if (k == 0 && kzi == 0) {
    for (int i = 0; i < n; ++i) { u[i] = r1[i]; v[i] = mk(0.0, 0.0); w[i] = r2[i]; th[i] = r3[i]; }
    return;
  }
  const double kxp = k * lmfxr, kzp = kzi * lmfzr;
  const double k2 = kzp * kzp + kxp * kxp;
  const OwnBand Db{dbuf + sys, plane, W};
  for (int q = 0; q < W; ++q) Db.at(0, q) = 0.0;
  Db.at(0, m) = -k2 * P.D0(0, m) + P.D2(0, m);
  for (int r = 1; r < n - 1; ++r)
    for (int q = 0; q < W; ++q) Db.at(r, q) = -k2 * P.D0(r, q) + P.D2(r, q);
  for (int q = 0; q < W; ++q) Db.at(n - 1, q) = 0.0;
  Db.at(n - 1, m) = -k2 * P.D0(n - 1, m) + P.D2(n - 1, m);
  band_lu(P.g, Db);
  for (int i = 0; i < n; ++i) cc[i] = r2[i];
  cc[0] = mk(0.0, 0.0);
  cc[n - 1] = mk(0.0, 0.0);
  band_bs(P.g, Db, cc);
  band_mv(P.g, P.D0, cc, v);
  band_mv(P.g, P.D1, cc, f);
  for (int i = 0; i < n; ++i) { u[i] = r1[i]; w[i] = f[i]; }
  solve_u_w(n, u, w, kxp, kzp);
  for (int i = 0; i < n; ++i) th[i] = r3[i];
}

// ------------------------------------------- LABS band operations, batched
// Packed bands (2bw+1, n) one after the other, Fortran order (slot fastest);
// one thread per matrix / right-hand side.  Same functions as the solver.
__global__ void __launch_bounds__(64) labs_lu_kernel(int n, int bw, int nmat, double *A) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmat) return;
  const BandGeom g{n, bw};
  band_lu(g, OwnBand{A + (size_t)m * n * g.W(), 1, g.W()});
}
// b: nrhs complex columns of n values, one after the other
__global__ void __launch_bounds__(64) labs_bs_kernel(int n, int bw, const double *LU, int nrhs, cplx *b) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nrhs) return;
  const BandGeom g{n, bw};
  band_bs(g, SharedBand{LU, g.W()}, SV{b + (size_t)m * n, 1});
}
__global__ void __launch_bounds__(64) labs_mv_kernel(int n, int bw, const double *A, int nrhs, const cplx *x,
                                                     cplx *b) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nrhs) return;
  const BandGeom g{n, bw};
  band_mv(g, SharedBand{A, g.W()}, SV{const_cast<cplx *>(x) + (size_t)m * n, 1}, SV{b + (size_t)m * n, 1});
}

}  // namespace pb
