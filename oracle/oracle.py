"""ctypes front end of the parity oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg.  The product never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from channel_passive_amd.config import ChannelConfig, PbConfig

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(PbConfig)]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)] + [_dp] * 7
        L.orc_init_stores.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_nonlinear.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_opp_planes.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_wavespace.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp,
                                    ctypes.c_int, ctypes.c_int, _dp]
        L.orc_substep.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, _dp, _dp]
        L.orc_set_control.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp]
        L.orc_pressure.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_statistics.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_spectra.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_labs_setup.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, _dp, _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int)]
        L.orc_bspline_row.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, _dp, ctypes.POINTER(ctypes.c_int)]
        L.orc_full_to_band.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_band_lu.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
        L.orc_band_solve.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_band_matvec.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_fft.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_hvhg.argtypes = [ctypes.c_void_p, _dp]
        L.orc_init_from_vg.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_nonlinear_mom.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, _dp, _dp]
        L.orc_statistics2.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_init_from_gphi.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_set_mutant.argtypes = [ctypes.c_int]
        L.orc_set_mutant.restype = None
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


# ------------------------------------------------------------ LABS helpers
def labs_setup(ny: int, labs_k: int, knot_map: int, stretch: float, ly: float = 2.0):
    k = labs_k + 1
    nbreak = ny - k + 2
    y = np.zeros(ny); brk = np.zeros(nbreak); integ = np.zeros(ny); center = np.zeros(ny)
    full = np.zeros((5, ny, ny))
    bw = ctypes.c_int(0)
    r = lib().orc_labs_setup(ny, labs_k, knot_map, stretch, ly, _p(y), _p(brk), _p(integ), _p(center),
                             _p(full), ctypes.byref(bw))
    if r:
        raise RuntimeError(f"orc_labs_setup failed: {r}")
    return dict(y=y, brk=brk, integ=integ, center=center, full=full, bw=bw.value)


def bspline_row(ny, labs_k, knot_map, stretch, yv, ly=2.0):
    vals = np.zeros(labs_k + 1)
    first = ctypes.c_int(0)
    r = lib().orc_bspline_row(ny, labs_k, knot_map, stretch, ly, yv, _p(vals), ctypes.byref(first))
    if r:
        raise RuntimeError(f"orc_bspline_row failed: {r}")
    return first.value, vals


def full_to_band(full: np.ndarray, bw: int) -> np.ndarray:
    n = full.shape[0]
    band = np.zeros((n, 2 * bw + 1))
    lib().orc_full_to_band(_p(np.ascontiguousarray(full)), _p(band), bw, n)
    return band


def band_lu(band: np.ndarray, bw: int) -> np.ndarray:
    A = np.ascontiguousarray(band).copy()
    lib().orc_band_lu(_p(A), A.shape[0], bw)
    return A


def band_solve(lu: np.ndarray, bw: int, b: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(b, dtype=np.complex128).copy()
    lib().orc_band_solve(_p(np.ascontiguousarray(lu)), lu.shape[0], bw, x.ctypes.data_as(_dp))
    return x


def band_matvec(A: np.ndarray, bw: int, x: np.ndarray) -> np.ndarray:
    xx = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(xx)
    lib().orc_band_matvec(_p(np.ascontiguousarray(A)), A.shape[0], bw, xx.ctypes.data_as(_dp),
                          out.ctypes.data_as(_dp))
    return out


def fft(x: np.ndarray, sign: int) -> np.ndarray:
    xx = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(xx)
    lib().orc_fft(xx.ctypes.data_as(_dp), out.ctypes.data_as(_dp), xx.shape[0], sign)
    return out


class mutant:
    """Context manager planting mistake number m in the oracle's per-line
    formulas (the list is in oracle.cpp); for tests that show a pin fails on
    a wrong oracle."""

    def __init__(self, m: int):
        self.m = m

    def __enter__(self):
        lib().orc_set_mutant(self.m)
        return self

    def __exit__(self, *exc):
        lib().orc_set_mutant(0)
        return False


# ----------------------------------------------------------------- solver
class Oracle:
    """One channel configuration on the CPU, reference layout throughout."""

    def __init__(self, cfg: ChannelConfig):
        self.cfg = cfg
        self._c = cfg.to_c()
        self.h = lib().orc_create(ctypes.byref(self._c))
        if not self.h:
            raise RuntimeError("orc_create failed")
        bw = ctypes.c_int(0)
        lib().orc_info(self.h, ctypes.byref(bw), *([None] * 7))
        self.bw = bw.value
        n, W = cfg.ny, 2 * self.bw + 1
        self.y = np.zeros(n); self.d0 = np.zeros((n, W)); self.d1 = np.zeros((n, W)); self.d2 = np.zeros((n, W))
        self.d0lu = np.zeros((n, W)); self.integ = np.zeros(n); self.dyr = np.zeros(n)
        lib().orc_info(self.h, ctypes.byref(bw), _p(self.y), _p(self.d0), _p(self.d1), _p(self.d2),
                       _p(self.d0lu), _p(self.integ), _p(self.dyr))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def init_stores(self, S: np.ndarray) -> np.ndarray:
        cfg = self.cfg
        RK = np.zeros((3, cfg.nz, cfg.nxh, cfg.ny), dtype=np.complex128)
        lib().orc_init_stores(self.h, S.ctypes.data_as(_dp), RK.ctypes.data_as(_dp))
        return RK

    def init_from_vg(self, v, g, theta, u_mean, w_mean, t_mean):
        """Restart from v, g, theta and the mean lines (set_initial_data_vg):
        returns (S, RK).  u_mean is the stored U (with u_gal added)."""
        cfg = self.cfg
        S = np.zeros((4, cfg.nz, cfg.nxh, cfg.ny), np.complex128)
        RK = np.zeros((3, cfg.nz, cfg.nxh, cfg.ny), np.complex128)
        S[1], RK[0], RK[2] = v, g, theta
        RK[0, 0, 0] = np.asarray(u_mean) - cfg.u_gal
        RK[1, 0, 0] = w_mean
        RK[2, 0, 0] = t_mean
        lib().orc_init_from_vg(self.h, S.ctypes.data_as(_dp), RK.ctypes.data_as(_dp))
        return S, RK

    def init_from_gphi(self, g, phi, theta, u_mean, w_mean, t_mean):
        """Restart from g, phi, theta and the mean lines
        (set_initial_data_gphi): returns (S, RK)."""
        cfg = self.cfg
        S = np.zeros((4, cfg.nz, cfg.nxh, cfg.ny), np.complex128)
        RK = np.zeros((3, cfg.nz, cfg.nxh, cfg.ny), np.complex128)
        RK[0], RK[1], RK[2] = g, phi, theta
        RK[0, 0, 0] = np.asarray(u_mean) - cfg.u_gal
        RK[1, 0, 0] = w_mean
        RK[2, 0, 0] = t_mean
        lib().orc_init_from_gphi(self.h, S.ctypes.data_as(_dp), RK.ctypes.data_as(_dp))
        return S, RK

    def nonlinear(self, S: np.ndarray, want_vv: bool = False, y0: int = 0, y1: int | None = None):
        cfg = self.cfg
        P = np.zeros((10, cfg.nz, cfg.nxh, cfg.ny), dtype=np.complex128)
        vm = ctypes.c_double(0.0)
        lib().orc_nonlinear(self.h, S.ctypes.data_as(_dp), P.ctypes.data_as(_dp), int(want_vv), y0,
                            cfg.ny if y1 is None else y1, ctypes.byref(vm))
        return P, vm.value

    STAT2_NAMES = ("uuu", "vvv", "www", "uuuu", "vvvv", "wwww", "u6", "v6", "w6", "u8", "v8", "w8", "u10", "v10",
                   "w10", "-duuvdy", "-dvvvdy", "-dvwwdy", "-duvvdy", "up", "-dpudy", "vp", "-2dpvdy", "wp", "pp",
                   "2pdudx", "2pdvdy", "2pdwdz", "pdudy+pdvdx")

    def statistics2(self, S: np.ndarray):
        """get_statistics_2's datasets for the state S (4 slots) at a
        statistics substep: products with the moments, the pressure, then
        the correlations.  Returns (dict name -> ny line, P with p in slot 5,
        the raw moment sums (19, ny))."""
        cfg = self.cfg
        P = np.zeros((10, cfg.nz, cfg.nxh, cfg.ny), dtype=np.complex128)
        mom = np.zeros((19, cfg.ny))
        vm = ctypes.c_double(0.0)
        Sc = np.ascontiguousarray(S)
        lib().orc_nonlinear_mom(self.h, Sc.ctypes.data_as(_dp), P.ctypes.data_as(_dp), 1, ctypes.byref(vm), _p(mom))
        self.pressure(Sc, P)
        out = np.zeros((29, cfg.ny))
        lib().orc_statistics2(self.h, Sc.ctypes.data_as(_dp), P.ctypes.data_as(_dp), _p(mom), _p(out))
        return {k: out[i] for i, k in enumerate(self.STAT2_NAMES)}, P, mom

    def opp_planes(self, S: np.ndarray) -> np.ndarray:
        cfg = self.cfg
        out = np.zeros((2, cfg.nz, cfg.nxh), dtype=np.complex128)
        lib().orc_opp_planes(self.h, S.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
        return out

    def wavespace(self, s: int, dt: float, P: np.ndarray, RK: np.ndarray, vd: np.ndarray | None = None,
                  usrc: np.ndarray | None = None):
        """wavespace on products P (10 slots) and stores RK, in place; vd the
        detection-plane v (opposition control), usrc the u of the substep's
        input state (mixed-type scalar condition).  Returns phi_u, phi_t,
        tau_w, q_w."""
        cfg = self.cfg
        mean = np.zeros(4)
        us = None if usrc is None else np.ascontiguousarray(usrc, dtype=np.complex128)
        r = lib().orc_wavespace(self.h, s, dt, P.ctypes.data_as(_dp), RK.ctypes.data_as(_dp),
                                None if vd is None else vd.ctypes.data_as(_dp),
                                None if us is None else us.ctypes.data_as(_dp), 0, cfg.nxh * cfg.nz, _p(mean))
        if r:
            raise RuntimeError(f"orc_wavespace failed: {r}")
        return mean

    def hvhg(self, P: np.ndarray) -> np.ndarray:
        """hvhg alone on products P (10 slots, reference layout) of every
        line except the mean: returns (h_g, h_v, h_theta), each (nz, nxh, ny)."""
        Q = np.ascontiguousarray(P, dtype=np.complex128).copy()
        lib().orc_hvhg(self.h, Q.ctypes.data_as(_dp))
        return Q[0].copy(), Q[4].copy(), Q[6].copy()

    def pressure(self, S: np.ndarray, P: np.ndarray) -> np.ndarray:
        """get_pressure on the products P (10 slots, v*v in slot 5) of the
        state S: slot 5 of P becomes p; returns <vv>(y)."""
        vv = np.zeros(self.cfg.ny)
        lib().orc_pressure(self.h, S.ctypes.data_as(_dp), P.ctypes.data_as(_dp), _p(vv))
        return vv

    STAT_NAMES = ("U", "W", "Theta", "dUdy", "dWdy", "dTdy", "uu", "vv", "ww", "tt", "uv", "uw", "vw", "ut",
                  "wt", "vt")

    def statistics(self, S: np.ndarray) -> dict:
        """get_statistics_1's profiles of the state S (4 slots): a dict of
        ny-vectors plus u_center and t_center."""
        n = self.cfg.ny
        out = np.zeros(16 * n + 2)
        lib().orc_statistics(self.h, np.ascontiguousarray(S).ctypes.data_as(_dp), _p(out))
        d = {k: out[i * n:(i + 1) * n] for i, k in enumerate(self.STAT_NAMES)}
        d["u_center"], d["t_center"] = out[16 * n], out[16 * n + 1]
        return d

    # write_energy_spectra.f90's datasets, in its order (fields: u v w t, vorticity ox oy oz)
    SPECTRA_NAMES = ("uu", "vv", "ww", "tt", "uv", "vw", "uw", "ut", "vt", "wt", "oxox", "oyoy", "ozoz",
                     "oxoy", "oxoz", "oyoz", "voz", "woy")

    def spectra(self, S: np.ndarray):
        """1-D energy (co)spectra of the state S (4 slots): (ex, ez) with
        ex[q, kx, y] summed over kz and ez[q, kz, y] summed over kx, q in
        SPECTRA_NAMES order ("ozoz" of ez is the reference's (ox, oz))."""
        c = self.cfg
        ex = np.zeros((18, c.nxh, c.ny))
        ez = np.zeros((18, c.nz, c.ny))
        lib().orc_spectra(self.h, np.ascontiguousarray(S).ctypes.data_as(_dp), _p(ex), _p(ez))
        return ex, ez

    def set_control(self, alpha: float, act: np.ndarray | None = None):
        """Opposition gain and wall actuation (nz, nx/2) complex, or None."""
        a = None if act is None else np.ascontiguousarray(act, dtype=np.complex128)
        if a is not None:
            assert a.shape == (self.cfg.nz, self.cfg.nx // 2)
        r = lib().orc_set_control(self.h, ctypes.c_double(alpha), None if a is None else a.ctypes.data_as(_dp))
        if r:
            raise RuntimeError(f"orc_set_control failed: {r}")

    def substep(self, S: np.ndarray, RK: np.ndarray, s: int, dt: float):
        out = np.zeros(6)
        d = ctypes.c_double(dt)
        r = lib().orc_substep(self.h, S.ctypes.data_as(_dp), RK.ctypes.data_as(_dp), s, ctypes.byref(d), _p(out))
        if r:
            raise RuntimeError(f"orc_substep failed: {r}")
        return dict(zip(["dt", "velmax", "phi_u", "phi_t", "tau_w", "q_w"], out))

    def step(self, S, RK, dt):
        """Three substeps; substep 1 may replace dt (CFL rule), and substeps
        2 and 3 run with that dt, as the reference's global dt does
        (dft_x_realspace.f90:492, time_march.f90:36-128)."""
        outs = []
        for s in (1, 2, 3):
            outs.append(self.substep(S, RK, s, dt))
            dt = outs[-1]["dt"]
        return outs


def sample_times(cfg: ChannelConfig, nplanes: int, nlines: int, s: int = 2):
    """Bounded CPU baseline: seconds for the nonlinear stage on `nplanes`
    y planes and for the wavespace on `nlines` (kx,kz) systems of the full
    grid; returns (t_nonlin, t_wave, threads)."""
    L = lib()
    L.orc_sample.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp]
    L.orc_sample.restype = ctypes.c_int
    o = Oracle(cfg)
    a, b = ctypes.c_double(0), ctypes.c_double(0)
    nt = L.orc_sample(o.h, nplanes, nlines, s, cfg.dt, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value, nt
