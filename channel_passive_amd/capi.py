"""ctypes binding of the HIP library (include/pb_capi.h).

This is the host-side mirror of the reference's time-step interface
(SOURCE_PASSOPB/time_march.f90:36, wavespace.f90:22, dft_x_realspace.f90:24):
``Channel.substep`` is ``time_march`` for one RK substep, ``Channel.step``
runs the three substeps of main.f90:134-156.  Arrays cross the boundary in
the reference's y-pencil layout (numpy complex128, shape (slot, nz, nx/2, ny),
i.e. Fortran ``U(ny, nx/2, nz, slot)``).

There is no CPU fallback: when ``libpoongback_hip.so`` is missing or no GPU
is present every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .config import ChannelConfig, PbConfig

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpoongback_hip.so")

_dp = ctypes.POINTER(ctypes.c_double)


class PbSubstepOut(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("dt", "velmax", "phi_u", "phi_t", "tau_w", "q_w")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class PbError(RuntimeError):
    pass


# name -> (restype, argtypes); every symbol of include/pb_capi.h
SIGNATURES = {
    "pb_create": (ctypes.c_int, [ctypes.POINTER(PbConfig), ctypes.POINTER(ctypes.c_void_p)]),
    "pb_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "pb_last_error": (ctypes.c_char_p, []),
    "pb_get_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)]),
    "pb_get_labs": (ctypes.c_int, [ctypes.c_void_p] + [_dp] * 6),
    "pb_set_state": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "pb_get_state": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "pb_set_state_slots": (ctypes.c_int, [ctypes.c_void_p] + [_dp] * 5),
    "pb_get_state_slots": (ctypes.c_int, [ctypes.c_void_p] + [_dp] * 5),
    "pb_init_stores": (ctypes.c_int, [ctypes.c_void_p]),
    "pb_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "pb_timing_read": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_int64)]),
    "pb_substep": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _dp, ctypes.POINTER(PbSubstepOut)]),
    "pb_substep_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _dp, ctypes.POINTER(PbSubstepOut), _dp,
                                        _dp]),
    "pb_statistics": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "pb_spectra": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "pb_step": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.POINTER(PbSubstepOut)]),
    "pb_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "pb_wavespace_host": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, _dp, _dp, _dp,
                                         ctypes.POINTER(PbSubstepOut)]),
    "pb_nonlinear_host": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, ctypes.c_int32, _dp]),
    "pb_y_to_x_host": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "pb_realspace_host": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, ctypes.c_int32, _dp]),
    "pb_x_to_y_host": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _dp, _dp, ctypes.c_double]),
    "pb_pressure_host": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp]),
    "pb_labs_lu": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp]),
    "pb_labs_bs": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _dp, ctypes.c_int32, _dp]),
    "pb_labs_mat_mul": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _dp, ctypes.c_int32, _dp, _dp]),
    "pb_labs_bs_noalloc": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _dp, ctypes.c_int32, _dp]),
    "pb_step_graph": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.POINTER(PbSubstepOut)]),
    "pb_set_state_vg": (ctypes.c_int, [ctypes.c_void_p] + [_dp] * 6),
    "pb_substep_stats2": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _dp, ctypes.POINTER(PbSubstepOut), _dp,
                                         _dp, _dp]),
    "pb_set_state_gphi": (ctypes.c_int, [ctypes.c_void_p] + [_dp] * 6),
    "pb_get_checkpoint": (ctypes.c_int, [ctypes.c_void_p] + [_dp] * 9),
    "pb_ensemble_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(PbConfig), ctypes.POINTER(ctypes.c_void_p)]),
    "pb_ensemble_channel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "pb_ensemble_step": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.POINTER(PbSubstepOut)]),
    "pb_ensemble_tune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _dp]),
    "pb_ensemble_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "pb_set_control": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, _dp]),
    "pb_adj_wavespace_host": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, _dp, _dp, _dp, _dp,
                                             _dp, _dp, ctypes.c_double, ctypes.c_double]),
    "pb_adj_nonlinear_host": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "pb_adj_substep": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, _dp, _dp, _dp,
                                      ctypes.c_double, ctypes.c_double, _dp, _dp]),
    "pb_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                   _dp, _dp, _dp]),
}

_lib = None


def load():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PbError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise PbError(f"{load().pb_last_error().decode()} (status {rc})")


def _ptr(a: np.ndarray):
    if a is None:
        return None
    if a.dtype != np.complex128 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError("arrays must be C-contiguous complex128")
    return a.ctypes.data_as(_dp)


class Channel:
    """One channel on the current GPU, state resident in HBM."""

    def __init__(self, cfg: ChannelConfig, handle: ctypes.c_void_p | None = None):
        """handle: a context owned by an Ensemble (not destroyed by close)."""
        self.cfg = cfg
        self._c = cfg.to_c()
        self._owned = handle is None
        if handle is None:
            handle = ctypes.c_void_p()
            _check(load().pb_create(ctypes.byref(self._c), ctypes.byref(handle)))
        self.h = handle
        bw, nbytes = ctypes.c_int32(0), ctypes.c_int64(0)
        _check(load().pb_get_info(self.h, ctypes.byref(bw), ctypes.byref(nbytes)))
        self.bw, self.device_bytes = bw.value, nbytes.value

    def close(self):
        if getattr(self, "h", None):
            if self._owned:
                load().pb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, nslot):
        c = self.cfg
        return (nslot, c.nz, c.nxh, c.ny)

    def labs(self):
        n, W = self.cfg.ny, 2 * self.bw + 1
        out = {k: np.zeros((n, W)) for k in ("d0", "d1", "d2", "d0lu")}
        y, integ = np.zeros(n), np.zeros(n)
        _check(load().pb_get_labs(self.h, y.ctypes.data_as(_dp), *[out[k].ctypes.data_as(_dp) for k in
                                  ("d0", "d1", "d2", "d0lu")], integ.ctypes.data_as(_dp)))
        out["y"], out["integ"] = y, integ
        return out

    def set_state(self, S: np.ndarray, RK: np.ndarray | None = None):
        assert S.shape == self.shape(4)
        if RK is not None:
            assert RK.shape == self.shape(3)
        _check(load().pb_set_state(self.h, _ptr(S), _ptr(RK)))

    def get_state(self):
        S = np.zeros(self.shape(4), np.complex128)
        RK = np.zeros(self.shape(3), np.complex128)
        _check(load().pb_get_state(self.h, _ptr(S), _ptr(RK)))
        return S, RK

    def init_stores(self):
        _check(load().pb_init_stores(self.h))

    def _field(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.shape == self.shape(1)[1:]
        return a

    def set_state_vg(self, v, g, theta, u_mean, w_mean, t_mean):
        """Restart from v, g, theta (nz, nx/2, ny) and the mean lines (ny)."""
        f = [self._field(x) for x in (v, g, theta)]
        m = [np.ascontiguousarray(x, dtype=np.float64) for x in (u_mean, w_mean, t_mean)]
        _check(load().pb_set_state_vg(self.h, *[_ptr(x) for x in f], *[x.ctypes.data_as(_dp) for x in m]))

    def set_state_gphi(self, g, phi, theta, u_mean, w_mean, t_mean):
        """Restart from g, phi, theta (nz, nx/2, ny) and the mean lines (ny)."""
        f = [self._field(x) for x in (g, phi, theta)]
        m = [np.ascontiguousarray(x, dtype=np.float64) for x in (u_mean, w_mean, t_mean)]
        _check(load().pb_set_state_gphi(self.h, *[_ptr(x) for x in f], *[x.ctypes.data_as(_dp) for x in m]))

    def checkpoint(self):
        """data_store's datasets: dict of u, v, w, theta, g, phi (nz, nx/2, ny)
        and U_mean, W_mean, T_mean (ny)."""
        shp = self.shape(1)[1:]
        f = {k: np.zeros(shp, np.complex128) for k in ("u", "v", "w", "theta", "g", "phi")}
        m = {k: np.zeros(self.cfg.ny) for k in ("U_mean", "W_mean", "T_mean")}
        _check(load().pb_get_checkpoint(self.h, *[_ptr(f[k]) for k in ("u", "v", "w", "theta", "g", "phi")],
                                        *[m[k].ctypes.data_as(_dp) for k in ("U_mean", "W_mean", "T_mean")]))
        f.update(m)
        return f

    def substep(self, s: int, dt: float):
        d = ctypes.c_double(dt)
        o = PbSubstepOut()
        _check(load().pb_substep(self.h, s, ctypes.byref(d), ctypes.byref(o)))
        return o.as_dict()

    def substep_stats(self, s: int, dt: float):
        """Substep s with the statistics-time pressure (on s == 1): returns
        (out, p (nz, nx/2, ny) complex128, <vv> (ny,))."""
        d = ctypes.c_double(dt)
        o = PbSubstepOut()
        p = np.zeros(self.shape(1)[1:], np.complex128)
        vv = np.zeros(self.cfg.ny)
        _check(load().pb_substep_stats(self.h, s, ctypes.byref(d), ctypes.byref(o), _ptr(p),
                                       vv.ctypes.data_as(_dp)))
        return o.as_dict(), p, vv

    STAT2_NAMES = ("uuu", "vvv", "www", "uuuu", "vvvv", "wwww", "u6", "v6", "w6", "u8", "v8", "w8", "u10", "v10",
                   "w10", "-duuvdy", "-dvvvdy", "-dvwwdy", "-duvvdy", "up", "-dpudy", "vp", "-2dpvdy", "wp", "pp",
                   "2pdudx", "2pdvdy", "2pdwdz", "pdudy+pdvdx")

    def substep_stats2(self, dt: float):
        """Substep 1 with get_statistics_1's pressure and get_statistics_2:
        returns (out, p, <vv>, dict name -> ny line)."""
        d = ctypes.c_double(dt)
        o = PbSubstepOut()
        p = np.zeros(self.shape(1)[1:], np.complex128)
        vv = np.zeros(self.cfg.ny)
        st = np.zeros((29, self.cfg.ny))
        _check(load().pb_substep_stats2(self.h, 1, ctypes.byref(d), ctypes.byref(o), _ptr(p),
                                        vv.ctypes.data_as(_dp), st.ctypes.data_as(_dp)))
        return o.as_dict(), p, vv, {k: st[i] for i, k in enumerate(self.STAT2_NAMES)}

    STAT_NAMES = ("U", "W", "Theta", "dUdy", "dWdy", "dTdy", "uu", "vv", "ww", "tt", "uv", "uw", "vw", "ut",
                  "wt", "vt")

    def statistics(self):
        """get_statistics_1's profiles of the current state: a dict of
        ny-vectors plus u_center and t_center."""
        n = self.cfg.ny
        out = np.zeros(16 * n + 2)
        _check(load().pb_statistics(self.h, out.ctypes.data_as(_dp)))
        d = {k: out[i * n:(i + 1) * n] for i, k in enumerate(self.STAT_NAMES)}
        d["u_center"], d["t_center"] = out[16 * n], out[16 * n + 1]
        return d

    # write_energy_spectra.f90's datasets in its order (pb_spectra)
    SPECTRA_NAMES = ("uu", "vv", "ww", "tt", "uv", "vw", "uw", "ut", "vt", "wt", "oxox", "oyoy", "ozoz",
                     "oxoy", "oxoz", "oyoz", "voz", "woy")

    def spectra(self):
        """1-D energy (co)spectra of the current state: (ex, ez) with
        ex[q, kx, y] summed over kz and ez[q, kz, y] summed over kx."""
        c = self.cfg
        ex = np.zeros((18, c.nxh, c.ny))
        ez = np.zeros((18, c.nz, c.ny))
        _check(load().pb_spectra(self.h, ex.ctypes.data_as(_dp), ez.ctypes.data_as(_dp)))
        return ex, ez

    def step(self, dt: float):
        d = ctypes.c_double(dt)
        o = (PbSubstepOut * self.cfg.rk_step)()
        _check(load().pb_step(self.h, ctypes.byref(d), o))
        return [x.as_dict() for x in o]

    def step_graph(self, dt: float):
        """pb_step_graph: the step captured in a hipGraph on the first call,
        replayed after it."""
        d = ctypes.c_double(dt)
        o = (PbSubstepOut * self.cfg.rk_step)()
        _check(load().pb_step_graph(self.h, ctypes.byref(d), o))
        return [x.as_dict() for x in o]

    def sync(self):
        _check(load().pb_sync(self.h))

    def nonlinear_host(self, S: np.ndarray, want_vv: bool = False):
        P = np.zeros(self.shape(10), np.complex128)
        vm = ctypes.c_double(0)
        _check(load().pb_nonlinear_host(self.h, _ptr(np.ascontiguousarray(S)), _ptr(P), int(want_vv),
                                        ctypes.byref(vm)))
        return P, vm.value

    def pshape(self, nslot):
        """x-pencil padded layout Upad(nx/2, 3nz/2, ny, slot) as numpy
        (slot, ny, 3nz/2, nx/2)."""
        c = self.cfg
        return (nslot, c.ny, 3 * c.nz // 2, c.nxh)

    def y_to_x_host(self, S: np.ndarray) -> np.ndarray:
        """y_to_x of u, v, w, theta: (4, nz, nx/2, ny) -> (4, ny, 3nz/2, nx/2)."""
        assert S.shape == self.shape(4)
        out = np.zeros(self.pshape(4), np.complex128)
        _check(load().pb_y_to_x_host(self.h, _ptr(np.ascontiguousarray(S)), _ptr(out)))
        return out

    def realspace_host(self, Upad: np.ndarray, want_vv: bool = False):
        """dft_x_realspace: 4 padded fields -> 10 padded slots (reference
        order) and velmax."""
        assert Upad.shape == self.pshape(4)
        out = np.zeros(self.pshape(10), np.complex128)
        vm = ctypes.c_double(0)
        _check(load().pb_realspace_host(self.h, _ptr(np.ascontiguousarray(Upad)), _ptr(out), int(want_vv),
                                        ctypes.byref(vm)))
        return out, vm.value

    def x_to_y_host(self, Upad: np.ndarray, scale: float = 1.0) -> np.ndarray:
        """x_to_y (times scale) of nfield padded fields -> (nfield, nz, nx/2, ny)."""
        nf = Upad.shape[0]
        assert Upad.shape == self.pshape(nf)
        out = np.zeros(self.shape(nf), np.complex128)
        _check(load().pb_x_to_y_host(self.h, nf, _ptr(np.ascontiguousarray(Upad)), _ptr(out), scale))
        return out

    def pressure_host(self, S: np.ndarray, P: np.ndarray):
        """get_pressure from the state S (4 slots) and its products P (10
        slots, v*v in slot 5): (p (nz, nx/2, ny), <vv> (ny,))."""
        p = np.zeros(self.shape(1)[1:], np.complex128)
        vv = np.zeros(self.cfg.ny)
        _check(load().pb_pressure_host(self.h, _ptr(np.ascontiguousarray(S)), _ptr(np.ascontiguousarray(P)),
                                       _ptr(p), vv.ctypes.data_as(_dp)))
        return p, vv

    def wavespace_host(self, s: int, dt: float, P: np.ndarray, RK: np.ndarray, v_old: np.ndarray | None = None):
        o = PbSubstepOut()
        _check(load().pb_wavespace_host(self.h, s, dt, _ptr(P), _ptr(RK),
                                        None if v_old is None else _ptr(np.ascontiguousarray(v_old)),
                                        ctypes.byref(o)))
        return o.as_dict()

    # ---- opposition gain, wall actuation, discrete adjoint (pb_capi.h)
    def set_control(self, alpha: float, act: np.ndarray | None = None):
        """Opposition gain and wall actuation act (nz, nx/2) complex (or None)."""
        a = None if act is None else np.ascontiguousarray(act, dtype=np.complex128)
        if a is not None:
            assert a.shape == (self.cfg.nz, self.cfg.nxh)
        _check(load().pb_set_control(self.h, float(alpha), _ptr(a)))
        self.cfg.opp_alpha = float(alpha)

    def adj_wavespace_host(self, s: int, dt: float, AU: np.ndarray, ARK: np.ndarray, vd: np.ndarray | None = None,
                           tau_bar: float = 0.0, q_bar: float = 0.0):
        """Transpose of the wall-normal stage: AU (9, nz, nx/2, ny) and ARK
        (3, ...) in place; returns (avd (2, nz, nx/2), act_bar (nz, nx/2),
        alpha terms (nz, nx/2))."""
        c = self.cfg
        avd = np.zeros((2, c.nz, c.nxh), np.complex128)
        act = np.zeros((c.nz, c.nxh), np.complex128)
        al = np.zeros((c.nz, c.nxh))
        _check(load().pb_adj_wavespace_host(self.h, s, dt, _ptr(AU), _ptr(ARK),
                                            None if vd is None else _ptr(np.ascontiguousarray(vd)), _ptr(avd),
                                            _ptr(act), al.ctypes.data_as(_dp), tau_bar, q_bar))
        return avd, act, al

    def adj_nonlinear_host(self, S: np.ndarray, aprod: np.ndarray) -> np.ndarray:
        """Transpose of the nonlinear stage at state S: aprod (9 or 10, ...)
        -> adjoint of S (4, nz, nx/2, ny)."""
        out = np.zeros(self.shape(4), np.complex128)
        _check(load().pb_adj_nonlinear_host(self.h, _ptr(np.ascontiguousarray(S)),
                                            _ptr(np.ascontiguousarray(aprod[:9])), _ptr(out)))
        return out

    def adj_substep(self, s: int, dt: float, S: np.ndarray, aS: np.ndarray, aRK: np.ndarray, tau_bar=0.0,
                    q_bar=0.0):
        """Transpose of substep s around input state S: aS, aRK in place (out
        -> in); returns (d/d alpha, act_bar (nz, nx/2))."""
        al = ctypes.c_double(0.0)
        act = np.zeros((self.cfg.nz, self.cfg.nxh), np.complex128)
        _check(load().pb_adj_substep(self.h, s, dt, _ptr(np.ascontiguousarray(S)), _ptr(aS), _ptr(aRK), tau_bar,
                                     q_bar, ctypes.byref(al), _ptr(act)))
        return al.value, act

    def gradient(self, nsteps: int, dt: float, ckpt_every: int = 20, wq: float = 0.5):
        """J = mean over nsteps of tau_w + wq q_w and its gradient: (J,
        dJ/dalpha, dJ/da (nz, nx/2) complex)."""
        J, da = ctypes.c_double(0.0), ctypes.c_double(0.0)
        g = np.zeros((self.cfg.nz, self.cfg.nxh), np.complex128)
        _check(load().pb_gradient(self.h, nsteps, ckpt_every, dt, wq, ctypes.byref(J), ctypes.byref(da), _ptr(g)))
        return J.value, da.value, g


# ---- LABS band operations (module_labs.f90:462, 499, 570) on the device
def labs_lu(A: np.ndarray, bw: int) -> np.ndarray:
    """LABS_lu of packed bands A (..., n, 2bw+1) (one or a stack)."""
    A = np.ascontiguousarray(A, dtype=np.float64).copy()
    n = A.shape[-2]
    nmat = int(np.prod(A.shape[:-2])) if A.ndim > 2 else 1
    _check(load().pb_labs_lu(n, bw, nmat, A.ctypes.data_as(_dp)))
    return A


def labs_bs(LU: np.ndarray, bw: int, b: np.ndarray) -> np.ndarray:
    """LABS_bs: LU^-1 b for columns b (..., n) complex."""
    x = np.ascontiguousarray(b, dtype=np.complex128).copy()
    n = LU.shape[0]
    _check(load().pb_labs_bs(n, bw, np.ascontiguousarray(LU).ctypes.data_as(_dp), x.size // n, _ptr(x)))
    return x


def labs_bs_noalloc(LU: np.ndarray, bw: int, b: np.ndarray) -> np.ndarray:
    """labs_bs through the library's device scratch (no allocation)."""
    x = np.ascontiguousarray(b, dtype=np.complex128).copy()
    n = LU.shape[0]
    _check(load().pb_labs_bs_noalloc(n, bw, np.ascontiguousarray(LU).ctypes.data_as(_dp), x.size // n, _ptr(x)))
    return x


def labs_mat_mul(A: np.ndarray, bw: int, x: np.ndarray) -> np.ndarray:
    """LABS_mat_mul: A x for columns x (..., n) complex."""
    xx = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(xx)
    n = A.shape[0]
    _check(load().pb_labs_mat_mul(n, bw, np.ascontiguousarray(A).ctypes.data_as(_dp), xx.size // n, _ptr(xx),
                                  _ptr(out)))
    return out


class Ensemble:
    """Independent channels of uneven size advanced together: every channel's
    step in one captured graph, one launch per ensemble step
    (pb_ensemble_*).  channel(i) gives a Channel view for the state exchange."""

    def __init__(self, cfgs):
        self.cfgs = list(cfgs)
        arr = (PbConfig * len(self.cfgs))(*[c.to_c() for c in self.cfgs])
        h = ctypes.c_void_p()
        _check(load().pb_ensemble_create(len(self.cfgs), arr, ctypes.byref(h)))
        self.h = h
        self.channels = []
        for i, c in enumerate(self.cfgs):
            ch = ctypes.c_void_p()
            _check(load().pb_ensemble_channel(self.h, i, ctypes.byref(ch)))
            self.channels.append(Channel(c, handle=ch))

    def channel(self, i: int) -> Channel:
        return self.channels[i]

    def tune(self, reps: int = 5):
        """The autotuner (before the states are set): per-channel step times
        and the ensemble's, in ms."""
        ms = np.zeros(len(self.cfgs) + 1)
        _check(load().pb_ensemble_tune(self.h, reps, ms.ctypes.data_as(_dp)))
        return ms

    def step(self, dts):
        dt = np.ascontiguousarray(dts, dtype=np.float64).copy()
        nout = sum(c.rk_step for c in self.cfgs)
        o = (PbSubstepOut * nout)()
        _check(load().pb_ensemble_step(self.h, dt.ctypes.data_as(_dp), o))
        outs, k = [], 0
        for c in self.cfgs:
            outs.append([o[k + q].as_dict() for q in range(c.rk_step)])
            k += c.rk_step
        return outs

    def close(self):
        if getattr(self, "h", None):
            for ch in self.channels:
                ch.h = None
            load().pb_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
