"""Derived observables of a long run: skin friction, Nusselt number and rms
profiles, from the library's outputs (the tau_w and q_w of pb_substep_out,
the profiles of pb_statistics) and its collocation operators (Channel.labs).

The reference computes none of these; the definitions are fixed here:

  Cf  = 2 tau_w / U_b^2, tau_w = nu <|dU/dy|> over both walls (wave00.f90:210-211)
  Nu  = 2 h (q_w / alpha) / (theta_m - theta_w), alpha = nu / Pr, h = 1,
        theta_w = 0, q_w = alpha <|dTheta/dy|> over both walls (wave00.f90:621-622)
  theta_m: "bulk"  = (1 / 2h) int Theta dy, the plain average that bulk_t holds
                     (calculate_phi, wave00.f90:608-619);
           "mixed" = int U Theta dy / int U dy, the velocity-weighted mean
  rms: u_rms = sqrt(<uu> - U0^2) with U0 the stored mean mode (get_uv's sums
       include it; pb_statistics' U adds u_gal), likewise v, w, theta.

Integrals use the reference's quadrature: int f dy = integ . (D0^-1 f)
(LABS_integration, module_labs.f90:401-413), on the collocation values of f.
"""
from __future__ import annotations

import numpy as np


def band_to_dense(band: np.ndarray, bw: int) -> np.ndarray:
    """Dense matrix of the reference's near-boundary packed band
    (LABS_full_to_band, module_labs.f90:301-328): row r holds columns
    first(r)..last(r); the first bw rows keep their far right entries, the
    last bw rows their far left entries, in the otherwise unused slots."""
    n = band.shape[0]
    A = np.zeros((n, n))
    for r in range(n):
        first = 0 if r < bw else (n - 1 - 2 * bw if r >= n - bw else r - bw)
        last = 2 * bw if r < bw else (n - 1 if r >= n - bw else r + bw)
        for c in range(first, last + 1):
            if r < bw and c > r + bw:
                s = 2 * bw - c
            elif r >= n - bw and c < r - bw:
                s = n - 1 - c
            else:
                s = bw + c - r
            A[r, c] = band[r, s]
    return A


class Quadrature:
    """int f dy over the channel for collocation values f."""

    def __init__(self, labs: dict, bw: int):
        self.d0 = band_to_dense(labs["d0"], bw)
        self.integ = np.asarray(labs["integ"], float)

    def __call__(self, f: np.ndarray) -> float:
        return float(self.integ @ np.linalg.solve(self.d0, f))


def skin_friction(tau_w: float, u_bulk: float) -> float:
    return 2.0 * tau_w / (u_bulk * u_bulk)


def nusselt(q_w: float, nu: float, pr: float, theta_m: float, theta_w: float = 0.0, h: float = 1.0) -> float:
    alpha = nu / pr
    return 2.0 * h * (q_w / alpha) / (theta_m - theta_w)


def theta_mean(stats: dict, quad: Quadrature, kind: str = "bulk", u_gal: float = 0.0, ly: float = 2.0) -> float:
    """theta_m of the Nusselt number from pb_statistics' profiles."""
    theta = np.asarray(stats["Theta"], float)
    if kind == "bulk":
        return quad(theta) / ly
    if kind == "mixed":
        u = np.asarray(stats["U"], float) - u_gal
        return quad(u * theta) / quad(u)
    raise ValueError(kind)


def rms_profiles(stats: dict, u_gal: float = 0.0) -> dict:
    """Fluctuation rms from the Parseval sums (which include the mean mode)."""
    means = {"u": np.asarray(stats["U"], float) - u_gal, "v": 0.0, "w": np.asarray(stats["W"], float),
             "t": np.asarray(stats["Theta"], float)}
    out = {}
    for f, key in (("u", "uu"), ("v", "vv"), ("w", "ww"), ("t", "tt")):
        var = np.asarray(stats[key], float) - np.square(means[f])
        out[f + "_rms"] = np.sqrt(np.maximum(var, 0.0))
    return out
