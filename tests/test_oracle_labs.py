"""Pins of the oracle's packed-band algebra (module_labs.f90:301-328 packing,
3487-3636 LU, 4252-4318 solve, 4395-4430 matvec) against dense linear
algebra on the unpacked matrices."""
import numpy as np
import pytest

from tests.helpers import unpack_band, rel


def operators(oracle_lib, ny, labs_k=7, kmap=1, stretch=2.0):
    r = oracle_lib.labs_setup(ny, labs_k, kmap, stretch, 2.0)
    return r["full"][:3], r["bw"]


def helmholtz(D0, D2, k2, b):
    H = (1 + b * k2) * D0 - b * D2
    H[0, :] = 0; H[0, 0] = (1 + b * k2) * D0[0, 0] - b * D2[0, 0]
    H[-1, :] = 0; H[-1, -1] = (1 + b * k2) * D0[-1, -1] - b * D2[-1, -1]
    return H


def poisson(D0, D2, k2):
    Q = D2 - k2 * D0
    Q[0, :] = 0; Q[0, 0] = D0[0, 0]
    Q[-1, :] = 0; Q[-1, -1] = D0[-1, -1]
    return Q


@pytest.mark.parametrize("ny", [33, 65, 129, 385])
def test_band_roundtrip(oracle_lib, ny):
    (D0, D1, D2), bw = operators(oracle_lib, ny)
    for M in (D0, D1, D2, helmholtz(D0, D2, 3.0, 1e-3), poisson(D0, D2, 7.0)):
        assert np.array_equal(unpack_band(oracle_lib.full_to_band(M, bw), bw), M)


@pytest.mark.parametrize("ny", [33, 65, 385])
@pytest.mark.parametrize("kind", ["D0", "H", "Q"])
def test_lu_factors(oracle_lib, ny, kind):
    """L (with diagonal) times unit-upper U reproduces A; the factors fit the
    packed pattern (no fill outside it)."""
    (D0, D1, D2), bw = operators(oracle_lib, ny)
    A = {"D0": D0, "H": helmholtz(D0, D2, 40.0, 2e-4), "Q": poisson(D0, D2, 40.0)}[kind]
    lu = oracle_lib.band_lu(oracle_lib.full_to_band(A, bw), bw)
    F = unpack_band(lu, bw)
    L = np.tril(F)
    U = np.triu(F, 1) + np.eye(ny)
    assert np.abs(L @ U - A).max() <= 1e-14 * np.abs(A).max()


@pytest.mark.parametrize("ny", [33, 65, 385])
def test_solve_and_matvec(oracle_lib, ny):
    rng = np.random.default_rng(ny)
    (D0, D1, D2), bw = operators(oracle_lib, ny)
    b = rng.standard_normal(ny) + 1j * rng.standard_normal(ny)
    for A in (D0, helmholtz(D0, D2, 1.0, 1e-3), helmholtz(D0, D2, 1e4, 1e-4)):
        x = oracle_lib.band_solve(oracle_lib.band_lu(oracle_lib.full_to_band(A, bw), bw), bw, b)
        assert rel(x, np.linalg.solve(A, b)) <= 1e-13
    for M in (D0, D1, D2):
        y = oracle_lib.band_matvec(oracle_lib.full_to_band(M, bw), bw, b)
        assert rel(y, M @ b) <= 1e-15


@pytest.mark.parametrize("ny", [65, 129])
def test_poisson_residual(oracle_lib, ny):
    """The wall-row Poisson operator is ill conditioned (2-norm ~1e7 at ny=129);
    the Crout solve must still give a small backward error."""
    rng = np.random.default_rng(1)
    (D0, D1, D2), bw = operators(oracle_lib, ny)
    Q = poisson(D0, D2, 2.0)
    b = rng.standard_normal(ny) + 0j
    x = oracle_lib.band_solve(oracle_lib.band_lu(oracle_lib.full_to_band(Q, bw), bw), bw, b)
    resid = np.abs(Q @ x - b).max() / (np.abs(Q).max() * np.abs(x).max())
    assert resid <= 1e-15


def test_planted_mistake_is_caught(oracle_lib):
    """A transposed operand in the matvec (a plausible mistake) would fail the
    dense comparison used above."""
    (D0, D1, D2), bw = operators(oracle_lib, 33)
    x = np.arange(33) + 1j
    good = oracle_lib.band_matvec(oracle_lib.full_to_band(D1, bw), bw, x)
    wrong = D1.T @ x
    assert rel(good, D1 @ x) <= 1e-15 and rel(wrong, D1 @ x) > 1e-3
