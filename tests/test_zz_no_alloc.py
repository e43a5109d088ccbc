"""GPU: pb_labs_bs_noalloc (no allocation; staging through the library's
file-scope device scratch) against pb_labs_bs, which allocates its device
buffers, and against the oracle's LABS_bs restatement: bit-identical, on
two calls in a row, the second smaller than the first; sizes beyond the
scratch are rejected with PB_E_ARG."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def band_case(n, bw, nrhs, seed):
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    F = np.zeros((n, n))
    for r in range(n):
        lo, hi = max(0, r - bw), min(n, r + bw + 1)
        F[r, lo:hi] = rng.standard_normal(hi - lo)
        F[r, r] += 2.0 * bw + 1.0
    A = O.full_to_band(F, bw)
    LU = O.band_lu(A, bw)
    B = rng.standard_normal((nrhs, n)) + 1j * rng.standard_normal((nrhs, n))
    return LU, B


def test_noalloc_matches_allocating_entry_point(oracle_lib):
    from channel_passive_amd import capi
    from oracle import oracle as O
    for n, bw, nrhs, seed in ((385, 4, 600, 1), (33, 3, 5, 2)):  # second call smaller
        LU, B = band_case(n, bw, nrhs, seed)
        got = capi.labs_bs_noalloc(LU, bw, B)
        ref = capi.labs_bs(LU, bw, B)
        assert np.array_equal(got, ref), (n, bw, nrhs)
        for q in (0, nrhs - 1):
            assert np.array_equal(got[q], O.band_solve(LU, bw, B[q]))


def test_noalloc_rejects_sizes_beyond_scratch():
    from channel_passive_amd import capi
    n, bw = 385, 4
    LU = np.zeros((n, 2 * bw + 1))
    LU[:, bw] = 1.0
    nrhs = (8 << 20) // (16 * n) + 1  # one column more than the 8 MiB scratch holds
    B = np.zeros((nrhs, n), complex)
    with pytest.raises(capi.PbError, match="scratch"):
        capi.labs_bs_noalloc(LU, bw, B)
