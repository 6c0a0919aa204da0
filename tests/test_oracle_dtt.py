"""Pins the oracle's real-to-real transforms (DCT / DST types 1-8) before the
GPU stages are checked against it.

The oracle restates FFTW's r2r kinds (REDFT00..11, RODFT00..11) as direct
sums; FFTW itself is not in this image.  scipy.fft.dct / dst with norm=None
compute the same unnormalised definitions through pocketfft, an independent
implementation, so they pin every type and length here.  The inverse pairs
of dtt_invert (src/fft_fftw3_f03.f90:424-444) must give the scaled identity
with the factors of examples/dtt/fftw_f03_dtt_x.f90:124-153, and the 3D
driver must pass that example's own check: every combination of types on
the 8 x 9 x 10 grid, forward then backward, back to the input within 100
machine epsilons on the transformed points."""
import itertools

import numpy as np
import pytest
import scipy.fft as sf

LENGTHS = [2, 3, 4, 5, 7, 8, 9, 12, 16, 17, 31]


def _scipy(t, x):
    return sf.dct(x, type=t) if t <= 4 else sf.dst(x, type=t - 4)


@pytest.mark.parametrize("L", LENGTHS)
@pytest.mark.parametrize("t", range(1, 9))
def test_dtt_matrix_matches_pocketfft(oracle, t, L):
    x = np.random.default_rng(100 * t + L).standard_normal(L)
    want = _scipy(t, x)
    got = oracle.dtt_matrix(t, L) @ x
    assert np.max(np.abs(got - want)) <= 1e-14 * max(1.0, np.max(np.abs(want)))


def test_dtt_check_catches_a_swapped_type(oracle):
    """A DCT-II / DCT-III or DST-II / DST-III mix-up (the planted mistake) fails
    the pocketfft check."""
    x = np.random.default_rng(7).standard_normal(9)
    assert np.max(np.abs(oracle.dtt_matrix(3, 9) @ x - _scipy(2, x))) > 1e-3
    assert np.max(np.abs(oracle.dtt_matrix(7, 9) @ x - _scipy(6, x))) > 1e-3
    # DST-III without its (-1)^k X_{L-1} term
    bad = oracle.dtt_matrix(7, 9).copy()
    bad[:, -1] = 0.0
    assert np.max(np.abs(bad @ x - _scipy(7, x))) > 1e-3


INV = {1: 1, 2: 3, 3: 2, 4: 4, 5: 5, 6: 7, 7: 6, 8: 8}


@pytest.mark.parametrize("L", [2, 5, 8, 9])
@pytest.mark.parametrize("t", range(1, 9))
def test_inverse_pairs_scaled_identity(oracle, t, L):
    """inv(t) after t is N * identity: 2(L-1) for DCT-I, 2(L+1) for DST-I, 2L
    for the others (the reference example's factors with L = n - ndismiss)."""
    A = oracle.dtt_matrix(INV[t], L) @ oracle.dtt_matrix(t, L)
    N = 2 * (L - 1) if t == 1 else 2 * (L + 1) if t == 5 else 2 * L
    assert np.max(np.abs(A - N * np.eye(L))) < 1e-12 * N


def test_dtt_config_defaults(oracle):
    """dtt_assign_default + dtt_invert (src/fft_fftw3_f03.f90:394-444)."""
    c = oracle.dtt_config([1, 5, 7])
    assert c[0:3] == (1, 5, 7)
    assert c[3:6] == (1, 2, 2)      # ifirst: DST-I and DST-III start at 2
    assert c[6:9] == (0, 2, 1)      # ndismiss: 0 for DCT-I, 2 for DST-I
    assert c[9:12] == c[3:6]        # ofirst = ifirst
    assert c[12:15] == (1, 5, 6)    # inv(DST-III) = DST-II
    c = oracle.dtt_config([2, 3, 6])
    assert c[3:6] == (1, 1, 1) and c[6:9] == (1, 1, 1) and c[12:15] == (3, 2, 7)
    full = [4, 8, 2, 3, 1, 2, 4, 0, 1, 3, 1, 2]
    assert oracle.dtt_config(full)[:12] == tuple(full)
    with pytest.raises(ValueError):
        oracle.dtt_config([0, 1, 1])
    with pytest.raises(ValueError):
        oracle.dtt_config([1, 9, 1])


def test_dtt_1d_passes_other_points_through(oracle):
    x = np.asfortranarray(np.random.default_rng(3).standard_normal((10, 4, 3)))
    y = oracle.dtt_1d(x, 0, 6, 2, 5)
    assert np.array_equal(y[:2], x[:2]) and np.array_equal(y[7:], x[7:])
    want = np.apply_along_axis(lambda v: sf.dst(v, type=2), 0, x[2:7])
    assert np.max(np.abs(y[2:7] - want)) < 1e-13


def test_dtt3d_separable(oracle):
    """A 3D transform of an outer product is the outer product of the 1D ones."""
    rng = np.random.default_rng(5)
    a, b, c = rng.standard_normal(8), rng.standard_normal(9), rng.standard_normal(10)
    x = np.asfortranarray(np.einsum("i,j,k->ijk", a, b, c))
    y = oracle.dtt3d(x, [2, 4, 1], -1)
    cfg = oracle.dtt_config([2, 4, 1])
    ya = a.copy(); ya[0:7] = sf.dct(a[0:7], type=2)
    yb = b.copy(); yb[0:8] = sf.dct(b[0:8], type=4)
    yc = sf.dct(c, type=1)
    assert cfg[6:9] == (1, 1, 0)
    want = np.einsum("i,j,k->ijk", ya, yb, yc)
    assert np.max(np.abs(y - want)) < 1e-12 * np.max(np.abs(want))


@pytest.mark.parametrize("fmt", [1, 3])
def test_reference_dtt_example_round_trip(oracle, fmt):
    """examples/dtt/fftw_f03_dtt_x.f90 (and _z): in(i,j,k) = i/nx * j/ny * k/nz
    (1-based) on 8 x 9 x 10, all 512 type combinations, forward then
    backward, divided by the example's factors, within 100 eps (Linf) on the
    transformed points."""
    nx, ny, nz = 8, 9, 10
    i, j, k = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), np.arange(1, nz + 1),
                          indexing="ij")
    x = np.asfortranarray(i / nx * j / ny * k / nz)
    eps = np.finfo(np.float64).eps
    for dtt in itertools.product(range(1, 9), repeat=3):
        c = oracle.dtt_config(dtt)
        y = oracle.dtt3d(oracle.dtt3d(x, dtt, -1, fmt), dtt, 1, fmt) / oracle.dtt_norm(dtt, x.shape)
        sl = tuple(slice(c[3 + a] - 1, c[3 + a] - 1 + x.shape[a] - c[6 + a]) for a in range(3))
        err = np.max(np.abs(y[sl] - x[sl]))
        assert err < 100 * eps, (dtt, err)
