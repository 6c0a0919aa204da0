"""Pins the CPU oracle (oracle/) before anything is checked against it.

The Glassman restatement (src/glassman.f90:29-108 and the generic drivers
of src/fft_generic.f90 / src/fft_common_3d.f90) is checked against
independent results: numpy pocketfft golden fixtures (tests/golden), a
long-double direct DFT, and closed forms (single Fourier modes, Hermitian
symmetry, Parseval, round trip).  The reference's generic backend is less
accurate than FFTW (examples/fft_physical_x/README.md:30-33): its 2*pi is
truncated (src/glassman.f90:84) and its twiddles are accumulated products
(:85-103), so the oracle is held to 2e-13 relative L2 where pocketfft holds
1e-15.
"""
import numpy as np
import pytest

from conftest import golden_files, load_golden, rel_l2

ORACLE_TOL = 2e-13


@pytest.mark.parametrize("name", golden_files())
def test_oracle_c2c_golden(oracle, name):
    d = load_golden(name)
    x = d["c_rand"]
    for fmt in (oracle.PHYSICAL_IN_X, oracle.PHYSICAL_IN_Z):
        assert rel_l2(oracle.fft3d_c2c(x, -1, fmt), d["c_rand_fwd"]) < ORACLE_TOL
        assert rel_l2(oracle.fft3d_c2c(x, 1, fmt), d["c_rand_bwd"]) < ORACLE_TOL
    assert rel_l2(oracle.fft3d_c2c(d["c_ref"], -1), d["c_ref_fwd"]) < ORACLE_TOL
    assert rel_l2(oracle.fft3d_c2c(d["c_mg"], -1), d["c_mg_fwd"]) < ORACLE_TOL


@pytest.mark.parametrize("name", golden_files())
def test_oracle_r2c_c2r_golden(oracle, name):
    d = load_golden(name)
    shape = tuple(int(v) for v in d["shape"])
    for fmt, tag in ((oracle.PHYSICAL_IN_X, "x"), (oracle.PHYSICAL_IN_Z, "z")):
        assert rel_l2(oracle.fft3d_r2c(d["r_rand"], fmt), d["r_rand_r2c_" + tag]) < ORACLE_TOL
        assert rel_l2(oracle.fft3d_r2c(d["r_ref"], fmt), d["r_ref_r2c_" + tag]) < ORACLE_TOL
        got = oracle.fft3d_c2r(d["s_rand_" + tag], shape, fmt)
        assert rel_l2(got, d["s_rand_c2r_" + tag]) < ORACLE_TOL


def test_golden_check_catches_a_wrong_sign(oracle):
    """A planted mistake (the opposite sign) must fail the golden comparison."""
    d = load_golden("fft_17x13x11.npz")
    assert rel_l2(oracle.fft3d_c2c(d["c_rand"], 1), d["c_rand_fwd"]) > 0.5


def test_golden_check_catches_a_transposed_axis(oracle):
    """A planted mistake (x and y swapped) must fail the golden comparison."""
    d = load_golden("fft_16x12x10.npz")
    x = np.asfortranarray(d["c_rand"])
    wrong = oracle.fft3d_c2c(np.asfortranarray(np.swapaxes(x, 0, 1)), -1)
    wrong = np.swapaxes(wrong, 0, 1)
    assert rel_l2(wrong, d["c_rand_fwd"]) < ORACLE_TOL  # swapping consistently is fine
    shifted = np.roll(oracle.fft3d_c2c(x, -1), 1, axis=0)
    assert rel_l2(shifted, d["c_rand_fwd"]) > 0.5


@pytest.mark.parametrize("shape", [(9, 8, 7), (5, 1, 3), (4, 6, 2), (11, 3, 2)])
def test_oracle_vs_long_double_dft(oracle, shape):
    rng = np.random.default_rng(7)
    x = np.asfortranarray(rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape))
    for isign in (-1, 1):
        ref = oracle.dft3d_ld(x, isign)
        assert rel_l2(oracle.fft3d_c2c(x, isign), ref) < ORACLE_TOL
    # the long-double DFT itself agrees with pocketfft to FP64 rounding
    assert rel_l2(oracle.dft3d_ld(x, -1), np.fft.fftn(x)) < 1e-15


@pytest.mark.parametrize("n", list(range(1, 70)) + [96, 128, 210, 243, 256, 289, 512])
def test_spcfft_every_length(oracle, n):
    """Glassman's smallest-prime-first factoring covers every n."""
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    assert rel_l2(oracle.spcfft(x, -1), np.fft.fft(x)) < ORACLE_TOL
    assert rel_l2(oracle.spcfft(x, 1), np.fft.ifft(x) * n) < ORACLE_TOL


def test_oracle_single_mode(oracle):
    """exp(2 pi i k.x / N) has the spectrum N * delta_k (closed form)."""
    nx, ny, nz = 12, 10, 9
    kx, ky, kz = 5, 3, 7
    i = np.arange(nx)[:, None, None]
    j = np.arange(ny)[None, :, None]
    k = np.arange(nz)[None, None, :]
    x = np.exp(2j * np.pi * (kx * i / nx + ky * j / ny + kz * k / nz))
    out = oracle.fft3d_c2c(np.asfortranarray(x), -1)
    ref = np.zeros_like(out)
    ref[kx, ky, kz] = nx * ny * nz
    assert np.abs(out - ref).max() < 1e-10


def test_oracle_r2c_hermitian_and_parseval(oracle):
    rng = np.random.default_rng(3)
    nx, ny, nz = 10, 7, 6
    x = np.asfortranarray(rng.uniform(-1, 1, (nx, ny, nz)))
    s = oracle.fft3d_r2c(x, oracle.PHYSICAL_IN_X)
    # kx = 0 plane is Hermitian in (ky, kz)
    p0 = s[0]
    assert np.abs(p0 - np.conj(np.roll(np.flip(p0, (0, 1)), 1, (0, 1)))).max() < 1e-12
    # Parseval over the full spectrum rebuilt from the half spectrum
    full = np.fft.fftn(x)
    assert rel_l2(s, full[: nx // 2 + 1]) < ORACLE_TOL
    e_x = np.sum(x ** 2) * nx * ny * nz
    assert abs(np.sum(np.abs(full) ** 2) - e_x) / e_x < 1e-13


def test_oracle_round_trip_reference_criterion(oracle):
    """The reference's own check (examples/fft_physical_x/fft_c2c_x.f90:101-154):
    forward, backward, 1/N, mean |err| per point <= 50 * eps * ntest."""
    nx, ny, nz = 17, 13, 11
    d = load_golden("fft_17x13x11.npz")
    x = d["c_ref"]
    ntest = 3
    y = x.copy()
    for _ in range(ntest):
        y = oracle.fft3d_c2c(oracle.fft3d_c2c(y, -1), 1) / (nx * ny * nz)
    err = np.abs(y - x).sum() / (nx * ny * nz)
    assert err < 50 * np.finfo(np.float64).eps * ntest
    r = d["r_ref"]
    yr = oracle.fft3d_c2r(oracle.fft3d_r2c(r), (nx, ny, nz)) / (nx * ny * nz)
    assert np.abs(yr - r).sum() / (nx * ny * nz) < 50 * np.finfo(np.float64).eps


def test_oracle_skip_flags(oracle):
    """opt_skip_XYZ_c2c leaves the skipped axis untransformed."""
    d = load_golden("fft_16x12x10.npz")
    x = d["c_rand"]
    out = oracle.fft3d_c2c(x, -1, skip=(True, False, True))
    ref = np.fft.fft(x, axis=1)
    assert rel_l2(out, ref) < ORACLE_TOL


def test_oracle_sampler_runs(oracle):
    t, used = oracle.sample_stage_times(64, 32, 16, 8, nthreads=2)
    assert used >= 1 and all(v > 0 for v in t)
