"""Parity of the gfx950 FFT path (through the C ABI) with the oracles.

Tolerances (relative L2 over the whole output):
  FP64 against numpy pocketfft / long-double DFT / closed forms: 1e-13
  FP64 against the Glassman restatement of the reference's generic backend:
      2e-13 (the generic backend's own error, see tests/test_oracle.py)
  FP32 against FP64 references: 5e-6
  split-BF16 (FP32 storage, matrix-core radix-16 passes) against FP64
      references of the FP32-rounded input: 1e-6
"""
import numpy as np
import pytest

from conftest import golden_files, load_golden, rel_l2

pytestmark = pytest.mark.gpu

F64_TOL = 1e-13
GLASSMAN_TOL = 2e-13
F32_TOL = 5e-6
SPLIT_TOL = 1e-6  # split-BF16 plans against FP64 references of the FP32-rounded input


def to_dev(torch, a, dtype=None):
    flat = np.ascontiguousarray(np.asarray(a).ravel(order="F"))
    if dtype is not None:
        flat = flat.astype(dtype)
    return torch.from_numpy(flat).to("cuda")


def from_dev(t, shape):
    return t.cpu().numpy().reshape(shape, order="F")


def spec_shape(fmt, nx, ny, nz):
    return (nx // 2 + 1, ny, nz) if fmt == 1 else (nx, ny, nz // 2 + 1)


def run_c2c(d2d, torch, x, isign, fmt, prec=64, **kw):
    nx, ny, nz = x.shape
    cdt = np.complex128 if prec == 64 else np.complex64
    e = d2d.FFTEngine(fmt, nx, ny, nz, prec=prec, **kw)
    xin = to_dev(torch, x, cdt)
    out = torch.empty_like(xin)
    e.fft_c2c(xin, out, isign)
    torch.cuda.synchronize()
    return from_dev(out, x.shape), from_dev(xin, x.shape)


def run_r2c(d2d, torch, x, fmt, prec=64, **kw):
    nx, ny, nz = x.shape
    rdt, cdt = (np.float64, torch.complex128) if prec == 64 else (np.float32, torch.complex64)
    e = d2d.FFTEngine(fmt, nx, ny, nz, prec=prec, **kw)
    xin = to_dev(torch, x, rdt)
    ss = spec_shape(fmt, nx, ny, nz)
    out = torch.empty(int(np.prod(ss)), dtype=cdt, device="cuda")
    e.fft_r2c(xin, out)
    torch.cuda.synchronize()
    return from_dev(out, ss)


def run_c2r(d2d, torch, s, shape, fmt, prec=64, **kw):
    nx, ny, nz = shape
    cdt, rdt = (np.complex128, torch.float64) if prec == 64 else (np.complex64, torch.float32)
    e = d2d.FFTEngine(fmt, nx, ny, nz, prec=prec, **kw)
    sin = to_dev(torch, s, cdt)
    keep = sin.clone()
    out = torch.empty(nx * ny * nz, dtype=rdt, device="cuda")
    e.fft_c2r(sin, out)
    torch.cuda.synchronize()
    return from_dev(out, shape), bool(torch.equal(sin, keep))


@pytest.mark.parametrize("name", golden_files())
@pytest.mark.parametrize("fmt", [1, 3])
def test_c2c_golden(d2d, torch_gpu, oracle, name, fmt):
    d = load_golden(name)
    x = d["c_rand"]
    print("c2c", name, "format", fmt, flush=True)
    got, kept = run_c2c(d2d, torch_gpu, x, -1, fmt)
    assert rel_l2(got, d["c_rand_fwd"]) < F64_TOL
    assert rel_l2(got, oracle.fft3d_c2c(x, -1, fmt)) < GLASSMAN_TOL
    np.testing.assert_array_equal(kept, x)  # non-inplace preserves the input
    got, _ = run_c2c(d2d, torch_gpu, x, 1, fmt)
    assert rel_l2(got, d["c_rand_bwd"]) < F64_TOL
    got, _ = run_c2c(d2d, torch_gpu, d["c_ref"], -1, fmt)
    assert rel_l2(got, d["c_ref_fwd"]) < F64_TOL


@pytest.mark.parametrize("name", golden_files())
@pytest.mark.parametrize("fmt,tag", [(1, "x"), (3, "z")])
def test_r2c_c2r_golden(d2d, torch_gpu, oracle, name, fmt, tag):
    d = load_golden(name)
    shape = tuple(int(v) for v in d["shape"])
    print("r2c/c2r", name, tag, flush=True)
    got = run_r2c(d2d, torch_gpu, d["r_rand"], fmt)
    assert rel_l2(got, d["r_rand_r2c_" + tag]) < F64_TOL
    assert rel_l2(got, oracle.fft3d_r2c(d["r_rand"], fmt)) < GLASSMAN_TOL
    got = run_r2c(d2d, torch_gpu, d["r_ref"], fmt)
    assert rel_l2(got, d["r_ref_r2c_" + tag]) < F64_TOL
    got, preserved = run_c2r(d2d, torch_gpu, d["s_rand_" + tag], shape, fmt)
    assert rel_l2(got, d["s_rand_c2r_" + tag]) < F64_TOL
    assert rel_l2(got, oracle.fft3d_c2r(d["s_rand_" + tag], shape, fmt)) < GLASSMAN_TOL
    assert preserved


def test_reference_round_trip_criterion(d2d, torch_gpu):
    """examples/fft_physical_x/fft_c2c_x.f90:101-154 on the 17x13x11 grid:
    ntest forward+backward pairs with 1/N, mean |err| <= 50 eps ntest."""
    d = load_golden("fft_17x13x11.npz")
    x = d["c_ref"]
    nx, ny, nz = x.shape
    ntest = 10
    for fmt in (1, 3):
        e = d2d.FFTEngine(fmt, nx, ny, nz)
        a = to_dev(torch_gpu, x)
        b = torch_gpu.empty_like(a)
        for _ in range(ntest):
            e.fft_c2c(a, b, -1)
            e.fft_c2c(b, a, 1)
            a /= nx * ny * nz
        y = from_dev(a, x.shape)
        assert np.abs(y - x).sum() / x.size < 50 * np.finfo(np.float64).eps * ntest
        r = d["r_ref"]
        er = d2d.FFTEngine(fmt, nx, ny, nz)
        ra = to_dev(torch_gpu, r)
        sp = torch_gpu.empty(int(np.prod(spec_shape(fmt, nx, ny, nz))), dtype=torch_gpu.complex128,
                             device="cuda")
        for _ in range(ntest):
            er.fft_r2c(ra, sp)
            er.fft_c2r(sp, ra)
            ra /= nx * ny * nz
        assert np.abs(from_dev(ra, r.shape) - r).sum() / r.size < 50 * np.finfo(np.float64).eps * ntest


def test_inplace_c2c(d2d, torch_gpu):
    d = load_golden("fft_18x14x12.npz")
    x = d["c_rand"]
    nx, ny, nz = x.shape
    for inplace_flag in (False, True):
        e = d2d.FFTEngine(1, nx, ny, nz, inplace=inplace_flag)
        a = to_dev(torch_gpu, x)
        e.fft_c2c(a, a, -1)
        torch_gpu.cuda.synchronize()
        assert rel_l2(from_dev(a, x.shape), d["c_rand_fwd"]) < F64_TOL


def test_inplace_c2r_overwrites_only_with_flag(d2d, torch_gpu):
    d = load_golden("fft_16x12x10.npz")
    shape = (16, 12, 10)
    got, preserved = run_c2r(d2d, torch_gpu, d["s_rand_x"], shape, 1, inplace=True)
    assert rel_l2(got, d["s_rand_c2r_x"]) < F64_TOL
    assert not preserved


@pytest.mark.parametrize("skip", [(True, False, True), (False, True, False), (True, True, True),
                                  (False, False, True)])
@pytest.mark.parametrize("fmt", [1, 3])
def test_skip_c2c(d2d, torch_gpu, oracle, skip, fmt):
    """opt_skip_XYZ_c2c (examples/fft_physical_x/fft_c2c_x_skip.f90)."""
    d = load_golden("fft_17x13x11.npz")
    x = d["c_rand"]
    got, _ = run_c2c(d2d, torch_gpu, x, -1, fmt, skip_xyz=skip)
    axes = tuple(a for a in range(3) if not skip[a])
    ref = np.fft.fftn(x, axes=axes) if axes else x
    assert rel_l2(got, ref) < F64_TOL
    assert rel_l2(got, oracle.fft3d_c2c(x, -1, fmt, skip=skip)) < GLASSMAN_TOL


def test_r2c_skip_warns_and_still_transforms(d2d, torch_gpu):
    """src/fft_generic.f90:222-223: the r2c axis is transformed regardless,
    with the warning 'r2c / c2r transform can not be skipped'."""
    d = load_golden("fft_16x12x10.npz")
    with pytest.warns(d2d.D2dWarning, match="can not be skipped"):
        got = run_r2c(d2d, torch_gpu, d["r_rand"], 1, skip_xyz=(True, False, False))
    assert rel_l2(got, d["r_rand_r2c_x"]) < F64_TOL


def test_config1_c2c_128_cube(d2d, torch_gpu, oracle):
    """C2C FP64 128^3, uniform[-1,1) from seed 0 (SplitMix64 stream),
    PHYSICAL_IN_X: forward vs pocketfft and the Glassman oracle, round trip
    after 1/N^3, and bit-exact one-rank transposes."""
    n = 128
    x = np.zeros((n, n, n), np.complex128, order="F")
    d2d.host_fill_uniform(x, d2d.C16, 0)
    e = d2d.FFTEngine(1, n, n, n)
    a = to_dev(torch_gpu, x)
    b = torch_gpu.empty_like(a)
    e.fft_c2c(a, b, -1)
    torch_gpu.cuda.synchronize()
    got = from_dev(b, x.shape)
    assert rel_l2(got, np.fft.fftn(x)) < F64_TOL
    assert rel_l2(got, oracle.fft3d_c2c(x, -1)) < GLASSMAN_TOL
    c = torch_gpu.empty_like(a)
    e.fft_c2c(b, c, 1)
    c /= n ** 3
    assert rel_l2(from_dev(c, x.shape), x) < F64_TOL
    dec = d2d.decomp_info_init(n, n, n)
    for fn in (d2d.transpose_x_to_y, d2d.transpose_y_to_z, d2d.transpose_z_to_y,
               d2d.transpose_y_to_x):
        t = torch_gpu.zeros_like(a)
        fn(b, t, dec)
        torch_gpu.cuda.synchronize()
        assert torch_gpu.equal(t, b)


def test_tuned_plan_cache(d2d, torch_gpu, tmp_path, monkeypatch):
    """D2DHIP_FLAG_TUNE times each stage's candidates (the general kernel and the
    specialised tiles of 128 and 256 points), keeps a candidate of that stage,
    writes the choice to the cache file, and the tuned plan transforms as the
    untuned one does (to rounding: the choices may differ)."""
    cache = tmp_path / "tune.tsv"
    monkeypatch.setenv("D2DHIP_TUNE_CACHE", str(cache))
    shape = (256, 128, 128)
    x = np.zeros(shape, np.complex128, order="F")
    d2d.host_fill_uniform(x, d2d.C16, 3)
    t = d2d.FFTEngine(1, *shape, tune=True)
    for st, name in t.stage_kernels().items():
        assert name in t.stage_candidates(st), (st, name)
    u = d2d.FFTEngine(1, *shape)
    a = to_dev(torch_gpu, x)
    bt = torch_gpu.empty_like(a)
    bu = torch_gpu.empty_like(a)
    t.fft_c2c(a, bt, -1)
    u.fft_c2c(a, bu, -1)
    torch_gpu.cuda.synchronize()
    assert rel_l2(from_dev(bt, shape), from_dev(bu, shape)) < F64_TOL
    assert rel_l2(from_dev(bt, shape), np.fft.fftn(x)) < F64_TOL
    # a new cache path is read afresh, so every stage was tuned and written
    text = cache.read_text()
    assert "io=0 ax=0 n=256" in text
    assert len(text.splitlines()) >= 5, text


@pytest.mark.parametrize("shape", [(90, 60, 48), (48, 80, 96), (160, 192, 32), (45, 27, 25),
                                   (7, 49, 98), (64, 1, 33), (1, 1, 1), (2, 3, 5)])
@pytest.mark.parametrize("fmt", [1, 3])
def test_mixed_radix_vs_pocketfft(d2d, torch_gpu, shape, fmt):
    rng = np.random.default_rng(sum(shape) + fmt)
    x = np.asfortranarray(rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape))
    got, _ = run_c2c(d2d, torch_gpu, x, -1, fmt)
    assert rel_l2(got, np.fft.fftn(x)) < F64_TOL
    r = np.asfortranarray(x.real)
    axes = (2, 1, 0) if fmt == 1 else (0, 1, 2)
    s = run_r2c(d2d, torch_gpu, r, fmt)
    assert rel_l2(s, np.fft.rfftn(r, axes=axes)) < F64_TOL
    back, _ = run_c2r(d2d, torch_gpu, s, shape, fmt)
    assert rel_l2(back / r.size, r) < F64_TOL


@pytest.mark.parametrize("fmt", [1, 3])
def test_fp32_plan(d2d, torch_gpu, fmt):
    shape = (64, 48, 40)
    rng = np.random.default_rng(11)
    x = np.asfortranarray(rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape))
    got, _ = run_c2c(d2d, torch_gpu, x, -1, fmt, prec=32)
    assert rel_l2(got, np.fft.fftn(x.astype(np.complex64).astype(np.complex128))) < F32_TOL
    r = np.asfortranarray(x.real)
    axes = (2, 1, 0) if fmt == 1 else (0, 1, 2)
    s = run_r2c(d2d, torch_gpu, r, fmt, prec=32)
    assert rel_l2(s, np.fft.rfftn(r.astype(np.float32).astype(np.float64), axes=axes)) < F32_TOL
    back, _ = run_c2r(d2d, torch_gpu, s, shape, fmt, prec=32)
    assert rel_l2(back / r.size, r) < F32_TOL


def _modal_peaks_error(d2d, torch, out, modes, fmt, nx, ny, nz):
    """Relative L2 distance of an r2c spectrum (device tensor) from the
    closed form of a noise-free modal field: amp*N/2 at +k and -k."""
    ss = spec_shape(fmt, nx, ny, nz)
    N = nx * ny * nz
    peaks = {}
    for q in range(modes.nmodes):
        for sgn in (1, -1):
            k = ((sgn * modes.kx[q]) % nx, (sgn * modes.ky[q]) % ny, (sgn * modes.kz[q]) % nz)
            peaks[k] = peaks.get(k, 0.0) + modes.amp[q] * N / 2
    o = out.view(ss[2], ss[1], ss[0])  # Fortran (i,j,k) -> torch [k][j][i]
    diff = o.clone()
    ref_norm2 = 0.0
    for (i, j, k), v in peaks.items():
        if i < ss[0] and j < ss[1] and k < ss[2]:
            diff[k, j, i] -= v
            ref_norm2 += abs(v) ** 2
    return float(torch.linalg.vector_norm(diff)) / np.sqrt(ref_norm2)


@pytest.mark.parametrize("n", [512])
def test_config2_r2c_512_cube_closed_form(d2d, torch_gpu, n):
    """R2C FP64 512^3 PHYSICAL_IN_X on a noise-free 32-mode field: the
    spectrum is known in closed form; then the noisy config-2 field
    (sigma 1e-3, seed 1) round trips through c2r."""
    m = d2d.modes_make(32, 64, 0, 0.0, 0.0, 0.0, 1)
    x = torch_gpu.empty(n ** 3, dtype=torch_gpu.float64, device="cuda")
    d2d.fill_modal(x, n, n, n, d2d.R8, m)
    e = d2d.FFTEngine(1, n, n, n)
    out = torch_gpu.empty((n // 2 + 1) * n * n, dtype=torch_gpu.complex128, device="cuda")
    e.fft_r2c(x, out)
    torch_gpu.cuda.synchronize()
    assert _modal_peaks_error(d2d, torch_gpu, out, m, 1, n, n, n) < F64_TOL
    mn = d2d.modes_make(32, 64, 0, 0.0, 0.0, 1e-3, 1)
    d2d.fill_modal(x, n, n, n, d2d.R8, mn)
    y = torch_gpu.empty_like(x)
    e.fft_r2c(x, out)
    e.fft_c2r(out, y)
    y /= n ** 3
    err = float(torch_gpu.linalg.vector_norm(y - x) / torch_gpu.linalg.vector_norm(x))
    assert err < F64_TOL


def test_config3a_r2c_z_720_540_432(d2d, torch_gpu):
    """R2C FP64 720x540x432 PHYSICAL_IN_Z (factors 2,3,5): closed-form
    spectrum of a modal field and a c2r round trip of the uniform seed-2 field."""
    nx, ny, nz = 720, 540, 432
    m = d2d.modes_make(16, 60, 0, 0.0, 0.0, 0.0, 2)
    x = torch_gpu.empty(nx * ny * nz, dtype=torch_gpu.float64, device="cuda")
    d2d.fill_modal(x, nx, ny, nz, d2d.R8, m)
    e = d2d.FFTEngine(3, nx, ny, nz)
    out = torch_gpu.empty(nx * ny * (nz // 2 + 1), dtype=torch_gpu.complex128, device="cuda")
    e.fft_r2c(x, out)
    torch_gpu.cuda.synchronize()
    assert _modal_peaks_error(d2d, torch_gpu, out, m, 3, nx, ny, nz) < F64_TOL
    d2d.fill_uniform(x, nx * ny * nz, d2d.R8, 2)
    y = torch_gpu.empty_like(x)
    e.fft_r2c(x, out)
    e.fft_c2r(out, y)
    y /= nx * ny * nz
    err = float(torch_gpu.linalg.vector_norm(y - x) / torch_gpu.linalg.vector_norm(x))
    assert err < F64_TOL


THIN = [(1024, 6, 5), (6, 1024, 5), (6, 5, 1024), (512, 6, 5), (6, 512, 5), (6, 5, 512),
        (1024, 5, 3), (5, 3, 1024), (512, 3, 5),
        (720, 5, 3), (3, 720, 5), (540, 3, 5), (5, 540, 3), (432, 5, 3), (5, 3, 432),
        (256, 5, 3), (5, 256, 3), (5, 3, 256), (128, 5, 3), (5, 128, 3), (3, 5, 128),
        (64, 5, 3), (5, 64, 3), (5, 3, 64), (32, 5, 3), (3, 32, 5), (5, 3, 32),
        (48, 5, 3), (5, 48, 3), (3, 5, 48), (80, 5, 3), (5, 80, 3), (3, 5, 80),
        (96, 5, 3), (5, 96, 3), (3, 5, 96), (160, 5, 3), (5, 160, 3), (3, 5, 160),
        (192, 5, 3), (5, 192, 3), (3, 5, 192), (160, 48, 20), (96, 80, 36), (2048, 6, 5),
        (384, 5, 3), (5, 384, 3), (3, 5, 384), (640, 5, 3), (5, 640, 3), (3, 5, 640),
        (768, 5, 3), (5, 768, 3), (3, 5, 768), (320, 5, 3), (5, 320, 3), (3, 5, 320),
        (576, 5, 3), (5, 576, 3), (3, 5, 576), (800, 5, 3), (5, 800, 3), (3, 5, 800),
        (960, 5, 3), (5, 960, 3), (3, 5, 960), (1000, 5, 3), (5, 1000, 3), (3, 5, 1000),
        (1152, 5, 3), (5, 1152, 3), (3, 5, 1152), (1280, 5, 3), (5, 1280, 3), (3, 5, 1280),
        (1536, 5, 3), (5, 1536, 3), (3, 5, 1536), (5, 2048, 3), (3, 5, 2048)] + [
        s for n in (100, 120, 200, 240, 300, 360, 400, 480, 500, 600, 896, 1200, 1600)
        for s in ((n, 5, 3), (5, n, 3), (3, 5, n))]
STAGE_USE = {"c2c_x": "c2c", "c2c_y": "c2c", "c2c_z": "c2c", "r2c": "r2c", "spec_y": "rt",
             "spec_last": "rt", "c2r": "c2r"}


@pytest.mark.parametrize("shape", THIN)
@pytest.mark.parametrize("fmt", [1, 3])
@pytest.mark.parametrize("prec", [64, 32, 16])
def test_every_stage_kernel(d2d, torch_gpu, shape, fmt, prec):
    """Every kernel the planner can pick for a stage (the general kernel and
    each specialised variant) gives the pocketfft result."""
    torch = torch_gpu
    nx, ny, nz = shape
    tol = {64: F64_TOL, 32: F32_TOL, 16: SPLIT_TOL}[prec]
    rng = np.random.default_rng(sum(shape) + fmt + prec)
    x = np.asfortranarray(rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape))
    r = np.asfortranarray(x.real)
    cdt = torch.complex128 if prec == 64 else torch.complex64
    rdt = torch.float64 if prec == 64 else torch.float32
    npc = np.complex128 if prec == 64 else np.complex64
    npr = np.float64 if prec == 64 else np.float32
    axes = (2, 1, 0) if fmt == 1 else (0, 1, 2)
    ss = spec_shape(fmt, nx, ny, nz)
    want_c = np.fft.fftn(x.astype(npc).astype(np.complex128))
    want_b = np.fft.ifftn(x.astype(npc).astype(np.complex128)) * x.size
    want_r = np.fft.rfftn(r.astype(npr).astype(np.float64), axes=axes)
    e = d2d.FFTEngine(fmt, nx, ny, nz, prec=prec)
    xin = to_dev(torch, x, npc)
    rin = to_dev(torch, r, npr)
    for stage in e.STAGES:
        for name in e.stage_candidates(stage):
            e.set_stage_kernel(stage, name)
            print(shape, fmt, prec, stage, name, flush=True)
            use = STAGE_USE[stage]
            if use == "c2c":
                out = torch.empty_like(xin)
                e.fft_c2c(xin, out, -1)
                assert rel_l2(from_dev(out, shape), want_c) < tol, (stage, name)
                e.fft_c2c(xin, out, 1)
                assert rel_l2(from_dev(out, shape), want_b) < tol, (stage, name)
            if use in ("r2c", "rt", "c2r"):
                sp = torch.empty(int(np.prod(ss)), dtype=cdt, device="cuda")
                e.fft_r2c(rin, sp)
                assert rel_l2(from_dev(sp, ss), want_r) < tol, (stage, name)
                back = torch.empty(nx * ny * nz, dtype=rdt, device="cuda")
                e.fft_c2r(sp, back)
                got = from_dev(back, shape) / (nx * ny * nz)
                assert rel_l2(got, r) < tol, (stage, name)
        e.set_stage_kernel(stage, e.stage_candidates(stage)[-1])


@pytest.mark.parametrize("shape", [(19, 23, 29), (37, 8, 6), (101, 7, 3), (8, 6, 211), (46, 31, 5),
                                   (1021, 3, 2), (3, 2, 1021)])
@pytest.mark.parametrize("fmt", [1, 3])
@pytest.mark.parametrize("prec", [64, 32])
def test_bluestein_lengths(d2d, torch_gpu, shape, fmt, prec):
    """Lengths with a prime factor above 17 run Bluestein's chirp-z path."""
    tol = F64_TOL if prec == 64 else F32_TOL
    rng = np.random.default_rng(sum(shape) * 7 + fmt)
    x = np.asfortranarray(rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape))
    npc = np.complex128 if prec == 64 else np.complex64
    npr = np.float64 if prec == 64 else np.float32
    got, _ = run_c2c(d2d, torch_gpu, x, -1, fmt, prec=prec)
    assert rel_l2(got, np.fft.fftn(x.astype(npc).astype(np.complex128))) < tol
    got, _ = run_c2c(d2d, torch_gpu, x, 1, fmt, prec=prec)
    assert rel_l2(got, np.fft.ifftn(x.astype(npc).astype(np.complex128)) * x.size) < tol
    r = np.asfortranarray(x.real)
    axes = (2, 1, 0) if fmt == 1 else (0, 1, 2)
    s = run_r2c(d2d, torch_gpu, r, fmt, prec=prec)
    assert rel_l2(s, np.fft.rfftn(r.astype(npr).astype(np.float64), axes=axes)) < tol
    back, _ = run_c2r(d2d, torch_gpu, s, shape, fmt, prec=prec)
    assert rel_l2(back / r.size, r) < tol


@pytest.mark.parametrize("shape,chunks", [((64, 48, 40), [(3, 7), (8, 1), (39, 2)]),
                                          ((33, 15, 11), [(2, 3), (4, 4)]),
                                          ((512, 24, 20), [(6, 5)]),
                                          ((1024, 12, 9), [(4, 2)])])
@pytest.mark.parametrize("prec", [64, 32])
def test_plane_chunking(d2d, torch_gpu, shape, chunks, prec):
    """Plane-chunked stage pairs (forward r2c-x + y, inverse y + c2r-x) give
    the same spectrum and round trip as whole stages, with a shorter last
    chunk, odd ny*chunk (a real pair straddling chunks) and the specialised
    x kernels (512, 1024); FP64 against pocketfft as well."""
    torch = torch_gpu
    nx, ny, nz = shape
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, shape)
    rdt, cdt = (torch.float64, torch.complex128) if prec == 64 else (torch.float32, torch.complex64)
    tol = F64_TOL if prec == 64 else F32_TOL
    e = d2d.FFTEngine(1, nx, ny, nz, prec=prec)
    assert e.chunking() == (0, 0)
    xin = to_dev(torch, x, np.float64 if prec == 64 else np.float32)
    ns = (nx // 2 + 1) * ny * nz
    s0 = torch.empty(ns, dtype=cdt, device="cuda")
    e.fft_r2c(xin, s0)
    y0 = torch.empty_like(xin)
    e.fft_c2r(s0, y0)
    ref = np.fft.fftn(x)[: nx // 2 + 1]
    for f, b in chunks:
        e.set_chunking(f, b)
        assert e.chunking() == (f, b)
        s1 = torch.empty_like(s0)
        e.fft_r2c(xin, s1)
        y1 = torch.empty_like(xin)
        e.fft_c2r(s1, y1)
        torch.cuda.synchronize()
        a0, a1 = s0.cpu().numpy(), s1.cpu().numpy()
        assert rel_l2(a1, a0) <= tol, (f, b)
        assert rel_l2(y1.cpu().numpy(), y0.cpu().numpy()) <= tol, (f, b)
        if prec == 64:
            assert rel_l2(from_dev(s1, (nx // 2 + 1, ny, nz)), ref) <= F64_TOL
            assert rel_l2(from_dev(y1, shape) / x.size, x) <= F64_TOL
    e.set_chunking(0, 0)
    assert e.chunking() == (0, 0)


@pytest.mark.parametrize("shape", [(1024, 16, 8), (16, 1024, 8), (16, 8, 1024), (1024, 6, 1024),
                                   (512, 12, 10), (10, 512, 12), (12, 10, 512)])
@pytest.mark.parametrize("fmt", [1, 3])
def test_split_bf16_plan(d2d, torch_gpu, shape, fmt):
    """Split-BF16 plans run their length-512/1024 axes on the matrix-core kernels
    and stay within 1e-6 (relative L2) of FP64 pocketfft on the FP32-rounded
    input, for c2c both ways, r2c and the c2r round trip."""
    torch = torch_gpu
    nx, ny, nz = shape
    rng = np.random.default_rng(sum(shape) + fmt)
    x = np.asfortranarray(rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(np.complex64)
    r = np.asfortranarray(x.real)
    e = d2d.FFTEngine(fmt, nx, ny, nz, prec=d2d.PREC_SPLIT_BF16)
    kern = e.stage_kernels()
    axes_len = {"c2c_x": nx, "c2c_y": ny, "c2c_z": nz}
    for st, n in axes_len.items():
        assert kern[st].startswith("mfma") == (n in (512, 1024)), (st, kern[st])
    xin = to_dev(torch, x)
    out = torch.empty_like(xin)
    e.fft_c2c(xin, out, -1)
    want = np.fft.fftn(x.astype(np.complex128))
    assert rel_l2(from_dev(out, shape), want) <= SPLIT_TOL
    e.fft_c2c(xin, out, 1)
    assert rel_l2(from_dev(out, shape), np.fft.ifftn(x.astype(np.complex128)) * x.size) <= SPLIT_TOL
    axes = (2, 1, 0) if fmt == 1 else (0, 1, 2)
    ss = spec_shape(fmt, nx, ny, nz)
    rin = to_dev(torch, r)
    sp = torch.empty(int(np.prod(ss)), dtype=torch.complex64, device="cuda")
    e.fft_r2c(rin, sp)
    assert rel_l2(from_dev(sp, ss), np.fft.rfftn(r.astype(np.float64), axes=axes)) <= SPLIT_TOL
    back = torch.empty_like(rin)
    e.fft_c2r(sp, back)
    assert rel_l2(from_dev(back, shape) / r.size, r) <= SPLIT_TOL
