"""Blocked spectral layouts of PHYSICAL_IN_X r2c / c2r plans (d2dhip_plan_set_layout:
blocks of 16 and of 32 i-columns).

The calls' inputs and outputs keep the reference's natural (i, j, k) order
(src/fft_common_3d.f90:121-296); only the pencil between the stages, in the
plan's workspace, is blocked, and the y and z passes run in the other order.
Checked against numpy pocketfft (tolerances as in test_gpu_fft.py: FP64 1e-13,
FP32 5e-6, split-BF16 1e-6 against the FP64 reference of the FP32-rounded
input), with the input of c2r preserved, the layout agreeing with the plain
one to rounding, every candidate kernel of the four stages in the blocked
form, the tuner's choice, and the calls that keep the plain layout.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = {64: 1e-13, 32: 5e-6, 16: 1e-6}


def to_dev(torch, a, dtype):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F")).astype(dtype)).to("cuda")


def from_dev(t, shape):
    return t.cpu().numpy().reshape(shape, order="F")


def field(shape, prec, seed):
    rng = np.random.default_rng(seed)
    x = np.asfortranarray(rng.uniform(-1, 1, shape))
    return x.astype(np.float32).astype(np.float64) if prec != 64 else x


def reference(x):
    return np.fft.fftn(np.fft.rfft(x, axis=0), axes=(1, 2))


def round_trip(d2d, torch, e, x, prec):
    nx, ny, nz = x.shape
    rdt = np.float64 if prec == 64 else np.float32
    cdt = torch.complex128 if prec == 64 else torch.complex64
    xin = to_dev(torch, x, rdt)
    spec = torch.empty((nx // 2 + 1) * ny * nz, dtype=cdt, device="cuda")
    e.fft_r2c(xin, spec)
    keep = spec.clone()
    back = torch.empty_like(xin)
    e.fft_c2r(spec, back)
    torch.cuda.synchronize()
    assert torch.equal(spec, keep), "c2r changed its input"
    return from_dev(spec, (nx // 2 + 1, ny, nz)), from_dev(back, x.shape) / x.size


# (nx, ny, nz): nx/2+1 = 11 (no whole block), 16 (whole blocks only), 18, 17 with
# odd nx, ny, nz (an unpaired last line per plane), 66, 33 with 1024-point axes
# (the 1024^3 kernels), 34 with 512/1024 axes
SHAPES = [(20, 12, 10), (30, 16, 24), (34, 20, 18), (33, 21, 19), (130, 64, 128),
          (64, 1024, 1024), (66, 512, 1024)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec", [64, 32, 16])
@pytest.mark.parametrize("layout", [1, 2])
def test_blocked_r2c_c2r(d2d, torch_gpu, shape, prec, layout):
    if prec == 16 and not (shape[1] in (512, 1024) and shape[2] in (512, 1024)):
        pytest.skip("split-BF16 plans use the matrix cores on 512/1024-point axes only")
    print("blocked r2c/c2r", shape, "prec", prec, "layout", layout, flush=True)
    torch = torch_gpu
    x = field(shape, prec, sum(shape) + prec)
    e = d2d.FFTEngine(d2d.PHYSICAL_IN_X, *shape, prec=prec)
    e.set_layout(layout)
    assert e.layout() == layout
    spec, back = round_trip(d2d, torch, e, x, prec)
    ref = reference(x)
    assert rel_l2(spec, ref) <= TOL[prec]
    assert rel_l2(back, x) <= TOL[prec]
    # the plain layout on the same engine: equal to rounding
    e.set_layout(d2d.LAYOUT_PLAIN)
    spec0, back0 = round_trip(d2d, torch, e, x, prec)
    assert rel_l2(spec, spec0) <= TOL[prec]
    assert rel_l2(back, back0) <= TOL[prec]


@pytest.mark.parametrize("shape,prec,layout", [((34, 64, 128), 64, 1), ((34, 1024, 1024), 64, 1),
                                               ((34, 1024, 1024), 32, 1), ((34, 1024, 1024), 16, 1),
                                               ((66, 64, 128), 64, 2), ((66, 1024, 1024), 16, 2)])
def test_blocked_every_kernel(d2d, torch_gpu, shape, prec, layout):
    """Every kernel able to run the x, y or z stage, in the blocked form."""
    torch = torch_gpu
    x = field(shape, prec, 7)
    ref = reference(x)
    e = d2d.FFTEngine(d2d.PHYSICAL_IN_X, *shape, prec=prec)
    e.set_layout(layout)
    for stage in ("spec_y", "spec_last", "r2c", "c2r"):
        keep = e.stage_kernels()[stage]
        for name in e.stage_candidates(stage):
            print("blocked", shape, prec, stage, name, flush=True)
            e.set_stage_kernel(stage, name)
            spec, back = round_trip(d2d, torch, e, x, prec)
            assert rel_l2(spec, ref) <= TOL[prec], (stage, name)
            assert rel_l2(back, x) <= TOL[prec], (stage, name)
        e.set_stage_kernel(stage, keep)


def test_tuned_layout(d2d, torch_gpu, tmp_path, monkeypatch):
    """A tuned plan races both layouts (each with its own kernels) and caches
    the pick; a second plan of the shape takes it from the cache."""
    torch = torch_gpu
    cache = tmp_path / "tune.tsv"
    monkeypatch.setenv("D2DHIP_TUNE_CACHE", str(cache))
    shape = (130, 64, 128)
    x = field(shape, 64, 5)
    ref = reference(x)
    e = d2d.FFTEngine(d2d.PHYSICAL_IN_X, *shape, tune=True)
    spec, back = round_trip(d2d, torch, e, x, 64)
    assert rel_l2(spec, ref) <= TOL[64] and rel_l2(back, x) <= TOL[64]
    lines = [ln for ln in cache.read_text().splitlines() if " layout " in ln]
    assert len(lines) == 1
    val = lines[0].split("\t")[1]
    assert val == "0" or (val[0] in "12" and val[1] == "|" and val.count("|") == 4), val
    e2 = d2d.FFTEngine(d2d.PHYSICAL_IN_X, *shape, tune=True)
    assert e2.layout() == e.layout() and e2.stage_kernels() == e.stage_kernels()
    spec2, back2 = round_trip(d2d, torch, e2, x, 64)
    assert np.array_equal(spec2, spec) and np.array_equal(back2, back)


def test_tuning_without_room(d2d, torch_gpu, tmp_path, monkeypatch):
    """A tuned plan whose tuning scratch does not fit device memory keeps its
    default kernels and the plain layout, caches nothing, and is exact."""
    torch = torch_gpu
    cache = tmp_path / "tune.tsv"
    monkeypatch.setenv("D2DHIP_TUNE_CACHE", str(cache))
    monkeypatch.setenv("D2DHIP_TUNE_MEM_LIMIT", "0")
    shape = (130, 64, 128)
    x = field(shape, 64, 13)
    e = d2d.FFTEngine(d2d.PHYSICAL_IN_X, *shape, tune=True)
    ref = d2d.FFTEngine(d2d.PHYSICAL_IN_X, *shape)
    assert e.stage_kernels() == ref.stage_kernels()
    assert e.layout() == d2d.LAYOUT_PLAIN and e.chunking() == (0, 0)
    assert not cache.exists() or not cache.read_text().strip()
    spec, back = round_trip(d2d, torch, e, x, 64)
    assert rel_l2(spec, reference(x)) <= TOL[64] and rel_l2(back, x) <= TOL[64]


def test_blocked_plain_fallbacks(d2d, torch_gpu):
    """Chunked or skipped stages and exec_c2r_noalloc keep the plain layout;
    plans without a workspace or of one rank of a process grid refuse it."""
    torch = torch_gpu
    shape = (34, 20, 18)
    x = field(shape, 64, 3)
    ref = reference(x)
    e = d2d.FFTEngine(d2d.PHYSICAL_IN_X, *shape)
    e.set_layout(d2d.LAYOUT_BLOCKED)
    e.set_chunking(4, 4)
    spec, back = round_trip(d2d, torch, e, x, 64)
    assert rel_l2(spec, ref) <= TOL[64] and rel_l2(back, x) <= TOL[64]
    e.set_chunking(0, 0)
    xin = to_dev(torch, x, np.float64)
    sp = torch.empty(18 * 20 * 18, dtype=torch.complex128, device="cuda")
    e.fft_r2c(xin, sp)
    out = torch.empty_like(xin)
    e.fft_c2r_noalloc(sp, out)
    torch.cuda.synchronize()
    assert rel_l2(from_dev(out, shape) / x.size, x) <= TOL[64]
    s = d2d.FFTEngine(d2d.PHYSICAL_IN_X, *shape, skip_xyz=(False, False, True))
    s.set_layout(d2d.LAYOUT_BLOCKED)
    s.fft_r2c(xin, sp)
    torch.cuda.synchronize()
    want = np.fft.fft(np.fft.rfft(x, axis=0), axis=1)  # z skipped
    assert rel_l2(from_dev(sp, (18, 20, 18)), want) <= TOL[64]
    inplace = d2d.FFTEngine(d2d.PHYSICAL_IN_X, *shape, inplace=True)
    with pytest.raises(d2d.D2dError):
        inplace.set_layout(d2d.LAYOUT_BLOCKED)
    rank = d2d.FFTEngine(d2d.PHYSICAL_IN_X, 32, 32, 32, rank=(2, 2, 0, 1))
    with pytest.raises(d2d.D2dError):
        rank.set_layout(d2d.LAYOUT_BLOCKED)
    z = d2d.FFTEngine(d2d.PHYSICAL_IN_Z, *shape)
    with pytest.raises(d2d.D2dError):
        z.set_layout(d2d.LAYOUT_BLOCKED)
    with pytest.raises(d2d.D2dError):
        e.set_layout(3)


def test_blocked_graph_replay(d2d, torch_gpu):
    """A captured blocked round trip replays equal to the eager calls."""
    torch = torch_gpu
    shape = (130, 64, 128)
    x = field(shape, 64, 11)
    e = d2d.FFTEngine(d2d.PHYSICAL_IN_X, *shape)
    e.set_layout(d2d.LAYOUT_BLOCKED)
    xin = to_dev(torch, x, np.float64)
    sp = torch.empty(66 * 64 * 128, dtype=torch.complex128, device="cuda")
    out = torch.empty_like(xin)
    e.fft_r2c(xin, sp)
    e.fft_c2r(sp, out)
    torch.cuda.synchronize()
    sp0, out0 = sp.clone(), out.clone()
    e.capture_begin()
    e.fft_r2c(xin, sp)
    e.fft_c2r(sp, out)
    e.capture_end()
    for rep in range(3):
        sp.zero_()
        out.zero_()
        torch.cuda.synchronize()
        e.graph_launch()
        torch.cuda.synchronize()
        assert torch.equal(sp, sp0) and torch.equal(out, out0), rep
