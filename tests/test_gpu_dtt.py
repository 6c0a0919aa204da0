"""Real-to-real transforms (decomp_2d_dtt_3d_r2r, src/fft_fftw3_f03.f90:1959-2023)
on the GPU against the oracle's direct DCT / DST sums.

FP64 is held to 1e-13 relative L2 against the oracle and to the reference
example's own round-trip bound (examples/dtt/fftw_f03_dtt_x.f90: Linf below
100 eps after forward, backward and the normalisation, all 512 type
combinations on 8 x 9 x 10).  FP32 is held to 2e-6 against the oracle."""
import itertools

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

FMTS = [1, 3]


def _dev(torch, x, rdt):
    return torch.from_numpy(np.ascontiguousarray(np.asfortranarray(x).ravel(order="F"))).to(
        rdt).cuda()


def _host(t, shape):
    return t.cpu().numpy().astype(np.float64).reshape(shape, order="F")


def _ref_input(shape):
    nx, ny, nz = shape
    i, j, k = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), np.arange(1, nz + 1),
                          indexing="ij")
    return np.asfortranarray(i / nx * j / ny * k / nz)


@pytest.mark.parametrize("fmt", FMTS)
def test_reference_example_all_types(d2d, oracle, torch_gpu, fmt):
    """examples/dtt/fftw_f03_dtt_x.f90 / _z.f90 at one rank: every type triple
    on 8 x 9 x 10; the forward result against the oracle, the round trip
    against the example's 100-eps bound."""
    torch = torch_gpu
    shape = (8, 9, 10)
    x = _ref_input(shape)
    xin = _dev(torch, x, torch.float64)
    fwd = torch.empty_like(xin)
    back = torch.empty_like(xin)
    eps = np.finfo(np.float64).eps
    e = d2d.FFTEngine(fmt, *shape)
    for dtt in itertools.product(range(1, 9), repeat=3):
        e.set_dtt(dtt)
        c = e.dtt
        assert c == oracle.dtt_config(dtt), dtt
        e.dtt_3d_r2r(xin, fwd, d2d.FORWARD)
        e.dtt_3d_r2r(fwd, back, d2d.BACKWARD)
        torch.cuda.synchronize()
        got = _host(fwd, shape)
        want = oracle.dtt3d(x, dtt, -1, fmt)
        assert rel_l2(got, want) < 1e-13, dtt
        y = _host(back, shape) / oracle.dtt_norm(dtt, shape)
        sl = tuple(slice(c[3 + a] - 1, c[3 + a] - 1 + shape[a] - c[6 + a]) for a in range(3))
        err = float(np.max(np.abs(y[sl] - x[sl])))
        assert err < 100 * eps, (dtt, err)
    assert np.array_equal(_host(xin, shape), x)  # input preserved


# (shape, dtt): larger lines, odd and prime lengths, a Bluestein extension
# (24 points, DCT-II: L = 23, m = 92 = 4 * 23), explicit ifirst / ndismiss
CASES = [
    ((64, 48, 40), (2, 2, 2)),
    ((64, 48, 40), (3, 6, 7)),
    ((33, 17, 65), (1, 5, 1)),
    ((100, 30, 20), (4, 8, 3)),
    ((24, 20, 12), (2, 3, 4)),
    ((17, 13, 11), (6, 7, 8)),
    ((40, 24, 18), (2, 6, 1, 3, 2, 1, 4, 3, 0, 3, 2, 1)),
    ((256, 16, 8), (5, 2, 8)),
    # (2^k + 1)-point axes: the specialised DCT-II / III tiles
    ((65, 129, 257), (2, 3, 6)),
    ((257, 65, 129), (7, 6, 2)),
    ((1025, 33, 16), (3, 2, 7)),
    ((129, 513, 65), (6, 7, 3)),
    ((66, 130, 40), (2, 3, 6, 2, 1, 1, 2, 2, 7, 2, 1, 1)),
    # extensions of 2^k points: the specialised DCT-I / IV, DST-I / IV tiles
    ((129, 65, 33), (4, 8, 5)),
    ((257, 513, 17), (1, 5, 1)),
    # DCT-IV / DST-IV of even length on L/2-point tiles
    ((257, 129, 513), (8, 4, 8)),
]


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("fmt", FMTS)
@pytest.mark.parametrize("case", CASES)
def test_dtt_against_oracle(d2d, oracle, torch_gpu, case, fmt, prec):
    torch = torch_gpu
    shape, dtt = case
    rdt = torch.float64 if prec == 64 else torch.float32
    tol = 1e-13 if prec == 64 else 2e-6
    x = np.asfortranarray(np.random.default_rng(sum(shape) + fmt).uniform(-1, 1, shape))
    e = d2d.FFTEngine(fmt, *shape, prec=prec, opt_DTT=dtt)
    xin = _dev(torch, x, rdt)
    xr = _host(xin, shape)  # the input as stored (rounded for FP32)
    out = torch.empty_like(xin)
    e.dtt_3d_r2r(xin, out, d2d.FORWARD)
    torch.cuda.synchronize()
    assert rel_l2(_host(out, shape), oracle.dtt3d(xr, dtt, -1, fmt)) < tol
    back = torch.empty_like(xin)
    e.dtt_3d_r2r(out, back, d2d.BACKWARD)
    torch.cuda.synchronize()
    want = oracle.dtt3d(_host(out, shape), dtt, 1, fmt)
    assert rel_l2(_host(back, shape), want) < tol


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("shape,dtt", [((513, 257, 129), (2, 7, 6)),
                                       ((65, 129, 1025), (3, 6, 2)),
                                       ((513, 129, 65), (1, 4, 5)),
                                       ((65, 257, 33), (8, 1, 4)),
                                       ((257, 129, 65), (8, 4, 4))])
def test_specialised_dtt_tiles_match_general(d2d, oracle, torch_gpu, prec, shape, dtt):
    """Every axis runs on a specialised real-to-real tile (DFTs of 64 to 1024
    points on lines and columns); the general kernel gives the same
    transform to rounding, and the round trip returns the input."""
    torch = torch_gpu
    rdt = torch.float64 if prec == 64 else torch.float32
    tol = 1e-13 if prec == 64 else 2e-6
    e = d2d.FFTEngine(1, *shape, prec=prec, opt_DTT=dtt)
    names = e.dtt_stage_kernels()
    assert all("dct" in n for n in names.values()), names
    x = np.asfortranarray(np.random.default_rng(9).uniform(-1, 1, shape))
    xin = _dev(torch, x, rdt)
    out = torch.empty_like(xin)
    back = torch.empty_like(xin)
    e.dtt_3d_r2r(xin, out, d2d.FORWARD)
    e.dtt_3d_r2r(out, back, d2d.BACKWARD)
    torch.cuda.synchronize()
    fast_out, fast_back = out.clone(), back.clone()
    for st in range(7, 13):
        e.set_stage_kernel(st, "general")
    e.dtt_3d_r2r(xin, out, d2d.FORWARD)
    e.dtt_3d_r2r(out, back, d2d.BACKWARD)
    torch.cuda.synchronize()
    assert rel_l2(_host(fast_out, shape), _host(out, shape)) < tol
    assert rel_l2(_host(fast_back, shape), _host(back, shape)) < tol
    y = _host(fast_back, shape) / oracle.dtt_norm(dtt, shape)
    c = oracle.dtt_config(dtt)
    sl = tuple(slice(c[3 + a] - 1, c[3 + a] - 1 + shape[a] - c[6 + a]) for a in range(3))
    assert rel_l2(y[sl], _host(xin, shape)[sl]) < tol


@pytest.mark.parametrize("shape,dtt,axis", [((9, 513, 17), (1, 1, 1), 1), ((9, 17, 513), (2, 2, 2), 2),
                                            ((9, 17, 513), (4, 6, 6), 2)])
def test_every_dtt_column_kernel(d2d, oracle, torch_gpu, shape, dtt, axis):
    """Every specialised kernel able to run the long axis's real-to-real
    stages (forward and backward; the 1024-point extension of DCT-I, the
    512-point tiles of DCT-II / III and DST-II / III) matches the oracle to
    1e-13 and returns the input."""
    torch = torch_gpu
    e = d2d.FFTEngine(1, *shape, opt_DTT=dtt)
    x = np.asfortranarray(np.random.default_rng(3).uniform(-1, 1, shape))
    xin = _dev(torch, x, torch.float64)
    out = torch.empty_like(xin)
    back = torch.empty_like(xin)
    want = oracle.dtt3d(x, dtt, -1, 1)
    fwd, bwd = 7 + axis, 10 + axis
    cands_f = [n for n in e.stage_candidates(fwd) if n != "general"]
    cands_b = [n for n in e.stage_candidates(bwd) if n != "general"]
    assert cands_f and cands_b, (cands_f, cands_b)
    for name_f in cands_f:
        for name_b in cands_b:
            if name_f.split(" dct")[0] != name_b.split(" dct")[0]:
                continue
            print("dtt", shape, dtt, name_f, "/", name_b, flush=True)
            e.set_stage_kernel(fwd, name_f)
            e.set_stage_kernel(bwd, name_b)
            e.dtt_3d_r2r(xin, out, d2d.FORWARD)
            e.dtt_3d_r2r(out, back, d2d.BACKWARD)
            torch.cuda.synchronize()
            assert rel_l2(_host(out, shape), want) < 1e-13, (name_f, name_b)
            y = _host(back, shape) / oracle.dtt_norm(dtt, shape)
            c = oracle.dtt_config(dtt)
            sl = tuple(slice(c[3 + a] - 1, c[3 + a] - 1 + shape[a] - c[6 + a]) for a in range(3))
            assert rel_l2(y[sl], x[sl]) < 1e-13, (name_f, name_b)


def test_dtt_on_split_bf16_plan(d2d, oracle, torch_gpu):
    """A split-BF16 plan stores FP32, so its real-to-real stages run the FP32
    tiles (the matrix-core kernels are complex-to-complex only)."""
    torch = torch_gpu
    shape, dtt = (513, 65, 33), (2, 3, 4)
    e = d2d.FFTEngine(1, *shape, prec=d2d.PREC_SPLIT_BF16, opt_DTT=dtt)
    names = e.dtt_stage_kernels()
    assert names[7].startswith("float") and "dct2" in names[7], names
    x = np.asfortranarray(np.random.default_rng(4).uniform(-1, 1, shape))
    xin = _dev(torch, x, torch.float32)
    out = torch.empty_like(xin)
    e.dtt_3d_r2r(xin, out, d2d.FORWARD)
    torch.cuda.synchronize()
    assert rel_l2(_host(out, shape), oracle.dtt3d(_host(xin, shape), dtt, -1)) < 2e-6


def test_dtt_in_place_and_skip(d2d, oracle, torch_gpu):
    torch = torch_gpu
    shape, dtt = (48, 40, 36), (4, 1, 6)
    x = np.asfortranarray(np.random.default_rng(11).uniform(-1, 1, shape))
    e = d2d.FFTEngine(1, *shape, opt_DTT=dtt)
    buf = _dev(torch, x, torch.float64)
    e.dtt_3d_r2r(buf, buf, d2d.FORWARD)
    torch.cuda.synchronize()
    assert rel_l2(_host(buf, shape), oracle.dtt3d(x, dtt, -1)) < 1e-13
    s = d2d.FFTEngine(3, *shape, skip_xyz=(False, True, False), opt_DTT=dtt)
    xin = _dev(torch, x, torch.float64)
    out = torch.empty_like(xin)
    s.dtt_3d_r2r(xin, out, d2d.BACKWARD)
    torch.cuda.synchronize()
    want = oracle.dtt3d(x, dtt, 1, 3, skip=(False, True, False))
    assert rel_l2(_host(out, shape), want) < 1e-13
    a = d2d.FFTEngine(1, *shape, skip_xyz=(True, True, True), opt_DTT=dtt)
    a.dtt_3d_r2r(xin, out, d2d.FORWARD)
    torch.cuda.synchronize()
    assert torch.equal(out, xin)


@pytest.mark.parametrize("fmt", FMTS)
def test_specialised_dtt_in_place(d2d, oracle, torch_gpu, fmt):
    """in == out on the specialised tiles of every kind (DCT-II / DCT-I / DST-III
    forward, their inverses backward), copied-through points included."""
    torch = torch_gpu
    shape, dtt = (129, 65, 257), (2, 1, 7)
    x = np.asfortranarray(np.random.default_rng(21).uniform(-1, 1, shape))
    e = d2d.FFTEngine(fmt, *shape, opt_DTT=dtt)
    assert all(("dct" in n) for n in e.dtt_stage_kernels().values())
    buf = _dev(torch, x, torch.float64)
    e.dtt_3d_r2r(buf, buf, d2d.FORWARD)
    torch.cuda.synchronize()
    fwd = _host(buf, shape)
    assert rel_l2(fwd, oracle.dtt3d(x, dtt, -1, fmt)) < 1e-13
    e.dtt_3d_r2r(buf, buf, d2d.BACKWARD)
    torch.cuda.synchronize()
    assert rel_l2(_host(buf, shape), oracle.dtt3d(fwd, dtt, 1, fmt)) < 1e-13


def test_tuned_dtt_plan(d2d, oracle, torch_gpu, tmp_path, monkeypatch):
    """A plan made to be tuned times its real-to-real stages' kernels too (the
    general kernel and the specialised tiles), caches the choices with the
    transform in the key, and computes the same transforms."""
    torch = torch_gpu
    cache = tmp_path / "tune.tsv"
    monkeypatch.setenv("D2DHIP_TUNE_CACHE", str(cache))
    shape, dtt = (129, 65, 257), (2, 4, 7)
    e = d2d.FFTEngine(1, *shape, tune=True, opt_DTT=dtt)
    names = e.dtt_stage_kernels()
    assert all(n in e.stage_candidates(st) for st, n in names.items()), names
    x = np.asfortranarray(np.random.default_rng(13).uniform(-1, 1, shape))
    xin = _dev(torch, x, torch.float64)
    out = torch.empty_like(xin)
    e.dtt_3d_r2r(xin, out, d2d.FORWARD)
    torch.cuda.synchronize()
    assert rel_l2(_host(out, shape), oracle.dtt3d(x, dtt, -1)) < 1e-13
    # a new cache path is read afresh, so every real-to-real stage was tuned
    # and written, each with its transform in the key
    text = cache.read_text()
    assert sum("dtt=" in line for line in text.splitlines()) >= 5, text


def test_dtt_graph_replays(d2d, torch_gpu):
    torch = torch_gpu
    shape = (32, 24, 20)
    e = d2d.FFTEngine(1, *shape, opt_DTT=(2, 5, 8))
    x = torch.rand(int(np.prod(shape)), dtype=torch.float64, device="cuda")
    f = torch.empty_like(x)
    b = torch.empty_like(x)
    e.dtt_3d_r2r(x, f, d2d.FORWARD)
    e.dtt_3d_r2r(f, b, d2d.BACKWARD)
    torch.cuda.synchronize()
    f0, b0 = f.clone(), b.clone()
    e.capture_begin()
    e.dtt_3d_r2r(x, f, d2d.FORWARD)
    e.dtt_3d_r2r(f, b, d2d.BACKWARD)
    e.capture_end()
    for rep in range(3):
        f.zero_()
        b.zero_()
        torch.cuda.synchronize()
        e.graph_launch()
        torch.cuda.synchronize()
        assert torch.equal(f, f0) and torch.equal(b, b0), rep


def test_dtt_errors(d2d, torch_gpu):
    torch = torch_gpu
    e = d2d.FFTEngine(1, 8, 9, 10)
    a = torch.zeros(720, dtype=torch.float64, device="cuda")
    with pytest.raises(d2d.D2dError):
        e.dtt_3d_r2r(a, a, d2d.FORWARD)  # no DTT set
    with pytest.raises(d2d.D2dError):
        _ = e.dtt
    for bad in ([0, 1, 1], [1, 9, 1], [1, 2], [2, 2, 2, 1, 1, 1, 1, 1, 1, 2, 1, 1]):
        with pytest.raises(d2d.D2dError):
            e.set_dtt(bad)
    with pytest.raises(d2d.D2dError):
        e.set_dtt([2, 2, 2, 1, 1, 1, 9, 1, 1, 1, 1, 1])  # ndismiss beyond the line
    e.set_dtt([2, 2, 2])
    with pytest.raises(d2d.D2dError):
        e.dtt_3d_r2r(a, a, 0)
    names = [e.stage_candidates(s) for s in range(7, 13)]
    assert all(n == ["general"] for n in names)  # no specialised tiles of 7, 8, 9 points
