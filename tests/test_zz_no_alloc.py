"""d2dhip_exec_c2r_noalloc: the c2r with its workspace in the library's
static scratch array equals d2dhip_exec_c2r bit for bit, preserves its input
whatever the plan's flags, and refuses spectra larger than the array.  Two
calls in a row, the second on a smaller grid, show that nothing of the first
call's scratch contents leaks into the second."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(d2d, torch, fmt, shape, prec, inplace=False, seed=0):
    nx, ny, nz = shape
    rdt = torch.float64 if prec == 64 else torch.float32
    cdt = torch.complex128 if prec == 64 else torch.complex64
    ref = d2d.FFTEngine(fmt, nx, ny, nz, prec=prec)
    e = d2d.FFTEngine(fmt, nx, ny, nz, prec=prec, inplace=inplace)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.rand(nx * ny * nz, dtype=rdt, device="cuda", generator=g) * 2 - 1
    ns = (nx // 2 + 1) * ny * nz if fmt == 1 else nx * ny * (nz // 2 + 1)
    sp = torch.empty(ns, dtype=cdt, device="cuda")
    ref.fft_r2c(x, sp)
    keep = sp.clone()
    want = torch.empty_like(x)
    ref.fft_c2r(sp, want)
    got = torch.full_like(x, float("nan"))
    e.fft_c2r_noalloc(sp, got)
    torch.cuda.synchronize()
    assert torch.equal(sp, keep), "input spectrum changed"
    assert torch.equal(got, want)
    err = float(torch.linalg.vector_norm(got / (nx * ny * nz) - x) / torch.linalg.vector_norm(x))
    assert err < (1e-13 if prec == 64 else 5e-6)


def test_noalloc_c2r_matches_and_second_call_smaller(d2d, torch_gpu):
    torch = torch_gpu
    # first call on a large grid (about 58 MB of the 64 MiB scratch array)
    _case(d2d, torch, 1, (256, 240, 118), 64, seed=1)
    # second call, smaller grid, other format, inplace plan (no workspace of its own)
    _case(d2d, torch, 3, (17, 13, 11), 64, inplace=True, seed=2)
    # FP32 and split-BF16 plans, chunked and unchunked, skip flags
    _case(d2d, torch, 1, (96, 80, 36), 32, seed=3)
    _case(d2d, torch, 1, (1024, 12, 10), d2d.PREC_SPLIT_BF16, seed=4)


def test_noalloc_c2r_with_chunking_and_skips(d2d, torch_gpu):
    torch = torch_gpu
    nx, ny, nz = 64, 48, 40
    for kw in ({"skip_xyz": (False, True, False)}, {"skip_xyz": (False, False, True)}, {}):
        ref = d2d.FFTEngine(1, nx, ny, nz, **kw)
        e = d2d.FFTEngine(1, nx, ny, nz, inplace=True, **kw)
        if not kw:
            ref.set_chunking(8, 8)
            e.set_chunking(8, 8)
        sp = torch.randn((nx // 2 + 1) * ny * nz, dtype=torch.complex128, device="cuda")
        keep = sp.clone()
        want = torch.empty(nx * ny * nz, dtype=torch.float64, device="cuda")
        got = torch.empty_like(want)
        ref.fft_c2r(sp, want)
        e.fft_c2r_noalloc(sp, got)
        torch.cuda.synchronize()
        assert torch.equal(sp, keep), kw
        assert torch.equal(got, want), kw


def test_noalloc_c2r_refuses_large_spectra(d2d, torch_gpu):
    torch = torch_gpu
    nx, ny, nz = 512, 256, 64  # 257*256*64*16 = 67.4 MB > 64 MiB
    assert (nx // 2 + 1) * ny * nz * 16 > d2d.SCRATCH_BYTES
    e = d2d.FFTEngine(1, nx, ny, nz, inplace=True)
    sp = torch.zeros((nx // 2 + 1) * ny * nz, dtype=torch.complex128, device="cuda")
    out = torch.empty(nx * ny * nz, dtype=torch.float64, device="cuda")
    with pytest.raises(d2d.D2dError, match="scratch"):
        e.fft_c2r_noalloc(sp, out)
    # the same grid in FP32 fits
    e32 = d2d.FFTEngine(1, nx, ny, nz, prec=32, inplace=True)
    sp32 = torch.randn((nx // 2 + 1) * ny * nz, dtype=torch.complex64, device="cuda")
    out32 = torch.empty(nx * ny * nz, dtype=torch.float32, device="cuda")
    ref = d2d.FFTEngine(1, nx, ny, nz, prec=32)
    want = torch.empty_like(out32)
    ref.fft_c2r(sp32, want)
    e32.fft_c2r_noalloc(sp32, out32)
    torch.cuda.synchronize()
    assert torch.equal(out32, want)
