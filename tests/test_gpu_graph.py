"""A plan's transforms recorded into a hipGraph (d2dhip_plan_capture_begin /
_end) and replayed: every replay, with eager work and synchronisation in
between, equals the eager calls bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fmt,shape,prec", [(1, (64, 48, 40), 64), (3, (96, 80, 36), 64),
                                             (1, (1024, 16, 8), 32), (3, (17, 13, 11), 64)])
def test_r2c_c2r_graph_replays(d2d, torch_gpu, fmt, shape, prec):
    torch = torch_gpu
    nx, ny, nz = shape
    rdt = torch.float64 if prec == 64 else torch.float32
    cdt = torch.complex128 if prec == 64 else torch.complex64
    e = d2d.FFTEngine(fmt, nx, ny, nz, prec=prec)
    ns = (nx // 2 + 1) * ny * nz if fmt == 1 else nx * ny * (nz // 2 + 1)
    x = torch.rand(nx * ny * nz, dtype=rdt, device="cuda") * 2 - 1
    sp = torch.empty(ns, dtype=cdt, device="cuda")
    y = torch.empty_like(x)
    e.fft_r2c(x, sp)
    e.fft_c2r(sp, y)
    torch.cuda.synchronize()
    sp_eager, y_eager = sp.clone(), y.clone()
    e.capture_begin()
    e.fft_r2c(x, sp)
    e.fft_c2r(sp, y)
    e.capture_end()
    for rep in range(4):
        sp.zero_()
        y.zero_()
        junk = torch.randn(1000, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        e.graph_launch()
        torch.cuda.synchronize()
        assert torch.equal(sp, sp_eager), rep
        assert torch.equal(y, y_eager), rep
        assert float(junk) == float(junk)
    err = float(torch.linalg.vector_norm(y / (nx * ny * nz) - x) / torch.linalg.vector_norm(x))
    assert err < (1e-13 if prec == 64 else 5e-6)


def test_c2c_graph_and_refusals(d2d, torch_gpu):
    torch = torch_gpu
    shape = (48, 32, 20)
    e = d2d.FFTEngine(1, *shape)
    a = torch.randn(int(np.prod(shape)), dtype=torch.complex128, device="cuda")
    b = torch.empty_like(a)
    with pytest.raises(d2d.D2dError):
        e.graph_launch()  # nothing captured yet
    e.fft_c2c(a, b, -1)
    e.fft_c2c(b, b, 1)
    torch.cuda.synchronize()
    eager = b.clone()
    e.capture_begin()
    with pytest.raises(d2d.D2dError):
        e.capture_begin()  # already capturing
    one = d2d.decomp_info_init(25, 32, 20)
    with pytest.raises(d2d.D2dError):
        e.fft_r2c_pack(a, b, one)  # fused calls are not captured
    e.fft_c2c(a, b, -1)
    e.fft_c2c(b, b, 1)
    e.capture_end()
    for rep in range(3):
        b.zero_()
        torch.cuda.synchronize()
        e.graph_launch()
        torch.cuda.synchronize()
        assert torch.equal(b, eager), rep
    # eager calls still work after a capture
    b.zero_()
    e.fft_c2c(a, b, -1)
    e.fft_c2c(b, b, 1)
    torch.cuda.synchronize()
    assert torch.equal(b, eager)


def test_reconfiguring_drops_the_graph(d2d, torch_gpu):
    """A captured graph holds the kernels and chunk offsets it was recorded
    with: changing a stage kernel or the chunking is refused while capturing
    and drops the graph afterwards, so a replay cannot run a stale
    configuration; after a new capture the replay equals the eager calls."""
    torch = torch_gpu
    nx, ny, nz = 64, 48, 40
    e = d2d.FFTEngine(1, nx, ny, nz)
    x = torch.rand(nx * ny * nz, dtype=torch.float64, device="cuda")
    sp = torch.empty((nx // 2 + 1) * ny * nz, dtype=torch.complex128, device="cuda")
    y = torch.empty_like(x)

    def capture():
        e.capture_begin()
        e.fft_r2c(x, sp)
        e.fft_c2r(sp, y)
        e.capture_end()

    capture()
    e.graph_launch()
    torch.cuda.synchronize()
    cands = e.stage_candidates(4)
    assert len(cands) >= 2
    e.capture_begin()
    with pytest.raises(d2d.D2dError):
        e.set_stage_kernel(4, cands[-1])
    with pytest.raises(d2d.D2dError):
        e.set_chunking(8, 8)
    e.fft_r2c(x, sp)
    e.capture_end()
    e.set_stage_kernel(4, "general")
    with pytest.raises(d2d.D2dError):
        e.graph_launch()  # dropped by the kernel change
    capture()
    e.set_chunking(8, 8)
    with pytest.raises(d2d.D2dError):
        e.graph_launch()  # dropped by the chunking change
    e.fft_r2c(x, sp)
    e.fft_c2r(sp, y)
    torch.cuda.synchronize()
    sp_eager, y_eager = sp.clone(), y.clone()
    capture()
    for rep in range(3):
        sp.zero_()
        y.zero_()
        torch.cuda.synchronize()
        e.graph_launch()
        torch.cuda.synchronize()
        assert torch.equal(sp, sp_eager), rep
        assert torch.equal(y, y_eager), rep
