"""The x stage of a PHYSICAL_IN_X transform fused with the x->y pack
(r2c_1m_x then mem_split_xy, src/fft_common_3d.f90:133-140 and
src/transpose_x_to_y.f90:93-101) and the y->x unpack fused with the inverse
x stage (mem_merge_yx then c2r_1m_x, src/transpose_y_to_x.f90:131-133 and
src/fft_common_3d.f90:259-266), for one rank of a process grid.

The fused calls must equal the unfused sequence bit for bit: the x stage
written as the plain spectral x-pencil (a one-block decomposition) followed
by d2dhip_pack, and d2dhip_unpack followed by the x stage.  The plain x
stage itself is checked against numpy's rfft / irfft along x."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

# (global real grid, p_row, p_col, coord_row, coord_col): the rank's x-pencil
# is nx x (its share of ny) x (its share of nz)
CASES = [
    ((64, 48, 40), 2, 3, 1, 2),
    ((1024, 64, 32), 4, 4, 0, 3),
    ((100, 30, 20), 3, 2, 2, 1),
    ((17, 13, 11), 2, 2, 1, 1),
    ((48, 40, 36), 5, 1, 4, 0),
    ((2048, 32, 16), 4, 2, 3, 1),
]


@pytest.mark.parametrize("even", [False, True])
@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("prec", [64, 32])
def test_fused_pack_and_unpack(d2d, torch_gpu, prec, case, even):
    torch = torch_gpu
    (nxg, nyg, nzg), pr, pc, cr, cc = case
    sp = d2d.decomp_info_init(nxg // 2 + 1, nyg, nzg, pr, pc, cr, cc)
    nx, ny, nz = nxg, sp.xsz[1], sp.xsz[2]
    nxc = nx // 2 + 1
    one = d2d.decomp_info_init(nxc, ny, nz)  # p_row = 1: the plain spectral x-pencil
    rdt = torch.float64 if prec == 64 else torch.float32
    cdt = torch.complex128 if prec == 64 else torch.complex64
    elem = d2d.C16 if prec == 64 else d2d.C8
    tol = 1e-13 if prec == 64 else 5e-6
    e = d2d.FFTEngine(d2d.PHYSICAL_IN_X, nx, ny, nz, prec=prec)
    rng = np.random.default_rng(nx + 7 * ny + pr)
    x = np.asfortranarray(rng.uniform(-1, 1, (nx, ny, nz)))
    xin = torch.from_numpy(np.ascontiguousarray(x.ravel(order="F"))).to(rdt).cuda()

    # the plain x stage against numpy
    spec = torch.empty(nxc * ny * nz, dtype=cdt, device="cuda")
    e.fft_r2c_pack(xin, spec, one)
    torch.cuda.synchronize()
    got = spec.cpu().numpy().reshape((nxc, ny, nz), order="F")
    xr = xin.cpu().numpy().astype(np.float64).reshape((nx, ny, nz), order="F")
    assert rel_l2(got, np.fft.rfft(xr, axis=0)) < tol

    # fused pack == x stage then d2dhip_pack
    nbuf = d2d.pack_buffer_elems(d2d.X_TO_Y, sp, even)
    send = torch.zeros(nbuf, dtype=cdt, device="cuda")
    ref = torch.zeros(nbuf, dtype=cdt, device="cuda")
    e.fft_r2c_pack(xin, send, sp, even)
    d2d.pack(d2d.X_TO_Y, elem, spec, ref, sp, even)
    torch.cuda.synchronize()
    assert torch.equal(send, ref)

    # fused unpack == d2dhip_unpack then x stage (the buffer read as if received)
    recv = ref
    spec2 = torch.empty_like(spec)
    d2d.unpack(d2d.Y_TO_X, elem, recv, spec2, sp, even)
    out_ref = torch.empty(nx * ny * nz, dtype=rdt, device="cuda")
    out = torch.empty_like(out_ref)
    e.fft_c2r_unpack(spec2, out_ref, one)
    e.fft_c2r_unpack(recv, out, sp, even)
    torch.cuda.synchronize()
    assert torch.equal(out, out_ref)
    back = out_ref.cpu().numpy().astype(np.float64).reshape((nx, ny, nz), order="F")
    assert rel_l2(back / nx, xr) < tol


def test_fused_pack_rejects_wrong_shape(d2d, torch_gpu):
    torch = torch_gpu
    e = d2d.FFTEngine(d2d.PHYSICAL_IN_X, 32, 8, 8)
    bad = d2d.decomp_info_init(16, 8, 8, 2, 1, 0, 0)  # not 32/2+1 points along x
    a = torch.zeros(32 * 8 * 8, dtype=torch.float64, device="cuda")
    b = torch.zeros(4096, dtype=torch.complex128, device="cuda")
    with pytest.raises(d2d.D2dError):
        e.fft_r2c_pack(a, b, bad)
    z = d2d.FFTEngine(d2d.PHYSICAL_IN_Z, 32, 8, 8)
    good = d2d.decomp_info_init(17, 8, 8, 2, 1, 0, 0)
    with pytest.raises(d2d.D2dError):
        z.fft_r2c_pack(a, b, good)
