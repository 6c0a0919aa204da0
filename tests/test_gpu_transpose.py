"""Bit-exact parity of the gfx950 transposes and alltoallv packs with the
oracle's restatement of mem_split_* / mem_merge_* (src/transpose_*.f90)."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = {0: np.float64, 1: np.complex128, 2: np.float32, 3: np.complex64}
CASES = [(17, 13, 11, 2, 2), (17, 13, 11, 3, 2), (33, 15, 11, 4, 3), (16, 12, 10, 2, 2),
         (9, 8, 7, 1, 3), (9, 8, 7, 3, 1), (64, 32, 16, 4, 4),
         # rows of more than 1024 16-byte vectors (chunked rows on the vector path)
         (4096, 6, 4, 2, 2), (2050, 6, 4, 2, 2)]


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).to("cuda")


def rand_pencil(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, shape)
    if np.issubdtype(dtype, np.complexfloating):
        a = a + 1j * rng.uniform(-1, 1, shape)
    return np.asfortranarray(a.astype(dtype))


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("direction", [0, 1, 2, 3])
@pytest.mark.parametrize("elem", [0, 1, 2, 3])
@pytest.mark.parametrize("even", [False, True])
def test_pack_unpack_bit_exact(d2d, torch_gpu, oracle, case, direction, elem, even):
    nx, ny, nz, pr, pc = case
    dt = DT[elem]
    src_p = {0: "x", 1: "y", 2: "z", 3: "y"}[direction]
    dst_p = {0: "y", 1: "z", 2: "y", 3: "x"}[direction]
    for r, c in itertools.product(range(pr), range(pc)):
        dl = d2d.decomp_info_init(nx, ny, nz, pr, pc, r, c)
        do = oracle.decomp_info(nx, ny, nz, pr, pc, r, c)
        src = rand_pencil(tuple(getattr(do, src_p + "sz")), dt, 100 * r + c)
        want = oracle.mem_split(direction, src, do, even)
        nbuf = d2d.pack_buffer_elems(direction, dl, even)
        assert nbuf >= src.size
        send = torch_gpu.zeros(max(nbuf, want.size), dtype=to_dev(torch_gpu, src[:1, :1, :1]).dtype,
                               device="cuda")
        d2d.pack(direction, elem, to_dev(torch_gpu, src), send, dl, even)
        got = send.cpu().numpy()[:want.size]
        # padded EVEN slots are not written by either side; compare the data
        np.testing.assert_array_equal(got.view(np.uint8), want.view(np.uint8))
        # unpack of a receive buffer laid out for this rank's destination pencil
        dshape = tuple(getattr(do, dst_p + "sz"))
        nrecv = max(int(np.prod(dshape)), d2d.pack_buffer_elems(direction, dl, even))
        recv = np.asarray(rand_pencil((nrecv, 1, 1), dt, 7 + r).ravel())
        want_u = oracle.mem_merge(direction, recv, dshape, dt, do, even)
        dst = torch_gpu.zeros(int(np.prod(dshape)), dtype=send.dtype, device="cuda")
        d2d.unpack(direction, elem, to_dev(torch_gpu, recv), dst, dl, even)
        got_u = dst.cpu().numpy().reshape(dshape, order="F")
        np.testing.assert_array_equal(got_u.ravel(order="F").view(np.uint8),
                                      want_u.ravel(order="F").view(np.uint8))


def test_one_rank_transposes_are_copies(d2d, torch_gpu):
    """dims = (1,1): transpose_* is dst = src (src/transpose_x_to_y.f90:41-50)."""
    d = d2d.decomp_info_init(17, 13, 11)
    for elem, dt in DT.items():
        x = rand_pencil((17, 13, 11), dt, elem)
        a = to_dev(torch_gpu, x)
        for fn in (d2d.transpose_x_to_y, d2d.transpose_y_to_z, d2d.transpose_z_to_y,
                   d2d.transpose_y_to_x):
            b = torch_gpu.zeros_like(a)
            fn(a, b, d, elem=elem)
            torch_gpu.cuda.synchronize()
            assert torch_gpu.equal(a, b)


def test_multi_rank_transpose_is_unsupported(d2d, torch_gpu):
    d = d2d.decomp_info_init(17, 13, 11, 2, 2)
    a = torch_gpu.zeros(10, dtype=torch_gpu.float64, device="cuda")
    with pytest.raises(d2d.D2dError) as e:
        d2d.transpose_x_to_y(a, a, d, elem=d2d.R8)
    assert e.value.status == d2d.ERR_UNSUPPORTED


def test_config3b_pack_2048_cube_4x4(d2d, torch_gpu, oracle):
    """One rank of a 2048^3 grid at p_row x p_col = 4 x 4: its x-pencil
    2048x512x512 real8 (uniform[-1,1), seed 2) packed in mem_split_xy order,
    and the 512x2048x512 y-pencil packed in mem_split_yz order; bit-exact."""
    n = 2048
    dl = d2d.decomp_info_init(n, n, n, 4, 4, 0, 0)
    do = oracle.decomp_info(n, n, n, 4, 4, 0, 0)
    assert tuple(dl.xsz) == (2048, 512, 512)
    nel = 2048 * 512 * 512
    x = torch_gpu.empty(nel, dtype=torch_gpu.float64, device="cuda")
    d2d.fill_uniform(x, nel, d2d.R8, 2)
    send = torch_gpu.empty_like(x)
    d2d.pack(d2d.X_TO_Y, d2d.R8, x, send, dl)
    torch_gpu.cuda.synchronize()
    xh = x.cpu().numpy().reshape((2048, 512, 512), order="F")
    want = oracle.mem_split(oracle.XY, xh, do)
    assert np.array_equal(send.cpu().numpy(), want)
    del want
    # the y-pencil of the same rank for the second exchange
    assert tuple(dl.ysz) == (512, 2048, 512)
    send2 = torch_gpu.empty_like(x)
    d2d.pack(d2d.Y_TO_Z, d2d.R8, x, send2, dl)
    torch_gpu.cuda.synchronize()
    yh = x.cpu().numpy().reshape((512, 2048, 512), order="F")
    assert np.array_equal(send2.cpu().numpy(), oracle.mem_split(oracle.YZ, yh, do))


@pytest.mark.parametrize("geo", [(37, 3, 37 * 5, 37, 5), (100, 1, 0, 100, 1030), (16, 4, 16 * 70, 16, 70)])
@pytest.mark.parametrize("elem", ["C16", "C8"])
def test_column_copy_pattern(d2d, torch_gpu, geo, elem):
    """The bandwidth reference copies exactly the columns it is given (bit
    for bit) and touches nothing else, for partial 16-column tiles and
    columns shorter and longer than one thread's share."""
    torch = torch_gpu
    C0, C1, s1, sj, n = geo
    total = C0 - 1 + (C1 - 1) * s1 + (n - 1) * sj + 1
    cdt = torch.complex128 if elem == "C16" else torch.complex64
    src = torch.randn(total + 7, dtype=cdt, device="cuda")
    dst = torch.full_like(src, complex(-7.0, 3.0))
    d2d.column_copy(src, dst, getattr(d2d, elem), C0, C1, s1, sj, n)
    torch.cuda.synchronize()
    idx = np.array([c0 + c1 * s1 + j * sj for c0 in range(C0) for c1 in range(C1) for j in range(n)])
    mask = np.zeros(total + 7, dtype=bool)
    mask[idx] = True
    s, d = src.cpu().numpy(), dst.cpu().numpy()
    assert np.array_equal(d[mask], s[mask])
    assert np.all(d[~mask] == complex(-7.0, 3.0))
