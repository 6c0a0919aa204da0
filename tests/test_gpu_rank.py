"""Plans of one rank of a process grid (d2dhip_plan_create_rank), their single
1D stages (d2dhip_exec_1m: c2c_1m_*, r2c_1m_*, c2r_1m_*, src/fft_generic.f90:112-384)
and the y and z stages fused with their exchanges' pack or unpack
(d2dhip_exec_1m_pack / _unpack).

- Every stage of every rank of a 2 x 3 grid against numpy along its axis.
- The distributed r2c / c2r of src/fft_common_3d.f90:121-296 run rank by rank
  on one GPU, the alltoallv between stages done on the host with the
  reference's counts and displacements (x1cnts/x1disp ... z2cnts/z2disp,
  src/transpose_x_to_y.f90:93-133 and the other three files): the gathered
  spectrum equals numpy's rfftn and the round trip returns N times the input.
- At the 1024^3 rank shape of a 4 x 2 grid: the fused y and z stages equal
  the stage followed by d2dhip_pack (d2dhip_unpack followed by the stage, and
  for the y stage unpack, stage and pack) bit for bit, for every candidate
  kernel of the stage (the fused forms of the specialised kernels, the general
  kernel, and the unfused fallback).
"""
import itertools

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

F64_TOL = 1e-13
PAIR = {0: ("x1", "y1", "col"), 3: ("y1", "x1", "col"), 1: ("y2", "z2", "row"),
        2: ("z2", "y2", "row")}


def ranks(pr, pc):
    return list(itertools.product(range(pr), range(pc)))


def sub(u, st, en):
    return np.asfortranarray(u[st[0] - 1:en[0], st[1] - 1:en[1], st[2] - 1:en[2]])


def put(u, st, en, v):
    u[st[0] - 1:en[0], st[1] - 1:en[1], st[2] - 1:en[2]] = v


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).to("cuda")


def host(t, shape):
    return t.cpu().numpy().reshape(shape, order="F")


def alltoallv(direction, send, D, pr, pc):
    """MPI_ALLTOALLV over the column (x<->y) or row (y<->z) communicator with
    the counts and displacements the reference passes."""
    s_name, r_name, comm = PAIR[direction]
    recv = {}
    for (r, c), d in D.items():
        peers = [(m, c) for m in range(pr)] if comm == "col" else [(r, m) for m in range(pc)]
        me = peers.index((r, c))
        total = sum(getattr(d, r_name + "cnts")[m] for m in range(len(peers)))
        buf = np.zeros(total, dtype=np.complex128)
        for m, peer in enumerate(peers):
            dp = D[peer]
            cnt = getattr(dp, s_name + "cnts")[me]
            assert cnt == getattr(d, r_name + "cnts")[m]
            s_off = getattr(dp, s_name + "disp")[me]
            r_off = getattr(d, r_name + "disp")[m]
            buf[r_off:r_off + cnt] = send[peer][s_off:s_off + cnt]
        recv[(r, c)] = buf
    return recv


@pytest.mark.parametrize("fmt", [1, 3])
def test_rank_stages_against_numpy(d2d, torch_gpu, fmt):
    torch = torch_gpu
    nx, ny, nz, pr, pc = 30, 24, 20, 2, 3
    rng = np.random.default_rng(fmt)
    for r, c in ranks(pr, pc):
        e = d2d.FFTEngine(fmt, nx, ny, nz, rank=(pr, pc, r, c))
        ph, sp = e.decomps()
        assert (ph.p_row, ph.p_col) == (pr, pc)
        with pytest.raises(d2d.D2dError):  # 3D calls need the exchange
            e.fft_c2c(torch.zeros(8, dtype=torch.complex128, device="cuda"),
                      torch.zeros(8, dtype=torch.complex128, device="cuda"), -1)
        for stage, axis, dec, pen in (("c2c_x", 0, ph, "x"), ("c2c_y", 1, ph, "y"),
                                      ("c2c_z", 2, ph, "z"), ("spec_y", 1, sp, "y"),
                                      ("spec_last", 2 if fmt == 1 else 0, sp,
                                       "z" if fmt == 1 else "x")):
            shape = tuple(getattr(dec, pen + "sz"))
            x = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
            a = dev(torch, x)
            b = torch.empty_like(a)
            e.exec_1m(stage, a, b, d2d.FORWARD)
            assert rel_l2(host(b, shape), np.fft.fft(x, axis=axis)) < F64_TOL, (stage, r, c)
            e.exec_1m(stage, b, b, d2d.BACKWARD)  # in place
            assert rel_l2(host(b, shape) / shape[axis], x) < F64_TOL, (stage, r, c)
        # r2c / c2r along x (X) or z (Z): physical pencil -> spectral pencil
        axis = 0 if fmt == 1 else 2
        pshape = tuple(ph.xsz) if fmt == 1 else tuple(ph.zsz)
        sshape = tuple(sp.xsz) if fmt == 1 else tuple(sp.zsz)
        xr = rng.uniform(-1, 1, pshape)
        a = dev(torch, xr)
        s = torch.empty(int(np.prod(sshape)), dtype=torch.complex128, device="cuda")
        e.exec_1m("r2c", a, s, d2d.FORWARD)
        assert rel_l2(host(s, sshape), np.fft.rfft(xr, axis=axis)) < F64_TOL
        back = torch.empty_like(a)
        e.exec_1m("c2r", s, back, d2d.BACKWARD)
        assert rel_l2(host(back, pshape) / pshape[axis], xr) < F64_TOL


@pytest.mark.parametrize("fmt,grid,pg", [(1, (30, 24, 20), (2, 3)), (3, (30, 24, 20), (3, 2)),
                                         (1, (64, 48, 40), (2, 2))])
def test_distributed_round_trip_on_one_gpu(d2d, torch_gpu, fmt, grid, pg):
    """The reference's distributed r2c and c2r, every rank in turn on this GPU."""
    torch = torch_gpu
    nx, ny, nz = grid
    pr, pc = pg
    C16 = d2d.C16
    u = np.asfortranarray(np.random.default_rng(7).uniform(-1, 1, grid))
    E = {rc: d2d.FFTEngine(fmt, nx, ny, nz, rank=(pr, pc) + rc) for rc in ranks(pr, pc)}
    PH = {rc: e.decomps()[0] for rc, e in E.items()}
    SP = {rc: e.decomps()[1] for rc, e in E.items()}
    first, last = ("x", "z") if fmt == 1 else ("z", "x")
    XY, YZ, ZY, YX = d2d.X_TO_Y, d2d.Y_TO_Z, d2d.Z_TO_Y, d2d.Y_TO_X
    to_y, from_y = (XY, YZ) if fmt == 1 else (ZY, YX)  # exchanges before / after the y stage
    back_to_y, back_from_y = (ZY, YX) if fmt == 1 else (XY, YZ)

    def pen(d, p):
        return tuple(getattr(d, p + "sz"))

    # ---- forward: r2c on the first axis, exchange, y, exchange, last axis --
    send = {}
    for i, (rc, e) in enumerate(E.items()):
        d = SP[rc]
        x = dev(torch, sub(u, getattr(PH[rc], first + "st"), getattr(PH[rc], first + "en")))
        buf = torch.empty(d2d.pack_buffer_elems(to_y, d), dtype=torch.complex128, device="cuda")
        if fmt == 1 and i % 2 == 0:  # the x stage fused with its pack
            e.fft_r2c_pack(x, buf, d)
        else:
            s = torch.empty(int(np.prod(pen(d, first))), dtype=torch.complex128, device="cuda")
            e.exec_1m("r2c", x, s, d2d.FORWARD)
            d2d.pack(to_y, C16, s, buf, d)
        send[rc] = buf.cpu().numpy()
    recv = alltoallv(to_y, send, SP, pr, pc)
    send = {}
    for i, (rc, e) in enumerate(E.items()):
        d = SP[rc]
        y = torch.empty(int(np.prod(pen(d, "y"))), dtype=torch.complex128, device="cuda")
        buf = torch.empty(d2d.pack_buffer_elems(from_y, d), dtype=torch.complex128, device="cuda")
        r = torch.from_numpy(recv[rc]).to("cuda")
        if i % 3 == 2:  # both fused into the y stage
            e.exec_1m_unpack_pack("spec_y", r, buf, to_y, from_y, d, d2d.FORWARD)
        elif i % 2 == 0:  # unpack fused into the y stage, then the plain pack
            e.exec_1m_unpack("spec_y", r, y, to_y, d, d2d.FORWARD)
            d2d.pack(from_y, C16, y, buf, d)
        else:  # plain unpack, then the y stage fused with the pack
            d2d.unpack(to_y, C16, r, y, d)
            e.exec_1m_pack("spec_y", y, buf, from_y, d, d2d.FORWARD)
        send[rc] = buf.cpu().numpy()
    recv = alltoallv(from_y, send, SP, pr, pc)
    spec = np.zeros((nx // 2 + 1, ny, nz) if fmt == 1 else (nx, ny, nz // 2 + 1), np.complex128)
    lastpen = {}
    for i, (rc, e) in enumerate(E.items()):
        d = SP[rc]
        z = torch.empty(int(np.prod(pen(d, last))), dtype=torch.complex128, device="cuda")
        r = torch.from_numpy(recv[rc]).to("cuda")
        if fmt == 1 and i % 2:  # the unpack fused into the z stage
            e.exec_1m_unpack("spec_last", r, z, from_y, d, d2d.FORWARD)
        else:
            d2d.unpack(from_y, C16, r, z, d)
            e.exec_1m("spec_last", z, z, d2d.FORWARD)
        lastpen[rc] = z
        put(spec, getattr(d, last + "st"), getattr(d, last + "en"), host(z, pen(d, last)))
    want = np.fft.rfftn(u, axes=(2, 1, 0) if fmt == 1 else (0, 1, 2))
    assert rel_l2(spec, want) < F64_TOL

    # ---- backward: last axis, exchange, y, exchange, c2r ----------------------
    send = {}
    for i, (rc, e) in enumerate(E.items()):
        d, z = SP[rc], lastpen[rc]
        buf = torch.empty(d2d.pack_buffer_elems(back_to_y, d), dtype=torch.complex128, device="cuda")
        if fmt == 1 and i % 2 == 0:  # the z stage fused with its pack
            e.exec_1m_pack("spec_last", z, buf, back_to_y, d, d2d.BACKWARD)
        else:
            e.exec_1m("spec_last", z, z, d2d.BACKWARD)
            d2d.pack(back_to_y, C16, z, buf, d)
        send[rc] = buf.cpu().numpy()
    recv = alltoallv(back_to_y, send, SP, pr, pc)
    send = {}
    for i, (rc, e) in enumerate(E.items()):
        d = SP[rc]
        y = torch.empty(int(np.prod(pen(d, "y"))), dtype=torch.complex128, device="cuda")
        buf = torch.empty(d2d.pack_buffer_elems(back_from_y, d), dtype=torch.complex128,
                          device="cuda")
        r = torch.from_numpy(recv[rc]).to("cuda")
        if i % 3 == 1:
            e.exec_1m_unpack_pack("spec_y", r, buf, back_to_y, back_from_y, d, d2d.BACKWARD)
        elif i % 2:
            e.exec_1m_unpack("spec_y", r, y, back_to_y, d, d2d.BACKWARD)
            d2d.pack(back_from_y, C16, y, buf, d)
        else:
            d2d.unpack(back_to_y, C16, r, y, d)
            e.exec_1m_pack("spec_y", y, buf, back_from_y, d, d2d.BACKWARD)
        send[rc] = buf.cpu().numpy()
    recv = alltoallv(back_from_y, send, SP, pr, pc)
    out = np.zeros(grid)
    for i, (rc, e) in enumerate(E.items()):
        d, dp = SP[rc], PH[rc]
        r = torch.from_numpy(recv[rc]).to("cuda")
        xr = torch.empty(int(np.prod(pen(dp, first))), dtype=torch.float64, device="cuda")
        if fmt == 1 and i % 2:  # the unpack fused into the c2r x stage
            e.fft_c2r_unpack(r, xr, d)
        else:
            s = torch.empty(int(np.prod(pen(d, first))), dtype=torch.complex128, device="cuda")
            d2d.unpack(back_from_y, C16, r, s, d)
            e.exec_1m("c2r", s, xr, d2d.BACKWARD)
        put(out, getattr(dp, first + "st"), getattr(dp, first + "en"), host(xr, pen(dp, first)))
    assert rel_l2(out / u.size, u) < F64_TOL


@pytest.mark.parametrize("prec", [64, 32, 16])
@pytest.mark.parametrize("stage", ["spec_y", "spec_last"])
def test_fused_stage_bit_exact_at_1024_rank(d2d, torch_gpu, prec, stage):
    """Rank (1, 1) of 1024^3 on a 4 x 2 grid, spectral y-pencil 128 x 1024 x 512
    and z-pencil 128 x 512 x 1024: every candidate kernel of the y stage fused
    with the pack of y->z and y->x and the unpack of z->y and x->y, and of the
    z stage fused with the pack of z->y and the unpack of y->z, plain and EVEN
    offsets."""
    torch = torch_gpu
    n = 1024
    e = d2d.FFTEngine(1, n, n, n, prec=prec, rank=(4, 2, 1, 1))
    ph, sp = e.decomps()
    XY, YZ, ZY, YX = d2d.X_TO_Y, d2d.Y_TO_Z, d2d.Z_TO_Y, d2d.Y_TO_X
    # pencil, exchanges after the stage (pack), exchanges before it (unpack)
    p, packs, unpacks = {
        "spec_y": ("y", ((YZ, False, -1), (YX, True, 1)), ((ZY, False, 1), (XY, True, -1))),
        "spec_last": ("z", ((ZY, False, 1), (ZY, True, -1)), ((YZ, False, -1), (YZ, True, 1))),
    }[stage]
    shape = tuple(getattr(sp, p + "sz"))
    assert shape == ((128, 1024, 512) if p == "y" else (128, 512, 1024))
    cdt = torch.complex128 if prec == 64 else torch.complex64
    elem = d2d.C16 if prec == 64 else d2d.C8
    g = torch.Generator(device="cuda")
    g.manual_seed(prec)
    y = torch.randn(int(np.prod(shape)), dtype=cdt, device="cuda", generator=g)
    tmp = torch.empty_like(y)
    cands = e.stage_candidates(stage)
    assert "general" in cands and len(cands) >= 2
    for name in cands:
        e.set_stage_kernel(stage, name)
        for direction, even, isign in packs:
            nb = d2d.pack_buffer_elems(direction, sp, even)
            # the same sentinel in both: positions neither writes (EVEN padding)
            # must keep it
            fused = torch.full((nb,), 7.0, dtype=cdt, device="cuda")
            plain = torch.full((nb,), 7.0, dtype=cdt, device="cuda")
            e.exec_1m_pack(stage, y, fused, direction, sp, isign, even)
            e.exec_1m(stage, y, tmp, isign)
            d2d.pack(direction, elem, tmp, plain, sp, even)
            torch.cuda.synchronize()
            assert torch.equal(fused, plain), (name, direction, even)
        for direction, even, isign in unpacks:
            # a receive buffer holds the pencil (padded under EVEN)
            nb = max(d2d.pack_buffer_elems(direction, sp, even), int(np.prod(shape)))
            recv = torch.randn(nb, dtype=cdt, device="cuda", generator=g)
            fused = torch.empty_like(y)
            plain = torch.empty_like(y)
            e.exec_1m_unpack(stage, recv, fused, direction, sp, isign, even)
            d2d.unpack(direction, elem, recv, tmp, sp, even)
            e.exec_1m(stage, tmp, plain, isign)
            torch.cuda.synchronize()
            assert torch.equal(fused, plain), (name, direction, even)
        if stage != "spec_y":
            continue
        # both sides at once: the forward's x->y unpack and y->z pack, the
        # backward's z->y unpack and y->x pack
        for d_in, d_out, even, isign in ((XY, YZ, False, -1), (ZY, YX, True, 1)):
            nb = max(d2d.pack_buffer_elems(d_in, sp, even), int(np.prod(shape)))
            recv = torch.randn(nb, dtype=cdt, device="cuda", generator=g)
            ns = d2d.pack_buffer_elems(d_out, sp, even)
            fused = torch.full((ns,), 7.0, dtype=cdt, device="cuda")
            plain = torch.full((ns,), 7.0, dtype=cdt, device="cuda")
            e.exec_1m_unpack_pack(stage, recv, fused, d_in, d_out, sp, isign, even)
            d2d.unpack(d_in, elem, recv, tmp, sp, even)
            e.exec_1m(stage, tmp, tmp, isign)
            d2d.pack(d_out, elem, tmp, plain, sp, even)
            torch.cuda.synchronize()
            assert torch.equal(fused, plain), (name, d_in, d_out, even)


def test_fused_stage_refusals(d2d, torch_gpu):
    """The fused calls take only the maps that border their stage."""
    torch = torch_gpu
    e = d2d.FFTEngine(3, 30, 24, 20, rank=(2, 3, 1, 2))  # PHYSICAL_IN_Z: stage 5 is x
    ph, sp = e.decomps()
    a = torch.zeros(int(np.prod(sp.zsz)) + 64, dtype=torch.complex128, device="cuda")
    b = torch.zeros_like(a)
    with pytest.raises(d2d.D2dError):
        e.exec_1m_pack("spec_last", a, b, d2d.X_TO_Y, sp, d2d.FORWARD)
    e1 = d2d.FFTEngine(1, 30, 24, 20, rank=(2, 3, 1, 2))
    ph, sp = e1.decomps()
    with pytest.raises(d2d.D2dError):  # a z stage packs only for z->y
        e1.exec_1m_pack("spec_last", a, b, d2d.Y_TO_X, sp, d2d.FORWARD)
    with pytest.raises(d2d.D2dError):  # and unpacks only y->z
        e1.exec_1m_unpack("spec_last", a, b, d2d.Z_TO_Y, sp, d2d.FORWARD)
    with pytest.raises(d2d.D2dError):  # the x stages have their own fused calls
        e1.exec_1m_pack("c2c_x", a, b, d2d.X_TO_Y, ph, d2d.FORWARD)
