"""Decomposition metadata and the alltoallv pack/unpack maps (CPU).

The oracle's maps (oracle/decomp.c, restating src/transpose_*.f90
mem_split_* / mem_merge_* and src/decomp_2d.f90:1016-1206) are pinned by
the reference's own known-answer tests: the integer index field of
examples/test2d/test2d.f90:77 must land exactly on every rank's target
pencil after a simulated multi-rank split -> alltoallv -> merge, and the
partition sizes must add up as examples/init_test/init_test.f90:84-112
checks.  The library's host-side metadata is then compared with the oracle.
"""
import itertools

import numpy as np
import pytest

CASES = [
    (17, 13, 11, 1, 1), (17, 13, 11, 2, 2), (17, 13, 11, 4, 1), (17, 13, 11, 1, 4),
    (17, 13, 11, 3, 2), (16, 12, 10, 2, 2), (9, 8, 7, 3, 3), (33, 15, 11, 4, 3),
    (8, 8, 8, 4, 2), (12, 12, 12, 2, 4), (32, 16, 8, 4, 4),
]


def test_distribute_reference_example(oracle):
    # "17 meshes across 4 processor would be distibuted as (4,4,4,5)"
    # (src/decomp_2d.f90:1104-1106)
    st, en, sz = oracle.distribute(17, 4)
    assert list(sz) == [4, 4, 4, 5]
    assert list(st) == [1, 5, 9, 13] and list(en) == [4, 8, 12, 17]


@pytest.mark.parametrize("case", CASES)
def test_partition_sums(oracle, case):
    """init_test: the pencils of all ranks tile the global grid."""
    nx, ny, nz, pr, pc = case
    tot = nx * ny * nz
    sx = sy = sz = 0
    for r, c in itertools.product(range(pr), range(pc)):
        d = oracle.decomp_info(nx, ny, nz, pr, pc, r, c)
        sx += np.prod(list(d.xsz))
        sy += np.prod(list(d.ysz))
        sz += np.prod(list(d.zsz))
        for st, en, s in ((d.xst, d.xen, d.xsz), (d.yst, d.yen, d.ysz), (d.zst, d.zen, d.zsz)):
            assert all(en[i] - st[i] + 1 == s[i] for i in range(3))
    assert sx == sy == sz == tot


def index_field(nx, ny, nz):
    i = np.arange(1, nx + 1)[:, None, None]
    j = np.arange(1, ny + 1)[None, :, None]
    k = np.arange(1, nz + 1)[None, None, :]
    return np.asfortranarray((i + (j - 1) * nx + (k - 1) * nx * ny).astype(np.float64))


def sub(u, st, en):
    return np.asfortranarray(u[st[0] - 1:en[0], st[1] - 1:en[1], st[2] - 1:en[2]])


def simulate(oracle, direction, nx, ny, nz, pr, pc, even, dtype=np.float64):
    """Every rank packs, the alltoallv is done by hand, every rank unpacks."""
    u = index_field(nx, ny, nz).astype(dtype)
    if np.iscomplexobj(u):
        u = u + 1j * (u - 1)  # timing2d_complex's (m, m-1) field
    D = {(r, c): oracle.decomp_info(nx, ny, nz, pr, pc, r, c)
         for r, c in itertools.product(range(pr), range(pc))}
    src_of = {oracle.XY: "x", oracle.YZ: "y", oracle.ZY: "z", oracle.YX: "y"}[direction]
    dst_of = {oracle.XY: "y", oracle.YZ: "z", oracle.ZY: "y", oracle.YX: "x"}[direction]
    send = {}
    for rc, d in D.items():
        src = sub(u, getattr(d, src_of + "st"), getattr(d, src_of + "en"))
        send[rc] = oracle.mem_split(direction, src, d, even)
    xy = direction in (oracle.XY, oracle.YX)
    # send counts / displacements on each side, per the reference's calls
    names = {oracle.XY: ("x1", "y1"), oracle.YX: ("y1", "x1"),
             oracle.YZ: ("y2", "z2"), oracle.ZY: ("z2", "y2")}[direction]
    for rc, d in D.items():
        r, c = rc
        peers = [(m, c) for m in range(pr)] if xy else [(r, m) for m in range(pc)]
        dst_shape = tuple(getattr(d, dst_of + "sz"))
        nbuf = max(int(np.prod(dst_shape)),
                   (d.x1count * pr if xy else d.y2count * pc) if even else 0)
        recv = np.zeros(nbuf, dtype=u.dtype)
        me = peers.index(rc)
        for m, peer in enumerate(peers):
            dp = D[peer]
            if even:
                cnt = getattr(dp, names[0] + "count")
                s_off, r_off = me * cnt, m * getattr(d, names[1] + "count")
            else:
                cnt = getattr(dp, names[0] + "cnts")[me]
                s_off = getattr(dp, names[0] + "disp")[me]
                r_off = getattr(d, names[1] + "disp")[m]
                assert cnt == getattr(d, names[1] + "cnts")[m]
            recv[r_off:r_off + cnt] = send[peer][s_off:s_off + cnt]
        got = oracle.mem_merge(direction, recv, dst_shape, u.dtype, d, even)
        want = sub(u, getattr(d, dst_of + "st"), getattr(d, dst_of + "en"))
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("direction", [0, 1, 2, 3])
def test_simulated_transpose_index_field(oracle, case, direction):
    simulate(oracle, direction, *case, even=False)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("direction", [0, 1, 2, 3])
def test_simulated_transpose_index_field_even(oracle, case, direction):
    simulate(oracle, direction, *case, even=True)


def test_simulated_transpose_complex(oracle):
    simulate(oracle, oracle.XY, 17, 13, 11, 3, 2, False, dtype=np.complex128)
    simulate(oracle, oracle.YZ, 17, 13, 11, 3, 2, False, dtype=np.complex128)


def test_planted_map_mistake_is_caught(oracle):
    """A wrong block order in the send buffer must break the index-field check."""
    d = oracle.decomp_info(17, 13, 11, 2, 2, 0, 0)
    u = index_field(17, 13, 11)
    src = sub(u, d.xst, d.xen)
    buf = oracle.mem_split(oracle.XY, src, d)
    n0 = d.x1cnts[0]
    swapped = np.concatenate([buf[n0:], buf[:n0]])
    got = oracle.mem_merge(oracle.YX, swapped, tuple(d.xsz), np.float64, d)
    assert not np.array_equal(got, src)
    got_ok = oracle.mem_merge(oracle.YX, buf, tuple(d.xsz), np.float64, d)
    # split_xy and merge_yx are inverse maps on one rank
    np.testing.assert_array_equal(got_ok, src)


FIELDS = ["xst", "xen", "xsz", "yst", "yen", "ysz", "zst", "zen", "zsz"]


@pytest.mark.parametrize("case", CASES + [(2048, 2048, 2048, 4, 4), (1025, 2048, 2048, 4, 4)])
def test_library_decomp_matches_oracle(d2d, oracle, case):
    nx, ny, nz, pr, pc = case
    for r, c in itertools.product(range(pr), range(pc)):
        a = d2d.decomp_info_init(nx, ny, nz, pr, pc, r, c)
        b = oracle.decomp_info(nx, ny, nz, pr, pc, r, c)
        for f in FIELDS:
            assert list(getattr(a, f)) == list(getattr(b, f)), f
        for f, n in (("x1dist", pr), ("y1dist", pr), ("y2dist", pc), ("z2dist", pc),
                     ("x1cnts", pr), ("y1cnts", pr), ("x1disp", pr), ("y1disp", pr),
                     ("y2cnts", pc), ("z2cnts", pc), ("y2disp", pc), ("z2disp", pc)):
            assert list(getattr(a, f))[:n] == list(getattr(b, f))[:n], f
        for f in ("x1count", "y1count", "y2count", "z2count", "even"):
            assert getattr(a, f) == getattr(b, f), f


def test_spectral_dist_example(d2d):
    """1025 over 4 = (256,256,256,257): the larger block goes last."""
    d = d2d.decomp_info_init(1025, 2048, 2048, 4, 4, 3, 0)
    assert list(d.x1dist)[:4] == [256, 256, 256, 257]
    assert d.ysz[0] == 257


def test_grid_too_small(d2d):
    with pytest.raises(d2d.D2dError) as e:
        d2d.decomp_info_init(3, 8, 8, 4, 2)
    assert e.value.status == d2d.ERR_GRID


@pytest.mark.parametrize("nproc,want", [(1, (1, 1)), (2, (1, 2)), (4, (2, 2)), (7, (1, 7)),
                                        (8, (2, 4)), (12, (3, 4)), (16, (4, 4)), (36, (6, 6)),
                                        (48, (6, 8))])
def test_best_2d_grid(d2d, nproc, want):
    """best_2d_grid (src/decomp_2d_init_fin.f90:270-300): p_col is the factor at
    position nfact/2+1 of the sorted factor list."""
    facs = [f for f in range(1, nproc + 1) if nproc % f == 0]
    col = facs[len(facs) // 2]
    assert (nproc // col, col) == want
    assert d2d.best_2d_grid(nproc) == want
