"""The pack / unpack order pinned a second way: the reference states every
mem_split_* / mem_merge_* map twice, as the OpenMP loop that the oracle
restates and, in its GPU build, as one cudaMemcpy2D per block.  Here the
cudaMemcpy2D form is restated in numpy (row r of `height` rows copies
`width` elements from src[soff + r*spitch] to dst[doff + r*dpitch], all in
elements, Fortran order) and the oracle's maps must equal it byte for byte
for all eight maps, alltoallv and EVEN offsets, real and complex data.

The calls restated, per block m with i1..i2 = the block's range along the
split axis, w = i2 - i1 + 1 and off = init_pos - 1:
  split_xy  src/transpose_x_to_y.f90:300-306  (out+off, w, in(i1,1,1), n1, w, n2*n3)
  merge_xy  src/transpose_x_to_y.f90:418-424  (out(1,i1,1), n1*n2, in+off, n1*w, n1*w, n3)
  split_yz  src/transpose_y_to_z.f90:404-410  (out+off, n1*w, in(1,i1,1), n1*n2, n1*w, n3)
  merge_yz  src/transpose_y_to_z.f90:524-530  (out(1,1,i1), n1, in+off, n1, n1, n2*w)
  split_zy  src/transpose_z_to_y.f90:404-410  (out+off, n1, in(1,1,i1), n1, n1, n2*w)
  merge_zy  src/transpose_z_to_y.f90:522-528  (out(1,i1,1), n1*n2, in+off, n1*w, n1*w, n3)
  split_yx  src/transpose_y_to_x.f90:300-306  (out+off, n1*w, in(1,i1,1), n1*n2, n1*w, n3)
  merge_yx  src/transpose_y_to_x.f90:418-424  (out(i1,1,1), n1, in+off, w, w, n2*n3)
with the block sizes and offsets each call site passes (x1dist/x1disp for
split_xy and merge_yx, y1dist/y1disp for merge_xy and split_yx,
y2dist/y2disp for split_yz and merge_zy, z2dist/z2disp for merge_yz and
split_zy; m*count under EVEN).
"""
import itertools

import numpy as np
import pytest

XY, YZ, ZY, YX = 0, 1, 2, 3


def memcpy2d(dst, doff, dpitch, src, soff, spitch, width, height):
    """cudaMemcpy2D in elements on flat arrays."""
    if width <= 0 or height <= 0:
        return
    r = np.arange(height)[:, None]
    c = np.arange(width)[None, :]
    dst[doff + r * dpitch + c] = src[soff + r * spitch + c]


def _blocks(dist, p):
    out, i1 = [], 1
    for m in range(p):
        out.append((m, i1, i1 + dist[m] - 1))
        i1 += dist[m]
    return out


def _side(d, direction, split):
    """(dist, disp, count, nproc) that the reference passes for this map."""
    key = {(XY, True): "x1", (XY, False): "y1", (YX, True): "y1", (YX, False): "x1",
           (YZ, True): "y2", (YZ, False): "z2", (ZY, True): "z2", (ZY, False): "y2"}[
        (direction, split)]
    p = d.p_row if direction in (XY, YX) else d.p_col
    return (list(getattr(d, key + "dist"))[:p], list(getattr(d, key + "disp"))[:p],
            int(getattr(d, key + "count")), p)


def split_memcpy2d(direction, src, d, even, nbuf):
    n1, n2, n3 = src.shape
    a = src.ravel(order="F")
    out = np.zeros(nbuf, dtype=src.dtype)
    dist, disp, count, p = _side(d, direction, True)
    for m, i1, i2 in _blocks(dist, p):
        w = i2 - i1 + 1
        off = m * count if even else disp[m]
        if direction == XY:
            memcpy2d(out, off, w, a, i1 - 1, n1, w, n2 * n3)
        elif direction in (YZ, YX):
            memcpy2d(out, off, n1 * w, a, (i1 - 1) * n1, n1 * n2, n1 * w, n3)
        else:  # ZY
            memcpy2d(out, off, n1, a, (i1 - 1) * n1 * n2, n1, n1, n2 * w)
    return out


def merge_memcpy2d(direction, buf, dst_shape, d, even):
    n1, n2, n3 = dst_shape
    out = np.zeros(n1 * n2 * n3, dtype=buf.dtype)
    dist, disp, count, p = _side(d, direction, False)
    for m, i1, i2 in _blocks(dist, p):
        w = i2 - i1 + 1
        off = m * count if even else disp[m]
        if direction in (XY, ZY):
            memcpy2d(out, (i1 - 1) * n1, n1 * n2, buf, off, n1 * w, n1 * w, n3)
        elif direction == YZ:
            memcpy2d(out, (i1 - 1) * n1 * n2, n1, buf, off, n1, n1, n2 * w)
        else:  # YX
            memcpy2d(out, i1 - 1, n1, buf, off, w, w, n2 * n3)
    return out.reshape(dst_shape, order="F")


SRC = {XY: "x", YZ: "y", ZY: "z", YX: "y"}
DST = {XY: "y", YZ: "z", ZY: "y", YX: "x"}
GRIDS = [((17, 13, 11), (2, 3)), ((16, 12, 10), (4, 2)), ((9, 8, 7), (3, 3)),
         ((33, 15, 11), (1, 4)), ((32, 16, 8), (4, 1)), ((25, 19, 6), (5, 2))]


def _ranks(p_row, p_col):
    return list(itertools.product(range(p_row), range(p_col)))


@pytest.mark.parametrize("grid,pg", GRIDS)
@pytest.mark.parametrize("direction", [XY, YZ, ZY, YX])
@pytest.mark.parametrize("even", [False, True])
@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_oracle_maps_equal_the_memcpy2d_form(oracle, grid, pg, direction, even, dtype):
    nx, ny, nz = grid
    rng = np.random.default_rng(hash((grid, direction, even)) % 2 ** 32)
    for r, c in _ranks(*pg):
        d = oracle.decomp_info(nx, ny, nz, pg[0], pg[1], r, c)
        shape = tuple(getattr(d, SRC[direction] + "sz"))
        src = rng.standard_normal(shape)
        if dtype == np.complex128:
            src = src + 1j * rng.standard_normal(shape)
        src = np.asfortranarray(src.astype(dtype))
        got = oracle.mem_split(direction, src, d, even)
        want = split_memcpy2d(direction, src, d, even, got.size)
        assert np.array_equal(got, want), ("split", r, c)
        # unpack: the receive buffer is the pack of the matching shape
        dshape = tuple(getattr(d, DST[direction] + "sz"))
        buf = rng.standard_normal(got.size * 2 + 64).astype(dtype)
        got_m = oracle.mem_merge(direction, buf, dshape, dtype, d, even)
        want_m = merge_memcpy2d(direction, buf, dshape, d, even)
        assert np.array_equal(got_m, want_m), ("merge", r, c)


def test_memcpy2d_form_rejects_planted_slips(oracle):
    """The comparison has teeth: split_xy with the block offsets taken from
    the wrong side's displacements, or with the rows of one plane only
    (height n2 for n2*n3), differs from the oracle."""
    d = oracle.decomp_info(17, 13, 11, 2, 3, 1, 2)
    src = np.asfortranarray(np.random.default_rng(1).standard_normal(tuple(d.xsz)))
    got = oracle.mem_split(XY, src, d)
    n1, n2, n3 = src.shape
    a = src.ravel(order="F")
    dist, disp, count, p = _side(d, XY, True)
    ydisp = _side(d, XY, False)[1]
    assert disp != ydisp
    for slip in ("disp", "height"):
        wrong = np.zeros_like(got)
        for m, i1, i2 in _blocks(dist, p):
            w = i2 - i1 + 1
            off = ydisp[m] if slip == "disp" else disp[m]
            h = n2 if slip == "height" else n2 * n3
            if off + (h - 1) * w + w <= wrong.size:
                memcpy2d(wrong, off, w, a, i1 - 1, n1, w, h)
        assert not np.array_equal(got, wrong), slip
