"""Pins the oracle's one-rank halo update (update_halo / halo_exchange,
src/halo.f90, src/halo_exchange_{x,y,z}_body.f90).

With every distributed axis periodic the halo array must be numpy's
wrap-around padding of the pencil (np.pad mode "wrap"), corners included;
a non-periodic axis must leave its halo cells exactly as they were; and the
reference example's use (examples/halo_test/halo_test.f90: a divergence by
central differences through the halo cells) must equal the same stencil
taken with periodic index arithmetic."""
import itertools

import numpy as np
import pytest

SHAPES = [(6, 5, 4), (1, 3, 7), (9, 2, 2)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("ipencil", [1, 2, 3])
@pytest.mark.parametrize("level", [1, 2])
def test_periodic_halo_is_wrap_padding(oracle, shape, ipencil, level):
    x = np.asfortranarray(np.random.default_rng(level + 3 * ipencil).standard_normal(shape))
    if min(s for a, s in enumerate(shape) if a != ipencil - 1) < level:
        pytest.skip("level above the pencil extent")
    out = oracle.update_halo(x, level, ipencil, (True, True, True))
    pad = [(level, level) if a != ipencil - 1 else (0, 0) for a in range(3)]
    assert np.array_equal(out, np.pad(x, pad, mode="wrap"))


@pytest.mark.parametrize("ipencil", [1, 2, 3])
def test_non_periodic_axes_keep_their_halo(oracle, ipencil):
    shape, level = (6, 5, 4), 1
    x = np.asfortranarray(np.random.default_rng(1).standard_normal(shape))
    hs = oracle.halo_shape(shape, level, ipencil)
    init = np.full(hs, -7.0, order="F")
    a1, a2 = oracle.HALO_ORDER[ipencil]
    for per in itertools.product([False, True], repeat=2):
        periodic = [False] * 3
        periodic[a1], periodic[a2] = per
        out = oracle.update_halo(x, level, ipencil, periodic, init)
        pad = [(level, level) if a != ipencil - 1 else (0, 0) for a in range(3)]
        wrapped = np.pad(x, pad, mode="wrap")
        interior = tuple(slice(level, -level) if a != ipencil - 1 else slice(None)
                         for a in range(3))
        assert np.array_equal(out[interior], x)
        for d, p in zip((a1, a2), per):
            lo = [slice(None)] * 3
            lo[d] = slice(0, level)
            # only the other axis's interior range: the corners depend on the order
            o = a2 if d == a1 else a1
            lo[o] = slice(level, -level)
            got = out[tuple(lo)]
            assert np.array_equal(got, wrapped[tuple(lo)] if p else init[tuple(lo)]), (d, p)
        if not any(per):
            halo = np.ones(hs, dtype=bool)
            halo[interior] = False
            assert np.array_equal(out[halo], init[halo])


def test_order_decides_the_corners(oracle):
    """y periodic, z not (x-pencil): the z halo keeps its cells; z periodic, y not:
    the z halo planes copy the y-halo rows too, which still hold their old cells."""
    shape, level = (3, 4, 5), 1
    x = np.asfortranarray(np.arange(60, dtype=np.float64).reshape(shape, order="F"))
    init = np.full(oracle.halo_shape(shape, level, 1), -1.0, order="F")
    out = oracle.update_halo(x, level, 1, (False, False, True), init)
    assert np.array_equal(out[:, 0, 0], init[:, 0, 0])          # corner: old y-halo cell
    assert np.array_equal(out[:, 1:-1, 0], x[:, :, -1])         # z halo wraps
    out = oracle.update_halo(x, level, 1, (False, True, False), init)
    assert np.array_equal(out[:, 0, 1:-1], x[:, -1, :])
    assert np.array_equal(out[:, 0, 0], init[:, 0, 0])          # z halo untouched


def test_halo_divergence_matches_periodic_stencil(oracle):
    """examples/halo_test/halo_test.f90:300-320 with periodic y and z."""
    rng = np.random.default_rng(2)
    shape = (8, 6, 5)
    u, v, w = (np.asfortranarray(rng.random(shape)) for _ in range(3))
    vh = oracle.update_halo(v, 1, 1, (False, True, True))
    wh = oracle.update_halo(w, 1, 1, (False, True, True))
    div = np.zeros(shape)
    div[1:-1] = ((u[2:] - u[:-2]) + (vh[1:-1, 2:, 1:-1] - vh[1:-1, :-2, 1:-1])
                 + (wh[1:-1, 1:-1, 2:] - wh[1:-1, 1:-1, :-2]))
    ref = np.zeros(shape)
    ref[1:-1] = ((u[2:] - u[:-2]) + (np.roll(v, -1, 1) - np.roll(v, 1, 1))[1:-1]
                 + (np.roll(w, -1, 2) - np.roll(w, 1, 2))[1:-1])
    assert np.array_equal(div, ref)
