"""Halo cells of one rank's pencil on the GPU (d2dhip_update_halo /
d2dhip_halo_exchange) against the oracle's restatement of update_halo and
halo_exchange (src/halo.f90, src/halo_exchange_{x,y,z}_body.f90), bit for
bit: every pencil orientation, levels 1-3, every periodic combination, all
four element kinds, with the halo array pre-filled so that cells the
exchange must not touch are checked too."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ELEMS = {"R8": np.float64, "C16": np.complex128, "R4": np.float32, "C8": np.complex64}


def _rand(rng, shape, dt):
    x = rng.standard_normal(shape)
    if np.dtype(dt).kind == "c":
        x = x + 1j * rng.standard_normal(shape)
    return np.asfortranarray(x.astype(dt))


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).cuda()


@pytest.mark.parametrize("ename", list(ELEMS))
@pytest.mark.parametrize("ipencil", [1, 2, 3])
@pytest.mark.parametrize("shape,level", [((17, 13, 11), 1), ((64, 48, 40), 2), ((9, 4, 3), 3)])
def test_update_halo_matches_oracle(d2d, oracle, torch_gpu, ename, ipencil, shape, level):
    torch = torch_gpu
    dt = ELEMS[ename]
    elem = getattr(d2d, ename)
    rng = np.random.default_rng(sum(shape) + 10 * ipencil + level)
    x = _rand(rng, shape, dt)
    hs = oracle.halo_shape(shape, level, ipencil)
    init = _rand(rng, hs, dt)
    xin = _dev(torch, x)
    for per in itertools.product([False, True], repeat=3):
        s = [shape[a] for a in range(3) if a != ipencil - 1 and per[a]]
        if s and min(s) < level:
            continue
        out = _dev(torch, init)
        d2d.update_halo(elem, xin, out, shape, level, ipencil, per)
        torch.cuda.synchronize()
        got = out.cpu().numpy().reshape(hs, order="F")
        want = oracle.update_halo(x, level, ipencil, per, init)
        assert np.array_equal(got, want), per


def test_halo_exchange_in_place_and_errors(d2d, oracle, torch_gpu):
    torch = torch_gpu
    shape, level, ipencil = (32, 24, 20), 2, 2
    rng = np.random.default_rng(5)
    hs = oracle.halo_shape(shape, level, ipencil)
    arr = _rand(rng, hs, np.float64)
    a = _dev(torch, arr)
    per = (True, False, True)
    d2d.halo_exchange(d2d.R8, a, shape, level, ipencil, per)
    torch.cuda.synchronize()
    want = oracle.halo_exchange(arr.copy(order="F"), level, ipencil, per)
    assert np.array_equal(a.cpu().numpy().reshape(hs, order="F"), want)
    with pytest.raises(d2d.D2dError):
        d2d.halo_exchange(d2d.R8, a, shape, level, 4, per)
    with pytest.raises(d2d.D2dError):  # level above the extent of a periodic axis
        d2d.halo_exchange(d2d.R8, a, (32, 24, 1), 2, 1, (False, False, True))


def test_halo_divergence_reference_example(d2d, torch_gpu):
    """examples/halo_test/halo_test.f90:280-330 at one rank with periodic y, z:
    div = du/dx + dv/dy + dw/dz by central differences through the halo cells
    of v and w (x-pencil), against the same stencil with periodic rolls."""
    torch = torch_gpu
    shape = (64, 48, 40)
    rng = np.random.default_rng(8)
    u, v, w = (np.asfortranarray(rng.random(shape)) for _ in range(3))
    hs = (64, 50, 42)
    vh = torch.zeros(int(np.prod(hs)), dtype=torch.float64, device="cuda")
    wh = torch.zeros_like(vh)
    per = (False, True, True)
    d2d.update_halo(d2d.R8, _dev(torch, v), vh, shape, 1, 1, per)
    d2d.update_halo(d2d.R8, _dev(torch, w), wh, shape, 1, 1, per)
    torch.cuda.synchronize()
    V = vh.cpu().numpy().reshape(hs, order="F")
    W = wh.cpu().numpy().reshape(hs, order="F")
    div = (u[2:] - u[:-2]) + (V[1:-1, 2:, 1:-1] - V[1:-1, :-2, 1:-1]) + \
        (W[1:-1, 1:-1, 2:] - W[1:-1, 1:-1, :-2])
    ref = (u[2:] - u[:-2]) + (np.roll(v, -1, 1) - np.roll(v, 1, 1))[1:-1] + \
        (np.roll(w, -1, 2) - np.roll(w, 1, 2))[1:-1]
    assert np.array_equal(div, ref)
