"""Many grids in one launch (the fft_multiple_grids use,
examples/fft_multiple_grids/fft_multiple_grids.f90), eager and captured in a
hipGraph, with each executor: one launch per stage phase and kernel (the
plans' own kernels: bit-identical to the plans' sequential decomp_2d_fft_3d
calls), one persistent launch on specialised one-wave tiles (bit-identical
to the plans on any specialised kernel of each length), and one persistent
launch on the general kernel's tiles (bit-identical to the plans on the
general kernel).  All match pocketfft to the FP64 / FP32 tolerances."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

SIZES = [32, 48, 64, 80, 96, 128, 160, 192, 256]


def grids(count, seed, pool):
    rng = np.random.default_rng(seed)
    return [tuple(int(v) for v in rng.choice(pool, 3)) for _ in range(count)]


def dev(torch, a, dt):
    return torch.from_numpy(np.ascontiguousarray(a.ravel(order="F")).astype(dt)).to("cuda")


MODES = [0, 1, 2]  # BATCH_PERSISTENT, BATCH_GROUPED, BATCH_PERSISTENT_GENERAL
C2C = ("c2c_x", "c2c_y", "c2c_z")


def match_plans(engines, b, mode):
    """Sets the plans' c2c stage kernels to what executor `mode` of batch b
    runs, so that their sequential calls are its bit-exact reference."""
    if mode == 1:
        return "plans"
    if mode == 0 and b.persistent_tiles():
        for e in engines:
            for st in C2C:
                e.set_stage_kernel(st, [n for n in e.stage_candidates(st) if n != "general"][0])
        return "specialised"
    for e in engines:
        for st in C2C:
            e.set_stage_kernel(st, "general")
    return "general"


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("prec", [64, 32])
def test_batch_matches_sequential_and_pocketfft(d2d, torch_gpu, prec, mode):
    torch = torch_gpu
    shapes = grids(6, 1, [8, 12, 16, 20, 24, 30, 32]) + [(17, 13, 11), (9, 8, 7)]
    npc = np.complex128 if prec == 64 else np.complex64
    tol = 1e-13 if prec == 64 else 5e-6
    rng = np.random.default_rng(2)
    xs = [np.asfortranarray(rng.uniform(-1, 1, s) + 1j * rng.uniform(-1, 1, s)) for s in shapes]
    engines = [d2d.FFTEngine(1 if i % 2 == 0 else 3, *s, prec=prec) for i, s in enumerate(shapes)]
    ins = [dev(torch, x, npc) for x in xs]
    outs = [torch.empty_like(a) for a in ins]
    seq = [torch.empty_like(a) for a in ins]
    b = d2d.Batch(engines)
    assert not b.persistent_tiles()  # lengths 8..30 have no one-wave tiles
    match_plans(engines, b, mode)
    for e, a, o in zip(engines, ins, seq):
        e.fft_c2c(a, o, -1)
    b.set_mode(mode)
    b.bind(ins, outs, mode=-1)
    b.launch()
    b.check()
    for x, o, q, s in zip(xs, outs, seq, shapes):
        assert torch.equal(o, q)
        got = o.cpu().numpy().reshape(s, order="F")
        assert rel_l2(got, np.fft.fftn(x.astype(npc).astype(np.complex128))) < tol
    # forward then backward in one launch: out = N * in
    b.bind(ins, outs, mode=0)
    b.launch()
    b.check()
    for x, o, s in zip(xs, outs, shapes):
        got = o.cpu().numpy().reshape(s, order="F") / np.prod(s)
        assert rel_l2(got, x.astype(npc)) < tol


@pytest.mark.parametrize("mode", MODES)
def test_batch_graph_replays(d2d, torch_gpu, mode):
    """A captured batch replayed several times, with eager work and
    synchronisation in between, gives the eager result every time."""
    torch = torch_gpu
    shapes = [(16, 12, 10), (32, 8, 6), (24, 20, 18)]
    rng = np.random.default_rng(5)
    xs = [np.asfortranarray(rng.uniform(-1, 1, s) + 1j * rng.uniform(-1, 1, s)) for s in shapes]
    engines = [d2d.FFTEngine(1, *s) for s in shapes]
    ins = [dev(torch, x, np.complex128) for x in xs]
    outs = [torch.empty_like(a) for a in ins]
    b = d2d.Batch(engines)
    b.set_mode(mode)
    b.bind(ins, outs, mode=0)
    b.launch()
    b.check()
    eager = [o.clone() for o in outs]
    b.capture_graph()
    for rep in range(4):
        for o in outs:
            o.zero_()
        junk = torch.randn(1000, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        b.graph_launch()
        b.check()
        for o, e in zip(outs, eager):
            assert torch.equal(o, e), rep
        assert float(junk) == float(junk)


def test_config5_64_uneven_grids(d2d, torch_gpu):
    """64 c2c FP64 grids, each dimension from {32,...,256} (seed 4), inputs
    uniform[-1,1): one launch for all forward+backward pairs, round trip to
    1/N, and the forward of the first grids against pocketfft."""
    torch = torch_gpu
    shapes = grids(64, 4, SIZES)
    engines = [d2d.FFTEngine(1, *s) for s in shapes]
    ins, outs = [], []
    for g, s in enumerate(shapes):
        a = torch.empty(int(np.prod(s)), dtype=torch.complex128, device="cuda")
        d2d.fill_uniform(a, a.numel(), d2d.C16, 4 + g)
        ins.append(a)
        outs.append(torch.empty_like(a))
    b = d2d.Batch(engines)
    b.bind(ins, outs, mode=0)
    b.launch()
    b.check()
    for a, o, s in zip(ins, outs, shapes):
        err = float(torch.linalg.vector_norm(o / np.prod(s) - a) / torch.linalg.vector_norm(a))
        assert err < 1e-13, s
    b.bind(ins, outs, mode=-1)
    b.launch()
    b.check()
    for g in range(3):
        x = ins[g].cpu().numpy().reshape(shapes[g], order="F")
        got = outs[g].cpu().numpy().reshape(shapes[g], order="F")
        assert rel_l2(got, np.fft.fftn(x)) < 1e-13
    # the per-phase executor runs the plans' own kernels: their bits
    seq = [torch.empty_like(a) for a in ins]
    for e, a, q in zip(engines, ins, seq):
        e.fft_c2c(a, q, -1)
    b.set_mode(1)
    b.launch()
    b.check()
    for o, q in zip(outs, seq):
        assert torch.equal(o, q)
    for o in outs:
        o.zero_()
    b.bind(ins, outs, mode=0)
    b.launch()
    b.check()
    for a, o, s in zip(ins, outs, shapes):
        err = float(torch.linalg.vector_norm(o / np.prod(s) - a) / torch.linalg.vector_norm(a))
        assert err < 1e-13, s
    # the single persistent launch runs specialised tiles for all 64 grids
    assert b.persistent_tiles()
    match_plans(engines, b, 0)
    for e, a, q in zip(engines, ins, seq):
        e.fft_c2c(a, q, -1)
    b.set_mode(0)
    b.bind(ins, outs, mode=-1)
    b.launch()
    b.check()
    for o, q in zip(outs, seq):
        assert torch.equal(o, q)


def test_config5_64_uneven_grids_fp32(d2d, torch_gpu):
    """The same 64 grids (seed 4) in FP32: the persistent launch (specialised
    tiles) round trips every grid to 1/N within the FP32 tolerance, its
    forward transforms are within it of FP64 pocketfft, and both executors
    equal the sequential plan calls on the kernels they run bit for bit."""
    torch = torch_gpu
    shapes = grids(64, 4, SIZES)
    engines = [d2d.FFTEngine(1, *s, prec=32) for s in shapes]
    ins, outs = [], []
    for g, s in enumerate(shapes):
        a = torch.empty(int(np.prod(s)), dtype=torch.complex64, device="cuda")
        d2d.fill_uniform(a, a.numel(), d2d.C8, 4 + g)
        ins.append(a)
        outs.append(torch.empty_like(a))
    b = d2d.Batch(engines)
    assert b.persistent_tiles()
    b.bind(ins, outs, mode=0)
    b.launch()
    b.check()
    for a, o, s in zip(ins, outs, shapes):
        err = float(torch.linalg.vector_norm(o / float(np.prod(s)) - a) / torch.linalg.vector_norm(a))
        assert err < 5e-6, s
    b.bind(ins, outs, mode=-1)
    b.launch()
    b.check()
    for g in (0, 21, 63):
        x = ins[g].cpu().numpy().reshape(shapes[g], order="F").astype(np.complex128)
        got = outs[g].cpu().numpy().reshape(shapes[g], order="F")
        assert rel_l2(got, np.fft.fftn(x)) < 5e-6, shapes[g]
    for mode in (1, 0):  # grouped: the plans' own kernels; persistent: specialised tiles
        match_plans(engines, b, mode)
        seq = [torch.empty_like(a) for a in ins]
        for e, a, q in zip(engines, ins, seq):
            e.fft_c2c(a, q, -1)
        b.set_mode(mode)
        b.bind(ins, outs, mode=-1)
        b.launch()
        b.check()
        for o, q, s in zip(outs, seq, shapes):
            assert torch.equal(o, q), (mode, s)


def test_default_executor_on_lengths_without_tiles(d2d, torch_gpu):
    """A batch with the reference's multi-grid shapes (fft_multiple_grids.f90:51-54:
    17x13x11, 17x15x27, 33x15x11) among specialised ones has no persistent
    tiles: the default executor then runs the plans' own kernels in grouped
    launches, bit-identical to the sequential calls, eager and replayed."""
    torch = torch_gpu
    shapes = [(17, 13, 11), (17, 15, 27), (33, 15, 11)] + grids(9, 5, SIZES)
    engines = [d2d.FFTEngine(1, *s) for s in shapes]
    ins, outs = [], []
    for g, s in enumerate(shapes):
        a = torch.empty(int(np.prod(s)), dtype=torch.complex128, device="cuda")
        d2d.fill_uniform(a, a.numel(), d2d.C16, 60 + g)
        ins.append(a)
        outs.append(torch.empty_like(a))
    seq = [torch.empty_like(a) for a in ins]
    for e, a, q in zip(engines, ins, seq):
        e.fft_c2c(a, q, -1)
        e.fft_c2c(q, q, 1)
    b = d2d.Batch(engines)
    assert not b.persistent_tiles()
    b.bind(ins, outs, mode=0)
    b.launch()
    b.check()
    for o, q in zip(outs, seq):
        assert torch.equal(o, q)
    b.capture_graph()
    for rep in range(3):
        for o in outs:
            o.zero_()
        torch.cuda.synchronize()
        b.graph_launch()
        b.check()
        for o, q in zip(outs, seq):
            assert torch.equal(o, q), rep
    x = ins[1].cpu().numpy().reshape(shapes[1], order="F")
    assert rel_l2(outs[1].cpu().numpy().reshape(shapes[1], order="F") / x.size, x) < 1e-13


def test_grouped_rejects_changed_plan(d2d, torch_gpu):
    """A plan whose stage kernel changes after bind must be bound again."""
    torch = torch_gpu
    shapes = [(64, 32, 16), (32, 32, 32)]
    engines = [d2d.FFTEngine(1, *s) for s in shapes]
    ins = [torch.zeros(int(np.prod(s)), dtype=torch.complex128, device="cuda") for s in shapes]
    outs = [torch.empty_like(a) for a in ins]
    b = d2d.Batch(engines)
    b.set_mode(1)
    b.bind(ins, outs, mode=-1)
    b.launch()
    b.check()
    kinds = engines[0].stage_kernels()
    names = [kinds[st] for st in ("c2c_x", "c2c_y", "c2c_z")]
    if all(n == "general" for n in names):
        pytest.skip("every stage already runs the general kernel")
    for st in ("c2c_x", "c2c_y", "c2c_z"):
        engines[0].set_stage_kernel(st, "general")
    with pytest.raises(d2d.D2dError):
        b.launch()
    b.bind(ins, outs, mode=-1)
    b.launch()
    b.check()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("budget", [0, 1, 8 << 20])
def test_cluster_order(d2d, torch_gpu, mode, budget):
    """Cluster-major queue order (one grid per cluster, an 8 MiB budget, and
    the whole batch phase by phase) gives the same bits as the plans' own
    sequential calls, forward then backward in one launch."""
    torch = torch_gpu
    shapes = grids(12, 7, [16, 32, 48, 64, 80, 96])
    engines = [d2d.FFTEngine(1, *s) for s in shapes]
    ins, outs, seq = [], [], []
    for g, s in enumerate(shapes):
        a = torch.empty(int(np.prod(s)), dtype=torch.complex128, device="cuda")
        d2d.fill_uniform(a, a.numel(), d2d.C16, 40 + g)
        ins.append(a)
        outs.append(torch.empty_like(a))
        seq.append(torch.empty_like(a))
    b = d2d.Batch(engines)
    match_plans(engines, b, mode)
    for e, a, q in zip(engines, ins, seq):
        e.fft_c2c(a, q, -1)
        e.fft_c2c(q, q, 1)
    b.set_mode(mode)
    b.set_cluster_bytes(budget)
    b.bind(ins, outs, mode=0)
    ncl, nph = b.clusters()
    if budget == 1:
        assert ncl == len(shapes) and nph == 6 * len(shapes)
    elif budget == 0:
        assert ncl == 1 and nph == 6
    else:
        assert 1 < ncl < len(shapes)
    b.launch()
    b.check()
    for o, q in zip(outs, seq):
        assert torch.equal(o, q)


@pytest.mark.parametrize("prec", [64, 32])
def test_persistent_specialised_tiles(d2d, torch_gpu, prec):
    """Grids whose every length has a one-wave specialised tile run the
    persistent kernel on those tiles: bit-identical to the plans on their
    specialised kernels, forward then backward in one launch and as hipGraph
    replays, with the round trip and the forward against pocketfft."""
    torch = torch_gpu
    shapes = grids(10, 11, [32, 48, 64, 80, 96, 128, 160, 192]) + [(256, 32, 48)]
    npc = np.complex128 if prec == 64 else np.complex64
    tol = 1e-13 if prec == 64 else 5e-6
    rng = np.random.default_rng(12)
    xs = [np.asfortranarray(rng.uniform(-1, 1, s) + 1j * rng.uniform(-1, 1, s)) for s in shapes]
    engines = [d2d.FFTEngine(1 if i % 3 else 3, *s, prec=prec) for i, s in enumerate(shapes)]
    ins = [dev(torch, x, npc) for x in xs]
    outs = [torch.empty_like(a) for a in ins]
    b = d2d.Batch(engines)
    assert b.persistent_tiles()
    assert match_plans(engines, b, 0) == "specialised"
    seq = [torch.empty_like(a) for a in ins]
    for e, a, q in zip(engines, ins, seq):
        e.fft_c2c(a, q, -1)
    b.set_mode(0)
    b.bind(ins, outs, mode=-1)
    b.launch()
    b.check()
    for x, o, q, s in zip(xs, outs, seq, shapes):
        assert torch.equal(o, q), s
        got = o.cpu().numpy().reshape(s, order="F")
        assert rel_l2(got, np.fft.fftn(x.astype(npc).astype(np.complex128))) < tol
    for e, q in zip(engines, seq):
        e.fft_c2c(q, q, 1)
    b.bind(ins, outs, mode=0)
    b.capture_graph()
    for rep in range(3):
        for o in outs:
            o.zero_()
        torch.cuda.synchronize()
        b.graph_launch()
        b.check()
        for o, q, s in zip(outs, seq, shapes):
            assert torch.equal(o, q), (rep, s)
    for x, o, s in zip(xs, outs, shapes):
        got = o.cpu().numpy().reshape(s, order="F") / np.prod(s)
        assert rel_l2(got, x.astype(npc)) < tol
    # one length without a one-wave tile: the general kernel's tiles
    b2 = d2d.Batch(engines[:2] + [d2d.FFTEngine(1, 40, 32, 32, prec=prec)])
    assert not b2.persistent_tiles()
