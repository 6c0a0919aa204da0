"""Benchmark of the single-GPU 3D FFT path: r2c + c2r round trip at 1024^3.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3] [--n 1024] [--no-cpu]

One step is decomp_2d_fft_3d(r2c) followed by decomp_2d_fft_3d(c2r) on a
1024^3 FP64 x-pencil (PHYSICAL_IN_X) that already sits in HBM: the modal
field of 32 Fourier modes with wavenumbers in [-128,128]^3, amplitudes
10^U(-6,2) with random signs, plus N(0, 1e-6) noise, seed 3 (synthetic).

Printed as one JSON line: ms per fwd+bwd (`value`), the 3-pass HBM model's
effective bandwidth and roofline fraction, the stricter compulsory-traffic
fraction (one read of the input and one write of the output per direction),
per-stage times from hipEvents recorded around every kernel launch of a
second set of round trips, the FP32 and split-BF16 variants, `configs` (the
other configurations, each timed as its own --workload line: c2c 128^3, r2c
512^3, r2c 720x540x432 PHYSICAL_IN_Z, the 2048^3 4x4 rank pack, 64 uneven
c2c grids in FP64 and FP32), and a CPU baseline: the Glassman restatement of
the reference's generic backend (oracle/) timed on a bounded sample of the
same workload's lines on all host cores and on one, scaled to the full round
trip.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))  # oracle/oracle.py (cpu_baseline only)

PEAK_HBM_GBS = 8000.0
# spectral pencil order between the stages (the tuner picks; see d2dhip_plan_set_layout)
LAYOUT_NAMES = {0: "plain", 1: "blocked", 2: "blocked32"}
PEAK_BF16_FLOPS = 2.5e15  # dense bf16 MFMA peak of MI355X

# HBM bytes per launch of the 1024^3 FP64 stage kernels, from two rocprofv3
# counter passes (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0, factors calibrated on a
# device-to-device copy of known size in the same run); the summary with the
# commands is profiles/hbm_traffic_r2c_c2r_1024cube.md.  Keyed by
# (n, stage, kernel label); a stage whose tuned kernel is not listed reports
# null.  The y stage moves 2.4-2.6 % more than its algorithmic bytes (rows of
# 513 complex values start off the 128-byte line grid); the others move
# theirs to within 0.1 %.
MEASURED_TRAFFIC = {
    (1024, "r2c", "double n=1024 tx=4 nt=256 r2c xs w3"): 17197322336,
    (1024, "spec_y", "double n=1024 tx=16 nt=1024 c2c xsh w4"): 17625138923,
    (1024, "spec_y", "double n=1024 tx=16 nt=1024 c2c xs w4"): 17660919919,
    (1024, "spec_last", "double n=1024 tx=16 nt=1024 c2c xsh w4"): 17225414907,
    (1024, "spec_last", "double n=1024 tx=8 nt=512 c2c xs w4"): 17214012992,
    (1024, "c2r", "double n=1024 tx=4 nt=256 c2r"): 17197203147,
    (1024, "c2r", "double n=1024 tx=4 nt=256 c2r xs w3"): 17197174955,
}
METRIC = "fwd+bwd 3D FFT ms & HBM-roofline % at 1024³ r2c, FP64 vs split-BF16 MFMA"


def stage_bytes(n, rb):
    """Algorithmic bytes of each stage launch (one read + one write of its
    pencils), r2c/c2r PHYSICAL_IN_X at n^3 with rb bytes per real."""
    nxc = n // 2 + 1
    s_real = n ** 3 * rb
    s_spec = nxc * n * n * 2 * rb
    return {"r2c": s_real + s_spec, "spec_y": 2 * s_spec, "spec_last": 2 * s_spec,
            "c2r": s_spec + s_real}


def round_trip_bytes(n, rb):
    b = stage_bytes(n, rb)
    return 2 * (b["r2c"] + b["spec_y"] + b["spec_last"])


def compulsory_bytes(n, rb):
    """One read of the input and one write of the output per direction."""
    b = stage_bytes(n, rb)
    return 2 * b["r2c"]


def time_round_trips(torch, eng, x, spec, y, steps, warmup, profile):
    """ms per r2c+c2r over `steps` timed steps (no per-launch events inside
    the timed region); with `profile`, the same number of further steps run
    with a hipEvent pair around every kernel launch for per-stage times."""
    for _ in range(warmup):
        eng.fft_r2c(x, spec)
        eng.fft_c2r(spec, y)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        eng.fft_r2c(x, spec)
        eng.fft_c2r(spec, y)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / steps
    rep = {}
    if profile:
        eng.profile(True)
        eng.profile_report()  # drop anything recorded before
        for _ in range(steps):
            eng.fft_r2c(x, spec)
            eng.fft_c2r(spec, y)
        torch.cuda.synchronize()
        rep = eng.profile_report()
        eng.profile(False)
    return ms, rep


def stage_ms(rep):
    """{stage: average ms per launch} from a profile report."""
    return {k: round(tot / cnt, 4) for k, (tot, cnt) in rep.items() if cnt}


def cpu_baseline(n, budget_lines):
    """The Glassman restatement timed on a bounded sample of the round trip's
    lines, once on every host core this process may run on (P) and once on
    one core, each scaled to the full round trip by lines per stage."""
    import oracle as O
    ncores = len(os.sched_getaffinity(0))
    nxc = n // 2 + 1
    lines = [n * n, nxc * n, nxc * n, nxc * n, nxc * n, n * n]  # x, y, z, z, y, x

    def run(nthreads, nlines):
        t0 = time.time()
        t, used = O.sample_stage_times(n, n, n, nlines, nthreads)
        return sum(tl * nl for tl, nl in zip(t, lines)) * 1e3, used, time.time() - t0

    ms_p, used, wall_p = run(ncores, budget_lines)
    lines1 = max(4096, budget_lines // max(1, used) // 2)
    ms_1, _, wall_1 = run(1, lines1)
    return {
        "value": ms_p,
        "unit": "ms per fwd+bwd (extrapolated)",
        "cores": used,
        "kind": "port",
        "P1_value": ms_1,
        "P1_over_P_speedup": round(ms_1 / ms_p, 2),
        "sample": ("Glassman restatement of the generic backend (oracle/glassman.c), "
                   "%d lines of each of the 6 stages of the %d^3 r2c+c2r round trip at their "
                   "real strides on P = %d OpenMP threads (all cores of this process's "
                   "affinity mask, %.1f s wall), and %d lines on P = 1 (%.1f s wall), each "
                   "scaled by lines per stage" % (budget_lines, n, used, wall_p, lines1, wall_1)),
    }


def config_summary(line):
    """The fields of a configuration's own bench line kept in the headline's
    `configs` entry."""
    keep = ("ms_per_step", "GBps_model", "hbm_frac_model", "GBps_3pass", "hbm_frac_3pass",
            "round_trip_rel_l2", "round_trip_rel_l2_max", "graph_ms", "sequential_ms",
            "persistent_ms", "persistent_graph_ms", "speedup_vs_sequential", "grouped_graph_ms",
            "executor", "value", "unit", "unfused_ms", "fused_speedup",
            "persistent_general_graph_ms")
    out = {"workload": line["config"]["workload"]}
    out.update({k: line[k] for k in keep if k in line})
    return out


def bench_batch(args, torch, d2d, batch_prec=None, mixed=False):
    """64 independent c2c grids, each dimension from {32,...,256} (seed 4),
    inputs uniform[-1,1): forward+backward of every grid as 64 sequential
    decomp_2d_fft_3d pairs, and as one batch with either executor (one launch
    per stage phase and kernel, or one persistent launch), eager and as a
    replay of the launches captured in a hipGraph."""
    import numpy as np
    rng = np.random.default_rng(4)
    sizes = [32, 48, 64, 80, 96, 128, 160, 192, 256]
    prec = d2d.PREC_F64 if (batch_prec or args.batch_prec) == 64 else d2d.PREC_F32
    cdt = torch.complex128 if prec == d2d.PREC_F64 else torch.complex64
    elem = d2d.C16 if prec == d2d.PREC_F64 else d2d.C8
    shapes = [tuple(int(v) for v in rng.choice(sizes, 3)) for _ in range(64)]
    if mixed:  # the reference's multi-grid shapes (fft_multiple_grids.f90:51-54) among them
        shapes[:4] = [(17, 13, 11), (17, 13, 11), (17, 15, 27), (33, 15, 11)]
    engines = [d2d.FFTEngine(d2d.PHYSICAL_IN_X, *s, prec=prec) for s in shapes]
    ins, outs = [], []
    for g, s in enumerate(shapes):
        a = torch.empty(int(np.prod(s)), dtype=cdt, device="cuda")
        d2d.fill_uniform(a, a.numel(), elem, 4 + g)
        ins.append(a)
        outs.append(torch.empty_like(a))

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps

    def sequential():
        for e, a, o in zip(engines, ins, outs):
            e.fft_c2c(a, o, -1)
            e.fft_c2c(o, o, 1)

    batch = d2d.Batch(engines)
    batch.bind(ins, outs, mode=0)
    ms_seq = timed(sequential)
    # the persistent executor: one launch, device work queue
    batch.set_mode(d2d.BATCH_PERSISTENT)
    ms_batch = timed(batch.launch)
    batch.check()
    batch.capture_graph()
    ms_graph = timed(batch.graph_launch)
    batch.check()
    # the default executor: one launch per stage phase and stage kernel
    batch.set_mode(d2d.BATCH_GROUPED)
    ms_grouped = timed(batch.launch)
    batch.capture_graph()
    ms_grouped_graph = timed(batch.graph_launch)
    batch.check()
    rt = max(float(torch.linalg.vector_norm(o / float(np.prod(s)) - a) / torch.linalg.vector_norm(a))
             for a, o, s in zip(ins, outs, shapes))
    rb = 16 if prec == d2d.PREC_F64 else 8
    nbytes = sum(2 * 6 * int(np.prod(s)) * rb for s in shapes)  # 3-pass model, both directions
    if mixed:  # no persistent tiles for 11..33: the default executor runs grouped launches
        assert not batch.persistent_tiles()
        batch.set_mode(d2d.BATCH_PERSISTENT)
        batch.capture_graph()
        ms_default = timed(batch.graph_launch)
        batch.check()
        batch.set_mode(d2d.BATCH_PERSISTENT_GENERAL)
        batch.capture_graph()
        ms_general = timed(batch.graph_launch)
        batch.check()
        rb = 16 if prec == d2d.PREC_F64 else 8
        nbytes = sum(2 * 6 * int(np.prod(s)) * rb for s in shapes)
        return {
            "metric": "grids/s, 64 c2c grids fwd+bwd incl. 17x13x11, 17x15x27, 33x15x11 "
                      "(default executor, hipGraph replay)",
            "value": round(64 / (ms_default * 1e-3), 1), "unit": "grids/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_default, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64" if prec == d2d.PREC_F64 else "f32", "data": "synthetic",
            "config": {"workload": "64 c2c grids: the reference's 3 multi-grid shapes + 60 with "
                                   "dims from {32..256} seed 4, fwd+bwd"},
            "executor": "grouped (no persistent tile for lengths 11-33)",
            "sequential_ms": round(ms_seq, 4), "persistent_general_graph_ms": round(ms_general, 4),
            "speedup_vs_sequential": round(ms_seq / ms_default, 3),
            "GBps_3pass": round(nbytes / (ms_default * 1e-3) / 1e9, 1),
            "round_trip_rel_l2_max": max(
                float(torch.linalg.vector_norm(o / float(np.prod(s)) - a) / torch.linalg.vector_norm(a))
                for a, o, s in zip(ins, outs, shapes)),
        }
    line = {
        "metric": "grids/s, 64 uneven c2c grids fwd+bwd (one persistent launch, hipGraph replay)",
        "value": round(64 / (ms_graph * 1e-3), 1), "unit": "grids/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_graph, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == d2d.PREC_F64 else "f32", "data": "synthetic",
        "config": {"workload": "64 c2c grids, dims from {32..256} seed 4, fwd+bwd",
                   "shapes": ["%dx%dx%d" % s for s in shapes[:4]] + ["..."]},
        "executor": "persistent (device work queues, %s tiles)"
                    % ("specialised" if batch.persistent_tiles() else "general"),
        "sequential_ms": round(ms_seq, 4),
        "persistent_ms": round(ms_batch, 4), "persistent_graph_ms": round(ms_graph, 4),
        "speedup_vs_sequential": round(ms_seq / ms_graph, 3),
        "grouped_ms": round(ms_grouped, 4), "grouped_graph_ms": round(ms_grouped_graph, 4),
        "grouped_speedup_vs_sequential": round(ms_seq / ms_grouped_graph, 3),
        "GBps_3pass": round(nbytes / (ms_graph * 1e-3) / 1e9, 1),
        "hbm_frac_3pass": round(nbytes / (ms_graph * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
        "round_trip_rel_l2_max": rt,
    }
    return line


def bench_config(args, torch, d2d, w):
    """One of the other configurations, timed like the headline (FP64, tuned
    plans, inputs in HBM): c2c 128^3 forward+backward (PHYSICAL_IN_X, uniform
    inputs seed 0); r2c+c2r 512^3 PHYSICAL_IN_X (modal field seed 1); r2c+c2r
    720x540x432 PHYSICAL_IN_Z (uniform seed 2); the mem_split_xy pack of one
    rank of 2048^3 on a 4x4 grid (x-pencil 2048x512x512 real8, seed 2)."""
    import numpy as np

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps

    extra = {}
    if w == "c2c128":
        n = 128
        x = torch.empty(n ** 3, dtype=torch.complex128, device="cuda")
        d2d.fill_uniform(x, x.numel(), d2d.C16, 0)
        y = torch.empty_like(x)
        e = d2d.FFTEngine(d2d.PHYSICAL_IN_X, n, n, n, tune=not args.no_tune)

        def step():
            e.fft_c2c(x, y, -1)
            e.fft_c2c(y, y, 1)
        ms = timed(step)
        nbytes = 2 * 6 * n ** 3 * 16
        err = float(torch.linalg.vector_norm(y / n ** 3 - x) / torch.linalg.vector_norm(x))
        cfg = {"workload": "c2c FP64 128^3 fwd+bwd PHYSICAL_IN_X, uniform seed 0"}
        extra["kernels"] = e.stage_kernels()
        # the same pair recorded once into a hipGraph and replayed
        e.capture_begin()
        step()
        e.capture_end()
        extra["graph_ms"] = round(timed(e.graph_launch), 4)
    elif w in ("r2c512", "r2cz720"):
        fmt = d2d.PHYSICAL_IN_X if w == "r2c512" else d2d.PHYSICAL_IN_Z
        nx, ny, nz = (512, 512, 512) if w == "r2c512" else (720, 540, 432)
        x = torch.empty(nx * ny * nz, dtype=torch.float64, device="cuda")
        if w == "r2c512":
            d2d.fill_modal(x, nx, ny, nz, d2d.R8, d2d.modes_make(32, 64, 0, 0.0, 0.0, 1e-3, 1))
        else:
            d2d.fill_uniform(x, x.numel(), d2d.R8, 2)
        ns = (nx // 2 + 1) * ny * nz if fmt == d2d.PHYSICAL_IN_X else nx * ny * (nz // 2 + 1)
        sp = torch.empty(ns, dtype=torch.complex128, device="cuda")
        y = torch.empty_like(x)
        e = d2d.FFTEngine(fmt, nx, ny, nz, tune=not args.no_tune)

        def step():
            e.fft_r2c(x, sp)
            e.fft_c2r(sp, y)
        ms = timed(step)
        s_real, s_spec = nx * ny * nz * 8, ns * 16
        nbytes = 2 * (s_real + 5 * s_spec)
        err = float(torch.linalg.vector_norm(y / (nx * ny * nz) - x) / torch.linalg.vector_norm(x))
        cfg = {"workload": "r2c+c2r FP64 %dx%dx%d %s" % (nx, ny, nz, "PHYSICAL_IN_X" if
                                                         fmt == d2d.PHYSICAL_IN_X else "PHYSICAL_IN_Z")}
        extra["kernels"] = e.stage_kernels()
    elif w == "r2cpack2048":
        # the same rank: its x stage (r2c along x, 2048 -> 1025) written
        # straight into the x->y send buffers of the spectral decomposition,
        # against the x stage (plain spectral x-pencil) followed by the pack
        dec = d2d.decomp_info_init(1025, 2048, 2048, 4, 4, 0, 0)
        nx, ny, nz = 2048, dec.xsz[1], dec.xsz[2]
        one = d2d.decomp_info_init(1025, ny, nz)
        x = torch.empty(nx * ny * nz, dtype=torch.float64, device="cuda")
        d2d.fill_uniform(x, x.numel(), d2d.R8, 2)
        spec = torch.empty(1025 * ny * nz, dtype=torch.complex128, device="cuda")
        buf = torch.empty(d2d.pack_buffer_elems(d2d.X_TO_Y, dec, 0), dtype=torch.complex128,
                          device="cuda")
        e = d2d.FFTEngine(d2d.PHYSICAL_IN_X, nx, ny, nz, tune=not args.no_tune)

        def step():
            e.fft_r2c_pack(x, buf, dec)
        ms = timed(step)

        def unfused():
            e.fft_r2c_pack(x, spec, one)
            d2d.pack(d2d.X_TO_Y, d2d.C16, spec, buf, dec, 0)
        ms_unfused = timed(unfused)
        nbytes = x.numel() * 8 + spec.numel() * 16
        err = 0.0
        cfg = {"workload": "r2c x stage fused with the mem_split_xy pack, one rank of 2048^3 on "
                           "4x4 (2048x512x512 real8 -> 1025x512x512 complex16 send buffers)"}
        extra["unfused_ms"] = round(ms_unfused, 4)
        extra["fused_speedup"] = round(ms_unfused / ms, 3)
        extra["kernels"] = {"r2c": e.stage_kernels()["r2c"]}
    elif w == "dtt513":
        # decomp_2d_dtt_3d_r2r forward (DCT-II on x, y, z with the default
        # ndismiss = 1: 512 transformed points of 513) then backward (DCT-III)
        n = 513
        x = torch.empty(n ** 3, dtype=torch.float64, device="cuda")
        d2d.fill_uniform(x, x.numel(), d2d.R8, 5)
        f = torch.empty_like(x)
        y = torch.empty_like(x)
        e = d2d.FFTEngine(d2d.PHYSICAL_IN_X, n, n, n, tune=not args.no_tune, opt_DTT=(2, 2, 2))

        def step():
            e.dtt_3d_r2r(x, f, d2d.FORWARD)
            e.dtt_3d_r2r(f, y, d2d.BACKWARD)
        ms = timed(step)
        nbytes = 6 * 2 * n ** 3 * 8
        # round trip on the transformed points, over (2 * 512)^3
        core = (slice(0, 512),) * 3
        xv = x.view(n, n, n)[core]
        yv = y.view(n, n, n)[core] / float(1024 ** 3)
        err = float(torch.linalg.vector_norm(yv - xv) / torch.linalg.vector_norm(xv))
        cfg = {"workload": "DCT-II fwd + DCT-III bwd FP64 513^3 PHYSICAL_IN_X (512 transformed "
                           "points per line), uniform seed 5"}
        e.profile(True)
        step()
        rep = e.profile_report()
        e.profile(False)
        extra["stage_ms"] = {k: round(v[0], 4) for k, v in rep.items()}
        extra["kernels"] = {k: v for k, v in e.dtt_stage_kernels().items()}
        # DCT-I both ways on the same grid (all 513 points, extension of 1024)
        e1 = d2d.FFTEngine(d2d.PHYSICAL_IN_X, n, n, n, tune=not args.no_tune, opt_DTT=(1, 1, 1))

        def step1():
            e1.dtt_3d_r2r(x, f, d2d.FORWARD)
            e1.dtt_3d_r2r(f, y, d2d.BACKWARD)
        extra["dct1_round_trip_ms"] = round(timed(step1), 4)
        err1 = torch.linalg.vector_norm(y / float(1024 ** 3) - x) / torch.linalg.vector_norm(x)
        extra["dct1_round_trip_rel_l2"] = float(err1)
        del e1, f, y
        # the DCT-II / DCT-III round trip in FP32 storage and arithmetic
        x32 = x.float()
        f32 = torch.empty_like(x32)
        y32 = torch.empty_like(x32)
        e32 = d2d.FFTEngine(d2d.PHYSICAL_IN_X, n, n, n, prec=d2d.PREC_F32, opt_DTT=(2, 2, 2))

        def step32():
            e32.dtt_3d_r2r(x32, f32, d2d.FORWARD)
            e32.dtt_3d_r2r(f32, y32, d2d.BACKWARD)
        extra["f32_round_trip_ms"] = round(timed(step32), 4)
        y32v = y32.view(n, n, n)[core].double() / float(1024 ** 3)
        extra["f32_round_trip_rel_l2"] = float(torch.linalg.vector_norm(y32v - xv) /
                                               torch.linalg.vector_norm(xv))
    elif w == "rank1024":
        # rank (1, 1) of 1024^3 on a 4 x 2 grid: the forward r2c's per-rank
        # stages and the packs / unpacks of its two exchanges (the exchanges
        # themselves need other GPUs; each receive buffer here is the send
        # buffer), unfused against the x and y stages fused with their packs
        n = 1024
        e = d2d.FFTEngine(d2d.PHYSICAL_IN_X, n, n, n, rank=(4, 2, 1, 1), tune=not args.no_tune)
        ph, sp = e.decomps()
        # the y and z stages on kernels with a fused pack / unpack form (or
        # the general kernel)
        fused_forms = ("double n=1024 tx=16 nt=1024 c2c xsh w4", "double n=1024 tx=8 nt=512 c2c xsh w4",
                       "double n=1024 tx=16 nt=1024 c2c xs w4", "double n=1024 tx=8 nt=512 c2c xs w4",
                       "double n=1024 tx=8 nt=512 c2c", "general")
        for st in ("spec_y", "spec_last"):
            if e.stage_kernels()[st] not in fused_forms:
                e.set_stage_kernel(st, fused_forms[0])
        nel = lambda v: int(np.prod(list(v)))
        x = torch.empty(nel(ph.xsz), dtype=torch.float64, device="cuda")
        d2d.fill_uniform(x, x.numel(), d2d.R8, 3)
        sx = torch.empty(nel(sp.xsz), dtype=torch.complex128, device="cuda")
        sy = torch.empty(nel(sp.ysz), dtype=torch.complex128, device="cuda")
        sz = torch.empty(nel(sp.zsz), dtype=torch.complex128, device="cuda")
        b1 = torch.empty(d2d.pack_buffer_elems(d2d.X_TO_Y, sp), dtype=torch.complex128, device="cuda")
        b2 = torch.empty(max(d2d.pack_buffer_elems(d2d.Y_TO_Z, sp), nel(sp.zsz)),
                         dtype=torch.complex128, device="cuda")
        C16 = d2d.C16

        def unfused():
            e.exec_1m("r2c", x, sx, d2d.FORWARD)
            d2d.pack(d2d.X_TO_Y, C16, sx, b1, sp)
            d2d.unpack(d2d.X_TO_Y, C16, b1, sy, sp)
            e.exec_1m("spec_y", sy, sy, d2d.FORWARD)
            d2d.pack(d2d.Y_TO_Z, C16, sy, b2, sp)
            d2d.unpack(d2d.Y_TO_Z, C16, b2, sz, sp)
            e.exec_1m("spec_last", sz, sz, d2d.FORWARD)

        def step():
            e.fft_r2c_pack(x, b1, sp)
            e.exec_1m_unpack_pack("spec_y", b1, b2, d2d.X_TO_Y, d2d.Y_TO_Z, sp, d2d.FORWARD)
            e.exec_1m_unpack("spec_last", b2, sz, d2d.Y_TO_Z, sp, d2d.FORWARD)
        ms = timed(step)
        ms_unfused = timed(unfused)
        rs, cs = x.numel() * 8, sx.numel() * 16
        # fused: r2c_pack (real + spectrum), the y stage from the x->y receive
        # buffer to the y->z send buffer, the z stage from the y->z receive
        # buffer: three passes
        nbytes = (rs + cs) + 2 * cs + 2 * cs
        err = 0.0
        cfg = {"workload": "forward r2c stages of rank (1,1) of 1024^3 on 4x2 (x-pencil "
                           "1024x256x512 real8, spectral y-pencil 128x1024x512, z-pencil "
                           "128x512x1024) with the packs and unpacks of its two exchanges, every "
                           "one fused into a stage (x: pack; y: unpack and pack; z: unpack); the "
                           "exchanges themselves omitted (one GPU)"}
        extra["unfused_ms"] = round(ms_unfused, 4)
        extra["fused_speedup"] = round(ms_unfused / ms, 3)
        extra["kernels"] = {k: v for k, v in e.stage_kernels().items() if k in ("r2c", "spec_y",
                                                                               "spec_last")}
    elif w == "host512":
        # the boundary handing over host arrays (the Fortran layer's
        # decomp_2d_fft_3d on host arrays): real input copied in, r2c + c2r,
        # real output copied back, PCIe included; pageable and pinned host
        # memory
        n = 512
        nxc = n // 2 + 1
        e = d2d.FFTEngine(d2d.PHYSICAL_IN_X, n, n, n, tune=not args.no_tune)
        xd = torch.empty(n ** 3, dtype=torch.float64, device="cuda")
        d2d.fill_uniform(xd, xd.numel(), d2d.R8, 1)
        sp = torch.empty(nxc * n * n, dtype=torch.complex128, device="cuda")
        yd = torch.empty_like(xd)
        res = {}
        for kind, pin in (("pageable", False), ("pinned", True)):
            xh = xd.cpu()
            yh = torch.empty(n ** 3, dtype=torch.float64)
            if pin:
                xh, yh = xh.pin_memory(), yh.pin_memory()

            def step():
                xd.copy_(xh, non_blocking=pin)
                e.fft_r2c(xd, sp)
                e.fft_c2r(sp, yd)
                yh.copy_(yd, non_blocking=pin)
            res[kind] = timed(step)
            torch.cuda.synchronize()
            res[kind + "_rel_l2"] = float(torch.linalg.vector_norm(yh / float(n ** 3) - xh) /
                                          torch.linalg.vector_norm(xh))

        def device_only():
            e.fft_r2c(xd, sp)
            e.fft_c2r(sp, yd)
        ms_dev = timed(device_only)
        ms = res["pinned"]
        xfer = 2 * xd.numel() * 8
        nbytes = round_trip_bytes(n, 8)
        err = max(res["pageable_rel_l2"], res["pinned_rel_l2"])
        cfg = {"workload": "r2c+c2r FP64 512^3 PHYSICAL_IN_X on host arrays: host-to-device copy "
                           "of the real input, the round trip, device-to-host copy of the real "
                           "output (value: pinned host memory)"}
        extra["pageable_ms"] = round(res["pageable"], 4)
        extra["pinned_ms"] = round(res["pinned"], 4)
        extra["device_only_ms"] = round(ms_dev, 4)
        extra["pcie_GBps_pinned"] = round(xfer / ((res["pinned"] - ms_dev) * 1e-3) / 1e9, 1)
        extra["pcie_GBps_pageable"] = round(xfer / ((res["pageable"] - ms_dev) * 1e-3) / 1e9, 1)
    elif w == "sizes":
        # r2c+c2r FP64 cubes of the mixed-radix lengths with specialised
        # kernels, each also timed with every stage on the general kernel
        sizes = {}
        ms, nbytes, err = 0.0, 0, 0.0
        every = (100, 120, 200, 240, 300, 320, 360, 384, 400, 480, 500, 576, 600, 640, 768, 800,
                 896, 960, 1000, 1152, 1200, 1280, 1536, 1600)
        pick = [int(v) for v in args.cubes.split(",") if v.strip()] if args.cubes else every
        bad = [v for v in pick if v not in every]
        if bad:
            raise SystemExit("bench.py: --cubes %s: not among the sizes workload's lengths %s"
                             % (bad, list(every)))
        for n in pick:
            nxc = n // 2 + 1
            x = torch.empty(n ** 3, dtype=torch.float64, device="cuda")
            d2d.fill_uniform(x, x.numel(), d2d.R8, 2)
            sp = torch.empty(nxc * n * n, dtype=torch.complex128, device="cuda")
            y = torch.empty_like(x)
            e = d2d.FFTEngine(d2d.PHYSICAL_IN_X, n, n, n, tune=not args.no_tune)

            def step():
                e.fft_r2c(x, sp)
                e.fft_c2r(sp, y)
            t = timed(step)
            b = 2 * (n ** 3 * 8 + 5 * nxc * n * n * 16)
            r = float(torch.linalg.vector_norm(y / n ** 3 - x) / torch.linalg.vector_norm(x))
            kern = e.stage_kernels()
            for st in ("r2c", "spec_y", "spec_last", "c2r"):
                e.set_stage_kernel(st, "general")
            tg = timed(step)
            sizes[str(n)] = {"ms": round(t, 4), "GBps_model": round(b / (t * 1e-3) / 1e9, 1),
                             "general_kernel_ms": round(tg, 4),
                             "speedup_vs_general": round(tg / t, 2), "round_trip_rel_l2": r,
                             "kernels": {k: kern[k] for k in ("r2c", "spec_y", "spec_last", "c2r")}}
            ms += t
            nbytes += b
            err = max(err, r)
            del e, x, sp, y
            torch.cuda.empty_cache()
        cfg = {"workload": "r2c+c2r FP64 n^3 PHYSICAL_IN_X, n in %s, uniform seed 2; one step = "
                           "one round trip of each" % " ".join(str(v) for v in pick)}
        extra["sizes"] = sizes
    elif w == "r2c2048":
        # the largest cube one MI355X holds in FP64: 2048^3 r2c+c2r, the
        # c2r working in place on its input (opt_inplace), so that input,
        # spectrum and output take 206 GB of the 288 GB; the tuner's scratch
        # does not fit beside them, so the plan keeps its default kernels
        n = 2048
        nxc = n // 2 + 1
        x = torch.empty(n ** 3, dtype=torch.float64, device="cuda")
        d2d.fill_uniform(x, x.numel(), d2d.R8, 2)
        sp = torch.empty(nxc * n * n, dtype=torch.complex128, device="cuda")
        y = torch.empty_like(x)
        e = d2d.FFTEngine(d2d.PHYSICAL_IN_X, n, n, n, inplace=True, tune=not args.no_tune)

        def step():
            e.fft_r2c(x, sp)
            e.fft_c2r(sp, y)
        ms = timed(step)
        nbytes = 2 * (n ** 3 * 8 + 5 * nxc * n * n * 16)
        cfg = {"workload": "r2c+c2r FP64 2048^3 PHYSICAL_IN_X, c2r in place on its input, "
                           "uniform seed 2"}
        e.profile(True)
        step()
        extra["stage_ms"] = {k: round(v[0] / v[1], 4) for k, v in e.profile_report().items() if v[1]}
        e.profile(False)
        extra["kernels"] = e.stage_kernels()
        extra["device_GB_in_use"] = round(torch.cuda.memory_allocated() / 1e9, 1)
        # the round trip's error in place (no 69 GB temporaries)
        y.div_(float(n) ** 3).sub_(x)
        err = float(torch.linalg.vector_norm(y) / torch.linalg.vector_norm(x))
    elif w == "halo512":
        # update_halo of a 512^3 real8 x-pencil, 2 halo levels, y and z periodic
        n, lev = 512, 2
        x = torch.empty(n ** 3, dtype=torch.float64, device="cuda")
        d2d.fill_uniform(x, x.numel(), d2d.R8, 6)
        hs = d2d.halo_shape((n, n, n), lev, 1)
        out = torch.zeros(hs[0] * hs[1] * hs[2], dtype=torch.float64, device="cuda")
        per = (False, True, True)

        def step():
            d2d.update_halo(d2d.R8, x, out, (n, n, n), lev, 1, per)
        ms = timed(step)
        # interior read + halo-array write, plus the halo slabs (read + write)
        halo = out.numel() - x.numel()
        nbytes = (x.numel() + out.numel() + 2 * halo) * 8
        err = 0.0
        cfg = {"workload": "update_halo, 512^3 real8 x-pencil, level 2, y and z periodic "
                           "(halo array 512x516x516)"}
    else:  # pack2048
        dec = d2d.decomp_info_init(2048, 2048, 2048, 4, 4, 0, 0)
        xs = dec.xsz
        x = torch.empty(xs[0] * xs[1] * xs[2], dtype=torch.float64, device="cuda")
        d2d.fill_uniform(x, x.numel(), d2d.R8, 2)
        buf = torch.empty(d2d.pack_buffer_elems(0, dec, 0), dtype=torch.float64, device="cuda")

        def step():
            d2d.pack(0, d2d.R8, x, buf, dec, 0)
        ms = timed(step)
        nbytes = 2 * x.numel() * 8
        err = 0.0
        cfg = {"workload": "mem_split_xy pack, one rank of 2048^3 on 4x4 (2048x512x512 real8)"}
    line = {"metric": "ms per step (%s)" % w, "value": round(ms, 4), "unit": "ms", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic", "config": cfg,
            "GBps_model": round(nbytes / (ms * 1e-3) / 1e9, 1),
            "hbm_frac_model": round(nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
            "round_trip_rel_l2": err}
    line.update(extra)
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="headline only: skip the other configurations' timings")
    ap.add_argument("--cpu-lines", type=int, default=524288,
                    help="lines of each stage in the CPU baseline sample (about 13 s on 16 cores)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the plan-time autotuner (use the default kernels)")
    ap.add_argument("--workload", choices=["r2c1024", "batch64", "c2c128", "r2c512", "r2cz720",
                                           "pack2048", "r2cpack2048", "dtt513", "halo512",
                                           "rank1024", "host512", "sizes", "r2c2048"],
                    default="r2c1024",
                    help="r2c1024: the headline round trip; batch64: 64 uneven c2c grids; "
                         "c2c128 / r2c512 / r2cz720 / pack2048: the other parity configurations, "
                         "timed; r2cpack2048: the pack2048 rank's x stage fused with its pack; "
                         "dtt513: 3D DCT-II + DCT-III round trip; halo512: update_halo; "
                         "rank1024: one rank's stages with their packs and unpacks; host512: the "
                         "512^3 round trip on host arrays, PCIe included; sizes: FP64 round "
                         "trips of 24 mixed-radix cubes against the general kernel; r2c2048: "
                         "the 2048^3 FP64 round trip (206 GB on one GPU)")
    ap.add_argument("--batch-prec", type=int, choices=[64, 32], default=64)
    ap.add_argument("--cubes", default="",
                    help="sizes workload: comma-separated subset of its cube lengths")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs the 3D FFT path on one GPU only; "
              "--gpus %d is not supported (multi-GPU pencils need the alltoallv exchange, "
              "which is out of scope)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import importlib
    import torch
    d2d = importlib.import_module("2decomp-fft_amd")
    if args.workload == "batch64":
        print(json.dumps(bench_batch(args, torch, d2d)), flush=True)
        return
    if args.workload in ("c2c128", "r2c512", "r2cz720", "pack2048", "r2cpack2048", "dtt513",
                         "halo512", "rank1024", "host512", "sizes", "r2c2048"):
        print(json.dumps(bench_config(args, torch, d2d, args.workload)), flush=True)
        return
    n = args.n
    nxc = n // 2 + 1
    print("bench: n=%d steps=%d warmup=%d" % (n, args.steps, args.warmup), file=sys.stderr,
          flush=True)

    # ---- FP64 drop-in -------------------------------------------------------
    modes = d2d.modes_make(32, 128, 1, -6.0, 2.0, 1e-6, 3)
    x = torch.empty(n ** 3, dtype=torch.float64, device="cuda")
    d2d.fill_modal(x, n, n, n, d2d.R8, modes)
    spec = torch.empty(nxc * n * n, dtype=torch.complex128, device="cuda")
    y = torch.empty_like(x)
    t0 = time.time()
    eng = d2d.FFTEngine(d2d.PHYSICAL_IN_X, n, n, n, prec=d2d.PREC_F64, tune=not args.no_tune)
    plan_s = time.time() - t0
    ms, rep = time_round_trips(torch, eng, x, spec, y, args.steps, args.warmup, profile=True)
    print("bench: fp64 %.3f ms per fwd+bwd" % ms, file=sys.stderr, flush=True)
    rt_err = float(torch.linalg.vector_norm(y / n ** 3 - x) / torch.linalg.vector_norm(x))
    spec64 = spec.clone()

    bytes_rt = round_trip_bytes(n, 8)
    gbs = bytes_rt / (ms * 1e-3) / 1e9
    sb = stage_bytes(n, 8)
    stages = {}
    for name, (tot, cnt) in rep.items():
        avg = tot / cnt
        stages[name] = {"avg_ms": round(avg, 4), "launches": cnt,
                        "GBps": round(sb[name] / (avg * 1e-3) / 1e9, 1)}
    dom = max(rep.items(), key=lambda kv: kv[1][0])[0]
    dom_avg = rep[dom][0] / rep[dom][1]
    achieved = sb[dom] / (dom_avg * 1e-3) / 1e9
    kernels64 = eng.stage_kernels()
    traffic = MEASURED_TRAFFIC.get((n, dom, kernels64.get(dom)))

    # device-to-device copy bandwidth of the same bytes, for context
    cpy = torch.empty_like(spec)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    cpy.copy_(spec)
    e0.record()
    for _ in range(5):
        cpy.copy_(spec)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 2 * spec.numel() * 16 / (e0.elapsed_time(e1) / 5 * 1e-3) / 1e9
    # the same bytes moved with the spectral stages' access pattern and no
    # arithmetic (16-column tiles of 1024 points at the z / y strides): the
    # bandwidth a strided stage can reach on this device
    pattern_gbs = {}
    for name, geo in (("spec_last", (nxc * n, 1, 0, nxc * n)), ("spec_y", (nxc, n, nxc * n, nxc))):
        d2d.column_copy(spec, cpy, d2d.C16, *geo, n)
        e0.record()
        for _ in range(5):
            d2d.column_copy(spec, cpy, d2d.C16, *geo, n)
        e1.record()
        torch.cuda.synchronize()
        pattern_gbs[name] = round(2 * spec.numel() * 16 / (e0.elapsed_time(e1) / 5 * 1e-3) / 1e9, 1)
    del cpy

    # ---- FP32 VALU variant ---------------------------------------------------
    variants = {}
    del y
    eng32 = d2d.FFTEngine(d2d.PHYSICAL_IN_X, n, n, n, prec=d2d.PREC_F32, tune=not args.no_tune)
    kernels32 = eng32.stage_kernels()
    x32 = x.float()
    spec32 = torch.empty(nxc * n * n, dtype=torch.complex64, device="cuda")
    y32 = torch.empty_like(x32)
    ms32, rep32 = time_round_trips(torch, eng32, x32, spec32, y32, args.steps, args.warmup, True)
    err32 = float(torch.linalg.vector_norm(spec32.to(torch.complex128) - spec64) /
                  torch.linalg.vector_norm(spec64))
    maxabs32 = float((spec32.to(torch.complex128) - spec64).abs().max())
    variants["fp32"] = {"ms": round(ms32, 3), "rel_l2_vs_fp64": err32, "max_abs_vs_fp64": maxabs32,
                        "GBps_3pass": round(round_trip_bytes(n, 4) / (ms32 * 1e-3) / 1e9, 1),
                        "speedup_vs_fp64": round(ms / ms32, 3),
                        "stage_ms": stage_ms(rep32),
                        "kernels": {k: kernels32[k] for k in ("r2c", "spec_y", "spec_last", "c2r")},
                        "chunk_planes": {"fwd": eng32.chunking()[0], "bwd": eng32.chunking()[1]},
                        "layout": LAYOUT_NAMES[eng32.layout()]}
    del spec32, y32, eng32

    # ---- split-BF16 matrix-core variant (FP32 storage) -------------------------
    engs = d2d.FFTEngine(d2d.PHYSICAL_IN_X, n, n, n, prec=d2d.PREC_SPLIT_BF16,
                         tune=not args.no_tune)
    kernels_s = engs.stage_kernels()
    specs = torch.empty(nxc * n * n, dtype=torch.complex64, device="cuda")
    ys = torch.empty_like(x32)
    mss, reps = time_round_trips(torch, engs, x32, specs, ys, args.steps, args.warmup, True)
    errs = float(torch.linalg.vector_norm(specs.to(torch.complex128) - spec64) /
                 torch.linalg.vector_norm(spec64))
    maxabs_s = float((specs.to(torch.complex128) - spec64).abs().max())
    rt_s = float(torch.linalg.vector_norm(ys.double() / n ** 3 - x) / torch.linalg.vector_norm(x))
    # matrix-core work: two radix-16 passes per 1024-point line, 12 MFMAs
    # (16x16x32, 16384 flops each) per 16 butterflies = 768 flops per point
    # and pass; r2c/c2r stages transform n^3/2 complex points, the spectral
    # stages nxc*n*n each
    pts = {"r2c": n ** 3 // 2, "spec_y": nxc * n * n, "spec_last": nxc * n * n, "c2r": n ** 3 // 2}
    mf_flops = sum(2 * 768 * pts[k] * (2 if k in ("spec_y", "spec_last") else 1)
                   for k in pts if kernels_s[k].startswith("mfma"))
    variants["split_bf16"] = {
        "ms": round(mss, 3), "rel_l2_vs_fp64": errs, "max_abs_vs_fp64": maxabs_s,
        "round_trip_rel_l2": rt_s, "speedup_vs_fp64": round(ms / mss, 3),
        "GBps_3pass": round(round_trip_bytes(n, 4) / (mss * 1e-3) / 1e9, 1),
        "mfma_tflops": round(mf_flops / (mss * 1e-3) / 1e12, 1),
        "mfma_util": round(mf_flops / (mss * 1e-3) / PEAK_BF16_FLOPS, 4),
        "stage_ms": stage_ms(reps),
        "kernels": {k: kernels_s[k] for k in ("r2c", "spec_y", "spec_last", "c2r")},
        "layout": LAYOUT_NAMES[engs.layout()],
    }
    del x32, specs, ys, engs

    line = {
        "metric": METRIC,
        "value": round(ms, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "r2c+c2r FP64 %d^3 PHYSICAL_IN_X, modal field seed 3" % n,
                   "grid": [n, n, n], "format": "PHYSICAL_IN_X", "precision": "f64"},
        "GBps_3pass": round(gbs, 1),
        "hbm_frac_3pass": round(gbs / PEAK_HBM_GBS, 4),
        "compulsory_GBps": round(compulsory_bytes(n, 8) / (ms * 1e-3) / 1e9, 1),
        "compulsory_frac": round(compulsory_bytes(n, 8) / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
        "copy_GBps": round(copy_gbs, 1),
        "pattern_copy_GBps": pattern_gbs,
        "round_trip_rel_l2": rt_err,
        "stages": stages,
        "kernels": {k: v for k, v in kernels64.items() if k in stages},
        "chunk_planes": {"fwd": eng.chunking()[0], "bwd": eng.chunking()[1]},
        "layout": LAYOUT_NAMES[eng.layout()],
        "plan_seconds": round(plan_s, 2),
        "roofline": {"bound": "hbm", "kernel": "%s (%s stage)" % (kernels64.get(dom), dom),
                     "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "traffic_source": ("rocprofv3 FETCH_SIZE/WRITE_SIZE of this kernel, "
                                        "profiles/hbm_traffic_r2c_c2r_1024cube.md"
                                        if traffic else "kernel not in the profiled table"),
                     "frac_of_pattern_copy": (round(achieved / pattern_gbs[dom], 4)
                                              if dom in pattern_gbs else None),
                     "algorithmic_bytes": sb[dom]},
        "variants": variants,
    }
    # the other configurations, each timed as its own --workload line does
    del x, spec, spec64, eng
    torch.cuda.empty_cache()
    if not args.no_configs:
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup = min(args.steps, 10), min(args.warmup, 3)
        cfgs = {}
        for key, w in (("C1_c2c128", "c2c128"), ("C2_r2c512", "r2c512"),
                       ("C3a_r2cz720", "r2cz720"), ("C3b_pack2048", "pack2048"),
                       ("C4_rank_4x2_fused", "rank1024")):
            cfgs[key] = config_summary(bench_config(sub, torch, d2d, w))
            torch.cuda.empty_cache()
        for prec in (64, 32):
            cfgs["C5_batch64_f%d" % prec] = config_summary(bench_batch(sub, torch, d2d, prec))
            torch.cuda.empty_cache()
        cfgs["C5_mixed_with_reference_grids_f64"] = config_summary(
            bench_batch(sub, torch, d2d, 64, mixed=True))
        torch.cuda.empty_cache()
        line["configs"] = cfgs
    if not args.no_cpu:
        try:
            line["cpu_baseline"] = cpu_baseline(n, args.cpu_lines)
        except Exception as exc:  # report, do not hide
            line["cpu_baseline"] = {"error": repr(exc)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
