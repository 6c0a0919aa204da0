"""The precision configuration at its full size: r2c / c2r of a 1024^3
x-pencil (PHYSICAL_IN_X) in FP64, FP32 and split-BF16.

The reference's own check of this use is the round trip of
examples/fft_physical_x/fft_r2c_x.f90:131-152 (forward, backward, divide by
nx*ny*nz, compare with the input).  At 1024^3 no oracle run fits in a test,
so the spectrum is held to the closed form of a noise-free modal field
instead: a field sum_m a_m cos(2 pi k_m . x / N) has the spectrum a_m*N/2 at
+k_m and -k_m and zero elsewhere.  The field is the generator of the
benchmark workload without its noise (32 modes, |k| <= 128, |a| = 10^U(-6,2)
with random signs, seed 3); the noisy field (sigma 1e-6) is then round
tripped.

Tolerances (relative L2): FP64 1e-13 against the closed form and on the
round trip; FP32 and split-BF16 1e-6 against the closed form of the FP64
field (the FP32 rounding of the field itself contributes about 4e-8) and on
the round trip.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 1024
F64_TOL = 1e-13
LOW_TOL = 1e-6


def _peaks(modes, n):
    """{(i, j, k) in the half spectrum: closed-form value} of a modal field."""
    nxc = n // 2 + 1
    out = {}
    for q in range(modes.nmodes):
        for sgn in (1, -1):
            key = ((sgn * modes.kx[q]) % n, (sgn * modes.ky[q]) % n, (sgn * modes.kz[q]) % n)
            out[key] = out.get(key, 0.0) + modes.amp[q] * n ** 3 / 2
    return {k: v for k, v in out.items() if k[0] < nxc}


def _closed_form_error(torch, spec, peaks, n):
    """Relative L2 distance of a (n/2+1, n, n) spectrum (device tensor, any
    complex dtype) from its closed form, computed one z-slab at a time so
    that no full-size FP64 copy of an FP32 spectrum is needed."""
    nxc = n // 2 + 1
    o = spec.view(n, n, nxc)  # Fortran (i, j, k) -> torch [k][j][i]
    by_k = {}
    for (i, j, k), v in peaks.items():
        by_k.setdefault(k, []).append((i, j, v))
    err2 = 0.0
    slab = 64
    for k0 in range(0, n, slab):
        d = o[k0:k0 + slab].to(torch.complex128)
        for k in range(k0, min(n, k0 + slab)):
            for i, j, v in by_k.get(k, ()):
                d[k - k0, j, i] -= v
        err2 += float(torch.linalg.vector_norm(d)) ** 2
    ref = np.sqrt(sum(abs(v) ** 2 for v in peaks.values()))
    return np.sqrt(err2) / ref


@pytest.fixture(scope="module")
def fields(d2d, torch_gpu):
    torch = torch_gpu
    clean = d2d.modes_make(32, 128, 1, -6.0, 2.0, 0.0, 3)
    noisy = d2d.modes_make(32, 128, 1, -6.0, 2.0, 1e-6, 3)
    # the generator draws the modes before the noise stream: same modes
    assert list(clean.kx[:32]) == list(noisy.kx[:32])
    assert list(clean.amp[:32]) == list(noisy.amp[:32])
    return clean, noisy


@pytest.mark.parametrize("prec", ["f64", "f32", "split_bf16"])
def test_c4_r2c_closed_form_and_round_trip(d2d, torch_gpu, fields, prec):
    torch = torch_gpu
    clean, noisy = fields
    p = {"f64": d2d.PREC_F64, "f32": d2d.PREC_F32, "split_bf16": d2d.PREC_SPLIT_BF16}[prec]
    rdt, cdt = (torch.float64, torch.complex128) if prec == "f64" else (torch.float32, torch.complex64)
    elem = d2d.R8 if prec == "f64" else d2d.R4
    tol = F64_TOL if prec == "f64" else LOW_TOL
    nxc = N // 2 + 1
    eng = d2d.FFTEngine(d2d.PHYSICAL_IN_X, N, N, N, prec=p)
    kern = eng.stage_kernels()
    if prec == "split_bf16":  # the 1024-point axes run on the matrix cores
        assert all(kern[s].startswith("mfma") for s in ("r2c", "spec_y", "spec_last", "c2r")), kern
    x = torch.empty(N ** 3, dtype=rdt, device="cuda")
    spec = torch.empty(nxc * N * N, dtype=cdt, device="cuda")
    d2d.fill_modal(x, N, N, N, elem, clean)
    eng.fft_r2c(x, spec)
    torch.cuda.synchronize()
    err = _closed_form_error(torch, spec, _peaks(clean, N), N)
    print("C4 %s: closed-form rel L2 %.3e" % (prec, err))
    assert err < tol
    # noisy field: forward, backward, divide by N^3 (fft_r2c_x.f90:131-152);
    # the spectrum is kept (c2r preserves its input)
    d2d.fill_modal(x, N, N, N, elem, noisy)
    y = torch.empty_like(x)
    eng.fft_r2c(x, spec)
    keep = spec[:: 4099].clone()
    eng.fft_c2r(spec, y)
    torch.cuda.synchronize()
    assert torch.equal(spec[:: 4099], keep)
    y /= N ** 3
    rt = float(torch.linalg.vector_norm((y - x).double()) / torch.linalg.vector_norm(x.double()))
    print("C4 %s: round-trip rel L2 %.3e" % (prec, rt))
    assert rt < tol
