"""The C ABI library on the CPU: it loads, exports every function that
include/d2dhip.h declares, rejects bad arguments before touching the GPU,
and its host-side input generators follow their published definitions."""
import ctypes
import subprocess

import numpy as np
import pytest


def test_library_exports_every_header_function(d2d):
    names = d2d.header_functions()
    assert len(names) >= 20
    L = d2d.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", d2d.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert set(names) <= exported


def test_status_strings(d2d):
    assert d2d.status_string(0) == "success"
    assert "skipped" in d2d.status_string(d2d.WARN_SKIP_IGNORED)
    assert "cache" in d2d.status_string(d2d.WARN_TUNE_CACHE)
    assert d2d.status_string(12345) == "unknown status"
    assert d2d.lib().d2dhip_version() >= 100


def _create(d2d, fmt, nx, ny, nz, prec=64, flags=0):
    p = ctypes.c_void_p()
    st = d2d.lib().d2dhip_plan_create(ctypes.byref(p), fmt, nx, ny, nz, prec, flags, None)
    return st, p


@pytest.mark.parametrize("args,msg", [
    ((1, 0, 4, 4), "Invalid value for nx"),
    ((1, 4, -1, 4), "Invalid value for ny"),
    ((1, 4, 4, 0), "Invalid value for nz"),
    ((2, 4, 4, 4), "Invalid value for pencil"),
])
def test_plan_argument_errors(d2d, args, msg):
    """The reference's checks of decomp_2d_fft_engine_init
    (src/fft_common.f90:157-163, :215) become status codes."""
    st, p = _create(d2d, *args)
    assert st == d2d.ERR_INVALID_VALUE
    assert d2d.last_message() == msg
    assert not p.value


def test_plan_bad_precision(d2d):
    st, _ = _create(d2d, 1, 8, 8, 8, prec=48)
    assert st == d2d.ERR_INVALID_VALUE


def test_null_arguments(d2d):
    L = d2d.lib()
    assert L.d2dhip_plan_create(None, 1, 4, 4, 4, 64, 0, None) == d2d.ERR_NULL
    assert L.d2dhip_exec_c2c(None, None, None, -1) == d2d.ERR_NULL
    assert L.d2dhip_exec_r2c(None, None, None) == d2d.ERR_NULL
    assert L.d2dhip_exec_c2r(None, None, None) == d2d.ERR_NULL
    three = (ctypes.c_int * 3)(2, 2, 2)
    assert L.d2dhip_plan_set_dtt(None, 3, three) == d2d.ERR_NULL
    assert L.d2dhip_plan_get_dtt(None, (ctypes.c_int * 15)()) == d2d.ERR_NULL
    assert L.d2dhip_exec_dtt(None, None, None, -1) == d2d.ERR_NULL
    assert L.d2dhip_plan_destroy(None) == 0
    d = d2d.decomp_info_init(8, 8, 8)
    assert L.d2dhip_pack(0, 0, None, None, ctypes.byref(d), 0, None) == d2d.ERR_NULL
    assert L.d2dhip_transpose(7, 0, 1, 1, ctypes.byref(d), None) == d2d.ERR_INVALID_VALUE


def _splitmix_np(key, idx):
    """SplitMix64 (Steele, Lea & Flood 2014), idx-th output of the stream
    seeded with key, in numpy's wrapping uint64 arithmetic."""
    with np.errstate(over="ignore"):
        z = np.uint64(key) + (idx.astype(np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def test_splitmix_known_answer():
    """First outputs of SplitMix64 seeded with 0 (the reference
    implementation's published sequence)."""
    got = _splitmix_np(0, np.arange(3))
    assert [int(v) for v in got] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


@pytest.mark.parametrize("elem", [0, 1, 2, 3])
def test_host_uniform_generator(d2d, elem):
    n = 1000
    dt = {0: np.float64, 1: np.complex128, 2: np.float32, 3: np.complex64}[elem]
    a = np.zeros(n, dt)
    d2d.host_fill_uniform(a, elem, 42)
    nv = 2 * n if elem in (1, 3) else n
    u = (_splitmix_np(42, np.arange(nv)) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    want = 2 * u - 1
    flat = a.view(np.float64 if elem in (0, 1) else np.float32)
    np.testing.assert_array_equal(flat, want.astype(flat.dtype))
    assert flat.min() >= -1 and flat.max() < 1


def test_modal_generator_closed_form(d2d):
    nx, ny, nz = 12, 10, 8
    m = d2d.modes_make(5, 3, 0, 0.0, 0.0, 0.0, 1)
    assert m.nmodes == 5
    kx = np.array(m.kx[:5]); ky = np.array(m.ky[:5]); kz = np.array(m.kz[:5])
    assert np.all(np.abs(np.concatenate([kx, ky, kz])) <= 3)
    a = np.zeros((nx, ny, nz), np.float64, order="F")
    d2d.host_fill_modal(a, nx, ny, nz, 0, m)
    i = np.arange(nx)[:, None, None]; j = np.arange(ny)[None, :, None]; k = np.arange(nz)[None, None, :]
    want = sum(m.amp[q] * np.cos(2 * np.pi * (kx[q] * i / nx + ky[q] * j / ny + kz[q] * k / nz))
               for q in range(5))
    assert np.abs(a - want).max() < 1e-13
    # its r2c spectrum is a sum of peaks amp*N/2 at +-k (closed form)
    s = np.fft.fftn(a)
    peaks = np.zeros((nx, ny, nz), complex)
    for q in range(5):
        peaks[kx[q] % nx, ky[q] % ny, kz[q] % nz] += m.amp[q] * nx * ny * nz / 2
        peaks[-kx[q] % nx, -ky[q] % ny, -kz[q] % nz] += m.amp[q] * nx * ny * nz / 2
    assert np.abs(s - peaks).max() < 1e-10


def test_modal_log_amplitudes_and_noise(d2d):
    m = d2d.modes_make(32, 128, 1, -6.0, 2.0, 1e-6, 3)
    amps = np.abs(np.array(m.amp[:32]))
    assert amps.min() >= 1e-6 and amps.max() <= 1e2
    assert m.sigma == 1e-6
    a = np.zeros((16, 8, 4), np.float64, order="F")
    d2d.host_fill_modal(a, 16, 8, 4, 0, m)
    assert np.isfinite(a).all()


def test_profiler_regions_host_api(d2d):
    """decomp_profiler_prep / _start / _end (src/profiler_caliper.f90:138-177)
    on the host: nested caller regions open and close, closing with none
    open is an error, and the category switches take the reference's four
    flags."""
    d2d.profiler_prep(transpose=False, io=False, fft=True, d2d=False)
    d2d.profiler_start("outer")
    d2d.profiler_start("inner")
    d2d.profiler_end("inner")
    d2d.profiler_end("outer")
    L = d2d.lib()
    assert L.d2dhip_profiler_end(b"none") == d2d.ERR_INVALID_VALUE
    assert "no region" in d2d.last_message()
    assert L.d2dhip_profiler_start(None) == d2d.ERR_NULL
    # host-only calls with regions on: the metadata and a refused transpose
    d2d.profiler_prep()
    d = d2d.decomp_info_init(17, 13, 11, 2, 3, 1, 2)
    assert L.d2dhip_transpose(0, d2d.R8, ctypes.c_void_p(8), ctypes.c_void_p(16),
                              ctypes.byref(d), None) == d2d.ERR_UNSUPPORTED
