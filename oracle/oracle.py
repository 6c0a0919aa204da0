"""ctypes front-end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, and only there as the checker.  The shipped
library never imports it.

Arrays follow the reference's Fortran layout: a pencil of global shape
(nx, ny, nz) is a numpy array of shape (nx, ny, nz) in Fortran order, i.e.
index i is fastest in memory.  Complex arrays are complex128.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PHYSICAL_IN_X = 1
PHYSICAL_IN_Z = 3


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        L = ctypes.CDLL(path)
        dp = ctypes.c_void_p
        ci = ctypes.c_int
        L.oracle_spcfft.argtypes = [dp, ci, ci, dp]
        for name in ("oracle_fft3d_c2c",):
            getattr(L, name).argtypes = [ci, ci, ci, ci, dp, dp, ci, dp]
            getattr(L, name).restype = ci
        for name in ("oracle_fft3d_r2c", "oracle_fft3d_c2r"):
            getattr(L, name).argtypes = [ci, ci, ci, ci, dp, dp, dp]
            getattr(L, name).restype = ci
        L.oracle_dft3d_ld.argtypes = [ci, ci, ci, dp, dp, ci]
        L.oracle_sample_stage_times.argtypes = [ci, ci, ci, ci, ci, ctypes.POINTER(ctypes.c_double)]
        L.oracle_sample_stage_times.restype = ci
        L.oracle_distribute.argtypes = [ci, ci, dp, dp, dp]
        L.oracle_distribute.restype = ci
        L.oracle_decomp_info_init.argtypes = [ci, ci, ci, ci, ci, ci, ci, dp]
        L.oracle_decomp_info_init.restype = ci
        L.oracle_mem_split.argtypes = [ci, dp, ci, ci, ci, dp, dp, ci, ci]
        L.oracle_mem_merge.argtypes = [ci, dp, ci, ci, ci, dp, dp, ci, ci]
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _fortran(a, dtype):
    return np.asfortranarray(a, dtype=dtype)


def spcfft(u, isign):
    """Glassman 1D FFT of a complex vector (src/glassman.f90:29-67)."""
    u = np.array(u, dtype=np.complex128)
    work = np.empty_like(u)
    lib().oracle_spcfft(_ptr(u), u.size, int(isign), _ptr(work))
    return u


def _skip_arr(skip):
    if skip is None:
        return None
    return (ctypes.c_int * 3)(*[int(bool(s)) for s in skip])


def fft3d_c2c(x, isign, fmt=PHYSICAL_IN_X, skip=None):
    """One-rank decomp_2d_fft_3d c2c (src/fft_common_3d.f90:9-116)."""
    x = _fortran(x, np.complex128)
    nx, ny, nz = x.shape
    out = np.empty((nx, ny, nz), dtype=np.complex128, order="F")
    rc = lib().oracle_fft3d_c2c(fmt, nx, ny, nz, _ptr(x), _ptr(out), int(isign), _skip_arr(skip))
    if rc:
        raise ValueError("oracle_fft3d_c2c failed: %d" % rc)
    return out


def spectral_shape(fmt, nx, ny, nz):
    """Shape of the r2c output (src/fft_common.f90:210-216)."""
    return (nx // 2 + 1, ny, nz) if fmt == PHYSICAL_IN_X else (nx, ny, nz // 2 + 1)


def fft3d_r2c(x, fmt=PHYSICAL_IN_X, skip=None):
    """One-rank decomp_2d_fft_3d r2c (src/fft_common_3d.f90:121-189)."""
    x = _fortran(x, np.float64)
    nx, ny, nz = x.shape
    out = np.empty(spectral_shape(fmt, nx, ny, nz), dtype=np.complex128, order="F")
    rc = lib().oracle_fft3d_r2c(fmt, nx, ny, nz, _ptr(x), _ptr(out), _skip_arr(skip))
    if rc:
        raise ValueError("oracle_fft3d_r2c failed: %d" % rc)
    return out


def fft3d_c2r(c, shape, fmt=PHYSICAL_IN_X, skip=None):
    """One-rank decomp_2d_fft_3d c2r (src/fft_common_3d.f90:194-296)."""
    nx, ny, nz = shape
    c = _fortran(c, np.complex128)
    assert c.shape == spectral_shape(fmt, nx, ny, nz)
    out = np.empty((nx, ny, nz), dtype=np.float64, order="F")
    rc = lib().oracle_fft3d_c2r(fmt, nx, ny, nz, _ptr(c), _ptr(out), _skip_arr(skip))
    if rc:
        raise ValueError("oracle_fft3d_c2r failed: %d" % rc)
    return out


def dft3d_ld(x, isign):
    """Direct 3D DFT accumulated in long double (tiny grids only)."""
    x = _fortran(x, np.complex128)
    out = np.empty_like(x, order="F")
    nx, ny, nz = x.shape
    lib().oracle_dft3d_ld(nx, ny, nz, _ptr(x), _ptr(out), int(isign))
    return out


def sample_stage_times(nx, ny, nz, lines, nthreads=0):
    t = (ctypes.c_double * 6)()
    used = lib().oracle_sample_stage_times(nx, ny, nz, lines, nthreads, t)
    if used < 0:
        raise MemoryError("oracle_sample_stage_times could not reserve the grid")
    return list(t), used


# ---- decomposition metadata -------------------------------------------------

class OracleDecomp(ctypes.Structure):
    _fields_ = [
        ("xst", ctypes.c_int * 3), ("xen", ctypes.c_int * 3), ("xsz", ctypes.c_int * 3),
        ("yst", ctypes.c_int * 3), ("yen", ctypes.c_int * 3), ("ysz", ctypes.c_int * 3),
        ("zst", ctypes.c_int * 3), ("zen", ctypes.c_int * 3), ("zsz", ctypes.c_int * 3),
        ("p_row", ctypes.c_int), ("p_col", ctypes.c_int),
        ("x1dist", ctypes.c_int * 64), ("y1dist", ctypes.c_int * 64),
        ("y2dist", ctypes.c_int * 64), ("z2dist", ctypes.c_int * 64),
        ("x1cnts", ctypes.c_int64 * 64), ("y1cnts", ctypes.c_int64 * 64),
        ("x1disp", ctypes.c_int64 * 64), ("y1disp", ctypes.c_int64 * 64),
        ("y2cnts", ctypes.c_int64 * 64), ("z2cnts", ctypes.c_int64 * 64),
        ("y2disp", ctypes.c_int64 * 64), ("z2disp", ctypes.c_int64 * 64),
        ("x1count", ctypes.c_int64), ("y1count", ctypes.c_int64),
        ("y2count", ctypes.c_int64), ("z2count", ctypes.c_int64),
        ("even", ctypes.c_int),
    ]


def distribute(data1, proc):
    st = np.zeros(proc, np.int32)
    en = np.zeros(proc, np.int32)
    sz = np.zeros(proc, np.int32)
    rc = lib().oracle_distribute(data1, proc, _ptr(st), _ptr(en), _ptr(sz))
    if rc:
        raise ValueError("distribute failed: %d" % rc)
    return st, en, sz


def decomp_info(nx, ny, nz, p_row, p_col, coord_row=0, coord_col=0):
    d = OracleDecomp()
    rc = lib().oracle_decomp_info_init(nx, ny, nz, p_row, p_col, coord_row, coord_col,
                                       ctypes.byref(d))
    if rc:
        raise ValueError("decomp_info_init failed: %d" % rc)
    return d


XY, YZ, ZY, YX = 0, 1, 2, 3


def mem_split(direction, src, d, even=False):
    """Send buffer of one rank (mem_split_* of src/transpose_*.f90)."""
    src = np.asfortranarray(src)
    n1, n2, n3 = src.shape
    nbuf = max(src.size, _padded_size(direction, d, 1) if even else 0)
    out = np.zeros(nbuf, dtype=src.dtype)
    lib().oracle_mem_split(direction, _ptr(src), n1, n2, n3, _ptr(out), ctypes.byref(d),
                           src.itemsize, int(bool(even)))
    return out


def mem_merge(direction, buf, dst_shape, dtype, d, even=False):
    """Receive-buffer unpack of one rank (mem_merge_* of src/transpose_*.f90)."""
    buf = np.ascontiguousarray(buf, dtype=dtype)
    n1, n2, n3 = dst_shape
    out = np.zeros(dst_shape, dtype=dtype, order="F")
    lib().oracle_mem_merge(direction, _ptr(buf), n1, n2, n3, _ptr(out), ctypes.byref(d),
                           out.itemsize, int(bool(even)))
    return out


def _padded_size(direction, d, split):
    if direction in (XY, YX):
        return int(d.x1count) * d.p_row
    return int(d.y2count) * d.p_col


# ---- real-to-real transforms (DCT / DST) ----------------------------------------
# decomp_2d_dtt_3d_r2r (src/fft_fftw3_f03.f90:1959-2023) runs FFTW r2r plans
# (r2r_1m_x/y/z_plan, :1208-1322) over the points [ifirst, ifirst + n -
# ndismiss) of every line (wrapper_r2r, :2071-2098).  FFTW 3.3.10 is not in
# this image; its r2r kinds are restated from their published definitions
# (FFTW manual, "What FFTW Really Computes", 1d real-even / real-odd DFTs),
# unnormalised, as direct sums with the angles reduced exactly in integers.

DTT_NAMES = {1: "DCT-I", 2: "DCT-II", 3: "DCT-III", 4: "DCT-IV",
             5: "DST-I", 6: "DST-II", 7: "DST-III", 8: "DST-IV"}


def dtt_matrix(t, L):
    """Y = A @ X for DTT type t (1-8 = REDFT00, 10, 01, 11, RODFT00, 10, 01, 11)."""
    j = np.arange(L)[None, :].astype(np.int64)
    k = np.arange(L)[:, None].astype(np.int64)
    if t == 1:  # X_0 + (-1)^k X_{L-1} + 2 sum_{j=1}^{L-2} X_j cos(pi j k / (L-1))
        if L < 2:
            raise ValueError("DCT-I needs 2 or more points")
        P = 2 * (L - 1)
        A = 2 * np.cos(np.pi * ((j * k) % P) / (L - 1))
        A[:, 0] = 1.0
        A[:, L - 1] = np.where(k[:, 0] % 2 == 0, 1.0, -1.0)
        return A
    if t == 2:  # 2 sum X_j cos(pi (2j+1) k / (2L))
        return 2 * np.cos(np.pi * (((2 * j + 1) * k) % (4 * L)) / (2 * L))
    if t == 3:  # X_0 + 2 sum_{j>=1} X_j cos(pi j (2k+1) / (2L))
        A = 2 * np.cos(np.pi * ((j * (2 * k + 1)) % (4 * L)) / (2 * L))
        A[:, 0] = 1.0
        return A
    if t == 4:  # 2 sum X_j cos(pi (2j+1)(2k+1) / (4L))
        return 2 * np.cos(np.pi * (((2 * j + 1) * (2 * k + 1)) % (8 * L)) / (4 * L))
    if t == 5:  # 2 sum X_j sin(pi (j+1)(k+1) / (L+1))
        return 2 * np.sin(np.pi * (((j + 1) * (k + 1)) % (2 * (L + 1))) / (L + 1))
    if t == 6:  # 2 sum X_j sin(pi (2j+1)(k+1) / (2L))
        return 2 * np.sin(np.pi * (((2 * j + 1) * (k + 1)) % (4 * L)) / (2 * L))
    if t == 7:  # (-1)^k X_{L-1} + 2 sum_{j<L-1} X_j sin(pi (j+1)(2k+1) / (2L))
        A = 2 * np.sin(np.pi * (((j + 1) * (2 * k + 1)) % (4 * L)) / (2 * L))
        A[:, L - 1] = np.where(k[:, 0] % 2 == 0, 1.0, -1.0)
        return A
    if t == 8:  # 2 sum X_j sin(pi (2j+1)(2k+1) / (4L))
        return 2 * np.sin(np.pi * (((2 * j + 1) * (2 * k + 1)) % (8 * L)) / (4 * L))
    raise ValueError("DTT type must be 1-8")


def dtt_config(dtt):
    """engine%dtt from opt_DTT: dtt_assign_default (src/fft_fftw3_f03.f90:394-420)
    for 3 values, then dtt_invert (:424-444).  Types stay 1-8."""
    v = [int(a) for a in dtt]
    if len(v) not in (3, 12) or min(v[:3]) < 1 or max(v[:3]) > 8:
        raise ValueError("Invalid argument")
    if len(v) == 3:
        ifirst = [2 if t in (5, 7) else 1 for t in v]
        ndismiss = [0 if t == 1 else (2 if t == 5 else 1) for t in v]
        v = v + ifirst + ndismiss + ifirst
    inv = {2: 3, 3: 2, 6: 7, 7: 6}
    return tuple(v + [inv.get(t, t) for t in v[:3]])


def dtt_1d(x, axis, t, first, L):
    """Type-t DTT of points [first, first + L) (0-based) of every line along
    axis; the other points are kept (FFTW's r2r on the sub-array)."""
    y = np.array(x, dtype=np.float64, order="F", copy=True)
    A = dtt_matrix(t, L)
    sl = [slice(None)] * 3
    sl[axis] = slice(first, first + L)
    seg = np.moveaxis(y[tuple(sl)], axis, 0)
    res = np.tensordot(A, seg, axes=(1, 0))
    y[tuple(sl)] = np.moveaxis(res, 0, axis)
    return y


def dtt3d(x, dtt, isign, fmt=PHYSICAL_IN_X, skip=None):
    """One-rank decomp_2d_dtt_3d_r2r (src/fft_fftw3_f03.f90:1959-2023): forward
    types and ifirst, or the inverse types and ofirst for isign = +1; x, y, z
    for PHYSICAL_IN_X forward / PHYSICAL_IN_Z backward, else z, y, x."""
    c = dtt_config(dtt)
    y = np.asfortranarray(x, dtype=np.float64)
    n = y.shape
    x_first = (fmt == PHYSICAL_IN_X) == (isign == -1)
    for ax in ((0, 1, 2) if x_first else (2, 1, 0)):
        if skip is not None and skip[ax]:
            continue
        t = c[ax] if isign == -1 else c[12 + ax]
        first = (c[3 + ax] if isign == -1 else c[9 + ax]) - 1
        y = dtt_1d(y, ax, t, first, n[ax] - c[6 + ax])
    return y


def dtt_norm(dtt, shape):
    """Round-trip factor of examples/dtt/fftw_f03_dtt_x.f90:124-153: 2n-2 for
    DCT-I and DCT-IV, 2n+2-2*ndismiss for DST-I, 2n-2*ndismiss otherwise."""
    c = dtt_config(dtt)
    f = 1.0
    for ax in range(3):
        t, n, nd = c[ax], shape[ax], c[6 + ax]
        if t in (1, 4):
            f *= 2 * n - 2
        elif t == 5:
            f *= 2 * n + 2 - 2 * nd
        else:
            f *= 2 * n - 2 * nd
    return f


# ---- halo cells ---------------------------------------------------------------
# update_halo / halo_exchange at one rank (src/halo.f90:112-260,
# src/halo_common.f90, src/halo_exchange_{x,y,z}_body.f90): the pencil goes
# into the interior of the halo array, then the two distributed axes are
# exchanged in the order of the body files (x-pencil: y then z; y-pencil: x
# then z; z-pencil: x then y), each over the whole extent of the other axes.
# MPI_CART_SHIFT on a communicator of size 1 returns the rank itself for a
# periodic axis (the periodic_bc flags of decomp_2d_init,
# src/decomp_2d_init_fin.f90:97-112) and MPI_PROC_NULL otherwise, so a
# periodic axis receives its own opposite interior slabs and a
# non-periodic one keeps whatever its halo cells held.

HALO_ORDER = {1: (1, 2), 2: (0, 2), 3: (0, 1)}


def halo_shape(sizes, level, ipencil):
    return tuple(s + (2 * level if a != ipencil - 1 else 0) for a, s in enumerate(sizes))


def halo_exchange(arr, level, ipencil, periodic):
    """In place on a halo array (numpy, Fortran order)."""
    for d in HALO_ORDER[ipencil]:
        if not periodic[d] or level == 0:
            continue
        n = arr.shape[d]
        lo = [slice(None)] * 3
        src = [slice(None)] * 3
        lo[d] = slice(0, level)
        src[d] = slice(n - 2 * level, n - level)   # MPI_IRECV(low halo) <- neighbour's last slabs
        hi = [slice(None)] * 3
        src2 = [slice(None)] * 3
        hi[d] = slice(n - level, n)
        src2[d] = slice(level, 2 * level)           # high halo <- neighbour's first slabs
        a = arr[tuple(src)].copy()
        b = arr[tuple(src2)].copy()
        arr[tuple(lo)] = a
        arr[tuple(hi)] = b
    return arr


def update_halo(x, level, ipencil, periodic, out_init=None):
    """update_halo(in, out, level, opt_pencil=ipencil) at one rank; out_init is
    what the halo array held before (the reference allocates it uninitialised)."""
    x = np.asfortranarray(x)
    shape = halo_shape(x.shape, level, ipencil)
    out = (np.zeros(shape, dtype=x.dtype, order="F") if out_init is None
           else np.array(out_init, dtype=x.dtype, order="F", copy=True))
    sl = tuple(slice(level, level + x.shape[a]) if a != ipencil - 1 else slice(None)
               for a in range(3))
    out[sl] = x
    return halo_exchange(out, level, ipencil, periodic)
