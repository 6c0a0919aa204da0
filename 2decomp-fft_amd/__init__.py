"""Python binding of libd2dhip, the gfx950 FFT backend of the 2DECOMP&FFT
single-GPU 3D FFT path.

The C ABI (include/d2dhip.h) is the product; this module only binds it with
ctypes so that tests and benchmarks can drive it.  Device memory comes from
PyTorch tensors (their data_ptr) and work is queued on PyTorch's current
stream.  Nothing here computes a transform on the host: if libd2dhip.so is
missing the import of the binding fails.

Names follow the reference's Fortran API (src/fft_common.f90:31-53,
src/decomp_2d.f90:98-309): ``decomp_2d_fft_init`` returns a plan whose
``fft_3d`` methods mirror ``decomp_2d_fft_3d``; ``transpose_*`` and the pack
entry points mirror the reference's transposes and mem_split / mem_merge.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libd2dhip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "d2dhip.h")

PHYSICAL_IN_X = 1
PHYSICAL_IN_Z = 3
FORWARD = -1
BACKWARD = 1
PREC_F64 = 64
PREC_F32 = 32
PREC_SPLIT_BF16 = 16
BATCH_PERSISTENT = 0
BATCH_GROUPED = 1
BATCH_PERSISTENT_GENERAL = 2
FLAG_INPLACE = 0x1
FLAG_SKIP_X = 0x2
FLAG_SKIP_Y = 0x4
FLAG_SKIP_Z = 0x8
FLAG_TUNE = 0x10
LAYOUT_PLAIN, LAYOUT_BLOCKED, LAYOUT_BLOCKED32 = 0, 1, 2
R8, C16, R4, C8 = 0, 1, 2, 3
X_TO_Y, Y_TO_Z, Z_TO_Y, Y_TO_X = 0, 1, 2, 3
SUCCESS = 0
ERR_INVALID_VALUE = 1
ERR_UNSUPPORTED = 2
ERR_ALLOC = 3
ERR_HIP = 4
ERR_NULL = 5
ERR_GRID = 6
WARN_SKIP_IGNORED = 100
WARN_TUNE_CACHE = 101
SCRATCH_BYTES = 64 << 20  # D2DHIP_SCRATCH_BYTES
MAX_PROCS = 64

_lib = None


class D2dError(RuntimeError):
    """A non-zero status of the C ABI (the reference would decomp_2d_abort)."""

    def __init__(self, status, message):
        super().__init__("d2dhip status %d (%s): %s" % (status, status_string(status), message))
        self.status = status


class D2dWarning(UserWarning):
    pass


class Decomp(ctypes.Structure):
    """Mirror of d2dhip_decomp (DECOMP_INFO, src/info.f90:19-47)."""
    _fields_ = [
        ("xst", ctypes.c_int * 3), ("xen", ctypes.c_int * 3), ("xsz", ctypes.c_int * 3),
        ("yst", ctypes.c_int * 3), ("yen", ctypes.c_int * 3), ("ysz", ctypes.c_int * 3),
        ("zst", ctypes.c_int * 3), ("zen", ctypes.c_int * 3), ("zsz", ctypes.c_int * 3),
        ("p_row", ctypes.c_int), ("p_col", ctypes.c_int),
        ("coord_row", ctypes.c_int), ("coord_col", ctypes.c_int),
        ("x1dist", ctypes.c_int * MAX_PROCS), ("y1dist", ctypes.c_int * MAX_PROCS),
        ("y2dist", ctypes.c_int * MAX_PROCS), ("z2dist", ctypes.c_int * MAX_PROCS),
        ("x1cnts", ctypes.c_longlong * MAX_PROCS), ("y1cnts", ctypes.c_longlong * MAX_PROCS),
        ("x1disp", ctypes.c_longlong * MAX_PROCS), ("y1disp", ctypes.c_longlong * MAX_PROCS),
        ("y2cnts", ctypes.c_longlong * MAX_PROCS), ("z2cnts", ctypes.c_longlong * MAX_PROCS),
        ("y2disp", ctypes.c_longlong * MAX_PROCS), ("z2disp", ctypes.c_longlong * MAX_PROCS),
        ("x1count", ctypes.c_longlong), ("y1count", ctypes.c_longlong),
        ("y2count", ctypes.c_longlong), ("z2count", ctypes.c_longlong),
        ("even", ctypes.c_int),
    ]


class Modes(ctypes.Structure):
    _fields_ = [
        ("nmodes", ctypes.c_int),
        ("kx", ctypes.c_int * 64), ("ky", ctypes.c_int * 64), ("kz", ctypes.c_int * 64),
        ("amp", ctypes.c_double * 64),
        ("sigma", ctypes.c_double),
        ("noise_seed", ctypes.c_ulonglong),
    ]


def lib():
    """Loads libd2dhip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libd2dhip.so is not built: run `python 2decomp-fft_amd/build.py`")
        L = ctypes.CDLL(LIB_PATH)
        vp, ci, cu, cll = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_longlong
        pl = ctypes.c_void_p
        sig = {
            "d2dhip_status_string": ([ci], ctypes.c_char_p),
            "d2dhip_last_message": ([], ctypes.c_char_p),
            "d2dhip_version": ([], ci),
            "d2dhip_profiler_prep": ([ci, ci, ci, ci], ci),
            "d2dhip_profiler_start": ([ctypes.c_char_p], ci),
            "d2dhip_profiler_end": ([ctypes.c_char_p], ci),
            "d2dhip_decomp_info_init": ([ci, ci, ci, ci, ci, ci, ci, ctypes.POINTER(Decomp)], ci),
            "d2dhip_best_2d_grid": ([ci, ctypes.POINTER(ci), ctypes.POINTER(ci)], ci),
            "d2dhip_plan_create": ([ctypes.POINTER(pl), ci, ci, ci, ci, ci, cu, vp], ci),
            "d2dhip_plan_create_rank": ([ctypes.POINTER(pl), ci, ci, ci, ci, ci, ci, ci, ci, ci,
                                         cu, vp], ci),
            "d2dhip_plan_get_rank": ([pl, ctypes.POINTER(Decomp), ctypes.POINTER(Decomp)], ci),
            "d2dhip_exec_1m": ([pl, ci, vp, vp, ci], ci),
            "d2dhip_exec_1m_pack": ([pl, ci, vp, vp, ci, ctypes.POINTER(Decomp), ci, ci], ci),
            "d2dhip_exec_1m_unpack": ([pl, ci, vp, vp, ci, ctypes.POINTER(Decomp), ci, ci], ci),
            "d2dhip_exec_1m_unpack_pack": ([pl, ci, vp, vp, ci, ci, ctypes.POINTER(Decomp), ci, ci], ci),
            "d2dhip_plan_destroy": ([pl], ci),
            "d2dhip_plan_set_stream": ([pl, vp], ci),
            "d2dhip_plan_get_sizes": ([pl, ctypes.POINTER(ci), ctypes.POINTER(ci)], ci),
            "d2dhip_plan_device_bytes": ([pl], cll),
            "d2dhip_plan_stage_kernel": ([pl, ci, ctypes.c_char_p, ci], ci),
            "d2dhip_plan_stage_candidate": ([pl, ci, ci, ctypes.c_char_p, ci], ci),
            "d2dhip_plan_set_stage_kernel": ([pl, ci, ctypes.c_char_p], ci),
            "d2dhip_plan_set_chunking": ([pl, ci, ci], ci),
            "d2dhip_plan_get_chunking": ([pl, ctypes.POINTER(ci), ctypes.POINTER(ci)], ci),
            "d2dhip_plan_set_layout": ([pl, ci], ci),
            "d2dhip_plan_get_layout": ([pl, ctypes.POINTER(ci)], ci),
            "d2dhip_plan_profile": ([pl, ci], ci),
            "d2dhip_plan_profile_report": ([pl, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ci)], ci),
            "d2dhip_exec_c2c": ([pl, vp, vp, ci], ci),
            "d2dhip_exec_r2c": ([pl, vp, vp], ci),
            "d2dhip_exec_c2r": ([pl, vp, vp], ci),
            "d2dhip_exec_c2r_noalloc": ([pl, vp, vp], ci),
            "d2dhip_plan_set_dtt": ([pl, ci, ctypes.POINTER(ci)], ci),
            "d2dhip_plan_get_dtt": ([pl, ctypes.POINTER(ci)], ci),
            "d2dhip_exec_dtt": ([pl, vp, vp, ci], ci),
            "d2dhip_exec_r2c_pack": ([pl, vp, vp, ctypes.POINTER(Decomp), ci], ci),
            "d2dhip_plan_capture_begin": ([pl], ci),
            "d2dhip_plan_capture_end": ([pl], ci),
            "d2dhip_plan_graph_launch": ([pl], ci),
            "d2dhip_exec_c2r_unpack": ([pl, vp, vp, ctypes.POINTER(Decomp), ci], ci),
            "d2dhip_transpose": ([ci, ci, vp, vp, ctypes.POINTER(Decomp), vp], ci),
            "d2dhip_batch_create": ([ctypes.POINTER(vp), ci, ctypes.POINTER(vp), vp], ci),
            "d2dhip_batch_bind_c2c": ([vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ci], ci),
            "d2dhip_batch_launch": ([vp], ci),
            "d2dhip_batch_capture_graph": ([vp], ci),
            "d2dhip_batch_graph_launch": ([vp], ci),
            "d2dhip_batch_check": ([vp], ci),
            "d2dhip_batch_set_stream": ([vp, vp], ci),
            "d2dhip_batch_set_mode": ([vp, ci], ci),
            "d2dhip_batch_set_cluster_bytes": ([vp, cll], ci),
            "d2dhip_batch_get_clusters": ([vp, ctypes.POINTER(ci), ctypes.POINTER(ci)], ci),
            "d2dhip_batch_persistent_tiles": ([vp, ctypes.POINTER(ci)], ci),
            "d2dhip_batch_set_item_points": ([vp, cll], ci),
            "d2dhip_batch_destroy": ([vp], ci),
            "d2dhip_pack": ([ci, ci, vp, vp, ctypes.POINTER(Decomp), ci, vp], ci),
            "d2dhip_unpack": ([ci, ci, vp, vp, ctypes.POINTER(Decomp), ci, vp], ci),
            "d2dhip_pack_buffer_elems": ([ci, ctypes.POINTER(Decomp), ci], cll),
            "d2dhip_update_halo": ([ci, vp, vp, ci, ci, ci, ci, ci, ctypes.POINTER(ci), vp], ci),
            "d2dhip_halo_exchange": ([ci, vp, ci, ci, ci, ci, ci, ctypes.POINTER(ci), vp], ci),
            "d2dhip_fill_uniform": ([vp, cll, ci, ctypes.c_ulonglong, vp], ci),
            "d2dhip_host_fill_uniform": ([vp, cll, ci, ctypes.c_ulonglong], None),
            "d2dhip_column_copy": ([vp, vp, ci, cll, cll, cll, cll, ci, vp], ci),
            "d2dhip_modes_make": ([ctypes.POINTER(Modes), ci, ci, ci, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong], ci),
            "d2dhip_fill_modal": ([vp, ci, ci, ci, ci, ctypes.POINTER(Modes), vp], ci),
            "d2dhip_host_fill_modal": ([vp, ci, ci, ci, ci, ctypes.POINTER(Modes)], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def header_functions():
    """Names of the functions include/d2dhip.h declares."""
    import re
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(d2dhip_\w+)\s*\(", text)))


def status_string(status):
    return lib().d2dhip_status_string(status).decode()


def last_message():
    return lib().d2dhip_last_message().decode()


def check(status):
    if status == SUCCESS:
        return status
    if status >= 100:
        import warnings
        warnings.warn(last_message(), D2dWarning, stacklevel=3)
        return status
    raise D2dError(status, last_message())


def _stream(stream):
    if stream is not None:
        return ctypes.c_void_p(stream)
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    """Device address of a torch tensor (or a plain int address)."""
    if isinstance(t, int):
        return ctypes.c_void_p(t)
    return ctypes.c_void_p(t.data_ptr())


# ---- decomposition ----------------------------------------------------------

def decomp_info_init(nx, ny, nz, p_row=1, p_col=1, coord_row=0, coord_col=0):
    """DECOMP_INFO of one rank (decomp_info_init, src/decomp_2d.f90:382-490)."""
    d = Decomp()
    check(lib().d2dhip_decomp_info_init(nx, ny, nz, p_row, p_col, coord_row, coord_col,
                                        ctypes.byref(d)))
    return d


def best_2d_grid(nproc):
    r, c = ctypes.c_int(), ctypes.c_int()
    check(lib().d2dhip_best_2d_grid(nproc, ctypes.byref(r), ctypes.byref(c)))
    return r.value, c.value


# ---- profiler regions (decomp_2d_profiler, src/profiler_caliper.f90) ------------

def profiler_prep(transpose=True, io=True, fft=True, d2d=True):
    """decomp_profiler_prep (src/profiler_caliper.f90:138-150): the region
    categories that the library opens (rocTX ranges)."""
    check(lib().d2dhip_profiler_prep(int(transpose), int(io), int(fft), int(d2d)))


def profiler_start(name):
    """decomp_profiler_start (src/profiler_caliper.f90:156-164)."""
    check(lib().d2dhip_profiler_start(name.encode()))


def profiler_end(name):
    """decomp_profiler_end (src/profiler_caliper.f90:169-177): closes the
    innermost region this thread opened."""
    check(lib().d2dhip_profiler_end(name.encode()))


# ---- FFT engine ----------------------------------------------------------------

class FFTEngine:
    """One decomp_2d_fft engine (src/fft_common.f90:141-242) at one rank."""

    def __init__(self, pencil, nx, ny, nz, prec=PREC_F64, inplace=False, skip_xyz=None,
                 stream=None, tune=False, opt_DTT=None, rank=None):
        """rank = (p_row, p_col, coord_row, coord_col): the engine of one rank
        of a process grid (its stages on that rank's pencils; run them with
        exec_1m)."""
        flags = FLAG_INPLACE if inplace else 0
        if tune:
            flags |= FLAG_TUNE
        if skip_xyz is not None:
            for bit, s in zip((FLAG_SKIP_X, FLAG_SKIP_Y, FLAG_SKIP_Z), skip_xyz):
                if s:
                    flags |= bit
        self._plan = ctypes.c_void_p()
        self.format = pencil
        self.nx, self.ny, self.nz = nx, ny, nz
        self.prec = prec
        self.flags = flags
        st = ctypes.c_void_p(stream) if stream is not None else ctypes.c_void_p()
        if rank is None:
            check(lib().d2dhip_plan_create(ctypes.byref(self._plan), pencil, nx, ny, nz, prec,
                                           flags, st))
        else:
            check(lib().d2dhip_plan_create_rank(ctypes.byref(self._plan), pencil, nx, ny, nz, prec,
                                                *rank, flags, st))
        self._explicit_stream = stream is not None
        if opt_DTT is not None:
            self.set_dtt(opt_DTT)

    def _sync_stream(self):
        if not self._explicit_stream:
            check(lib().d2dhip_plan_set_stream(self._plan, _stream(None)))

    def decomps(self):
        """(ph, sp): the plan's physical and spectral decompositions."""
        ph, sp = Decomp(), Decomp()
        check(lib().d2dhip_plan_get_rank(self._plan, ctypes.byref(ph), ctypes.byref(sp)))
        return ph, sp

    def exec_1m(self, stage, inp, out, isign):
        """One 1D stage (STAGE_*) on the plan's pencil: c2c_1m_*, r2c_1m_*, c2r_1m_*."""
        self._sync_stream()
        i = self.STAGES.index(stage) if isinstance(stage, str) else stage
        return check(lib().d2dhip_exec_1m(self._plan, i, _ptr(inp), _ptr(out), isign))

    def exec_1m_pack(self, stage, inp, send, direction, decomp, isign, even=False):
        """A y stage with its output in the send order of `direction` (Y_TO_Z,
        Y_TO_X), or a z stage in that of Z_TO_Y."""
        self._sync_stream()
        i = self.STAGES.index(stage) if isinstance(stage, str) else stage
        return check(lib().d2dhip_exec_1m_pack(self._plan, i, _ptr(inp), _ptr(send), direction,
                                               ctypes.byref(decomp), int(bool(even)), isign))

    def exec_1m_unpack(self, stage, recv, out, direction, decomp, isign, even=False):
        """A y stage reading its input in the receive order of `direction`
        (Z_TO_Y, X_TO_Y), or a z stage in that of Y_TO_Z."""
        self._sync_stream()
        i = self.STAGES.index(stage) if isinstance(stage, str) else stage
        return check(lib().d2dhip_exec_1m_unpack(self._plan, i, _ptr(recv), _ptr(out), direction,
                                                 ctypes.byref(decomp), int(bool(even)), isign))

    def exec_1m_unpack_pack(self, stage, recv, send, dir_in, dir_out, decomp, isign,
                            even=False):
        """A y stage reading the receive order of `dir_in` (X_TO_Y, Z_TO_Y) and
        writing the send order of `dir_out` (Y_TO_Z, Y_TO_X)."""
        self._sync_stream()
        i = self.STAGES.index(stage) if isinstance(stage, str) else stage
        return check(lib().d2dhip_exec_1m_unpack_pack(self._plan, i, _ptr(recv), _ptr(send),
                                                      dir_in, dir_out, ctypes.byref(decomp),
                                                      int(bool(even)), isign))

    def get_sizes(self):
        ph = (ctypes.c_int * 3)()
        sp = (ctypes.c_int * 3)()
        check(lib().d2dhip_plan_get_sizes(self._plan, ph, sp))
        return tuple(ph), tuple(sp)

    def stage_kernels(self):
        """{stage: kernel name} of the seven stage slots."""
        out = {}
        buf = ctypes.create_string_buffer(128)
        for i, name in enumerate(self.STAGES):
            check(lib().d2dhip_plan_stage_kernel(self._plan, i, buf, 128))
            out[name] = buf.value.decode()
        return out

    def stage_candidates(self, stage):
        """Kernel names able to run a stage slot ("general" first)."""
        i = self.STAGES.index(stage) if isinstance(stage, str) else stage
        names = []
        buf = ctypes.create_string_buffer(128)
        while lib().d2dhip_plan_stage_candidate(self._plan, i, len(names), buf, 128) == 0:
            names.append(buf.value.decode())
        return names

    def set_stage_kernel(self, stage, name):
        i = self.STAGES.index(stage) if isinstance(stage, str) else stage
        check(lib().d2dhip_plan_set_stage_kernel(self._plan, i, name.encode()))

    def device_bytes(self):
        return lib().d2dhip_plan_device_bytes(self._plan)

    def set_chunking(self, fwd_planes, bwd_planes):
        """Alternate the (r2c x, y) and (y, c2r x) stage pairs over chunks of
        this many z-planes (PHYSICAL_IN_X; 0 = whole stages)."""
        check(lib().d2dhip_plan_set_chunking(self._plan, int(fwd_planes), int(bwd_planes)))

    def chunking(self):
        f, b = ctypes.c_int(), ctypes.c_int()
        check(lib().d2dhip_plan_get_chunking(self._plan, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value

    def set_layout(self, layout):
        """LAYOUT_PLAIN, LAYOUT_BLOCKED or LAYOUT_BLOCKED32: the spectral pencil's order between
        the stages of fft_r2c / fft_c2r (PHYSICAL_IN_X whole-grid engines)."""
        check(lib().d2dhip_plan_set_layout(self._plan, int(layout)))

    def layout(self):
        v = ctypes.c_int()
        check(lib().d2dhip_plan_get_layout(self._plan, ctypes.byref(v)))
        return v.value

    STAGES = ("c2c_x", "c2c_y", "c2c_z", "r2c", "spec_y", "spec_last", "c2r")

    def profile(self, enable=True):
        """Bracket every kernel launch with a hipEvent pair (stage timing)."""
        check(lib().d2dhip_plan_profile(self._plan, int(bool(enable))))

    def profile_report(self):
        """{stage: (total_ms, launches)} since the last report."""
        ms = (ctypes.c_double * 7)()
        n = (ctypes.c_int * 7)()
        check(lib().d2dhip_plan_profile_report(self._plan, ms, n))
        return {self.STAGES[i]: (ms[i], n[i]) for i in range(7) if n[i]}

    def fft_c2c(self, inp, out, isign):
        self._sync_stream()
        return check(lib().d2dhip_exec_c2c(self._plan, _ptr(inp), _ptr(out), isign))

    def fft_r2c(self, inp, out):
        self._sync_stream()
        return check(lib().d2dhip_exec_r2c(self._plan, _ptr(inp), _ptr(out)))

    def fft_c2r(self, inp, out):
        self._sync_stream()
        return check(lib().d2dhip_exec_c2r(self._plan, _ptr(inp), _ptr(out)))

    def fft_c2r_noalloc(self, inp, out):
        """fft_c2r with the workspace in the library's static scratch array
        (no allocation; spectra up to SCRATCH_BYTES; concurrent calls on one
        GPU race)."""
        self._sync_stream()
        return check(lib().d2dhip_exec_c2r_noalloc(self._plan, _ptr(inp), _ptr(out)))

    def set_dtt(self, dtt):
        """DCT / DST types of the x, y, z axes (3 values, 1-8), or 12 values
        with ifirst, ndismiss and ofirst (decomp_2d_fft_engine_dtt_init,
        src/fft_fftw3_f03.f90:341-390)."""
        v = [int(t) for t in dtt]
        arr = (ctypes.c_int * len(v))(*v)
        check(lib().d2dhip_plan_set_dtt(self._plan, len(v), arr))

    @property
    def dtt(self):
        """engine%dtt: types, ifirst, ndismiss, ofirst, backward types (15 values)."""
        v = (ctypes.c_int * 15)()
        check(lib().d2dhip_plan_get_dtt(self._plan, v))
        return tuple(v)

    def dtt_stage_kernels(self):
        """{stage: kernel name} of the real-to-real stages (7-9 forward x, y, z;
        10-12 backward)."""
        out = {}
        buf = ctypes.create_string_buffer(128)
        for i in range(7, 13):
            check(lib().d2dhip_plan_stage_kernel(self._plan, i, buf, 128))
            out[i] = buf.value.decode()
        return out

    def dtt_input(self):
        """decomp_2d_fft_get_dtt_input (src/fft_fftw3_f03.f90:629-643)."""
        return self.dtt[0:3]

    def dtt_3d_r2r(self, inp, out, isign):
        """decomp_2d_dtt_3d_r2r (src/fft_fftw3_f03.f90:1959-2023); inp is preserved."""
        self._sync_stream()
        return check(lib().d2dhip_exec_dtt(self._plan, _ptr(inp), _ptr(out), isign))

    def capture_begin(self):
        """Record the following exec calls on this engine into a hipGraph."""
        self._sync_stream()
        check(lib().d2dhip_plan_capture_begin(self._plan))

    def capture_end(self):
        check(lib().d2dhip_plan_capture_end(self._plan))

    def graph_launch(self):
        """Replay the recorded exec calls on the current stream."""
        self._sync_stream()
        check(lib().d2dhip_plan_graph_launch(self._plan))

    def fft_r2c_pack(self, inp, send, decomp, even=False):
        """The x stage (r2c along x) of a PHYSICAL_IN_X engine, written in the
        transpose_x_to_y send-buffer order of `decomp` (the spectral
        decomposition; pack fused with the FFT stage)."""
        self._sync_stream()
        return check(lib().d2dhip_exec_r2c_pack(self._plan, _ptr(inp), _ptr(send),
                                                ctypes.byref(decomp), int(bool(even))))

    def fft_c2r_unpack(self, recv, out, decomp, even=False):
        """The x stage (c2r along x) of a PHYSICAL_IN_X engine, read from the
        transpose_y_to_x receive-buffer order of `decomp` (unpack fused with
        the FFT stage)."""
        self._sync_stream()
        return check(lib().d2dhip_exec_c2r_unpack(self._plan, _ptr(recv), _ptr(out),
                                                  ctypes.byref(decomp), int(bool(even))))

    def fin(self):
        if self._plan:
            lib().d2dhip_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.fin()
        except Exception:
            pass


def decomp_2d_fft_init(pencil, nx, ny, nz, **kw):
    return FFTEngine(pencil, nx, ny, nz, **kw)


class Batch:
    """c2c stages of many engines run together: one persistent-kernel launch
    fed by device work queues (BATCH_PERSISTENT, default) or one launch per
    stage phase and stage kernel (BATCH_GROUPED)."""

    def __init__(self, engines, stream=None):
        self.engines = list(engines)
        n = len(self.engines)
        arr = (ctypes.c_void_p * n)(*[e._plan.value for e in self.engines])
        self._b = ctypes.c_void_p()
        check(lib().d2dhip_batch_create(ctypes.byref(self._b), n, arr,
                                        ctypes.c_void_p(stream) if stream is not None
                                        else ctypes.c_void_p()))
        self._explicit_stream = stream is not None

    def _sync_stream(self):
        if not self._explicit_stream:
            check(lib().d2dhip_batch_set_stream(self._b, _stream(None)))

    def bind(self, ins, outs, mode=0):
        """mode -1 forward, +1 backward, 0 forward then backward (in place on outs)."""
        n = len(self.engines)
        a = (ctypes.c_void_p * n)(*[_ptr(t).value for t in ins])
        b = (ctypes.c_void_p * n)(*[_ptr(t).value for t in outs])
        check(lib().d2dhip_batch_bind_c2c(self._b, a, b, mode))

    def launch(self):
        self._sync_stream()
        check(lib().d2dhip_batch_launch(self._b))

    def set_mode(self, mode):
        """BATCH_PERSISTENT (default: one persistent launch, on specialised
        tiles where every stage has one), BATCH_GROUPED (one launch per stage
        phase and stage kernel, the plans' own kernels) or
        BATCH_PERSISTENT_GENERAL (one persistent launch on the general
        kernel's tiles)."""
        check(lib().d2dhip_batch_set_mode(self._b, int(mode)))

    def set_cluster_bytes(self, nbytes):
        """Byte budget of a cluster of consecutive grids that runs all its
        stages before the next (0 = the whole batch phase by phase); takes
        effect at the next bind."""
        check(lib().d2dhip_batch_set_cluster_bytes(self._b, int(nbytes)))

    def set_item_points(self, points):
        """Complex points per queue item of the persistent executors (whole
        tiles); takes effect at the next bind."""
        check(lib().d2dhip_batch_set_item_points(self._b, int(points)))

    def persistent_tiles(self):
        """True if BATCH_PERSISTENT runs specialised tiles for this batch."""
        v = ctypes.c_int()
        check(lib().d2dhip_batch_persistent_tiles(self._b, ctypes.byref(v)))
        return bool(v.value)

    def clusters(self):
        """(clusters, stage phases) of the current binding."""
        c, p = ctypes.c_int(), ctypes.c_int()
        check(lib().d2dhip_batch_get_clusters(self._b, ctypes.byref(c), ctypes.byref(p)))
        return c.value, p.value

    def capture_graph(self):
        check(lib().d2dhip_batch_capture_graph(self._b))

    def graph_launch(self):
        self._sync_stream()
        check(lib().d2dhip_batch_graph_launch(self._b))

    def check(self):
        check(lib().d2dhip_batch_check(self._b))

    def destroy(self):
        if self._b:
            lib().d2dhip_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ---- transposes and packs -------------------------------------------------------

def transpose(direction, elem, src, dst, decomp, stream=None):
    return check(lib().d2dhip_transpose(direction, elem, _ptr(src), _ptr(dst),
                                        ctypes.byref(decomp), _stream(stream)))


def transpose_x_to_y(src, dst, decomp, elem=C16, stream=None):
    return transpose(X_TO_Y, elem, src, dst, decomp, stream)


def transpose_y_to_z(src, dst, decomp, elem=C16, stream=None):
    return transpose(Y_TO_Z, elem, src, dst, decomp, stream)


def transpose_z_to_y(src, dst, decomp, elem=C16, stream=None):
    return transpose(Z_TO_Y, elem, src, dst, decomp, stream)


def transpose_y_to_x(src, dst, decomp, elem=C16, stream=None):
    return transpose(Y_TO_X, elem, src, dst, decomp, stream)


def pack(direction, elem, src, send, decomp, even=False, stream=None):
    return check(lib().d2dhip_pack(direction, elem, _ptr(src), _ptr(send), ctypes.byref(decomp),
                                   int(bool(even)), _stream(stream)))


def unpack(direction, elem, recv, dst, decomp, even=False, stream=None):
    return check(lib().d2dhip_unpack(direction, elem, _ptr(recv), _ptr(dst),
                                     ctypes.byref(decomp), int(bool(even)), _stream(stream)))


def pack_buffer_elems(direction, decomp, even=False):
    return lib().d2dhip_pack_buffer_elems(direction, ctypes.byref(decomp), int(bool(even)))


# ---- synthetic inputs ------------------------------------------------------------

def halo_shape(sizes, level, ipencil):
    """Extents of the halo array of a (s1, s2, s3) pencil (halo_extents_t,
    src/halo.f90:222-292, without its Fortran lower bounds)."""
    n = list(sizes)
    for a in range(3):
        if a != ipencil - 1:
            n[a] += 2 * level
    return tuple(n)


def update_halo(elem, src, out, sizes, level, ipencil, periodic=(False, False, False),
                stream=None):
    """update_halo (src/halo.f90:112-260) at one rank into the preallocated halo
    array `out` (halo_shape(sizes, level, ipencil) elements, i fastest)."""
    per = (ctypes.c_int * 3)(*[int(bool(p)) for p in periodic])
    s1, s2, s3 = sizes
    check(lib().d2dhip_update_halo(elem, _ptr(src), _ptr(out), s1, s2, s3, level, ipencil, per,
                                   _stream(stream)))


def halo_exchange(elem, arr, sizes, level, ipencil, periodic=(False, False, False), stream=None):
    """halo_exchange (src/halo_exchange_{x,y,z}_body.f90) at one rank, in place."""
    per = (ctypes.c_int * 3)(*[int(bool(p)) for p in periodic])
    s1, s2, s3 = sizes
    check(lib().d2dhip_halo_exchange(elem, _ptr(arr), s1, s2, s3, level, ipencil, per,
                                     _stream(stream)))


def column_copy(src, dst, elem, C0, C1, s1, sj, n, stream=None):
    """Copy complex columns with an axis stage's access pattern (bandwidth reference)."""
    check(lib().d2dhip_column_copy(_ptr(src), _ptr(dst), elem, C0, C1, s1, sj, n, _stream(stream)))


def fill_uniform(dst, n, elem, seed, stream=None):
    return check(lib().d2dhip_fill_uniform(_ptr(dst), n, elem, seed, _stream(stream)))


def host_fill_uniform(np_array, elem, seed):
    n = np_array.size
    lib().d2dhip_host_fill_uniform(np_array.ctypes.data_as(ctypes.c_void_p), n, elem, seed)


def modes_make(nmodes, kmax, amp_kind, amp_lo, amp_hi, sigma, seed):
    m = Modes()
    check(lib().d2dhip_modes_make(ctypes.byref(m), nmodes, kmax, amp_kind, amp_lo, amp_hi, sigma,
                                  seed))
    return m


def fill_modal(dst, nx, ny, nz, elem, modes, stream=None):
    return check(lib().d2dhip_fill_modal(_ptr(dst), nx, ny, nz, elem, ctypes.byref(modes),
                                         _stream(stream)))


def host_fill_modal(np_array, nx, ny, nz, elem, modes):
    lib().d2dhip_host_fill_modal(np_array.ctypes.data_as(ctypes.c_void_p), nx, ny, nz, elem,
                                 ctypes.byref(modes))
