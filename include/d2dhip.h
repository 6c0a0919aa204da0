/*
 * d2dhip.h -- C ABI of libd2dhip, the MI355X (gfx950) FFT backend for the
 * 2DECOMP&FFT single-GPU 3D FFT path.
 *
 * Every entry point returns an int status: 0 on success, 1..99 on error (the
 * caller maps it to decomp_2d_abort, as the reference does with
 * `istat /= 0` after each vendor call, src/fft_cufft.f90:86-94), and >= 100
 * for warnings after a completed call (the caller maps them to
 * decomp_2d_warning, src/decomp_2d_mpi.f90:259-299).
 *
 * Data layout is the reference's: Fortran order, i fastest; complex values
 * are interleaved (re, im).  Device pointers are plain HIP device addresses.
 * Streams are hipStream_t passed as void* (NULL = the null stream).
 * Calls are stream-ordered and return without synchronising.
 */
#ifndef D2DHIP_H
#define D2DHIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants (src/decomp_2d_constants.f90:77-93) ---------------------- */
#define D2DHIP_FFT_BACKEND 5 /* new id beside generic=0 .. cufft=4 (:77-81) */
#define D2DHIP_PHYSICAL_IN_X 1 /* PHYSICAL_IN_X (:92) */
#define D2DHIP_PHYSICAL_IN_Z 3 /* PHYSICAL_IN_Z (:93) */
#define D2DHIP_FORWARD (-1)    /* DECOMP_2D_FFT_FORWARD (:86) */
#define D2DHIP_BACKWARD 1      /* DECOMP_2D_FFT_BACKWARD (:87) */

/* arithmetic/storage precision of a plan */
#define D2DHIP_PREC_F64 64        /* real64 storage and arithmetic (drop-in) */
#define D2DHIP_PREC_F32 32        /* real32 storage and arithmetic */
#define D2DHIP_PREC_SPLIT_BF16 16 /* real32 storage; radix-16 passes on the matrix cores
                                     with operands split into three bf16 parts (axes of
                                     length 512 and 1024; other lengths run the FP32
                                     VALU kernels) */

/* plan flags (optional arguments of decomp_2d_fft_init, src/fft_common.f90:141-193) */
#define D2DHIP_FLAG_INPLACE 0x1u /* opt_inplace: c2c/c2r may overwrite input */
#define D2DHIP_FLAG_SKIP_X 0x2u  /* opt_skip_XYZ_c2c(1) */
#define D2DHIP_FLAG_SKIP_Y 0x4u  /* opt_skip_XYZ_c2c(2) */
#define D2DHIP_FLAG_SKIP_Z 0x8u  /* opt_skip_XYZ_c2c(3) */
#define D2DHIP_FLAG_TUNE 0x10u   /* time the candidate kernels of each stage at plan
                                    creation (also with env D2DHIP_TUNE=1); choices
                                    are cached in $D2DHIP_TUNE_CACHE or
                                    ~/.cache/d2dhip/tuning.tsv, read again
                                    whenever that path changes */

/* element kinds for transposes and packs */
#define D2DHIP_R8 0  /* real(real64)    */
#define D2DHIP_C16 1 /* complex(real64) */
#define D2DHIP_R4 2  /* real(real32)    */
#define D2DHIP_C8 3  /* complex(real32) */

/* transpose / pack directions */
#define D2DHIP_X_TO_Y 0
#define D2DHIP_Y_TO_Z 1
#define D2DHIP_Z_TO_Y 2
#define D2DHIP_Y_TO_X 3

/* status codes */
#define D2DHIP_SUCCESS 0
#define D2DHIP_ERR_INVALID_VALUE 1 /* bad argument (nx <= 0, bad pencil, ...) */
#define D2DHIP_ERR_UNSUPPORTED 2   /* valid but not implemented (see message) */
#define D2DHIP_ERR_ALLOC 3         /* device allocation failed */
#define D2DHIP_ERR_HIP 4           /* a HIP runtime call failed */
#define D2DHIP_ERR_NULL 5          /* NULL plan or pointer */
#define D2DHIP_ERR_GRID 6          /* grid too small for p_row x p_col (errorcode 6,
                                      src/decomp_2d.f90:399-405) */
#define D2DHIP_ERR_PLAN_MISMATCH 7 /* exec kind does not match the plan */
#define D2DHIP_WARN_SKIP_IGNORED 100 /* "r2c / c2r transform can not be skipped"
                                        (src/fft_generic.f90:222-223) */
#define D2DHIP_WARN_TUNE_CACHE 101 /* plan created (or DTT set up) and tuned, but the
                                      choices could not be written to the tuning cache */

#define D2DHIP_MAX_PROCS 64

typedef struct d2dhip_plan_s *d2dhip_plan;

/* Mirror of DECOMP_INFO (src/info.f90:19-47) for one rank of a
 * p_row x p_col grid.  Starts/ends are 1-based as in the reference. */
typedef struct {
  int xst[3], xen[3], xsz[3];
  int yst[3], yen[3], ysz[3];
  int zst[3], zen[3], zsz[3];
  int p_row, p_col;       /* dims(1), dims(2) */
  int coord_row, coord_col; /* coord(1), coord(2) of this rank */
  int x1dist[D2DHIP_MAX_PROCS], y1dist[D2DHIP_MAX_PROCS];
  int y2dist[D2DHIP_MAX_PROCS], z2dist[D2DHIP_MAX_PROCS];
  long long x1cnts[D2DHIP_MAX_PROCS], y1cnts[D2DHIP_MAX_PROCS];
  long long x1disp[D2DHIP_MAX_PROCS], y1disp[D2DHIP_MAX_PROCS];
  long long y2cnts[D2DHIP_MAX_PROCS], z2cnts[D2DHIP_MAX_PROCS];
  long long y2disp[D2DHIP_MAX_PROCS], z2disp[D2DHIP_MAX_PROCS];
  long long x1count, y1count, y2count, z2count; /* EVEN padded counts */
  int even;                                     /* evenly divisible grid */
} d2dhip_decomp;

/* ---- status ------------------------------------------------------------- */
const char *d2dhip_status_string(int status);
/* Detail text of the last non-zero status returned on this host thread. */
const char *d2dhip_last_message(void);
/* Library version as major*10000 + minor*100 + patch. */
int d2dhip_version(void);

/* ---- profiler regions (host only) -----------------------------------------
 * The reference's decomp_2d_profiler (src/profiler_caliper.f90): regions
 * around its entry points, here rocTX ranges (rocprofv3 --marker-trace
 * records them) with the reference's names: fft_init / fft_fin around plan
 * creation and destruction, fft_c2c / fft_r2c / fft_c2r around the 3D
 * transforms (src/fft_common_3d.f90:22,130,206), decomp_2d_dtt_3d_r2r
 * around d2dhip_exec_dtt, transp_x_y_r / _c ... around d2dhip_transpose
 * (src/transpose_x_to_y.f90:39,162 and the other three files), and
 * decomp_2d_init around d2dhip_decomp_info_init.
 * d2dhip_profiler_prep replaces decomp_profiler_prep
 * (src/profiler_caliper.f90:138-150): which categories open regions, in the
 * reference's order; all four are on until it is called.
 * d2dhip_profiler_start / _end replace decomp_profiler_start / _end
 * (:156-177): a caller's own region; regions nest and _end closes the
 * innermost one the calling thread opened (D2DHIP_ERR_INVALID_VALUE when it
 * opened none). */
int d2dhip_profiler_prep(int transpose, int io, int fft, int d2d);
int d2dhip_profiler_start(const char *name);
int d2dhip_profiler_end(const char *name);

/* ---- decomposition metadata (host only) ----------------------------------
 * Replaces decomp_info_init / get_dist / partition / prepare_buffer,
 * src/decomp_2d.f90:382-490 and :1016-1206, for rank (coord_row, coord_col). */
int d2dhip_decomp_info_init(int nx, int ny, int nz, int p_row, int p_col,
                            int coord_row, int coord_col, d2dhip_decomp *d);
/* Replaces best_2d_grid, src/decomp_2d_init_fin.f90:270-300 (the process
 * grid the reference picks when p_row = p_col = 0). */
int d2dhip_best_2d_grid(int nproc, int *p_row, int *p_col);

/* ---- FFT plans -----------------------------------------------------------
 * d2dhip_plan_create replaces decomp_2d_fft_engine_init + init_fft_engine
 * (src/fft_common.f90:141-242; cuFFT plan set-up src/fft_cufft.f90:263-431)
 * for one rank (p_row = p_col = 1).  format is PHYSICAL_IN_X or _Z; prec one
 * of D2DHIP_PREC_*; flags D2DHIP_FLAG_*.  The plan owns its twiddles and
 * workspace. */
int d2dhip_plan_create(d2dhip_plan *plan, int format, int nx, int ny, int nz,
                       int prec, unsigned flags, void *stream);
/* Replaces decomp_2d_fft_engine_fin / finalize_fft_engine
 * (src/fft_common.f90:274-295). */
int d2dhip_plan_destroy(d2dhip_plan plan);
int d2dhip_plan_set_stream(d2dhip_plan plan, void *stream);
/* Physical (ph) x-pencil size and spectral (sp) size of the r2c/c2r output:
 * (nx/2+1,ny,nz) for PHYSICAL_IN_X, (nx,ny,nz/2+1) for PHYSICAL_IN_Z
 * (decomp_2d_fft_get_size, src/fft_common.f90:301-327). */
int d2dhip_plan_get_sizes(d2dhip_plan plan, int ph_sz[3], int sp_sz[3]);
/* Stage timing, the counterpart of the reference's profiler regions
 * (decomp_profiler_start/end, src/profiler_caliper.f90:156-177): while
 * enabled, every kernel launch of the plan is bracketed by a hipEvent pair
 * on the plan's stream.  The report synchronises on those events, returns
 * the summed milliseconds and launch counts per stage and clears them.
 * Stages: 0-2 c2c along x, y, z (with the real-to-real stages along that
 * axis); 3 the r2c stage; 4 the spectral y stage; 5 the other spectral c2c
 * stage (z for PHYSICAL_IN_X, x for _Z); 6 c2r.  The stage-kernel calls
 * below also take 7-9 (forward real-to-real along x, y, z) and 10-12
 * (backward) once d2dhip_plan_set_dtt has run. */
int d2dhip_plan_profile(d2dhip_plan plan, int enable);
int d2dhip_plan_profile_report(d2dhip_plan plan, double ms[7], int launches[7]);
/* Name of the kernel a stage runs ("general" or a specialised variant). */
int d2dhip_plan_stage_kernel(d2dhip_plan plan, int stage, char *name, int len);
/* idx-th kernel able to run a stage: 0 = "general", then the specialised
 * variants compiled for its length; D2DHIP_ERR_INVALID_VALUE past the end. */
int d2dhip_plan_stage_candidate(d2dhip_plan plan, int stage, int idx, char *name, int len);
/* Makes a stage run the named candidate (what the autotuner does). */
int d2dhip_plan_set_stage_kernel(d2dhip_plan plan, int stage, const char *name);
/* Plane chunking of a PHYSICAL_IN_X plan (no counterpart in the reference,
 * whose stages always cover the whole pencil): the forward r2c-x and y
 * stages, and the inverse y and c2r-x stages, alternate over chunks of
 * fwd_planes / bwd_planes z-planes so that the second stage of each pair
 * re-reads a chunk from the memory-side cache instead of HBM; 0 = whole
 * stages (the default).  Plans created with D2DHIP_FLAG_TUNE pick both
 * sizes by timing; the environment variable D2DHIP_CHUNK_PLANES="f,b"
 * fixes them at plan creation.  Results agree with unchunked execution to
 * rounding (real pairs in the x stages can differ at a chunk edge when ny
 * times the chunk size is odd). */
int d2dhip_plan_set_chunking(d2dhip_plan plan, int fwd_planes, int bwd_planes);
int d2dhip_plan_get_chunking(d2dhip_plan plan, int *fwd_planes, int *bwd_planes);
/* Layout of the spectral pencil inside exec_r2c / exec_c2r of a
 * PHYSICAL_IN_X plan (no counterpart in the reference, whose stages work
 * on the natural (i, j, k) order; the inputs and outputs of the calls keep
 * that order either way).  D2DHIP_LAYOUT_BLOCKED (_BLOCKED32) keeps the
 * pencil between the x stage and the last stage in the plan's workspace as
 * blocks of 16 (32) i-columns and runs the stages in the order x, z, y
 * (forward) and y, z, x
 * (inverse), so that no stage walks rows a whole i-j plane apart; the
 * result agrees with the plain layout to rounding (the order of the y and z
 * passes differs).  It needs a whole-grid plan with a workspace (not
 * D2DHIP_FLAG_INPLACE), and exec_c2r_noalloc and chunked or skipped stages
 * keep the plain layout.  Plans created with D2DHIP_FLAG_TUNE pick the
 * layout by timing; the environment variable
 * D2DHIP_LAYOUT=plain|blocked|blocked32 fixes it at plan creation. */
#define D2DHIP_LAYOUT_PLAIN 0
#define D2DHIP_LAYOUT_BLOCKED 1
#define D2DHIP_LAYOUT_BLOCKED32 2
int d2dhip_plan_set_layout(d2dhip_plan plan, int layout);
int d2dhip_plan_get_layout(d2dhip_plan plan, int *layout);
/* Device bytes the plan holds (twiddles + workspace). */
long long d2dhip_plan_device_bytes(d2dhip_plan plan);

/* decomp_2d_fft_3d(in, out, isign): fft_3d_c2c, src/fft_common_3d.f90:9-116.
 * in is preserved unless the plan has D2DHIP_FLAG_INPLACE; in == out allowed. */
int d2dhip_exec_c2c(d2dhip_plan plan, void *in, void *out, int isign);
/* decomp_2d_fft_3d(in_r, out_c): fft_3d_r2c, src/fft_common_3d.f90:121-189. */
int d2dhip_exec_r2c(d2dhip_plan plan, const void *in_r, void *out_c);
/* decomp_2d_fft_3d(in_c, out_r): fft_3d_c2r, src/fft_common_3d.f90:194-296.
 * in_c is preserved unless the plan has D2DHIP_FLAG_INPLACE. */
int d2dhip_exec_c2r(d2dhip_plan plan, void *in_c, void *out_r);
/* d2dhip_exec_c2r that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory at call time (code
 * under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc), and for plans created with D2DHIP_FLAG_INPLACE, which hold no
 * workspace, when in_c must survive.  The spectral workspace of the c2r
 * stages is a static device array of D2DHIP_SCRATCH_BYTES that the library
 * holds on each GPU it runs on until the process ends; in_c is preserved
 * whatever the plan's flags, and the results equal d2dhip_exec_c2r's bit
 * for bit.  Returns D2DHIP_ERR_UNSUPPORTED when the spectral pencil
 * ((nx/2+1)*ny*nz or nx*ny*(nz/2+1) complex values of the plan's precision)
 * exceeds D2DHIP_SCRATCH_BYTES, and D2DHIP_ERR_INVALID_VALUE while stage
 * profiling is on.  Every call on one GPU shares that GPU's array, which
 * keeps the last call's contents (the stages write it before reading it):
 * calls that overlap on one GPU, from any plan or stream, race. */
#define D2DHIP_SCRATCH_BYTES (64ll << 20)
int d2dhip_exec_c2r_noalloc(d2dhip_plan plan, void *in_c, void *out_r);
/* Real-to-real transforms (DCT / DST types 1-8 = DCT-I..IV, DST-I..IV, FFTW's
 * REDFT00..11, RODFT00..11, unnormalised).  d2dhip_plan_set_dtt replaces
 * decomp_2d_fft_engine_dtt_init (src/fft_fftw3_f03.f90:341-390, the opt_DTT
 * argument of decomp_2d_fft_init): dtt holds ndtt = 3 types for x, y, z, or
 * 12 values (types, ifirst, ndismiss, ofirst; 1-based, ofirst == ifirst)
 * whose last nine default as in dtt_assign_default (:394-420).  Along an axis
 * of n points the transform covers points [ifirst, ifirst + n - ndismiss)
 * of each line; the other points are copied through.  The backward types
 * are the inverses (dtt_invert, :424-444: DCT-II <-> DCT-III, DST-II <->
 * DST-III, the others their own).  The stages run on the Stockham kernels
 * through the DFT of an even / odd extension of the line. */
int d2dhip_plan_set_dtt(d2dhip_plan plan, int ndtt, const int *dtt);
/* The resolved 15 values: types, ifirst, ndismiss, ofirst, backward types
 * (engine%dtt; types as 1-8, decomp_2d_fft_get_dtt_input :629-643). */
int d2dhip_plan_get_dtt(d2dhip_plan plan, int dtt[15]);
/* decomp_2d_dtt_3d_r2r(in, out, isign), src/fft_fftw3_f03.f90:1959-2023, on
 * the (nx,ny,nz) real array of the plan's precision.  in is preserved; in ==
 * out allowed. */
int d2dhip_exec_dtt(d2dhip_plan plan, const void *in, void *out, int isign);
/* hipGraph of a plan's transforms (no reference counterpart; the reference's
 * GPU backend launches every stage eagerly): between capture_begin and
 * capture_end the plan records the exec calls made on it, with their buffers
 * (for example the forward and inverse transforms of a time step), on a
 * capturing stream of its own that starts after the work already queued on
 * the plan's stream; graph_launch replays the recording on the plan's
 * stream.  Stream-ordered like the exec calls; a new capture replaces the
 * old graph.  The fused pack / unpack calls and stage profiling are refused
 * while capturing. */
int d2dhip_plan_capture_begin(d2dhip_plan plan);
int d2dhip_plan_capture_end(d2dhip_plan plan);
int d2dhip_plan_graph_launch(d2dhip_plan plan);
/* Transposes fused with the adjacent FFT stage, for one rank of a process
 * grid (PHYSICAL_IN_X plans of the rank's x-pencil, nx x ny x nz real):
 * d2dhip_exec_r2c_pack runs the forward x stage (r2c_1m_x) and writes its
 * (nx/2+1) x ny x nz spectrum straight into the send buffers of
 * transpose_x_to_y for the spectral decomposition d (mem_split_xy,
 * src/fft_common_3d.f90:133-140 then src/transpose_x_to_y.f90:93-101);
 * d2dhip_exec_c2r_unpack reads the receive buffers of transpose_y_to_x
 * (mem_merge_yx, src/transpose_y_to_x.f90:131-133) and runs the inverse x
 * stage (c2r_1m_x, src/fft_common_3d.f90:259-266).  d->xsz must be
 * (nx/2+1, ny, nz); blocks follow x1dist / x1disp (even: m * x1count).
 * Results equal the x stage followed by d2dhip_pack(D2DHIP_X_TO_Y, ...)
 * (resp. d2dhip_unpack(D2DHIP_Y_TO_X, ...) then the x stage) bit for bit,
 * with one read and one write of the data instead of two.  With p_row = 1
 * the buffer order is the plain spectral x-pencil. */
int d2dhip_exec_r2c_pack(d2dhip_plan plan, const void *in_r, void *send,
                         const d2dhip_decomp *d, int even);
int d2dhip_exec_c2r_unpack(d2dhip_plan plan, const void *recv, void *out_r,
                           const d2dhip_decomp *d, int even);

/* ---- one rank of a process grid ------------------------------------------
 * d2dhip_plan_create_rank is decomp_2d_fft_init for rank (coord_row,
 * coord_col) of a p_row x p_col grid (src/fft_common.f90:141-242 after
 * decomp_2d_init, src/decomp_2d_init_fin.f90:15-184): its stages run on the
 * rank's pencils of the physical decomposition of (nx, ny, nz) and the
 * spectral one of (nx/2+1, ny, nz) (PHYSICAL_IN_X) or (nx, ny, nz/2+1)
 * (PHYSICAL_IN_Z), as decomp_info_init gives them.  With p_row*p_col > 1
 * the 3D calls (exec_c2c / r2c / c2r / dtt) return D2DHIP_ERR_UNSUPPORTED:
 * between their stages the reference exchanges pencils over MPI
 * (src/fft_common_3d.f90:43-68, 143-159); the caller runs the stages one
 * at a time with d2dhip_exec_1m and moves the alltoallv buffers that
 * d2dhip_pack / d2dhip_unpack and the fused calls below write and read. */
int d2dhip_plan_create_rank(d2dhip_plan *plan, int format, int nx, int ny, int nz, int prec,
                            int p_row, int p_col, int coord_row, int coord_col, unsigned flags,
                            void *stream);
/* The plan's physical and spectral decompositions (one rank: the whole
 * grid). */
int d2dhip_plan_get_rank(d2dhip_plan plan, d2dhip_decomp *ph, d2dhip_decomp *sp);
/* Stage slots of d2dhip_exec_1m, as d2dhip_plan_profile numbers them. */
#define D2DHIP_STAGE_C2C_X 0     /* c2c_1m_x on the physical x-pencil */
#define D2DHIP_STAGE_C2C_Y 1     /* c2c_1m_y on the physical y-pencil */
#define D2DHIP_STAGE_C2C_Z 2     /* c2c_1m_z on the physical z-pencil */
#define D2DHIP_STAGE_R2C 3       /* r2c_1m_x (X) / r2c_1m_z (Z): physical -> spectral pencil */
#define D2DHIP_STAGE_SPEC_Y 4    /* c2c_1m_y on the spectral y-pencil */
#define D2DHIP_STAGE_SPEC_LAST 5 /* c2c_1m_z (X) / c2c_1m_x (Z) on the spectral pencil */
#define D2DHIP_STAGE_C2R 6       /* c2r_1m_x (X) / c2r_1m_z (Z): spectral -> physical pencil */
/* One 1D stage on the plan's pencil: c2c_1m_x / _y / _z (isign -1 forward,
 * +1 backward, unnormalised; in == out allowed), r2c_1m_x / _z (isign -1),
 * c2r_1m_x / _z (isign +1) -- src/fft_generic.f90:112-384, FFTW execs
 * src/fft_fftw3_f03.f90:1414-1552. */
int d2dhip_exec_1m(d2dhip_plan plan, int stage, const void *in, void *out, int isign);
/* A y stage (D2DHIP_STAGE_C2C_Y or _SPEC_Y) fused with the pack of the
 * exchange that follows it (dir D2DHIP_Y_TO_Z: mem_split_yz,
 * src/transpose_y_to_z.f90:370-428; D2DHIP_Y_TO_X: mem_split_yx,
 * src/transpose_y_to_x.f90:266-323), or with the unpack of the exchange
 * before it (dir D2DHIP_Z_TO_Y: mem_merge_zy,
 * src/transpose_z_to_y.f90:489-547; D2DHIP_X_TO_Y: mem_merge_xy,
 * src/transpose_x_to_y.f90:384-441); or a z stage (D2DHIP_STAGE_C2C_Z, and
 * D2DHIP_STAGE_SPEC_LAST of a PHYSICAL_IN_X plan) fused with the pack after
 * it (dir D2DHIP_Z_TO_Y: mem_split_zy, src/transpose_z_to_y.f90:370-427) or
 * the unpack before it (dir D2DHIP_Y_TO_Z: mem_merge_yz,
 * src/transpose_y_to_z.f90:490-548).  d->ysz (d->zsz) must be the stage's
 * pencil; even selects the padded offsets.  Results equal d2dhip_exec_1m
 * followed by d2dhip_pack (d2dhip_unpack then d2dhip_exec_1m) bit for bit.
 * A stage kernel without a fused form runs unfused through a workspace the
 * plan allocates on first use. */
int d2dhip_exec_1m_pack(d2dhip_plan plan, int stage, const void *in, void *send, int dir,
                        const d2dhip_decomp *d, int even, int isign);
int d2dhip_exec_1m_unpack(d2dhip_plan plan, int stage, const void *recv, void *out, int dir,
                          const d2dhip_decomp *d, int even, int isign);
/* A y stage between two exchanges with both fused: its input read in the
 * receive order of dir_in (D2DHIP_X_TO_Y: mem_merge_xy, or D2DHIP_Z_TO_Y:
 * mem_merge_zy) and its output written in the send order of dir_out
 * (D2DHIP_Y_TO_Z: mem_split_yz, or D2DHIP_Y_TO_X: mem_split_yx), as
 * src/fft_common_3d.f90:145-158 does with a transpose on each side of
 * c2c_1m_y.  Equals d2dhip_unpack, d2dhip_exec_1m and d2dhip_pack bit for
 * bit. */
int d2dhip_exec_1m_unpack_pack(d2dhip_plan plan, int stage, const void *recv, void *send,
                               int dir_in, int dir_out, const d2dhip_decomp *d, int even,
                               int isign);

/* ---- many grids in one launch ---------------------------------------------
 * The multiple-grid use of the reference (decomp_2d_fft_set_ngrid /
 * decomp_2d_fft_use_grid, src/fft_common.f90:381-443, exercised by
 * examples/fft_multiple_grids/fft_multiple_grids.f90) runs one
 * decomp_2d_fft_3d per grid.  A batch runs the c2c stages of many plans (same
 * precision) as stage phases (stage 0 of every grid, then stage 1, ...) with
 * one of two executors (d2dhip_batch_set_mode below): by default one
 * persistent-kernel launch fed by device work queues, in which a phase
 * starts when every tile of the previous phase has completed (one
 * release/acquire per workgroup and phase); or one ordinary launch per phase
 * and stage kernel.  Calls that overlap on one batch object race: the queue
 * counters belong to the batch.  The persistent launch assumes that at least
 * 8 of its workgroups can be resident at once (one per queue); otherwise a
 * phase wait times out after 5 s and check reports it.  bind lays out the
 * phases for the given device buffers (mode -1 forward, +1 backward, 0
 * forward then backward in place on out); launch is stream-ordered and
 * capturable; capture_graph records it into a hipGraph replayed by
 * graph_launch; check synchronises and reports a phase-wait timeout as
 * D2DHIP_ERR_HIP. */
typedef struct d2dhip_batch_s *d2dhip_batch;
int d2dhip_batch_create(d2dhip_batch *batch, int nplans, const d2dhip_plan *plans, void *stream);
int d2dhip_batch_bind_c2c(d2dhip_batch batch, void *const *in, void *const *out, int mode);
int d2dhip_batch_launch(d2dhip_batch batch);
int d2dhip_batch_capture_graph(d2dhip_batch batch);
int d2dhip_batch_graph_launch(d2dhip_batch batch);
int d2dhip_batch_check(d2dhip_batch batch);
int d2dhip_batch_set_stream(d2dhip_batch batch, void *stream);
/* Executor of a batch: D2DHIP_BATCH_PERSISTENT (default) is the single
 * persistent launch when every c2c stage of every plan has a length with a
 * persistent tile (32, 48, 64, 80, 96, 128, 160, 192, 256): its workgroups
 * run those specialised tiles (d2dhip_batch_persistent_tiles reports 1), and
 * results are bit-identical to the plans' calls on any specialised kernel of
 * each length.  Otherwise the default runs as D2DHIP_BATCH_GROUPED (one
 * persistent launch on general-kernel tiles is about 3x slower than the
 * plans' own kernels).  D2DHIP_BATCH_PERSISTENT_GENERAL forces the
 * persistent launch on the general kernel's tiles, bit-identical to the
 * plans' calls with every stage on the general kernel.
 * D2DHIP_BATCH_GROUPED runs, per stage phase, one ordinary launch per
 * distinct stage kernel of the grids' own plans, covering the tiles of every
 * grid that runs that kernel in that phase (the kernel boundaries order the
 * phases; no device queue).  Its results are bit-identical to the plans' own
 * d2dhip_exec_c2c calls; it uses the plans' stage kernels as of bind, and
 * launch fails with D2DHIP_ERR_INVALID_VALUE if one has changed since (bind
 * again).  Changing the mode drops a captured graph. */
#define D2DHIP_BATCH_PERSISTENT 0
#define D2DHIP_BATCH_GROUPED 1
#define D2DHIP_BATCH_PERSISTENT_GENERAL 2
int d2dhip_batch_set_mode(d2dhip_batch batch, int mode);
/* Work per queue item of the persistent executors, in complex points
 * (whole tiles, at least one; default 2048).  Takes effect at the next bind. */
int d2dhip_batch_set_item_points(d2dhip_batch batch, long long points);
/* 1 if D2DHIP_BATCH_PERSISTENT runs specialised tiles for this batch. */
int d2dhip_batch_persistent_tiles(d2dhip_batch batch, int *specialised);
/* Queue order of both executors: consecutive grids whose input and output
 * buffers total at most `bytes` form a cluster, and a cluster runs all of its
 * stage phases before the next cluster starts, so that its later stages read
 * data the earlier ones left in the memory-side cache (256 MiB on MI355X).
 * 0 = one cluster, every grid's stage k before any grid's stage k+1.
 * Default 0.  Takes effect at the next bind (the batch is unbound). */
int d2dhip_batch_set_cluster_bytes(d2dhip_batch batch, long long bytes);
/* Clusters and stage phases of the current binding. */
int d2dhip_batch_get_clusters(d2dhip_batch batch, int *nclusters, int *nphases);
int d2dhip_batch_destroy(d2dhip_batch batch);

/* ---- transposes and alltoallv packing ------------------------------------
 * d2dhip_transpose replaces transpose_x_to_y / y_to_z / z_to_y / y_to_x for
 * one rank (src/transpose_x_to_y.f90:41-50 and the matching branches of the
 * other three files): with dims = (1,1) it is the copy dst = src of the
 * source pencil.  For p_row*p_col > 1 it returns D2DHIP_ERR_UNSUPPORTED (the
 * exchange itself needs a communicator). */
int d2dhip_transpose(int dir, int elem, const void *src, void *dst,
                     const d2dhip_decomp *d, void *stream);
/* Send-buffer pack of one rank: mem_split_xy / _yz / _zy / _yx
 * (src/transpose_x_to_y.f90:266-323, src/transpose_y_to_z.f90:370-428,
 * src/transpose_z_to_y.f90:370-428, src/transpose_y_to_x.f90:266-323).
 * even != 0 selects the padded MPI_ALLTOALL offsets m*count (EVEN build). */
int d2dhip_pack(int dir, int elem, const void *src, void *send,
                const d2dhip_decomp *d, int even, void *stream);
/* Receive-buffer unpack of one rank: mem_merge_xy / _yz / _zy / _yx
 * (src/transpose_x_to_y.f90:384-441, src/transpose_y_to_z.f90:489-547,
 * src/transpose_z_to_y.f90:489-547, src/transpose_y_to_x.f90:384-441). */
int d2dhip_unpack(int dir, int elem, const void *recv, void *dst,
                  const d2dhip_decomp *d, int even, void *stream);
/* Elements a pack/unpack buffer must hold for (dir, d, even). */
long long d2dhip_pack_buffer_elems(int dir, const d2dhip_decomp *d, int even);

/* ---- halo cells (one rank) -----------------------------------------------
 * update_halo(in, out, level, opt_pencil=ipencil) (src/halo.f90:112-260,
 * src/halo_common.f90): in is the local (s1, s2, s3) pencil, out the halo
 * array with `level` extra cells on both sides of the pencil's two
 * distributed axes ((s1, s2+2l, s3+2l) for ipencil 1 = x-pencil, (s1+2l, s2,
 * s3+2l) for 2, (s1+2l, s2+2l, s3) for 3), i fastest; opt_global only moves
 * Fortran index bounds, not memory.  The interior is copied, then
 * halo_exchange (src/halo_exchange_{x,y,z}_body.f90) runs axis by axis in
 * the reference's order at one rank: an axis with periodic[axis] set
 * (decomp_2d_init's periodic_bc) is its own neighbour, so its halos wrap
 * round; other axes have no neighbour and their halo cells are left as they
 * were.  Rejects level > the pencil's extent along a periodic axis. */
int d2dhip_update_halo(int elem, const void *in, void *out, int s1, int s2, int s3,
                       int level, int ipencil, const int periodic[3], void *stream);
/* halo_exchange(arr, ipencil, halo_extents, level, [s1, s2, s3]) on a halo
 * array already holding its interior.  arr must hold exactly the halo
 * array's elements: (s1, s2, s3) plus 2*level on each axis other than
 * ipencil (there is no size argument to check it against).  Likewise `out`
 * of d2dhip_update_halo. */
int d2dhip_halo_exchange(int elem, void *arr, int s1, int s2, int s3, int level,
                         int ipencil, const int periodic[3], void *stream);

/* ---- synthetic inputs (no reference counterpart; used by tests/benchmarks) -
 * SplitMix64 stream keyed by seed, counter = 0-based Fortran linear index;
 * uniform values 2u-1 in [-1,1).  Complex elements use counters 2*idx (re)
 * and 2*idx+1 (im).  elem is D2DHIP_R8/C16/R4/C8. */
int d2dhip_fill_uniform(void *dst, long long n, int elem,
                        unsigned long long seed, void *stream);
void d2dhip_host_fill_uniform(void *dst, long long n, int elem,
                              unsigned long long seed);
/* Bandwidth reference (no reference counterpart; used by the benchmark):
 * copies the n points (sj elements apart) of the complex columns
 * c0 + c1*s1, c0 < C0, c1 < C1, from src to the same positions of dst, in
 * the tile order and access pattern of a strided axis stage (16 columns per
 * 1024-thread block) but without arithmetic.  elem D2DHIP_C16 or _C8. */
int d2dhip_column_copy(const void *src, void *dst, int elem, long long C0, long long C1,
                       long long s1, long long sj, int n, void *stream);
/* Modal field sum_m a_m cos(2 pi (kx i/nx + ky j/ny + kz k/nz)) + sigma*xi,
 * i,j,k 0-based, with nmodes integer wavevectors uniform in [-kmax,kmax]^3;
 * amp_kind 0: a_m ~ N(0,1); amp_kind 1: |a_m| = 10^U(amp_lo,amp_hi) with a
 * random sign.  xi ~ N(0,1) white noise.  Real output (elem R8 or R4). */
typedef struct {
  int nmodes;
  int kx[64], ky[64], kz[64];
  double amp[64];
  double sigma;
  unsigned long long noise_seed;
} d2dhip_modes;
int d2dhip_modes_make(d2dhip_modes *m, int nmodes, int kmax, int amp_kind,
                      double amp_lo, double amp_hi, double sigma,
                      unsigned long long seed);
int d2dhip_fill_modal(void *dst, int nx, int ny, int nz, int elem,
                      const d2dhip_modes *m, void *stream);
void d2dhip_host_fill_modal(void *dst, int nx, int ny, int nz, int elem,
                            const d2dhip_modes *m);

#ifdef __cplusplus
}
#endif
#endif /* D2DHIP_H */
