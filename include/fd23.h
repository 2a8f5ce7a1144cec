/*
 * fd23.h -- C ABI of the MI355X-native (gfx950) staggered-grid elastic
 * time-stepping library.
 *
 * The library replaces, behind plain C entry points, the calls that the
 * reference's per-GPU driver `singlegpu()` makes into its CUDA code
 * (PSV/clean_nobox/main.cpp:41-181, SH/main.cpp:38-142).  Each entry point
 * below names the reference function it stands in for.
 *
 * Conventions
 *   - Every function returns FD23_OK (0) or a negative status; none exits.
 *     fd23_last_error() returns the message of the last failure on the
 *     calling thread.
 *   - The caller owns host buffers; a context owns its device buffers.
 *   - A context is bound to one HIP stream and is not thread-safe; several
 *     contexts may run concurrently on separate streams.
 *   - Field values are physical velocities / stresses as in the reference;
 *     material arrays are pre-scaled by dt/h as the reference's get_model
 *     leaves them.
 */
#ifndef FD23_H
#define FD23_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define FD23_OK 0
#define FD23_EINVAL -1   /* bad argument or call order                   */
#define FD23_ENOMEM -2   /* device allocation failed / over the HBM cap  */
#define FD23_EHIP -3     /* HIP runtime error                            */
#define FD23_EIO -4      /* file missing or unreadable                   */
#define FD23_ESTAB -5    /* CFL stability number > 1                     */

/* ---- enumerations ---------------------------------------------------- */
#define FD23_WAVE_PSV 0 /* vx=U, vz=W, sxx=Txx, szz=Tzz, sxz=Txz         */
#define FD23_WAVE_SH 1  /* vy=V, sxy=Tx, szy=Tz                          */

#define FD23_STORE_F32 0
#define FD23_STORE_F16 1  /* FP16 storage, per-field power-of-two scale  */
#define FD23_STORE_BF16 2 /* BF16 storage, per-field power-of-two scale  */
#define FD23_STORE_F64 3  /* double fields and material                  */

/* source kinds (main.cpp:124-136 srctype, SH/main.cpp:100-112 stype)    */
#define FD23_SRC_LINE 0  /* '1' / stype 3: stf[it] while it < lsrc       */
#define FD23_SRC_P 1     /* 'p' / stype 13: running-sum stf, clamped     */
#define FD23_SRC_FORCE 2 /* SH point force on V (cu_point_force)         */

/* field ids for fd23_get_field / fd23_set_field                         */
#define FD23_F_U 0
#define FD23_F_W 1
#define FD23_F_TXX 2
#define FD23_F_TXZ 3
#define FD23_F_TZZ 4
#define FD23_F_V 0
#define FD23_F_TX 1
#define FD23_F_TZ 2

/* block of the first-arrival window (SH/param/const.h:20-21 BDIMX, BDIMY) */
#define FD23_PUNCH_BX 32
#define FD23_PUNCH_BZ 16

typedef struct fd23_ctx fd23_ctx;

typedef struct fd23_desc {
    int wave;       /* FD23_WAVE_*                                        */
    int storage;    /* FD23_STORE_*                                       */
    int nx, nz;     /* grid; any size >= 16 (no nx%32 / nz%16 rule)       */
    int npml;       /* C-PML width in cells                              */
    int topabsorb;  /* 1 = C-PML on top instead of the free surface       */
    int nshots;     /* shots sharing one model, stepped in one launch     */
    int max_frames; /* capacity of the on-device seismogram buffer        */
    int max_rec;    /* receivers per component per shot                   */
    void *stream;   /* hipStream_t to run on, or NULL for an own stream   */
} fd23_desc;

/* ---- context --------------------------------------------------------
 * Replaces FIELD::init_gpu_full (PSV/clean_nobox/field/cu_field.cu:7-23),
 * MATERIAL::init_gpu_full (material/cu_material.cu:8-50), the CPML device
 * constructor (cpml/cu_cpml.cu:5-90) and RECORD::cu_init_record
 * (record/cu_record.cu:79-110).  Fields start at zero.                  */
int fd23_ctx_create(const fd23_desc *desc, fd23_ctx **out);
int fd23_ctx_destroy(fd23_ctx *ctx);
const char *fd23_last_error(void);
/* device bytes the context holds */
size_t fd23_ctx_device_bytes(const fd23_ctx *ctx);
/* device bytes fd23_ctx_create would allocate for desc (no allocation)  */
size_t fd23_desc_device_bytes(const fd23_desc *desc);

/* Material, already scaled by dt/h (material.cpp:203-211).  P-SV order
 * BU, BW, MU, MUA, LAM; SH order BV, MUX, MUZ.  Each nx*nz, x fastest,
 * float (or double for FD23_STORE_F64).  Replaces init_gpu_full's H2D.  */
int fd23_set_material(fd23_ctx *ctx, const void *const *arrays);
/* The material computed on the device from the model: vp, vs, den nz*nx
 * (double for an FP64 P-SV context, float otherwise; SH reads vs and den
 * and takes FP32, FP16 or BF16 storage).  The arrays equal those of
 * fd23_psv_material[_f64] / fd23_sh_material bitwise, with the same
 * stability number (FD23_ESTAB above 1, the material still set); one
 * pointwise kernel replaces the host pass over the grid.  Replaces
 * MATERIAL::get_model + init_gpu_full (PSV/clean_nobox/material/
 * material.cpp:35-240, cu_material.cu:8-50; SH/material/material.cpp:
 * 31-135).                                                              */
int fd23_set_model(fd23_ctx *ctx, const void *vp, const void *vs, const void *den, double h, double dt,
                   double *stab);
/* As fd23_set_model, but it allocates no device memory: no counterpart in
 * the reference; for callers that cannot call hipMalloc (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * The staging of the model arrays and the two reduction keys live in a
 * fixed per-GPU scratch array of the library (256 bytes of keys plus
 * fd23_set_model_noalloc_capacity() = 32 MiB of staging, held by each GPU
 * the library runs on until the process ends); a model needing more
 * staging (nin * nx * nz * element bytes, nin = 3 for P-SV, 2 for SH) is
 * rejected with FD23_EINVAL.  Calls that overlap on one GPU share that
 * scratch and race; serialise them.  Same results as fd23_set_model,
 * bitwise.                                                              */
int fd23_set_model_noalloc(fd23_ctx *ctx, const void *vp, const void *vs, const void *den, double h,
                           double dt, double *stab);
size_t fd23_set_model_noalloc_capacity(void);
/* The model read from a file and its material computed on the device: a
 * .mdl polygon model (SH/model/model.cpp:175-237) is rasterised by a
 * kernel (each cell takes the last polygon containing it), raster files
 * (model.vp/.vs/.den) are read on the host; then as fd23_set_model.  The
 * model equals fd23_model_read's bitwise (widened to double for FP64).  */
int fd23_set_model_file(fd23_ctx *ctx, const char *model, double h, double dt, double *stab);
/* One material array (index in the fd23_set_material order) copied back
 * as nz*nx values (float, or double for F64); no reference counterpart. */
int fd23_get_material(fd23_ctx *ctx, int index, void *out);

/* C-PML profiles as b, c, k(=1/kappa) triples, double, packed in order
 * P-SV: Txx_x, Txz_x, U_x, W_x (nx each), Tzz_z, Txz_z, U_z, W_z (nz each);
 * SH:   Tx_x, V_x (nx each), Tz_z, V_z (nz each);
 * each profile stored as b[n], c[n], k[n].  Replaces CPML(deviceid, ...). */
int fd23_set_cpml(fd23_ctx *ctx, const double *packed);

/* Point source of one shot.  P-SV M = {Mxx, Mxz, Mzz}, SH M = {Mxy, Mzy},
 * already multiplied by the reference's factor (cu_point_source.cu:37-40).
 * stf has lsrc samples (the running sum for FD23_SRC_P, as main.cpp:94-99
 * passes psrc).  Replaces cu_point_source / cu_point_source_p /
 * cu_point_force (PSV/clean_nobox/point_source/point_source.h:6-7,
 * SH/point_source/point_source.h:6-8).                                  */
int fd23_set_source(fd23_ctx *ctx, int shot, int kind, int xs, int zs,
                    const double *stf, int lsrc, const double *M);

/* Receivers of one shot and component (P-SV 0=U 1=W, SH 0=V), sampled at
 * the flat index rx + rz*nx (cu_record.cu:142).  Replaces cu_init_record. */
int fd23_set_receivers(fd23_ctx *ctx, int shot, int comp, const int *rx,
                       const int *rz, int n);

/* Shot slots, for callers that start shots while a context runs (no
 * counterpart in the reference, whose shots all start at step 0).  A shot
 * with first step t0 injects its source sample i at step t0 + i and
 * records its frame f at step t0 + f*itrecord, so its seismograms read
 * back from frame 0 as if it had run alone from step 0.  An idle shot
 * (active 0; shots 0..63) is not stepped by fd23_run.  Defaults: active,
 * t0 = 0.  A change takes effect on the next fd23_run (its graphs are
 * captured again).  Contexts with a box or a first-arrival window run
 * every shot from step 0: fd23_run refuses other slots there.           */
int fd23_set_shot_slot(fd23_ctx *ctx, int shot, int active, int t0);
/* Zero one shot's fields, C-PML memory (both buffer sets) and seismogram
 * entries, enqueued on the context stream: the state of a shot that has
 * not started.                                                           */
int fd23_reset_shot(fd23_ctx *ctx, int shot);

/* ---- stepping --------------------------------------------------------
 * One stage each, in the reference's order (main.cpp:121-150):
 * cu_step_stress (fd/cu_fd.cu:391-483), source injection, cu_step_velocity
 * (:485-588), RECORD::cu_record (record/cu_record.cu:154-170).  All shots
 * of the context advance together.  Asynchronous on the context stream. */
int fd23_step_stress(fd23_ctx *ctx);
int fd23_inject(fd23_ctx *ctx, int it);
int fd23_step_velocity(fd23_ctx *ctx);
int fd23_record(fd23_ctx *ctx, int it, int itrecord);

/* The whole loop for it in [it0, it0+nt): stress, inject, velocity,
 * record every itrecord steps -- captured in HIP graphs and replayed
 * (128-step chunks plus one graph of the remaining nt % 128 steps).
 * Equivalent to the singlegpu() loop body (main.cpp:115-168).           */
int fd23_run(fd23_ctx *ctx, int it0, int nt, int itrecord);
/* Captures and instantiates the HIP graphs fd23_run(ctx, *, nt, itrecord)
 * replays (a 128-step chunk and one of nt % 128 steps), without running
 * them, so that run spends no time on capture; no reference counterpart
 * (the reference launches every kernel from the host).                 */
int fd23_prepare(fd23_ctx *ctx, int nt, int itrecord);
int fd23_sync(fd23_ctx *ctx);

/* First-arrival window ("punch", usepunch=1): times is the nz*nx float
 * travel-time grid (row-major, the timefile layout); per block of 32x16
 * cells, first = (int)((min t - before)/dt), last = (int)((max t +
 * after)/dt), in float, and both passes update the block's cells only
 * at steps first < it < last; sources and receivers are unaffected.
 * Replaces SLIDE::SLIDE (SH/slide/cu_slide.cu:7-61) and the kernels'
 * CHECK_IF_NEED_UPDATE (SH/fd/cu_fd.cu:18-25, PSV/psv_clean_double/fd/
 * cu_fd.cu:16-23).  Needs nx % 32 == 0, nz % 16 == 0 and FP32 or FP64
 * storage (else FD23_EINVAL); applies to fd23_run (the staged entry
 * points return FD23_EINVAL while a window is set).  times = NULL
 * removes the window.  Workgroups whose blocks are all outside their
 * window and hold no source tap or receiver return at once.           */
int fd23_set_punch(fd23_ctx *ctx, const float *times, float before,
                   float after, float dt);
/* The FP64 tree's window: times is the nz*nx double travel-time grid of
 * its timefile and the quotients are formed in double
 * (PSV/psv_clean_double/slide/cu_slide.cu:7-52); otherwise as
 * fd23_set_punch.                                                       */
int fd23_set_punch_f64(fd23_ctx *ctx, const double *times, double before,
                       double after, double dt);
/* The per-block window tables alone, on the host: first[k*nx/32 + i] and
 * last[...] for block row k and block column i, as SLIDE::SLIDE builds
 * them (float: SH/slide/cu_slide.cu:27-53; double: PSV/psv_clean_double/
 * slide/cu_slide.cu:27-44).  Quotients beyond the int range are clamped
 * to it.  FD23_EINVAL unless nx % 32 == 0, nz % 16 == 0 and dt > 0.     */
int fd23_punch_blocks(const float *times, int nx, int nz, float before,
                      float after, float dt, int *first, int *last);
int fd23_punch_blocks_f64(const double *times, int nx, int nz,
                          double before, double after, double dt,
                          int *first, int *last);

/* ---- outputs ---------------------------------------------------------
 * Seismograms of one shot/component, time-major [nframes][n] floats
 * (the raw {output}_U file layout, cu_record.cu:148-152).               */
int fd23_get_seismograms(fd23_ctx *ctx, int shot, int comp, float *out,
                         int nframes);
/* One field of one shot as nx*nz values (float, or double for F64).
 * Replaces cpy_backhost (record/cu_record.cu:119-135).                  */
int fd23_get_field(fd23_ctx *ctx, int shot, int field, void *out);
int fd23_set_field(fd23_ctx *ctx, int shot, int field, const void *in);
/* Checkpoint in the state%06d layout (main.cpp:253-318): fields then the
 * x-psi and z-psi arrays; values are float (double for F64).
 * Replaces FIELD/CPML::cu_save_state and cu_load_restart.               */
size_t fd23_state_bytes(const fd23_ctx *ctx);
int fd23_get_state(fd23_ctx *ctx, int shot, void *buf, size_t bytes);
int fd23_set_state(fd23_ctx *ctx, int shot, const void *buf, size_t bytes);

/* ---- hybrid box coupling ---------------------------------------------
 * Ring add/sub + residual sub-grid of PSV/psv_clean_double/box/box.cpp,
 * generalised to a bx*bz rectangle whose drive is an analytic plane wave
 * (incident=0 P, 1 SV; Ricker f0, delay t0; angle from vertical, waves
 * travelling upward).  Box cells [x0, x0+bx) x [z0, z0+bz).  P-SV only. */
int fd23_box_configure(fd23_ctx *ctx, int shot, int x0, int z0, int bx, int bz,
                       int incident, double angle_deg, double f0, double t0,
                       double amp, double vp, double vs, double rho,
                       double h, double dt);
/* The reference's own box, driven by its sourcefile: a bw x bw square
 * around (xs, zs) (cell (i,j) at (xs+s0+i, zs+s0+j), s0 = (bw+1)/2+1-bw),
 * frames = nframes displacement frames, each U then W on (bw+8)^2 cells
 * (x fastest, the sourcefile layout), step it driven by frames it and
 * it+1: velocities (u[it+1]-u[it])/dt and stresses C grad u[it+1] / dt
 * with the box material and the derivative order reduced toward the box
 * centre.  The source medium (srcpvel, srcsvel, srcden) replaces the
 * global material at box cells differing from it by more than 1e-4
 * relative.  fd23_run returns FD23_EINVAL for steps whose frame it+1 is
 * missing.  Replaces BOX::init_box, BOX::setboxsource and
 * BOX::cu_copy_source (PSV/psv_clean_double/box/box.cpp:10-197,
 * box/cu_box.cu:35-44) and MATERIAL::init_for_box_from_full
 * (material/material.cpp:39-99); cu_copy_source's half-size copy (it
 * copies sizeof(int)*bw*bw bytes of each double array) is not
 * reproduced.                                                           */
int fd23_box_configure_frames(fd23_ctx *ctx, int xs, int zs, int bw, const double *frames, int nframes,
                              double srcpvel, double srcsvel, double srcden, double h, double dt);
/* box cells whose material fd23_box_configure_frames set to the source
 * medium (the reference prints one line per such cell)                  */
int fd23_box_forced_cells(fd23_ctx *ctx);
/* residual-grid fields of one shot, U W Txx Tzz Txz (bz*bx each; float, or
 * double for F64).  Replaces reading BOX's residual FIELD back.          */
int fd23_box_get_residual(fd23_ctx *ctx, int shot, void *out);

/* ---- tuning and introspection ---------------------------------------- */
/* tile shape index (0 = default); fd23_autotune times the candidates on
 * this context's own arrays and keeps the fastest.  State is preserved. */
int fd23_set_tile(fd23_ctx *ctx, int tile);
int fd23_get_tile(const fd23_ctx *ctx);
int fd23_num_tiles(void);
/* shape of a tile index.  Kinds: FD23_TILE_TILED, two-pass kernels on 2D
 * tiles of 64 columns x 4*rz rows (rz rows per thread, spb shots per
 * workgroup) staged through LDS; FD23_TILE_FUSED, the fused P-SV step
 * (stress and velocity in one pass over two buffer sets, segments of
 * seg_rows rows in blocks of block_rows, or block_rows 0: the
 * tile-parallel step, FP32/FP64 storage only; 0 for tiled kinds);
 * FD23_TILE_BLOCK, the SH block step (seg_rows = K steps per launch over
 * regions of 64 columns x block_rows rows, rz rows per thread, each
 * region a tile plus a 7*K-cell halo that the launch recomputes, over two
 * buffer sets).  The fused step applies to fd23_run of P-SV contexts
 * without a box, the block step to fd23_run of SH contexts with FP32/FP64
 * storage, both without a first-arrival window; the staged entry points,
 * the box and the window run the tiled kernels (tile 0's shape).  No
 * reference counterpart: the reference launches two kernels per step
 * (SH/main.cpp:93-141).                                                 */
enum { FD23_TILE_TILED = 0, FD23_TILE_FUSED = 1, FD23_TILE_BLOCK = 2 };
int fd23_tile_shape(int tile, int *rz, int *spb, int *seg_rows, int *block_rows);
/* FD23_TILE_* of a tile index, negative for a bad index                 */
int fd23_tile_kind(int tile);
/* 1 when a tiled shape launches the interior tiles and the edge tiles of
 * each P-SV pass as two kernels, each compiled for one path (fewer
 * registers per lane, one more launch per pass); 0 otherwise           */
int fd23_tile_split(int tile);
int fd23_autotune(fd23_ctx *ctx, int steps, double *best_ms);
/* Chains (no reference counterpart): fd23_run captures each chunk's steps
 * as n = 2 or 4 independent chains of passes, chain i over the i-th of n
 * parts of the shot groups, on the context stream and n-1 more streams,
 * so that one chain's kernels fill the others' gaps between dependent
 * passes.  Applies to the two-pass path with at least n shot groups and
 * FP32/FP64 storage, no box and no first-arrival window; results are
 * bitwise the same.  n = 0 or 1: one chain.  fd23_autotune chooses n with
 * the tile shape.  fd23_get_chains: the number fd23_run uses.           */
int fd23_set_chains(fd23_ctx *ctx, int n);
int fd23_get_chains(const fd23_ctx *ctx);
/* FP16/BF16 storage: per-array power-of-two scales (stored = value/scale);
 * P-SV order U W Txx Tzz Txz then psi Txx_x Txz_x U_x W_x Tzz_z Txz_z U_z
 * W_z; SH order V Tx Tz then psi Tx_x V_x Tz_z V_z.  Set before stepping
 * to override the estimate from the sources (each scale a power of two in
 * [2^-100, 2^100], else FD23_EINVAL); get returns the scales and the max
 * |value| per array gathered for the next scale check (on the step before
 * each check).                                                          */
int fd23_set_scales(fd23_ctx *ctx, const float *scales, int n);
int fd23_get_scales(fd23_ctx *ctx, float *scales, unsigned *amax_bits, int n);
/* algorithmic HBM bytes of one stress+velocity step for this context   */
double fd23_bytes_per_step(const fd23_ctx *ctx);
/* Device-to-device copy of `bytes` (a multiple of 16) by a grid-stride
 * kernel of `grid` workgroups with 16-byte lanes, enqueued on `stream`
 * (hipStream_t, NULL = default): the practical HBM bandwidth the passes
 * are measured against, and the copy that scales the memory counters.
 * No reference counterpart.                                             */
int fd23_device_copy(void *dst, const void *src, size_t bytes, int grid, void *stream);
/* The passes' memory pattern without their arithmetic, enqueued on
 * `stream`: nreads (1..8) FP32 arrays read and nwrites (1..3) written,
 * nx x nz each (nx a multiple of 64, nz of 16), in the passes' 64 x 16
 * tiles and XCD band order; writes[k] = sum of the reads + k.  The
 * bandwidth a pass's access pattern sustains (the stress pass reads 8
 * arrays and writes 3, the velocity pass reads 7 and writes 2).  No
 * reference counterpart.                                                */
int fd23_device_stream_tiles(const float *const *reads, int nreads, float *const *writes, int nwrites,
                             int nx, int nz, void *stream);
/* Time the passes with HIP events on the context stream: one untimed
 * stress+velocity step, then nsteps (1..4096) stress passes back to back,
 * then nsteps velocity passes; fills the average milliseconds of each
 * pass.  The passes change the fields (no injection, no recording): use
 * it on a context whose state is no longer needed, or save the state.   */
int fd23_time_passes(fd23_ctx *ctx, int nsteps, double *stress_ms,
                     double *velocity_ms);

/* ---- host precompute (boundary code, no GPU) --------------------------
 * C-PML profile (cpml.cpp:7-83), computed in double.                    */
int fd23_cpml_profile(int n, int npml, double startpos, double pml_dt,
                      double pml_r, double pml_v, double pml_fc, double *b,
                      double *c, double *k);
/* packed profiles for fd23_set_cpml, P-SV or SH (cpml.cpp:84-186)       */
int fd23_cpml_pack(int wave, int nx, int nz, int npml, double pml_dt,
                   double pml_r, double pml_v, double pml_fc, double *out);
size_t fd23_cpml_pack_len(int wave, int nx, int nz);
/* P-SV material (PSV/clean_nobox/material/material.cpp:35-240), float   */
int fd23_psv_material(int nx, int nz, double h, double dt, const float *vp,
                      const float *vs, const float *den, float *BU, float *BW,
                      float *MU, float *MUA, float *LAM, double *stab);
int fd23_psv_material_f64(int nx, int nz, double h, double dt,
                          const double *vp, const double *vs,
                          const double *den, double *BU, double *BW,
                          double *MU, double *MUA, double *LAM, double *stab);
/* SH material (SH/material/material.cpp:31-135)                         */
int fd23_sh_material(int nx, int nz, double h, double dt, const float *vs,
                     const float *den, float *BV, float *MUX, float *MUZ,
                     double *stab);
/* moment-tensor components incl. scaling (cu_point_source.cu:16-40);
 * explosion=1 gives the -DEXPLOSION_SOURCE tensor                       */
int fd23_psv_moment(int kind, int explosion, double strike, double dip,
                    double rake, double azimuth, double h, double dt,
                    double *M);
int fd23_sh_moment(int kind, double strike, double dip, double rake,
                   double azimuth, double h, double dt, double *M);
/* source-time functions (source/waveform.cpp:8-47) and a Ricker        */
int fd23_stf_gaussian(int alpha, double *out, int cap, int *lsrc);
int fd23_stf_trapezoid(double dt, double trap1, double trap2, double trap3,
                       double *out, int cap, int *lsrc);
int fd23_stf_ricker(int n, double dt, double f0, double t0, double *out);
/* the FP64 tree's wavelets and moment tensor, in double throughout
 * (PSV/psv_clean_double/source/waveform.cpp:8-47,
 * point_source/cu_point_source.cu:11-27,90; no explosion variant)       */
int fd23_stf_gaussian_f64(int alpha, double *out, int cap, int *lsrc);
int fd23_stf_trapezoid_f64(double dt, double trap1, double trap2, double trap3,
                           double *out, int cap, int *lsrc);
int fd23_psv_moment_f64(int kind, double strike, double dip, double rake,
                        double azimuth, double h, double dt, double *M);
/* model: "{name}.mdl" polygons (SH/model/model.cpp:175-469) or raster
 * "{name}.vp/.vs/.den" float32 files (SH/model/util.cpp:12-29)          */
int fd23_model_read(const char *name, int nx, int nz, double h, float *vp,
                    float *vs, float *den);
/* the FP64 tree's reader: .mdl vertices, values and cell bounds in double
 * (PSV/psv_clean_double/model/model.cpp:175-237); raster files widened  */
int fd23_model_read_f64(const char *name, int nx, int nz, double h, double *vp,
                        double *vs, double *den);
/* ISIS trace file (PSV/clean_nobox/record/isis.cpp:8-61): raw is
 * time-major [ntrecord][nrec]                                           */
int fd23_isis_write(const char *path, const float *raw, int nrec,
                    int ntrecord, const int *rx, const int *rz, double shiftx,
                    double shiftz, double h, double dt, int itrecord, int xs,
                    int zs);
/* the FP64 tree's ISIS file ({output}_U.db.isis, PSV/psv_clean_double/
 * record/isis.cpp:8-61): 136-byte double header (source position shifted
 * by src_shift_x/z, user[1] = bw), double samples                       */
int fd23_isis_write_f64(const char *path, const double *raw, int nrec, int ntrecord,
                        const int *rx, const int *rz, double shiftx, double shiftz,
                        double h, double dt, int itrecord, int xs, int zs,
                        double src_shift_x, double src_shift_z, int bw);

/* receiver pairs (x, z), (x-1, z) per station: a line of nsta stations
 * (RECORD::init_line, PSV/clean_nobox/record/cu_record.cu:12-28) or a
 * station file of "z x" pairs (RECORD::init_list, :30-76); rx/rz hold
 * 2*nsta entries; fd23_station_read with rx == NULL only counts         */
int fd23_line_receivers(int nsta, int ix0, int iz0, int idx, int idz, int *rx,
                        int *rz);
int fd23_station_read(const char *path, int *rx, int *rz, int cap, int *nrec);

/* ---- par files (Caltech getpar semantics, getpar/getpar.c) ----------- */
typedef struct fd23_par fd23_par;
int fd23_par_open(int argc, const char *const *argv, fd23_par **out);
void fd23_par_close(fd23_par *p);
/* return 1 if found (last occurrence wins), 0 if not, <0 on error       */
int fd23_par_int(fd23_par *p, const char *name, int *val);
int fd23_par_float(fd23_par *p, const char *name, float *val);
int fd23_par_double(fd23_par *p, const char *name, double *val);
int fd23_par_string(fd23_par *p, const char *name, char *buf, int cap);

#ifdef __cplusplus
}
#endif
#endif /* FD23_H */
