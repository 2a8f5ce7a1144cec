// fd23 context: device buffers, launchers, graph-captured loop, and the
// C ABI declared in include/fd23.h.
#include "fd23.h"
#include "fd_kernels.hpp"
#include "fd_fused.hpp"
#include "fd_block.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <tuple>
#include <string>
#include <vector>

using namespace fd23;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
#define HIPCHK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess)                                                      \
            return fail(FD23_EHIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_),    \
                        __FILE__, __LINE__);                                       \
    } while (0)

extern "C" const char *fd23_last_error(void) { return g_err.c_str(); }
extern "C" void fd23_set_error_(const char *msg) { g_err = msg; }

// ---------------------------------------------------------------------------
// tile shapes: two-pass tiles (rows per thread RZ, tile height 4*RZ, shots
// per workgroup SPB), fused P-SV steps (segment height FH rows) and SH
// block steps (BK steps per launch over regions of 64 x BW*BCZ cells, BW
// wavefronts per workgroup)
// ---------------------------------------------------------------------------
// split: P-SV passes launch their interior tiles and their edge tiles as
// two kernels, each compiled for one path (fewer registers, more waves per
// SIMD; one more launch per pass, which small grids feel)
struct TileCfg {
    int rz, spb, fh, fr, split; // fused: segment height fh, rows per block fr
    int bk, bcz, bw;            // block: steps per launch, rows per thread, wavefronts
};
// tile 0, the shape untuned contexts run (the serving engine, the driver),
// is the one the autotuner picks on most workloads measured
static constexpr TileCfg kTiles[] = {
    {2, 1, 0, 0, 0},   {4, 1, 0, 0, 0},   {8, 1, 0, 0, 0},    {4, 2, 0, 0, 0},   {4, 4, 0, 0, 0},
    {4, 8, 0, 0, 0},   {8, 2, 0, 0, 0},   {8, 4, 0, 0, 0},    {2, 4, 0, 0, 0},   {2, 8, 0, 0, 0},
    {0, 1, 32, 4, 0},  {0, 1, 64, 4, 0},  {0, 1, 128, 4, 0},  {0, 1, 64, 2, 0},  {2, 1, 0, 0, 1},
    {4, 1, 0, 0, 1},   {2, 2, 0, 0, 1},   {4, 2, 0, 0, 1},
    {0, 1, 0, 0, 0, 1, 3, 8},  {0, 1, 0, 0, 0, 1, 3, 16}, {0, 1, 0, 0, 0, 2, 3, 16},
    {0, 1, 0, 0, 0, 2, 4, 16}, {0, 1, 0, 0, 0, 2, 6, 8},
    {0, 1, 16, 0, 0}, {0, 1, 32, 0, 0}, {1, 1, 0, 0, 0},
    {0, 1, 0, 0, 0, 2, 4, 10}, {0, 1, 0, 0, 0, 2, 3, 12}};
static const int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);

extern "C" int fd23_num_tiles(void) { return kNumTiles; }
extern "C" int fd23_tile_shape(int tile, int *rz, int *spb, int *seg_rows, int *block_rows) {
    if (tile < 0 || tile >= kNumTiles) return fail(FD23_EINVAL, "bad tile %d", tile);
    const TileCfg &t = kTiles[tile];
    if (rz) *rz = t.bk > 0 ? t.bcz : t.rz;
    if (spb) *spb = t.spb;
    if (seg_rows) *seg_rows = t.bk > 0 ? t.bk : t.fh;
    if (block_rows) *block_rows = t.bk > 0 ? t.bw * t.bcz : t.fr;
    return FD23_OK;
}
extern "C" int fd23_tile_split(int tile) {
    if (tile < 0 || tile >= kNumTiles) return fail(FD23_EINVAL, "bad tile %d", tile);
    return kTiles[tile].split;
}
extern "C" int fd23_tile_kind(int tile) {
    if (tile < 0 || tile >= kNumTiles) return fail(FD23_EINVAL, "bad tile %d", tile);
    return kTiles[tile].bk > 0 ? FD23_TILE_BLOCK : kTiles[tile].fh > 0 ? FD23_TILE_FUSED : FD23_TILE_TILED;
}

constexpr int kMaxTZ = 32;     // largest tile height (BY * max RZ)
constexpr int kGuard = 256;    // zero elements before shot 0 / after the last

static int round_up(int a, int b) { return (a + b - 1) / b * b; }

struct HostSrc {
    int kind = -1, xs = 0, zs = 0;
    int t0 = 0; // the shot's first step (fd23_set_shot_slot)
    std::vector<double> stf;
    double M[3] = {0, 0, 0};
};

struct fd23_ctx {
    fd23_desc d{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // chain forks and joins (used under stream capture only)
    hipEvent_t ev_fork = nullptr;
    // chains (fd23_set_chains, chosen by the autotuner): the launches of
    // chain i of n cover the i-th of n parts of the shot groups, on the
    // context stream (i = 0) or chain_stream[i-1]; chain -1: all groups
    hipStream_t chain_stream[3] = {};
    hipEvent_t chain_join[3] = {};
    int chain = -1;
    int chains = 1;
    bool timing = false; // graphs captured for timing: no injection, no recording
    int esize = 4;     // bytes per stored field/psi element
    int csize = 4;     // bytes per compute/material element
    int px = 0, rows = 0;
    long sstride = 0;  // elements between shots
    long xpsi = 0, zpsi = 0;
    int nfield = 0, npsi = 0, nmat = 0, ncomp = 0, nscale = 0;

    void *field_alloc[5] = {};
    void *fields[5] = {};   // shot 0, row 0, col 0
    void *psi[8] = {};
    // the other buffer set of the fused step (allocated on first use); the
    // fused step reads the current set and writes this one, then they swap
    void *alt_alloc[5] = {};
    void *alt_fields[5] = {};
    void *alt_psi[8] = {};
    int cur = 0;            // which set is current (keys the captured graphs)
    bool tables_dirty = false;
    void *mat[5] = {};      // row 0, col 0 (4 zero rows above and below, zero guards)
    void *mat_alloc[5] = {};
    void *prof = nullptr;   // [npsi][3][n] compute type
    long prof_off[8] = {};  // element offset of profile p (b, then c, then k)
    int prof_n[8] = {};
    void **d_fieldtab = nullptr;
    void **d_comptab = nullptr;
    int *d_compfield = nullptr;
    float *scl = nullptr;
    unsigned *amax = nullptr;
    int staged_steps = 0; // steps taken through the staged entry points
    // first-arrival window per 32 x 16 block (fd23_set_punch)
    bool punch = false;
    int *d_pfirst = nullptr, *d_plast = nullptr;
    size_t pfirst_cap = 0, plast_cap = 0;
    std::vector<int> h_pfirst, h_plast; // per block, row-major over nx/32 blocks
    int4 *d_pwin = nullptr;             // per (tile, shot), built with the event lists
    size_t pwin_cap = 0;
    int *clock = nullptr;
    bool scales_init = false;
    double bmax = 0;   // largest buoyancy-type material value (scale estimate)

    // sources
    std::vector<HostSrc> src;
    Tap *d_taps = nullptr;
    SrcInfo *d_src = nullptr;
    double *d_stf = nullptr;
    bool src_dirty = true;
    size_t taps_cap = 0, stf_cap = 0;
    std::vector<Tap> h_taps;
    std::vector<SrcInfo> h_srcinfo;

    // per-strip lists of the fused step
    int *d_ftap_start = nullptr, *d_frec_start = nullptr;
    TapF *d_ftaps = nullptr;
    RecF *d_frecs = nullptr;
    size_t ftap_start_cap = 0, frec_start_cap = 0, ftaps_cap = 0, frecs_cap = 0;
    bool fevents_dirty = true;
    int fevents_tile = -1; // tile the lists were built for

    // per-(tile, shot) lists of the SH block step: taps of every region cell,
    // receivers of the tile's own cells
    int *d_btap_start = nullptr, *d_brec_start = nullptr;
    TapT *d_btaps = nullptr;
    RecT *d_brecs = nullptr;
    size_t btap_start_cap = 0, brec_start_cap = 0, btaps_cap = 0, brecs_cap = 0;
    bool bevents_dirty = true;
    int bevents_tile = -1;

    // per-tile lists of the fused graph loop
    int *d_tap_start = nullptr, *d_rec_start = nullptr;
    TapT *d_tapsT = nullptr;
    RecT *d_recs = nullptr;
    size_t tap_start_cap = 0, rec_start_cap = 0, tapsT_cap = 0, recs_cap = 0;
    bool events_dirty = true;
    int events_tile = -1;

    // receivers
    long *d_roff = nullptr;
    int *d_nrec = nullptr;
    std::vector<long> h_roff;
    std::vector<int> h_nrec;
    void *seis = nullptr;

    int tile = 0;
    // idle shot slots 0..63 (fd23_set_shot_slot): a kernel argument of the
    // graph loop's launches, so a change invalidates the captured graphs
    unsigned long long idle = 0;
    bool have_mat = false, have_cpml = false;

    // graphs keyed by (itrecord, steps, current set); flips: the chunk
    // swaps the buffer sets an odd number of times
    struct Chunk {
        hipGraphExec_t exec;
        bool flips;
    };
    std::map<std::tuple<int, int, int>, Chunk> graphs;

    // box coupling state (see fd23_box.hip)
    void *box = nullptr;
};

// implemented in fd23_box.hip
int fd23_box_launch_pre_stress(fd23_ctx *c, int k);
int fd23_box_launch_post_stress(fd23_ctx *c, int k);
int fd23_box_launch_post_velocity(fd23_ctx *c, int k);
void fd23_box_free(fd23_ctx *c);
void fd23_box_set_step(fd23_ctx *c, int it);
int fd23_box_check_run(fd23_ctx *c, int it0, int nt);

// accessors used by the box and material modules
hipStream_t fd23_ctx_stream(fd23_ctx *c) { return c->stream; }
void *fd23_ctx_field(fd23_ctx *c, int internal) { return c->fields[internal]; }
void *fd23_ctx_material(fd23_ctx *c, int i) { return c->mat[i]; }
int fd23_ctx_px(fd23_ctx *c) { return c->px; }
long fd23_ctx_sstride(fd23_ctx *c) { return c->sstride; }
void *&fd23_ctx_box(fd23_ctx *c) { return c->box; }
const fd23_desc &fd23_ctx_desc(fd23_ctx *c) { return c->d; }
int *fd23_ctx_clock(fd23_ctx *c) { return c->clock; }
// captured graphs bake in the per-step launches; the box changes them.  A
// replay may still run on the context stream: it ends before its graph is
// destroyed.
void fd23_ctx_invalidate_graphs(fd23_ctx *c) {
    if (!c->graphs.empty()) hipStreamSynchronize(c->stream);
    for (auto &kv : c->graphs) hipGraphExecDestroy(kv.second.exec);
    c->graphs.clear();
}
const float *fd23_ctx_scl(fd23_ctx *c) { return c->scl; }
// the material module computed the arrays on the device
void fd23_ctx_material_ready(fd23_ctx *c, double bmax) {
    c->bmax = bmax;
    c->have_mat = true;
}

static bool is_psv(const fd23_ctx *c) { return c->d.wave == FD23_WAVE_PSV; }
// material arrays: nz rows of pitch px with HALO zero rows above and
// HALO+1 below and kGuard zero elements at both ends, so that clamped row
// indices in [-HALO, nz+HALO] and any column of the strips read zeros
static size_t mat_elems_for(int nz, int px) { return (size_t)(nz + 2 * HALO + 1) * px + 2 * kGuard; }
static size_t mat_elems(const fd23_ctx *c) { return mat_elems_for(c->d.nz, c->px); }

static size_t layout(const fd23_desc *d, int *px, int *rows, long *sstride, long *xpsi, long *zpsi, int *esize,
                     int *csize) {
    *px = round_up(d->nx + HALO, 64);
    *rows = round_up(d->nz, kMaxTZ) + 2 * HALO + 1;
    *sstride = (long)(*rows) * (*px);
    *xpsi = 2L * d->npml * d->nz;
    *zpsi = 2L * d->npml * d->nx;
    switch (d->storage) {
    case FD23_STORE_F16:
    case FD23_STORE_BF16: *esize = 2; *csize = 4; break;
    case FD23_STORE_F64: *esize = 8; *csize = 8; break;
    default: *esize = 4; *csize = 4; break;
    }
    const bool psv = d->wave == FD23_WAVE_PSV;
    const int nf = psv ? 5 : 3, np = psv ? 4 : 2, nm = psv ? 5 : 3, nc = psv ? 2 : 1;
    size_t S = (size_t)d->nshots;
    size_t bytes = 0;
    bytes += (size_t)nf * ((size_t)S * (*sstride) + 2 * kGuard) * (*esize);
    bytes += (size_t)np * S * ((*xpsi) + (*zpsi)) * (*esize);
    bytes += (size_t)nm * ((size_t)(d->nz + 2 * HALO + 1) * (*px) + 2 * kGuard) * (*csize);
    bytes += (size_t)2 * np * 3 * (d->nx + d->nz) * (*csize);
    bytes += (size_t)d->max_frames * S * nc * d->max_rec * (*csize);
    bytes += (size_t)S * nc * d->max_rec * sizeof(long);
    // the second field/psi set of the fused (P-SV) and block (SH) steps
    {
        bytes += (size_t)nf * ((size_t)S * (*sstride) + 2 * kGuard) * (*esize);
        bytes += (size_t)np * S * ((*xpsi) + (*zpsi)) * (*esize);
    }
    return bytes;
}

extern "C" size_t fd23_desc_device_bytes(const fd23_desc *d) {
    int px, rows, es, cs;
    long ss, xp, zp;
    if (!d) return 0;
    return layout(d, &px, &rows, &ss, &xp, &zp, &es, &cs);
}

extern "C" size_t fd23_ctx_device_bytes(const fd23_ctx *c) { return c ? fd23_desc_device_bytes(&c->d) : 0; }

static void free_ctx(fd23_ctx *c) {
    if (!c) return;
    fd23_ctx_invalidate_graphs(c);
    fd23_box_free(c);
    for (auto p : c->field_alloc) if (p) hipFree(p);
    for (auto p : c->psi) if (p) hipFree(p);
    for (auto p : c->alt_alloc) if (p) hipFree(p);
    for (auto p : c->alt_psi) if (p) hipFree(p);
    for (auto p : c->mat_alloc) if (p) hipFree(p);
    void *others[] = {c->prof, c->d_fieldtab, c->d_comptab, c->d_compfield, c->scl, c->amax, c->clock,
                      c->d_taps, c->d_src, c->d_stf, c->d_roff, c->d_nrec, c->seis, c->d_tap_start,
                      c->d_rec_start, c->d_tapsT, c->d_recs, c->d_ftap_start, c->d_frec_start, c->d_ftaps,
                      c->d_frecs, c->d_pfirst, c->d_plast, c->d_pwin, c->d_btap_start, c->d_brec_start,
                      c->d_btaps, c->d_brecs};
    for (auto p : others) if (p) hipFree(p);
    if (c->own_stream && c->stream) hipStreamDestroy(c->stream);
    for (int i = 0; i < 3; i++) {
        if (c->chain_stream[i]) hipStreamDestroy(c->chain_stream[i]);
        if (c->chain_join[i]) hipEventDestroy(c->chain_join[i]);
    }
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    delete c;
}

extern "C" int fd23_ctx_create(const fd23_desc *desc, fd23_ctx **out) {
    if (!desc || !out) return fail(FD23_EINVAL, "null argument");
    *out = nullptr;
    const fd23_desc &d = *desc;
    if (d.wave != FD23_WAVE_PSV && d.wave != FD23_WAVE_SH) return fail(FD23_EINVAL, "bad wave type %d", d.wave);
    if (d.storage < 0 || d.storage > 3) return fail(FD23_EINVAL, "bad storage type %d", d.storage);
    if (d.nx < 16 || d.nz < 16) return fail(FD23_EINVAL, "grid %dx%d too small (min 16)", d.nx, d.nz);
    if (d.npml < 0 || 2 * d.npml > d.nx || 2 * d.npml > d.nz)
        return fail(FD23_EINVAL, "npml %d does not fit grid %dx%d", d.npml, d.nx, d.nz);
    if (d.nshots < 1) return fail(FD23_EINVAL, "nshots must be >= 1");
    if (d.max_frames < 0 || d.max_rec < 0) return fail(FD23_EINVAL, "negative record capacity");
    if ((long)d.nx * d.nz > (1L << 30)) return fail(FD23_EINVAL, "grid too large");

    fd23_ctx *c = new fd23_ctx();
    c->d = d;
    size_t bytes = layout(&d, &c->px, &c->rows, &c->sstride, &c->xpsi, &c->zpsi, &c->esize, &c->csize);
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) == hipSuccess && bytes + (64u << 20) > freeb) {
        delete c;
        return fail(FD23_ENOMEM, "context needs %zu bytes, %zu free", bytes, freeb);
    }
    const bool psv = is_psv(c);
    c->nfield = psv ? 5 : 3;
    c->npsi = psv ? 8 : 4;
    c->nmat = psv ? 5 : 3;
    c->ncomp = psv ? 2 : 1;
    c->nscale = psv ? SC_PSV_N : SC_SH_N;
    c->src.resize(d.nshots);
    c->h_nrec.assign((size_t)d.nshots * c->ncomp, 0);
    c->h_roff.assign((size_t)d.nshots * c->ncomp * std::max(d.max_rec, 1), -1);

#define ALLOC(ptr, n)                                                           \
    do {                                                                        \
        hipError_t e_ = hipMalloc((void **)&(ptr), (n));                         \
        if (e_ != hipSuccess) {                                                 \
            free_ctx(c);                                                        \
            return fail(FD23_ENOMEM, "hipMalloc(%zu): %s", (size_t)(n), hipGetErrorString(e_)); \
        }                                                                       \
        hipMemset((ptr), 0, (n));                                               \
    } while (0)

    if (d.stream) {
        c->stream = (hipStream_t)d.stream;
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return fail(FD23_EHIP, "stream creation failed");
        }
        c->own_stream = true;
    }
    if (hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipStreamCreateWithFlags(&c->chain_stream[0], hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->chain_stream[1], hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->chain_stream[2], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->chain_join[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->chain_join[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->chain_join[2], hipEventDisableTiming) != hipSuccess) {
        free_ctx(c);
        return fail(FD23_EHIP, "stream creation failed");
    }
    const size_t S = d.nshots;
    for (int f = 0; f < c->nfield; f++) {
        size_t n = ((size_t)S * c->sstride + 2 * kGuard) * c->esize;
        ALLOC(c->field_alloc[f], n);
        c->fields[f] = (char *)c->field_alloc[f] + (size_t)(kGuard + (long)HALO * c->px) * c->esize;
    }
    for (int p = 0; p < c->npsi; p++) {
        const bool xdir = psv ? p < 4 : p < 2;
        size_t n = std::max<size_t>(1, (size_t)S * (xdir ? c->xpsi : c->zpsi)) * c->esize;
        ALLOC(c->psi[p], n);
    }
    for (int m = 0; m < c->nmat; m++) {
        ALLOC(c->mat_alloc[m], mat_elems(c) * c->csize);
        c->mat[m] = (char *)c->mat_alloc[m] + (size_t)(kGuard + (long)HALO * c->px) * c->csize;
    }
    {
        long off = 0;
        for (int p = 0; p < c->npsi; p++) {
            const bool xdir = psv ? p < 4 : p < 2;
            c->prof_n[p] = xdir ? d.nx : d.nz;
            c->prof_off[p] = off;
            off += 3L * c->prof_n[p];
        }
        ALLOC(c->prof, (size_t)off * c->csize);
    }
    ALLOC(c->d_fieldtab, sizeof(void *) * 5);
    ALLOC(c->d_comptab, sizeof(void *) * 2);
    ALLOC(c->d_compfield, sizeof(int) * 2);
    ALLOC(c->scl, sizeof(float) * 2 * kNextScale);
    ALLOC(c->amax, sizeof(unsigned) * 16 * kAmaxSpread);
    ALLOC(c->clock, sizeof(int) * 2);
    ALLOC(c->d_src, sizeof(SrcInfo) * S);
    ALLOC(c->d_nrec, sizeof(int) * S * c->ncomp);
    ALLOC(c->d_roff, sizeof(long) * S * c->ncomp * std::max(d.max_rec, 1));
    if (d.max_frames > 0 && d.max_rec > 0)
        ALLOC(c->seis, (size_t)d.max_frames * S * c->ncomp * d.max_rec * c->csize);
#undef ALLOC
    // pointer tables
    hipMemcpy(c->d_fieldtab, c->fields, sizeof(void *) * 5, hipMemcpyHostToDevice);
    void *comps[2] = {c->fields[0], psv ? c->fields[1] : nullptr};
    hipMemcpy(c->d_comptab, comps, sizeof(void *) * 2, hipMemcpyHostToDevice);
    int cf[2] = {0, 1};
    hipMemcpy(c->d_compfield, cf, sizeof cf, hipMemcpyHostToDevice);
    std::vector<float> ones(2 * kNextScale, 1.0f);
    hipMemcpy(c->scl, ones.data(), sizeof(float) * ones.size(), hipMemcpyHostToDevice);
    hipMemcpy(c->d_roff, c->h_roff.data(), sizeof(long) * c->h_roff.size(), hipMemcpyHostToDevice);
    c->tile = (d.nshots >= 4) ? 4 : 0;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        free_ctx(c);
        return fail(FD23_EHIP, "device init: %s", hipGetErrorString(e));
    }
    *out = c;
    return FD23_OK;
}

extern "C" int fd23_ctx_destroy(fd23_ctx *c) {
    if (!c) return FD23_OK;
    hipStreamSynchronize(c->stream);
    free_ctx(c);
    return FD23_OK;
}

// ---------------------------------------------------------------------------
// material, C-PML, scales
// ---------------------------------------------------------------------------
extern "C" int fd23_set_material(fd23_ctx *c, const void *const *arrays) {
    if (!c || !arrays) return fail(FD23_EINVAL, "null argument");
    for (int m = 0; m < c->nmat; m++) {
        if (!arrays[m]) return fail(FD23_EINVAL, "material array %d is null", m);
        HIPCHK(hipMemcpy2DAsync(c->mat[m], (size_t)c->px * c->csize, arrays[m], (size_t)c->d.nx * c->csize,
                                (size_t)c->d.nx * c->csize, c->d.nz, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    // largest BU/BW (P-SV) or BV (SH): a velocity is about bmax times a stress
    const int nb = is_psv(c) ? 2 : 1;
    double bm = 0;
    const size_t N = (size_t)c->d.nx * c->d.nz;
    for (int m = 0; m < nb; m++)
        for (size_t q = 0; q < N; q++)
            bm = std::max(bm, std::fabs(c->csize == 8 ? ((const double *)arrays[m])[q] : (double)((const float *)arrays[m])[q]));
    c->bmax = bm;
    c->have_mat = true;
    return FD23_OK;
}

extern "C" int fd23_get_material(fd23_ctx *c, int index, void *out) {
    if (!c || !out || index < 0 || index >= c->nmat) return fail(FD23_EINVAL, "bad arguments");
    HIPCHK(hipMemcpy2DAsync(out, (size_t)c->d.nx * c->csize, c->mat[index], (size_t)c->px * c->csize,
                            (size_t)c->d.nx * c->csize, c->d.nz, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FD23_OK;
}

extern "C" int fd23_set_cpml(fd23_ctx *c, const double *packed) {
    if (!c || !packed) return fail(FD23_EINVAL, "null argument");
    long total = c->prof_off[c->npsi - 1] + 3L * c->prof_n[c->npsi - 1];
    std::vector<char> buf((size_t)total * c->csize);
    for (long i = 0; i < total; i++) {
        if (c->csize == 8)
            ((double *)buf.data())[i] = packed[i];
        else
            ((float *)buf.data())[i] = (float)packed[i];
    }
    HIPCHK(hipMemcpyAsync(c->prof, buf.data(), buf.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_cpml = true;
    return FD23_OK;
}

// per-field power-of-two scales for FP16/BF16 storage
extern "C" int fd23_set_scales(fd23_ctx *c, const float *scales, int n) {
    if (!c || !scales || n < 1 || n > 16) return fail(FD23_EINVAL, "bad scales");
    // the kernels invert a scale by negating its exponent, so scales are
    // powers of two in [2^-kScaleExpMax, 2^kScaleExpMax]
    for (int i = 0; i < n; i++) {
        int e = 0;
        const float m = std::frexp(scales[i], &e);
        if (!(scales[i] > 0) || m != 0.5f || e - 1 < -kScaleExpMax || e - 1 > kScaleExpMax)
            return fail(FD23_EINVAL, "scale %d (%g) is not a power of two in [2^-%d, 2^%d]", i, scales[i],
                        kScaleExpMax, kScaleExpMax);
    }
    // current and next scales alike
    HIPCHK(hipMemcpyAsync(c->scl, scales, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->scl + kNextScale, scales, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->scales_init = true;
    return FD23_OK;
}
extern "C" int fd23_get_scales(fd23_ctx *c, float *scales, unsigned *amax_bits, int n) {
    if (!c || n < 1 || n > 16) return fail(FD23_EINVAL, "bad scales");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (scales) HIPCHK(hipMemcpy(scales, c->scl, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (amax_bits) {
        std::vector<unsigned> all((size_t)16 * kAmaxSpread);
        HIPCHK(hipMemcpy(all.data(), c->amax, sizeof(unsigned) * all.size(), hipMemcpyDeviceToHost));
        for (int f = 0; f < n; f++) {
            unsigned m = 0;
            for (int i = 0; i < kAmaxSpread; i++) m = std::max(m, all[(size_t)f * kAmaxSpread + i]);
            amax_bits[f] = m;
        }
    }
    return FD23_OK;
}

// ---------------------------------------------------------------------------
// sources
// ---------------------------------------------------------------------------
static long cell_off(const fd23_ctx *c, int x, int z) { return (long)z * c->px + x; }

extern "C" int fd23_set_source(fd23_ctx *c, int shot, int kind, int xs, int zs, const double *stf, int lsrc,
                               const double *M) {
    if (!c || shot < 0 || shot >= c->d.nshots) return fail(FD23_EINVAL, "bad shot %d", shot);
    if (kind < FD23_SRC_LINE || kind > FD23_SRC_FORCE || (kind == FD23_SRC_FORCE && is_psv(c)))
        return fail(FD23_EINVAL, "bad source kind %d", kind);
    if (lsrc < 0 || (lsrc > 0 && (!stf || !M))) return fail(FD23_EINVAL, "bad source samples");
    HostSrc &h = c->src[shot];
    h.kind = kind;
    h.xs = xs;
    h.zs = zs;
    h.stf.assign(stf, stf + lsrc);
    for (int i = 0; i < (is_psv(c) ? 3 : 2); i++) h.M[i] = M ? M[i] : 0.0;
    c->src_dirty = true;
    return FD23_OK;
}

// Build the tap list of one shot in the reference's order.  Taps outside
// the domain are skipped, as the reference's nz1/nz2 test would (its x
// range is never checked; such sources are outside its valid inputs).
static void build_taps(const fd23_ctx *c, const HostSrc &h, std::vector<Tap> &taps) {
    const int nx = c->d.nx, nz = c->d.nz;
    auto add = [&](int field, int x, int z, double m, double q, int mode) {
        if (x < 0 || x >= nx || z < 0 || z >= nz) return;
        taps.push_back(Tap{field, cell_off(c, x, z), m, q, mode});
    };
    const int xs = h.xs, zs = h.zs;
    if (is_psv(c)) {
        // internal field order U W Txx Tzz Txz
        const double Mxx = h.M[0], Mxz = h.M[1], Mzz = h.M[2];
        if (h.kind == FD23_SRC_LINE) {
            add(2, xs, zs, Mxx, -1.0, 0);
            add(3, xs, zs, Mzz, -1.0, 0);
            for (int x = xs - 1; x <= xs; x++)
                for (int z = zs; z <= zs + 1; z++) add(4, x, z, Mxz, -0.25, 0);
        } else {
            add(2, xs, zs, Mxx, -1.0, 0);
            add(2, xs - 1, zs, Mxx, 1.0, 0);
            add(3, xs, zs, Mzz, -1.0, 0);
            add(3, xs - 1, zs, Mzz, 1.0, 0);
            for (int x = xs - 1; x <= xs; x++)
                for (int z = zs; z <= zs + 1; z++) {
                    add(4, x, z, Mxz, -0.25, 0);
                    add(4, x - 1, z, Mxz, 0.25, 0);
                }
        }
    } else {
        // internal field order V Tx Tz
        const double Mxy = h.M[0], Mzy = h.M[1];
        if (h.kind == FD23_SRC_FORCE) {
            add(0, xs, zs, Mxy, -0.5, 1);
            return;
        }
        add(1, xs, zs, Mxy, -0.5, 1);
        add(1, xs - 1, zs, Mxy, -0.5, 1);
        add(2, xs, zs, Mzy, -0.5, 1);
        add(2, xs, zs + 1, Mzy, -0.5, 1);
        if (h.kind == FD23_SRC_P) {
            add(1, xs - 1, zs, Mxy, 0.5, 1);
            add(1, xs - 2, zs, Mxy, 0.5, 1);
            add(2, xs - 1, zs, Mzy, 0.5, 1);
            add(2, xs - 1, zs + 1, Mzy, 0.5, 1);
        }
    }
}

static int upload_sources(fd23_ctx *c) {
    if (!c->src_dirty) return FD23_OK;
    std::vector<Tap> taps;
    std::vector<SrcInfo> info(c->d.nshots);
    std::vector<double> stf;
    for (int s = 0; s < c->d.nshots; s++) {
        const HostSrc &h = c->src[s];
        SrcInfo si{};
        si.tap0 = (int)taps.size();
        si.stf0 = (long)stf.size();
        si.t0 = h.t0;
        if (h.kind >= 0) {
            build_taps(c, h, taps);
            si.lsrc = (int)h.stf.size();
            si.clamp = h.kind == FD23_SRC_P ? 1 : 0;
            stf.insert(stf.end(), h.stf.begin(), h.stf.end());
        }
        si.ntap = (int)taps.size() - si.tap0;
        info[s] = si;
    }
    c->h_taps = taps;
    c->h_srcinfo = info;
    if (taps.size() > c->taps_cap || !c->d_taps) {
        if (c->d_taps) hipFree(c->d_taps);
        c->taps_cap = std::max<size_t>(taps.size(), 16);
        HIPCHK(hipMalloc((void **)&c->d_taps, sizeof(Tap) * c->taps_cap));
    }
    if (stf.size() > c->stf_cap || !c->d_stf) {
        if (c->d_stf) hipFree(c->d_stf);
        c->stf_cap = std::max<size_t>(stf.size(), 16);
        HIPCHK(hipMalloc((void **)&c->d_stf, sizeof(double) * c->stf_cap));
    }
    if (!taps.empty())
        HIPCHK(hipMemcpyAsync(c->d_taps, taps.data(), sizeof(Tap) * taps.size(), hipMemcpyHostToDevice, c->stream));
    if (!stf.empty())
        HIPCHK(hipMemcpyAsync(c->d_stf, stf.data(), sizeof(double) * stf.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_src, info.data(), sizeof(SrcInfo) * info.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->src_dirty = false;
    c->events_dirty = true;
    c->fevents_dirty = true;
    c->bevents_dirty = true;
    return FD23_OK;
}

// ---------------------------------------------------------------------------
// receivers
// ---------------------------------------------------------------------------
extern "C" int fd23_set_receivers(fd23_ctx *c, int shot, int comp, const int *rx, const int *rz, int n) {
    if (!c || shot < 0 || shot >= c->d.nshots) return fail(FD23_EINVAL, "bad shot %d", shot);
    if (comp < 0 || comp >= c->ncomp) return fail(FD23_EINVAL, "bad component %d", comp);
    if (n < 0 || n > c->d.max_rec) return fail(FD23_EINVAL, "%d receivers exceed max_rec %d", n, c->d.max_rec);
    if (n > 0 && (!rx || !rz)) return fail(FD23_EINVAL, "null receiver arrays");
    const long nx = c->d.nx, N = nx * (long)c->d.nz;
    const long base = ((long)shot * c->ncomp + comp) * c->d.max_rec;
    for (int i = 0; i < c->d.max_rec; i++) {
        long off = -1;
        if (i < n) {
            // flat index rx + rz*nx (cu_record.cu:144): a receiver at x=-1
            // lands on the previous row's last cell
            long q = (long)rx[i] + (long)rz[i] * nx;
            if (q >= 0 && q < N) off = (q / nx) * c->px + (q % nx);
        }
        c->h_roff[base + i] = off;
    }
    c->h_nrec[(size_t)shot * c->ncomp + comp] = n;
    HIPCHK(hipMemcpyAsync(c->d_roff + base, c->h_roff.data() + base, sizeof(long) * c->d.max_rec,
                          hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_nrec, c->h_nrec.data(), sizeof(int) * c->h_nrec.size(), hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->events_dirty = true;
    c->fevents_dirty = true;
    c->bevents_dirty = true;
    return FD23_OK;
}

// ---------------------------------------------------------------------------
// shot slots: shots that start while the context runs
// ---------------------------------------------------------------------------
static bool slots_default(const fd23_ctx *c) {
    if (c->idle) return false;
    for (int s = 0; s < c->d.nshots; s++)
        if (c->src[s].t0 != 0) return false;
    return true;
}

extern "C" int fd23_set_shot_slot(fd23_ctx *c, int shot, int active, int t0) {
    if (!c || shot < 0 || shot >= c->d.nshots) return fail(FD23_EINVAL, "bad shot %d", shot);
    if (t0 < 0) return fail(FD23_EINVAL, "a shot's first step must be >= 0");
    if (!active && shot >= 64) return fail(FD23_EINVAL, "only shots 0..63 can be idle");
    if ((c->box || c->punch) && (!active || t0 != 0))
        return fail(FD23_EINVAL, "the box and the first-arrival window run every shot from step 0");
    if (c->src[shot].t0 != t0) {
        c->src[shot].t0 = t0;
        c->src_dirty = true; // rebuilds the lists, and with them the graphs
    }
    if (shot < 64) {
        const unsigned long long bit = 1ull << shot, was = c->idle;
        c->idle = active ? was & ~bit : was | bit;
        if (c->idle != was) fd23_ctx_invalidate_graphs(c); // a kernel argument
    }
    return FD23_OK;
}

extern "C" int fd23_reset_shot(fd23_ctx *c, int shot) {
    if (!c || shot < 0 || shot >= c->d.nshots) return fail(FD23_EINVAL, "bad shot %d", shot);
    const size_t es = c->esize;
    for (int set = 0; set < 2; set++) {
        void *const *f = set ? c->alt_fields : c->fields;
        void *const *q = set ? c->alt_psi : c->psi;
        if (set && !c->alt_alloc[0]) break; // no second buffer set yet
        for (int i = 0; i < c->nfield; i++)
            HIPCHK(hipMemset2DAsync((char *)f[i] + (size_t)shot * c->sstride * es, (size_t)c->px * es, 0,
                                    (size_t)c->d.nx * es, c->d.nz, c->stream));
        const int half = c->npsi / 2;
        for (int p = 0; p < c->npsi; p++) {
            const size_t n = p < half ? c->xpsi : c->zpsi;
            if (n) HIPCHK(hipMemsetAsync((char *)q[p] + (size_t)shot * n * es, 0, n * es, c->stream));
        }
    }
    // the shot's seismogram entries of every frame (receivers off the grid
    // are never written)
    if (c->seis && c->d.max_rec > 0) {
        const size_t frame = (size_t)c->d.nshots * c->ncomp * c->d.max_rec * c->csize;
        const size_t mine = (size_t)c->ncomp * c->d.max_rec * c->csize;
        HIPCHK(hipMemset2DAsync((char *)c->seis + shot * mine, frame, 0, mine, c->d.max_frames, c->stream));
    }
    return FD23_OK;
}

// ---------------------------------------------------------------------------
// per-tile event lists for the fused injection / gather of the graph loop
// ---------------------------------------------------------------------------
static bool tile_fused(int t) { return kTiles[t].fh > 0; }
static bool tile_block(int t) { return kTiles[t].bk > 0; }
static bool scaled_storage(const fd23_ctx *c) {
    return c->d.storage == FD23_STORE_F16 || c->d.storage == FD23_STORE_BF16;
}
// the fused step runs in fd23_run for P-SV contexts without a box, the
// block step in fd23_run for SH contexts with FP32/FP64 storage; both
// without a first-arrival window.  The staged entry points, the box and
// the window use two-pass tiles (tile 0's shape in place of the others).
static bool use_fused(const fd23_ctx *c) { return tile_fused(c->tile) && is_psv(c) && !c->box && !c->punch; }
static bool use_block(const fd23_ctx *c) {
    return tile_block(c->tile) && !is_psv(c) && !scaled_storage(c) && !c->punch;
}
static int pass_tile(const fd23_ctx *c) { return tile_fused(c->tile) || tile_block(c->tile) ? 0 : c->tile; }
static int tile_rows(const fd23_ctx *c) { return BY * kTiles[pass_tile(c)].rz; }

template <class T> static int upload_vec(fd23_ctx *c, T *&dev, size_t &cap, const std::vector<T> &v) {
    if (v.size() > cap || !dev) {
        if (dev) hipFree(dev);
        cap = std::max<size_t>(v.size(), 16);
        HIPCHK(hipMalloc((void **)&dev, sizeof(T) * cap));
    }
    if (!v.empty()) HIPCHK(hipMemcpyAsync(dev, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, c->stream));
    return FD23_OK;
}

static int build_events(fd23_ctx *c) {
    if (!c->events_dirty && c->events_tile == pass_tile(c)) return FD23_OK;
    const int TZ = tile_rows(c);
    const int ntx = (c->d.nx + TX - 1) / TX, ntz = (c->d.nz + TZ - 1) / TZ;
    const int S = c->d.nshots;
    const size_t nkey = (size_t)ntx * ntz * S;
    std::vector<std::vector<TapT>> tb(nkey);
    std::vector<std::vector<RecT>> rb(nkey);
    auto locate = [&](long off, int &key_tile, int &cell) {
        const int z = (int)(off / c->px), x = (int)(off % c->px);
        const int tx_ = x / TX, tz_ = z / TZ;
        key_tile = tz_ * ntx + tx_;
        cell = (z - tz_ * TZ) * TX + (x - tx_ * TX);
    };
    for (int s = 0; s < S; s++) {
        const SrcInfo &si = c->h_srcinfo[s];
        for (int i = 0; i < si.ntap; i++) {
            const Tap &tp = c->h_taps[si.tap0 + i];
            int tile, cell;
            locate(tp.off, tile, cell);
            tb[(size_t)tile * S + s].push_back(TapT{cell, tp.field, tp.m, tp.q, tp.mode});
        }
        for (int comp = 0; comp < c->ncomp; comp++) {
            const int n = c->h_nrec[(size_t)s * c->ncomp + comp];
            const long base = ((long)s * c->ncomp + comp) * c->d.max_rec;
            for (int r = 0; r < n; r++) {
                const long off = c->h_roff[base + r];
                if (off < 0) continue; // outside the grid: the frame keeps 0
                int tile, cell;
                locate(off, tile, cell);
                rb[(size_t)tile * S + s].push_back(RecT{cell, comp, (int)(base + r)});
            }
        }
    }
    std::vector<int> ts(nkey + 1), rs(nkey + 1);
    std::vector<TapT> taps;
    std::vector<RecT> recs;
    for (size_t k = 0; k < nkey; k++) {
        ts[k] = (int)taps.size();
        rs[k] = (int)recs.size();
        taps.insert(taps.end(), tb[k].begin(), tb[k].end());
        recs.insert(recs.end(), rb[k].begin(), rb[k].end());
    }
    ts[nkey] = (int)taps.size();
    rs[nkey] = (int)recs.size();
    int r;
    if (c->punch) {
        // per (tile, shot): x < it < y when a block of the tile may be open
        // (always, when the shot has a tap or receiver there), z < it < w
        // when every block of the tile is open
        const int nxb = c->d.nx / kPunchBX;
        std::vector<int4> win(nkey);
        for (int tz_ = 0; tz_ < ntz; tz_++)
            for (int tx_ = 0; tx_ < ntx; tx_++) {
                int lo = INT_MAX, hi = INT_MIN, olo = INT_MIN, ohi = INT_MAX;
                const int bz1 = (std::min((tz_ + 1) * TZ, c->d.nz) - 1) / kPunchBZ;
                const int bx1 = (std::min((tx_ + 1) * TX, c->d.nx) - 1) / kPunchBX;
                for (int bz = tz_ * TZ / kPunchBZ; bz <= bz1; bz++)
                    for (int bx = tx_ * TX / kPunchBX; bx <= bx1; bx++) {
                        const int f = c->h_pfirst[(size_t)bz * nxb + bx], l = c->h_plast[(size_t)bz * nxb + bx];
                        lo = std::min(lo, f);
                        hi = std::max(hi, l);
                        olo = std::max(olo, f);
                        ohi = std::min(ohi, l);
                    }
                for (int s = 0; s < S; s++) {
                    const size_t k = (size_t)(tz_ * ntx + tx_) * S + s;
                    const bool events = ts[k + 1] > ts[k] || rs[k + 1] > rs[k];
                    win[k] = events ? make_int4(INT_MIN, INT_MAX, olo, ohi) : make_int4(lo, hi, olo, ohi);
                }
            }
        if ((r = upload_vec(c, c->d_pwin, c->pwin_cap, win))) return r;
    }
    if ((r = upload_vec(c, c->d_tap_start, c->tap_start_cap, ts))) return r;
    if ((r = upload_vec(c, c->d_rec_start, c->rec_start_cap, rs))) return r;
    if ((r = upload_vec(c, c->d_tapsT, c->tapsT_cap, taps))) return r;
    if ((r = upload_vec(c, c->d_recs, c->recs_cap, recs))) return r;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->events_dirty = false;
    c->events_tile = pass_tile(c);
    // graphs hold the old list pointers
    fd23_ctx_invalidate_graphs(c);
    return FD23_OK;
}

// Lists of the fused step: source taps per (strip, shot) -- every strip
// whose 64 stress columns (its 56 own columns and 4 on each side) hold
// the tap -- and receivers per (segment, strip, shot) sorted by row.
static int build_fused_events(fd23_ctx *c) {
    if (!c->fevents_dirty && c->fevents_tile == c->tile) return FD23_OK;
    const int H = kTiles[c->tile].fh;
    const int nstrip = (c->d.nx + FO - 1) / FO, nseg = (c->d.nz + H - 1) / H;
    const int S = c->d.nshots;
    std::vector<std::vector<TapF>> tb((size_t)nstrip * S);
    std::vector<std::vector<RecF>> rb((size_t)nseg * nstrip * S);
    for (int s = 0; s < S; s++) {
        const SrcInfo &si = c->h_srcinfo[s];
        for (int i = 0; i < si.ntap; i++) {
            const Tap &tp = c->h_taps[si.tap0 + i];
            const int z = (int)(tp.off / c->px), x = (int)(tp.off % c->px);
            for (int st = 0; st < nstrip; st++) {
                const int lane = x - (st * FO - HALO);
                if (lane >= 0 && lane < FW) tb[(size_t)st * S + s].push_back(TapF{z, lane, tp.field, tp.mode, tp.m, tp.q});
            }
        }
        for (int comp = 0; comp < c->ncomp; comp++) {
            const int n = c->h_nrec[(size_t)s * c->ncomp + comp];
            const long base = ((long)s * c->ncomp + comp) * c->d.max_rec;
            for (int r = 0; r < n; r++) {
                const long off = c->h_roff[base + r];
                if (off < 0) continue; // outside the grid: the frame keeps 0
                const int z = (int)(off / c->px), x = (int)(off % c->px);
                const int st = x / FO, sg = z / H;
                rb[((size_t)sg * nstrip + st) * S + s].push_back(RecF{z, x - st * FO + HALO, comp, (int)(base + r)});
            }
        }
    }
    std::vector<int> ts(tb.size() + 1), rs(rb.size() + 1);
    std::vector<TapF> taps;
    std::vector<RecF> recs;
    for (size_t k = 0; k < tb.size(); k++) {
        ts[k] = (int)taps.size();
        taps.insert(taps.end(), tb[k].begin(), tb[k].end());
    }
    ts[tb.size()] = (int)taps.size();
    for (size_t k = 0; k < rb.size(); k++) {
        std::stable_sort(rb[k].begin(), rb[k].end(), [](const RecF &a, const RecF &b) { return a.row < b.row; });
        rs[k] = (int)recs.size();
        recs.insert(recs.end(), rb[k].begin(), rb[k].end());
    }
    rs[rb.size()] = (int)recs.size();
    if (taps.empty()) taps.push_back(TapF{-1, -1, 0, 0, 0.0, 0.0});
    if (recs.empty()) recs.push_back(RecF{-1, -1, 0, 0});
    int r;
    if ((r = upload_vec(c, c->d_ftap_start, c->ftap_start_cap, ts))) return r;
    if ((r = upload_vec(c, c->d_frec_start, c->frec_start_cap, rs))) return r;
    if ((r = upload_vec(c, c->d_ftaps, c->ftaps_cap, taps))) return r;
    if ((r = upload_vec(c, c->d_frecs, c->frecs_cap, recs))) return r;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->fevents_dirty = false;
    c->fevents_tile = c->tile;
    fd23_ctx_invalidate_graphs(c);
    return FD23_OK;
}

// Lists of the SH block step per (tile, shot): the source taps of every
// cell of the tile's region (the halo recomputes them too), in the order of
// the shot's tap table, and the receivers of the tile's own cells.  Cells
// are numbered lz*64 + lx inside the region.
static int build_block_events(fd23_ctx *c) {
    if (!c->bevents_dirty && c->bevents_tile == c->tile) return FD23_OK;
    const int K = kTiles[c->tile].bk, RH = kTiles[c->tile].bw * kTiles[c->tile].bcz;
    const int H = kBlockReach * K, OX = TX - 2 * H, OZ = RH - 2 * H;
    const int ntx = (c->d.nx + OX - 1) / OX, ntz = (c->d.nz + OZ - 1) / OZ;
    const int S = c->d.nshots;
    const size_t nkey = (size_t)ntx * ntz * S;
    std::vector<std::vector<TapT>> tb(nkey);
    std::vector<std::vector<RecT>> rb(nkey);
    for (int s = 0; s < S; s++) {
        const SrcInfo &si = c->h_srcinfo[s];
        for (int i = 0; i < si.ntap; i++) {
            const Tap &tp = c->h_taps[si.tap0 + i];
            const int z = (int)(tp.off / c->px), x = (int)(tp.off % c->px);
            // tiles whose region [t*O - H, t*O - H + extent) holds the cell
            for (int tz_ = std::max(0, (z - RH + H) / OZ); tz_ < ntz && tz_ * OZ - H <= z; tz_++)
                for (int tx_ = std::max(0, (x - TX + H) / OX); tx_ < ntx && tx_ * OX - H <= x; tx_++) {
                    const int lx = x - (tx_ * OX - H), lz = z - (tz_ * OZ - H);
                    if (lx < 0 || lx >= TX || lz < 0 || lz >= RH) continue;
                    tb[(size_t)(tz_ * ntx + tx_) * S + s].push_back(TapT{lz * TX + lx, tp.field, tp.m, tp.q, tp.mode});
                }
        }
        for (int comp = 0; comp < c->ncomp; comp++) {
            const int n = c->h_nrec[(size_t)s * c->ncomp + comp];
            const long base = ((long)s * c->ncomp + comp) * c->d.max_rec;
            for (int r = 0; r < n; r++) {
                const long off = c->h_roff[base + r];
                if (off < 0) continue; // outside the grid: the frame keeps 0
                const int z = (int)(off / c->px), x = (int)(off % c->px);
                const int tx_ = x / OX, tz_ = z / OZ;
                const int lx = x - (tx_ * OX - H), lz = z - (tz_ * OZ - H);
                rb[(size_t)(tz_ * ntx + tx_) * S + s].push_back(RecT{lz * TX + lx, comp, (int)(base + r)});
            }
        }
    }
    std::vector<int> ts(nkey + 1), rs(nkey + 1);
    std::vector<TapT> taps;
    std::vector<RecT> recs;
    for (size_t k = 0; k < nkey; k++) {
        // the kernel holds a list in the lanes of a wavefront
        if (tb[k].size() > (size_t)kBlockTaps)
            return fail(FD23_EINVAL, "block step: %zu source taps in one tile (at most %d)", tb[k].size(),
                        kBlockTaps);
        ts[k] = (int)taps.size();
        rs[k] = (int)recs.size();
        taps.insert(taps.end(), tb[k].begin(), tb[k].end());
        recs.insert(recs.end(), rb[k].begin(), rb[k].end());
    }
    ts[nkey] = (int)taps.size();
    rs[nkey] = (int)recs.size();
    // one spare entry at the end of each array: the kernel reads entry
    // min(i, n-1) of a list, which for an empty last list is this one
    taps.push_back(TapT{0, 0, 0.0, 0.0, 0});
    recs.push_back(RecT{0, 0, 0});
    int r;
    if ((r = upload_vec(c, c->d_btap_start, c->btap_start_cap, ts))) return r;
    if ((r = upload_vec(c, c->d_brec_start, c->brec_start_cap, rs))) return r;
    if ((r = upload_vec(c, c->d_btaps, c->btaps_cap, taps))) return r;
    if ((r = upload_vec(c, c->d_brecs, c->brecs_cap, recs))) return r;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->bevents_dirty = false;
    c->bevents_tile = c->tile;
    fd23_ctx_invalidate_graphs(c);
    return FD23_OK;
}

// second field/psi set of the fused step, zeroed (pads stay zero)
static int ensure_alt(fd23_ctx *c) {
    if (c->alt_alloc[0]) return FD23_OK;
    const size_t S = c->d.nshots;
    for (int f = 0; f < c->nfield; f++) {
        const size_t n = ((size_t)S * c->sstride + 2 * kGuard) * c->esize;
        if (hipMalloc(&c->alt_alloc[f], n) != hipSuccess) return fail(FD23_ENOMEM, "fused buffer set");
        HIPCHK(hipMemsetAsync(c->alt_alloc[f], 0, n, c->stream));
        c->alt_fields[f] = (char *)c->alt_alloc[f] + (size_t)(kGuard + (long)HALO * c->px) * c->esize;
    }
    for (int p = 0; p < c->npsi; p++) {
        const bool xdir = p < c->npsi / 2; // x-psi arrays first (P-SV 4, SH 2)
        const size_t n = std::max<size_t>(1, S * (xdir ? c->xpsi : c->zpsi)) * c->esize;
        if (hipMalloc(&c->alt_psi[p], n) != hipSuccess) return fail(FD23_ENOMEM, "fused buffer set");
        HIPCHK(hipMemsetAsync(c->alt_psi[p], 0, n, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return FD23_OK;
}

static void swap_sets(fd23_ctx *c) {
    for (int f = 0; f < 5; f++) {
        std::swap(c->field_alloc[f], c->alt_alloc[f]);
        std::swap(c->fields[f], c->alt_fields[f]);
    }
    for (int p = 0; p < 8; p++) std::swap(c->psi[p], c->alt_psi[p]);
    c->cur ^= 1;
    c->tables_dirty = true;
}

// the device pointer tables of the staged inject/record kernels, rewritten
// in stream order after the buffer sets have swapped
static void refresh_tables(fd23_ctx *c) {
    if (!c->tables_dirty) return;
    PtrTable5 t{};
    for (int f = 0; f < 5; f++) t.p[f] = c->fields[f];
    hipLaunchKernelGGL(k_set_ptrs, dim3(1), dim3(64), 0, c->stream, (void **)c->d_fieldtab, t, 5);
    PtrTable5 cp{};
    cp.p[0] = c->fields[0];
    cp.p[1] = is_psv(c) ? c->fields[1] : nullptr;
    hipLaunchKernelGGL(k_set_ptrs, dim3(1), dim3(64), 0, c->stream, (void **)c->d_comptab, cp, 2);
    c->tables_dirty = false;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static Grid make_grid(const fd23_ctx *c) {
    Grid g;
    g.nx = c->d.nx;
    g.nz = c->d.nz;
    g.px = c->px;
    g.sstride = c->sstride;
    g.npml = c->d.npml;
    g.top = c->d.topabsorb ? 1 : 0;
    g.nshots = c->d.nshots;
    g.xpsi_stride = c->xpsi;
    g.zpsi_stride = c->zpsi;
    g.pfirst = c->punch ? c->d_pfirst : nullptr;
    g.plast = c->punch ? c->d_plast : nullptr;
    g.pnxb = c->d.nx / kPunchBX;
    return g;
}

template <class C> static Prof<C> prof_of(const fd23_ctx *c, int p) {
    const C *base = (const C *)c->prof + c->prof_off[p];
    const int n = c->prof_n[p];
    return Prof<C>{base, base + n, base + 2 * n};
}

// Work of one launch over a region of the tile grid.  The interior
// rectangle holds the tiles whose cells are all at full order with no C-PML
// and no free surface (the test of tile_interior, separable in x and z);
// an empty rectangle leaves every tile to the edge launch.
static Work make_work(const fd23_ctx *c, int RZ, int SPB, dim3 &grid, int region = REGION_ALL) {
    Work w;
    const int TZ = BY * RZ;
    w.ntx = (c->d.nx + TX - 1) / TX;
    w.ntz = (c->d.nz + TZ - 1) / TZ;
    w.ngroups = (c->d.nshots + SPB - 1) / SPB;
    w.g0 = 0;
    if (c->chain >= 0) { // part c->chain of c->chains of the shot groups
        const int h = (w.ngroups + c->chains - 1) / c->chains;
        w.g0 = std::min(c->chain * h, w.ngroups);
        w.ngroups = std::min(h, w.ngroups - w.g0);
    }
    const int lo = std::max(c->d.npml, 4), ztop = c->d.topabsorb ? lo : 4;
    w.ix0 = (lo + TX - 1) / TX;
    w.ix1 = c->d.nx - lo - TX >= 0 ? (c->d.nx - lo - TX) / TX + 1 : 0;
    w.iz0 = (ztop + TZ - 1) / TZ;
    w.iz1 = c->d.nz - lo - TZ >= 0 ? (c->d.nz - lo - TZ) / TZ + 1 : 0;
    if (w.ix1 <= w.ix0 || w.iz1 <= w.iz0) w.ix0 = w.ix1 = w.iz0 = w.iz1 = 0;
    const int inner = (w.ix1 - w.ix0) * (w.iz1 - w.iz0);
    w.ntiles = region == REGION_INTERIOR ? inner : region == REGION_EDGE ? w.ntx * w.ntz - inner : w.ntx * w.ntz;
    w.nwork = w.ntiles * w.ngroups;
    grid = dim3(round_up(std::max(w.nwork, 1), 8), 1, 1);
    return w;
}

// fused events for step k of a captured chunk (k < 0: none)
static Events make_events(const fd23_ctx *c, int k, int itrecord) {
    Events e{};
    if (k < 0 || c->timing) return e;
    e.clock = c->clock;
    e.k = k;
    e.itrecord = itrecord;
    e.tap_start = c->d_tap_start;
    e.taps = c->d_tapsT;
    e.src = c->d_src;
    e.stf = c->d_stf;
    e.rec_start = c->d_rec_start;
    e.recs = c->d_recs;
    e.seis = c->seis;
    e.max_frames = c->seis ? c->d.max_frames : 0;
    e.frame_size = (long)c->d.nshots * c->ncomp * c->d.max_rec;
    e.pwin = c->punch ? c->d_pwin : nullptr;
    e.idle = c->idle;
    return e;
}

template <class S, int RZ, int SPB, bool SPLIT>
static void launch_stress_t(fd23_ctx *c, const Events &ev, unsigned *amax) {
    using C = typename Store<S>::C;
    const Grid g = make_grid(c);
    dim3 grid;
    const Work w = make_work(c, RZ, SPB, grid);
    dim3 block(TX, BY);
    if (is_psv(c)) {
        PsvFields<S> f{(S *)c->fields[0], (S *)c->fields[1], (S *)c->fields[2], (S *)c->fields[3],
                       (S *)c->fields[4]};
        PsvStressMat<C> m{(const C *)c->mat[4], (const C *)c->mat[2], (const C *)c->mat[3]};
        PsvStressPsi<S> ps{(S *)c->psi[2], (S *)c->psi[3], (S *)c->psi[6], (S *)c->psi[7]};
        PsvStressProf<C> pr{prof_of<C>(c, 2), prof_of<C>(c, 3), prof_of<C>(c, 6), prof_of<C>(c, 7)};
        bool split_done = false;
        if constexpr (SPLIT) {
            // the interior-only kernel has no first-arrival window check: with a
            // window the whole-grid kernel runs
            if (!c->punch) {
                // interior tiles, then the edge tiles (each kernel compiled for one path)
                dim3 gi, ge;
                const Work wi = make_work(c, RZ, SPB, gi, REGION_INTERIOR);
                const Work we = make_work(c, RZ, SPB, ge, REGION_EDGE);
                if (wi.nwork > 0)
                    hipLaunchKernelGGL((k_psv_stress<S, RZ, SPB, REGION_INTERIOR>), gi, block, 0, c->stream, f, m, ps,
                                       pr, g, wi, ev, c->scl, amax);
                if (we.nwork > 0)
                    hipLaunchKernelGGL((k_psv_stress<S, RZ, SPB, REGION_EDGE>), ge, block, 0, c->stream, f, m, ps,
                                       pr, g, we, ev, c->scl, amax);
                split_done = true;
            }
        }
        if (!split_done) {
            hipLaunchKernelGGL((k_psv_stress<S, RZ, SPB, REGION_ALL>), grid, block, 0, c->stream, f, m, ps, pr, g, w,
                               ev, c->scl, amax);
        }
    } else {
        ShFields<S> f{(S *)c->fields[0], (S *)c->fields[1], (S *)c->fields[2]};
        ShStressMat<C> m{(const C *)c->mat[1], (const C *)c->mat[2]};
        ShPsi2<S> ps{(S *)c->psi[1], (S *)c->psi[3]};
        ShProf2<C> pr{prof_of<C>(c, 1), prof_of<C>(c, 3)};
        hipLaunchKernelGGL((k_sh_stress<S, RZ, SPB>), grid, block, 0, c->stream, f, m, ps, pr, g, w, ev, c->scl,
                           amax);
    }
}

template <class S, int RZ, int SPB, bool SPLIT>
static void launch_velocity_t(fd23_ctx *c, const Events &ev, unsigned *amax) {
    using C = typename Store<S>::C;
    const Grid g = make_grid(c);
    dim3 grid;
    const Work w = make_work(c, RZ, SPB, grid);
    dim3 block(TX, BY);
    if (is_psv(c)) {
        PsvFields<S> f{(S *)c->fields[0], (S *)c->fields[1], (S *)c->fields[2], (S *)c->fields[3],
                       (S *)c->fields[4]};
        PsvVelMat<C> m{(const C *)c->mat[0], (const C *)c->mat[1]};
        PsvVelPsi<S> ps{(S *)c->psi[0], (S *)c->psi[1], (S *)c->psi[5], (S *)c->psi[4]};
        PsvVelProf<C> pr{prof_of<C>(c, 0), prof_of<C>(c, 1), prof_of<C>(c, 5), prof_of<C>(c, 4)};
        bool split_done = false;
        if constexpr (SPLIT) {
            // the interior-only kernel has no first-arrival window check: with a
            // window the whole-grid kernel runs
            if (!c->punch) {
                dim3 gi, ge;
                const Work wi = make_work(c, RZ, SPB, gi, REGION_INTERIOR);
                const Work we = make_work(c, RZ, SPB, ge, REGION_EDGE);
                if (wi.nwork > 0)
                    hipLaunchKernelGGL((k_psv_velocity<S, RZ, SPB, REGION_INTERIOR>), gi, block, 0, c->stream, f, m,
                                       ps, pr, g, wi, ev, c->scl, amax);
                if (we.nwork > 0)
                    hipLaunchKernelGGL((k_psv_velocity<S, RZ, SPB, REGION_EDGE>), ge, block, 0, c->stream, f, m, ps,
                                       pr, g, we, ev, c->scl, amax);
                split_done = true;
            }
        }
        if (!split_done) {
            hipLaunchKernelGGL((k_psv_velocity<S, RZ, SPB, REGION_ALL>), grid, block, 0, c->stream, f, m, ps, pr, g,
                               w, ev, c->scl, amax);
        }
    } else {
        ShFields<S> f{(S *)c->fields[0], (S *)c->fields[1], (S *)c->fields[2]};
        ShPsi2<S> ps{(S *)c->psi[0], (S *)c->psi[2]};
        ShProf2<C> pr{prof_of<C>(c, 0), prof_of<C>(c, 2)};
        hipLaunchKernelGGL((k_sh_velocity<S, RZ, SPB>), grid, block, 0, c->stream, f, (const C *)c->mat[0], ps,
                           pr, g, w, ev, c->scl, amax);
    }
}

template <class S> static void launch_pass(fd23_ctx *c, bool stress, const Events &ev, unsigned *amax) {
    const TileCfg t = kTiles[pass_tile(c)];
    const bool split = t.split && is_psv(c);
#define TILE_CASE(R, P, X)                                                                                 \
    if (t.rz == R && t.spb == P && split == X) {                                                           \
        if (stress) launch_stress_t<S, R, P, X>(c, ev, amax); else launch_velocity_t<S, R, P, X>(c, ev, amax); \
        return;                                                                                            \
    }
    TILE_CASE(4, 1, false) TILE_CASE(2, 1, false) TILE_CASE(8, 1, false) TILE_CASE(4, 2, false)
    TILE_CASE(4, 4, false) TILE_CASE(4, 8, false) TILE_CASE(8, 2, false) TILE_CASE(8, 4, false)
    TILE_CASE(2, 4, false) TILE_CASE(2, 8, false) TILE_CASE(2, 1, true) TILE_CASE(4, 1, true)
    TILE_CASE(2, 2, true) TILE_CASE(4, 2, true) TILE_CASE(1, 1, false)
#undef TILE_CASE
}

// track: gather the arrays' maxima for the scale check that follows this
// step (scaled storage only)
static void dispatch_pass(fd23_ctx *c, bool stress, const Events &ev, bool track = false) {
    unsigned *amax = track ? c->amax : nullptr;
    switch (c->d.storage) {
    case FD23_STORE_F16: launch_pass<__half>(c, stress, ev, amax); break;
    case FD23_STORE_BF16: launch_pass<__hip_bfloat16>(c, stress, ev, amax); break;
    case FD23_STORE_F64: launch_pass<double>(c, stress, ev, nullptr); break;
    default: launch_pass<float>(c, stress, ev, nullptr); break;
    }
}

template <class S, int H, int R>
static void launch_fused_t(fd23_ctx *c, int k, int itrecord, bool record, unsigned *amax) {
    using C = typename Store<S>::C;
    const Grid g = make_grid(c);
    FusedWork w;
    w.nstrip = (c->d.nx + FO - 1) / FO;
    w.nseg = (c->d.nz + H - 1) / H;
    w.H = H;
    w.nwork = w.nstrip * w.nseg * c->d.nshots;
    const dim3 grid(round_up(w.nwork, 8));
    FusedEvents ev{};
    ev.clock = c->clock;
    ev.k = k;
    ev.itrecord = itrecord;
    ev.tap_start = c->d_ftap_start;
    ev.taps = c->d_ftaps;
    ev.src = c->d_src;
    ev.stf = c->d_stf;
    ev.rec_start = c->d_frec_start;
    ev.recs = c->d_frecs;
    ev.seis = c->seis;
    ev.max_frames = (c->seis && record) ? c->d.max_frames : 0;
    ev.frame_size = (long)c->d.nshots * c->ncomp * c->d.max_rec;
    ev.idle = c->idle;
    auto fields = [](void *const *f) {
        return PsvFields<S>{(S *)f[0], (S *)f[1], (S *)f[2], (S *)f[3], (S *)f[4]};
    };
    PsvMat5<C> m{(const C *)c->mat[0], (const C *)c->mat[1], (const C *)c->mat[2], (const C *)c->mat[3],
                 (const C *)c->mat[4]};
    PsvStressPsi<S> spa{(S *)c->psi[2], (S *)c->psi[3], (S *)c->psi[6], (S *)c->psi[7]};
    PsvStressPsi<S> spb{(S *)c->alt_psi[2], (S *)c->alt_psi[3], (S *)c->alt_psi[6], (S *)c->alt_psi[7]};
    PsvVelPsi<S> vpa{(S *)c->psi[0], (S *)c->psi[1], (S *)c->psi[5], (S *)c->psi[4]};
    PsvVelPsi<S> vpb{(S *)c->alt_psi[0], (S *)c->alt_psi[1], (S *)c->alt_psi[5], (S *)c->alt_psi[4]};
    PsvStressProf<C> spr{prof_of<C>(c, 2), prof_of<C>(c, 3), prof_of<C>(c, 6), prof_of<C>(c, 7)};
    PsvVelProf<C> vpr{prof_of<C>(c, 0), prof_of<C>(c, 1), prof_of<C>(c, 5), prof_of<C>(c, 4)};
    hipLaunchKernelGGL((k_psv_fused<S, H, R>), grid, dim3(FW), 0, c->stream, fields(c->fields), fields(c->alt_fields),
                       m, spa, spb, vpa, vpb, spr, vpr, g, w, ev, c->scl, amax);
}

// tile-parallel fused step (block rows 0 in the tile table): segments of
// 4*CZ rows, FP32/FP64 storage only
template <class S, int CZ> static void launch_fused_tile_t(fd23_ctx *c, int k, int itrecord, bool record) {
    using C = typename Store<S>::C;
    const Grid g = make_grid(c);
    // strips x segments; the interior rectangle holds the tiles whose stress
    // cells (the strip's 64 columns, the segment's rows and 4 on each side)
    // are all at full order with no C-PML and no free surface
    const int OZ = 4 * CZ;
    Work w;
    w.ntx = (c->d.nx + FO - 1) / FO;
    w.ntz = (c->d.nz + OZ - 1) / OZ;
    w.ngroups = c->d.nshots;
    w.g0 = 0;
    const int lo = std::max(c->d.npml, 4), ztop = c->d.topabsorb ? lo : 4;
    w.ix0 = (lo + HALO + FO - 1) / FO; // x0 - 4 >= lo
    w.ix1 = c->d.nx - lo - HALO - FO >= 0 ? (c->d.nx - lo - HALO - FO) / FO + 1 : 0; // x0 + 60 <= nx - lo
    w.iz0 = (ztop + HALO + OZ - 1) / OZ;          // z0 - 4 >= ztop
    w.iz1 = c->d.nz - lo - HALO - OZ >= 0 ? (c->d.nz - lo - HALO - OZ) / OZ + 1 : 0; // z0 + OZ + 4 <= nz - lo
    if (w.ix1 <= w.ix0 || w.iz1 <= w.iz0) w.ix0 = w.ix1 = w.iz0 = w.iz1 = 0;
    const int inner = (w.ix1 - w.ix0) * (w.iz1 - w.iz0);
    FusedEvents ev{};
    ev.clock = c->clock;
    ev.k = k;
    ev.itrecord = itrecord;
    ev.tap_start = c->d_ftap_start;
    ev.taps = c->d_ftaps;
    ev.src = c->d_src;
    ev.stf = c->d_stf;
    ev.rec_start = c->d_frec_start;
    ev.recs = c->d_frecs;
    ev.seis = c->seis;
    ev.max_frames = (c->seis && record) ? c->d.max_frames : 0;
    ev.frame_size = (long)c->d.nshots * c->ncomp * c->d.max_rec;
    ev.idle = c->idle;
    auto fields = [](void *const *f) {
        return PsvFields<S>{(S *)f[0], (S *)f[1], (S *)f[2], (S *)f[3], (S *)f[4]};
    };
    PsvMat5<C> m{(const C *)c->mat[0], (const C *)c->mat[1], (const C *)c->mat[2], (const C *)c->mat[3],
                 (const C *)c->mat[4]};
    PsvStressPsi<S> spa{(S *)c->psi[2], (S *)c->psi[3], (S *)c->psi[6], (S *)c->psi[7]};
    PsvStressPsi<S> spb{(S *)c->alt_psi[2], (S *)c->alt_psi[3], (S *)c->alt_psi[6], (S *)c->alt_psi[7]};
    PsvVelPsi<S> vpa{(S *)c->psi[0], (S *)c->psi[1], (S *)c->psi[5], (S *)c->psi[4]};
    PsvVelPsi<S> vpb{(S *)c->alt_psi[0], (S *)c->alt_psi[1], (S *)c->alt_psi[5], (S *)c->alt_psi[4]};
    PsvStressProf<C> spr{prof_of<C>(c, 2), prof_of<C>(c, 3), prof_of<C>(c, 6), prof_of<C>(c, 7)};
    PsvVelProf<C> vpr{prof_of<C>(c, 0), prof_of<C>(c, 1), prof_of<C>(c, 5), prof_of<C>(c, 4)};
    Work wi = w, we = w;
    wi.ntiles = inner;
    wi.nwork = inner * w.ngroups;
    we.ntiles = w.ntx * w.ntz - inner;
    we.nwork = we.ntiles * w.ngroups;
    if (wi.nwork > 0)
        hipLaunchKernelGGL((k_psv_fused_tile<S, CZ, REGION_INTERIOR>), dim3(round_up(wi.nwork, 8)), dim3(FW, 4), 0,
                           c->stream, fields(c->fields), fields(c->alt_fields), m, spa, spb, vpa, vpb, spr, vpr, g,
                           wi, ev);
    if (we.nwork > 0)
        hipLaunchKernelGGL((k_psv_fused_tile<S, CZ, REGION_EDGE>), dim3(round_up(we.nwork, 8)), dim3(FW, 4), 0,
                           c->stream, fields(c->fields), fields(c->alt_fields), m, spa, spb, vpa, vpb, spr, vpr, g,
                           we, ev);
}

template <class S> static void launch_fused_h(fd23_ctx *c, int k, int itrecord, bool record, unsigned *amax) {
    const TileCfg t = kTiles[c->tile];
    if (t.fr == 0) {
        if constexpr (!Store<S>::scaled) {
            if (t.fh == 16) launch_fused_tile_t<S, 4>(c, k, itrecord, record);
            else launch_fused_tile_t<S, 8>(c, k, itrecord, record);
        }
        return;
    }
    if (t.fr == 2) launch_fused_t<S, 64, 2>(c, k, itrecord, record, amax);
    else if (t.fh == 32) launch_fused_t<S, 32, 4>(c, k, itrecord, record, amax);
    else if (t.fh == 64) launch_fused_t<S, 64, 4>(c, k, itrecord, record, amax);
    else launch_fused_t<S, 128, 4>(c, k, itrecord, record, amax);
}

// one fused step: reads the current set, writes the other, then swaps
static void dispatch_fused(fd23_ctx *c, int k, int itrecord, bool record = true, bool track = false) {
    unsigned *amax = track ? c->amax : nullptr;
    switch (c->d.storage) {
    case FD23_STORE_F16: launch_fused_h<__half>(c, k, itrecord, record, amax); break;
    case FD23_STORE_BF16: launch_fused_h<__hip_bfloat16>(c, k, itrecord, record, amax); break;
    case FD23_STORE_F64: launch_fused_h<double>(c, k, itrecord, record, nullptr); break;
    default: launch_fused_h<float>(c, k, itrecord, record, nullptr); break;
    }
    swap_sets(c);
}

// SH block step: nsub steps (at most the tile's bk) over the current set,
// written to the other; events: the sources and receivers of the chunk's
// steps k .. k+nsub-1 (off for timing)
template <class S, int K, int CZ, int NW>
static void launch_block_t(fd23_ctx *c, int k, int nsub, int itrecord, bool events) {
    using C = typename Store<S>::C;
    using B = BlockShape<K, CZ, NW>;
    const Grid g = make_grid(c);
    BlockWork w;
    w.ntx = (c->d.nx + B::OX - 1) / B::OX;
    w.ntz = (c->d.nz + B::OZ - 1) / B::OZ;
    w.nwork = w.ntx * w.ntz * c->d.nshots;
    w.nsub = nsub;
    Events ev{};
    if (events) {
        ev.clock = c->clock;
        ev.k = k;
        ev.itrecord = itrecord;
        ev.tap_start = c->d_btap_start;
        ev.taps = c->d_btaps;
        ev.src = c->d_src;
        ev.stf = c->d_stf;
        ev.rec_start = c->d_brec_start;
        ev.recs = c->d_brecs;
        ev.seis = c->seis;
        ev.max_frames = c->seis ? c->d.max_frames : 0;
        ev.frame_size = (long)c->d.nshots * c->ncomp * c->d.max_rec;
        ev.idle = c->idle;
    }
    auto fields = [](void *const *f) { return ShFields<S>{(S *)f[0], (S *)f[1], (S *)f[2]}; };
    auto psi = [](void *const *p) { return ShPsi4<S>{(S *)p[0], (S *)p[1], (S *)p[2], (S *)p[3]}; };
    ShMat3<C> m{(const C *)c->mat[0], (const C *)c->mat[1], (const C *)c->mat[2]};
    ShProf4<C> pr{prof_of<C>(c, 0), prof_of<C>(c, 1), prof_of<C>(c, 2), prof_of<C>(c, 3)};
    hipLaunchKernelGGL((k_sh_block<S, K, CZ, NW>), dim3(round_up(w.nwork, 8)), dim3(TX, NW), 0, c->stream,
                       fields(c->fields), fields(c->alt_fields), m, psi(c->psi), psi(c->alt_psi), pr, g, w, ev);
}

template <class S> static void launch_block_h(fd23_ctx *c, int k, int nsub, int itrecord, bool events) {
    const TileCfg t = kTiles[c->tile];
    if (t.bk == 1 && t.bw == 8) launch_block_t<S, 1, 3, 8>(c, k, nsub, itrecord, events);
    else if (t.bk == 1) launch_block_t<S, 1, 3, 16>(c, k, nsub, itrecord, events);
    else if (t.bw == 8) launch_block_t<S, 2, 6, 8>(c, k, nsub, itrecord, events);
    else if (t.bw == 10) launch_block_t<S, 2, 4, 10>(c, k, nsub, itrecord, events);
    else if (t.bw == 12) launch_block_t<S, 2, 3, 12>(c, k, nsub, itrecord, events);
    else if (t.bcz == 3) launch_block_t<S, 2, 3, 16>(c, k, nsub, itrecord, events);
    else launch_block_t<S, 2, 4, 16>(c, k, nsub, itrecord, events);
}

static void dispatch_block(fd23_ctx *c, int k, int nsub, int itrecord, bool events = true) {
    if (c->d.storage == FD23_STORE_F64) launch_block_h<double>(c, k, nsub, itrecord, events);
    else launch_block_h<float>(c, k, nsub, itrecord, events);
    swap_sets(c);
}

template <class S> static void launch_inject_t(fd23_ctx *c, int it) {
    const int nthr = std::min(256, round_up(c->d.nshots, 64));
    hipLaunchKernelGGL((k_inject<S>), dim3(1), dim3(nthr), 0, c->stream, (S *const *)c->d_fieldtab, c->sstride,
                       c->d_taps, c->d_src, c->d_stf, c->d.nshots, it, c->scl);
}

static void dispatch_inject(fd23_ctx *c, int it) {
    refresh_tables(c);
    switch (c->d.storage) {
    case FD23_STORE_F16: launch_inject_t<__half>(c, it); break;
    case FD23_STORE_BF16: launch_inject_t<__hip_bfloat16>(c, it); break;
    case FD23_STORE_F64: launch_inject_t<double>(c, it); break;
    default: launch_inject_t<float>(c, it); break;
    }
}

template <class S> static void launch_record_t(fd23_ctx *c, int it, int itrecord) {
    using C = typename Store<S>::C;
    const long total = (long)c->d.nshots * c->ncomp * c->d.max_rec;
    if (total == 0 || !c->seis) return;
    const int nblk = (int)std::min<long>((total + 255) / 256, 1024);
    hipLaunchKernelGGL((k_record<S>), dim3(nblk), dim3(256), 0, c->stream, (S *const *)c->d_comptab, c->ncomp,
                       c->sstride, c->d_roff, c->d_nrec, c->d.max_rec, c->d.nshots, (C *)c->seis,
                       c->d.max_frames, it, itrecord, c->scl, c->d_compfield, c->d_src);
}

static void dispatch_record(fd23_ctx *c, int it, int itrecord) {
    refresh_tables(c);
    switch (c->d.storage) {
    case FD23_STORE_F16: launch_record_t<__half>(c, it, itrecord); break;
    case FD23_STORE_BF16: launch_record_t<__hip_bfloat16>(c, it, itrecord); break;
    case FD23_STORE_F64: launch_record_t<double>(c, it, itrecord); break;
    default: launch_record_t<float>(c, it, itrecord); break;
    }
}

static int init_scales(fd23_ctx *c);
static int ready(fd23_ctx *c) {
    if (!c) return fail(FD23_EINVAL, "null context");
    if (!c->have_mat) return fail(FD23_EINVAL, "material not set");
    if (!c->have_cpml) return fail(FD23_EINVAL, "C-PML profiles not set");
    int r = upload_sources(c);
    return r ? r : init_scales(c);
}

static int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FD23_EHIP, "kernel launch: %s", hipGetErrorString(e));
    return FD23_OK;
}

static const Events kNoEvents{};

// ---------------------------------------------------------------------------
// FP16/BF16 scales
// ---------------------------------------------------------------------------
static const int kScaleEvery = 8;            // steps between scale checks
// stored max band (log2): grow the scale when the stored maximum passes
// 2^11 (FP16 overflows at 2^16, so a field may grow 32x between checks),
// putting the maximum at 2^8.  Scales never shrink: a wavefront's leading
// edge (values many decades below its peak) would otherwise pull a scale
// down just before the peak arrives and overflow; the first scales come
// from the sources, and FP16 keeps full relative precision 22 binades below
// the 2^8 target.
static const int kScaleLo = -1000, kScaleHi = 11, kScaleTarget = 8;

static float pow2_scale(double est) {
    if (!(est > 0) || !std::isfinite(est)) return 1.0f;
    int e;
    std::frexp(est, &e);
    return std::ldexp(1.0f, std::max(-kScaleExpMax, std::min(kScaleExpMax, e - kScaleTarget)));
}

// First scales from the sources: stresses are of the order of the running
// sum of the injected samples times |M|, velocities bmax times that.  The
// control kernel corrects them within kScaleEvery steps.
static int init_scales(fd23_ctx *c) {
    if (!scaled_storage(c) || c->scales_init) return FD23_OK;
    double t_est = 0;
    for (const HostSrc &h : c->src) {
        if (h.kind < 0) continue;
        double run = 0, peak = 0;
        for (double v : h.stf) {
            run = h.kind == FD23_SRC_P ? v : run + v;
            peak = std::max(peak, std::fabs(run));
        }
        const double mm = std::max(std::fabs(h.M[0]), std::max(std::fabs(h.M[1]), std::fabs(h.M[2])));
        t_est = std::max(t_est, peak * mm);
    }
    const double v_est = t_est * std::max(c->bmax, 1e-30);
    float sc[16];
    for (int i = 0; i < 16; i++) sc[i] = 1.0f;
    if (is_psv(c)) {
        sc[SC_U] = sc[SC_W] = pow2_scale(v_est);
        sc[SC_TXX] = sc[SC_TZZ] = sc[SC_TXZ] = pow2_scale(t_est);
        sc[SC_PTXX_X] = sc[SC_PTXZ_X] = sc[SC_PTZZ_Z] = sc[SC_PTXZ_Z] = pow2_scale(t_est);
        sc[SC_PU_X] = sc[SC_PW_X] = sc[SC_PU_Z] = sc[SC_PW_Z] = pow2_scale(v_est);
    } else {
        // a point force drives V directly
        sc[SC_V] = pow2_scale(std::max(v_est, t_est));
        sc[SC_TX] = sc[SC_TZ] = pow2_scale(t_est);
        sc[SC_PTX_X] = sc[SC_PTZ_Z] = pow2_scale(t_est);
        sc[SC_PV_X] = sc[SC_PV_Z] = pow2_scale(std::max(v_est, t_est));
    }
    HIPCHK(hipMemcpyAsync(c->scl, sc, sizeof sc, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->scl + kNextScale, sc, sizeof sc, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->scales_init = true;
    return FD23_OK;
}

// FP16/BF16 storage: every kScaleEvery steps the control kernel picks the
// next scale of each array from its maximum, which the passes of the step
// before the check gather (the other steps skip it); each pass reads arrays in
// the current scales and writes what it updates in the next ones, and a
// one-thread commit after the pass makes those the current scales.  The
// arrays are never rescaled in a pass of their own.
static void enqueue_scale_control(fd23_ctx *c) {
    hipLaunchKernelGGL(k_scale_control, dim3(1), dim3(kAmaxSpread), 0, c->stream, c->scl, c->amax, c->nscale, kScaleLo,
                       kScaleHi, kScaleTarget);
}

// scale slots a pass rewrites
static unsigned pass_mask(const fd23_ctx *c, bool stress) {
    if (is_psv(c))
        return stress ? (1u << SC_TXX) | (1u << SC_TZZ) | (1u << SC_TXZ) | (1u << SC_PU_X) | (1u << SC_PW_X) |
                            (1u << SC_PU_Z) | (1u << SC_PW_Z)
                      : (1u << SC_U) | (1u << SC_W) | (1u << SC_PTXX_X) | (1u << SC_PTXZ_X) | (1u << SC_PTZZ_Z) |
                            (1u << SC_PTXZ_Z);
    return stress ? (1u << SC_TX) | (1u << SC_TZ) | (1u << SC_PV_X) | (1u << SC_PV_Z)
                  : (1u << SC_V) | (1u << SC_PTX_X) | (1u << SC_PTZ_Z);
}

static void enqueue_commit(fd23_ctx *c, bool stress) {
    hipLaunchKernelGGL(k_commit_scales, dim3(1), dim3(64), 0, c->stream, c->scl, pass_mask(c, stress));
}

// a step bound truncated toward zero as the reference's int() does, held
// inside the int range (where that conversion would be undefined)
// First-arrival window, SH/slide/cu_slide.cu:7-61 (float) and
// PSV/psv_clean_double/slide/cu_slide.cu:7-52 (double): the per-block step
// tables come from the host functions fd23_punch_blocks[_f64].
static int set_punch_tables(fd23_ctx *c, std::vector<int> &first, std::vector<int> &last) {
    int r;
    if ((r = upload_vec(c, c->d_pfirst, c->pfirst_cap, first)) || (r = upload_vec(c, c->d_plast, c->plast_cap, last)))
        return r;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->h_pfirst = std::move(first);
    c->h_plast = std::move(last);
    c->punch = true;
    c->events_dirty = true; // the per-tile window table is built with them
    fd23_ctx_invalidate_graphs(c);
    return FD23_OK;
}

// shared checks; returns 1 when times is null (the window was removed)
static int punch_prologue(fd23_ctx *c, const void *times) {
    if (!c) return fail(FD23_EINVAL, "null context");
    if (!times) {
        c->punch = false;
        fd23_ctx_invalidate_graphs(c);
        return 1;
    }
    // FP16/BF16 storage rewrites every cell each step in the next step's
    // scale; a cell left alone would be read back in the wrong one
    if (scaled_storage(c)) return fail(FD23_EINVAL, "a first-arrival window needs FP32 or FP64 storage");
    return FD23_OK;
}

extern "C" int fd23_set_punch(fd23_ctx *c, const float *times, float before, float after, float dt) {
    int r = punch_prologue(c, times);
    if (r) return r > 0 ? FD23_OK : r;
    const size_t nb = (size_t)(c->d.nx / kPunchBX) * (c->d.nz / kPunchBZ);
    std::vector<int> first(nb), last(nb);
    if ((r = fd23_punch_blocks(times, c->d.nx, c->d.nz, before, after, dt, first.data(), last.data()))) return r;
    return set_punch_tables(c, first, last);
}

extern "C" int fd23_set_punch_f64(fd23_ctx *c, const double *times, double before, double after, double dt) {
    int r = punch_prologue(c, times);
    if (r) return r > 0 ? FD23_OK : r;
    const size_t nb = (size_t)(c->d.nx / kPunchBX) * (c->d.nz / kPunchBZ);
    std::vector<int> first(nb), last(nb);
    if ((r = fd23_punch_blocks_f64(times, c->d.nx, c->d.nz, before, after, dt, first.data(), last.data())))
        return r;
    return set_punch_tables(c, first, last);
}

// the first-arrival window needs the step index, which only the graph
// loop keeps on the device
static int no_punch(const fd23_ctx *c) {
    return c->punch ? fail(FD23_EINVAL, "a first-arrival window is set: step with fd23_run") : FD23_OK;
}

extern "C" int fd23_step_stress(fd23_ctx *c) {
    int r = ready(c);
    if (r || (r = no_punch(c))) return r;
    if (c->box && (r = fd23_box_launch_pre_stress(c, -1))) return r;
    // a scale check follows the velocity pass of every kScaleEvery-th step
    dispatch_pass(c, true, kNoEvents, (c->staged_steps + 1) % kScaleEvery == 0);
    if (scaled_storage(c)) enqueue_commit(c, true);
    if (c->box && (r = fd23_box_launch_post_stress(c, -1))) return r;
    return check_launch();
}
extern "C" int fd23_inject(fd23_ctx *c, int it) {
    int r = ready(c);
    if (r) return r;
    dispatch_inject(c, it);
    return check_launch();
}
extern "C" int fd23_step_velocity(fd23_ctx *c) {
    int r = ready(c);
    if (r || (r = no_punch(c))) return r;
    dispatch_pass(c, false, kNoEvents, (c->staged_steps + 1) % kScaleEvery == 0);
    if (scaled_storage(c)) {
        enqueue_commit(c, false);
        if (++c->staged_steps % kScaleEvery == 0) enqueue_scale_control(c);
    }
    if (c->box && (r = fd23_box_launch_post_velocity(c, -1))) return r;
    return check_launch();
}
extern "C" int fd23_record(fd23_ctx *c, int it, int itrecord) {
    if (!c) return fail(FD23_EINVAL, "null context");
    if (itrecord < 1) return fail(FD23_EINVAL, "itrecord must be >= 1");
    int r = upload_sources(c); // the shots' first steps
    if (r) return r;
    dispatch_record(c, it, itrecord);
    return check_launch();
}

// step k of a chunk: the step index is clock[0] + k on the device
static int enqueue_step(fd23_ctx *c, int k, int itrecord) {
    int r;
    const Events ev = make_events(c, k, itrecord);
    // a scale check ends every kScaleEvery-th step of a chunk, so the passes
    // of steps k = 0, kScaleEvery, ... commit the scales they wrote
    const bool commit = scaled_storage(c) && k % kScaleEvery == 0;
    const bool track = (k + 1) % kScaleEvery == 0;
    if (use_fused(c)) {
        dispatch_fused(c, k, itrecord, true, track);
        if (commit) {
            enqueue_commit(c, true);
            enqueue_commit(c, false);
        }
        if (scaled_storage(c) && (k + 1) % kScaleEvery == 0) enqueue_scale_control(c);
        return FD23_OK;
    }
    if (c->box && (r = fd23_box_launch_pre_stress(c, k))) return r;
    dispatch_pass(c, true, ev, track);
    if (commit) enqueue_commit(c, true);
    if (c->box && (r = fd23_box_launch_post_stress(c, k))) return r;
    dispatch_pass(c, false, ev, track);
    if (commit) enqueue_commit(c, false);
    if (c->box && (r = fd23_box_launch_post_velocity(c, k))) return r;
    if (scaled_storage(c) && (k + 1) % kScaleEvery == 0) enqueue_scale_control(c);
    return FD23_OK;
}

// steps per captured chunk: each replay ends with the clock's tick, a
// dependent one-thread launch; 128 steps per chunk take 2.4% off a small
// grid's step time against 32 (the 400x200 SH loop), 512 no more
static const int kGraphSteps = 128;

// the launches of steps k .. of a chunk, `left` steps remaining: a block
// step takes up to the tile's bk steps, every other path one
static int enqueue_launch(fd23_ctx *c, int k, int left, int itrecord, int *taken) {
    if (use_block(c)) {
        const int n = std::min(kTiles[c->tile].bk, left);
        dispatch_block(c, k, n, itrecord);
        *taken = n;
        return FD23_OK;
    }
    *taken = 1;
    return enqueue_step(c, k, itrecord);
}

// The graph of a chunk of `steps` steps starting on the current buffer set.
// Capturing swaps the host's view of the sets as the launches do; the view
// is put back afterwards, so that the caller swaps once per replay of a
// chunk that flips.
// Two chains: the steps of a captured chunk as two independent chains of
// passes, each over half the shot groups, on the context stream and a
// second one, forked at the chunk's start and joined before its tick; one
// chain's kernels fill the other's gaps between dependent passes (the
// 64-shot 1024x512 batch: +3.7%; the 8-shot 4096x2048 loop loses 3.5%, so
// the autotuner decides).  The two-pass path with more than one shot group
// and no state shared by the halves (scales, box, window).
static bool chains_possible(const fd23_ctx *c, int n = 2) {
    if (use_fused(c) || use_block(c) || scaled_storage(c) || c->box || c->punch) return false;
    const int spb = kTiles[pass_tile(c)].spb;
    return (c->d.nshots + spb - 1) / spb >= n; // a shot group per chain at least
}
static bool use_chains(const fd23_ctx *c) { return c->chains > 1 && chains_possible(c, c->chains); }

static int get_graph(fd23_ctx *c, int itrecord, int steps, fd23_ctx::Chunk *out) {
    const auto key = std::make_tuple(itrecord, steps, c->cur);
    auto it = c->graphs.find(key);
    if (it != c->graphs.end()) {
        *out = it->second;
        return FD23_OK;
    }
    const int cur0 = c->cur;
    hipGraph_t graph;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int r = FD23_OK, taken = 1;
    if (use_chains(c)) {
        // independent chains of passes over parts of the shot groups,
        // forked at the chunk's start and joined before its tick
        const int n = c->chains;
        hipEventRecord(c->ev_fork, c->stream);
        for (int i = 0; i < n - 1; i++) hipStreamWaitEvent(c->chain_stream[i], c->ev_fork, 0);
        const hipStream_t main = c->stream;
        for (int s = 0; s < steps && r == FD23_OK; s++)
            for (int i = 0; i < n && r == FD23_OK; i++) {
                c->chain = i;
                c->stream = i ? c->chain_stream[i - 1] : main;
                r = enqueue_step(c, s, itrecord);
                c->stream = main;
            }
        c->chain = -1;
        for (int i = 0; i < n - 1; i++) {
            hipEventRecord(c->chain_join[i], c->chain_stream[i]);
            hipStreamWaitEvent(c->stream, c->chain_join[i], 0);
        }
    } else {
        for (int s = 0; s < steps && r == FD23_OK; s += taken) r = enqueue_launch(c, s, steps - s, itrecord, &taken);
    }
    if (r == FD23_OK) hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, c->stream, c->clock, steps);
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    const bool flips = c->cur != cur0;
    if (flips) swap_sets(c);
    if (r) return r;
    if (e != hipSuccess) return fail(FD23_EHIP, "graph capture: %s", hipGetErrorString(e));
    hipGraphExec_t exec;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(FD23_EHIP, "graph instantiate: %s", hipGetErrorString(e));
    c->graphs[key] = fd23_ctx::Chunk{exec, flips};
    *out = c->graphs[key];
    return FD23_OK;
}

static int launch_chunk(fd23_ctx *c, int itrecord, int steps) {
    fd23_ctx::Chunk ch;
    int r = get_graph(c, itrecord, steps, &ch);
    if (r) return r;
    HIPCHK(hipGraphLaunch(ch.exec, c->stream));
    if (ch.flips) swap_sets(c);
    return FD23_OK;
}

// what fd23_run needs before it captures or replays
static int run_prologue(fd23_ctx *c, int it0, int nt, int itrecord) {
    int r = ready(c);
    if (r) return r;
    if (nt < 0 || it0 < 0 || itrecord < 1) return fail(FD23_EINVAL, "bad run range");
    if ((c->box || c->punch) && !slots_default(c))
        return fail(FD23_EINVAL, "the box and the first-arrival window run every shot from step 0");
    if (c->box && (r = fd23_box_check_run(c, it0, nt))) return r;
    if ((r = build_events(c))) return r;
    if (use_fused(c) && ((r = ensure_alt(c)) || (r = build_fused_events(c)))) return r;
    if (use_block(c) && ((r = ensure_alt(c)) || (r = build_block_events(c)))) return r;
    return init_scales(c);
}

// nt steps run as full kGraphSteps-step graphs plus one graph of the
// remaining steps, so every step of a run is replayed from a graph
extern "C" int fd23_run(fd23_ctx *c, int it0, int nt, int itrecord) {
    int r = run_prologue(c, it0, nt, itrecord);
    if (r) return r;
    hipLaunchKernelGGL(k_set_clock, dim3(1), dim3(64), 0, c->stream, c->clock, it0);
    const int full = nt / kGraphSteps, rest = nt % kGraphSteps;
    for (int i = 0; i < full; i++)
        if ((r = launch_chunk(c, itrecord, kGraphSteps))) return r;
    if (rest > 0 && (r = launch_chunk(c, itrecord, rest))) return r;
    if (c->box) fd23_box_set_step(c, it0 + nt);
    return check_launch();
}

// the graphs fd23_run(c, it0, nt, itrecord) replays (a chunk that flips the
// buffer sets is replayed from both), captured and instantiated now without
// running, so that run starts with no capture
extern "C" int fd23_prepare(fd23_ctx *c, int nt, int itrecord) {
    int r = run_prologue(c, 0, nt, itrecord);
    if (r) return r;
    fd23_ctx::Chunk ch;
    const int full = nt / kGraphSteps, rest = nt % kGraphSteps;
    bool flips = false;
    if (full > 0) {
        if ((r = get_graph(c, itrecord, kGraphSteps, &ch))) return r;
        flips = ch.flips;
        if (flips && full > 1) {
            swap_sets(c);
            r = get_graph(c, itrecord, kGraphSteps, &ch);
            swap_sets(c);
            if (r) return r;
        }
    }
    // the rest starts on the set the full chunks end on
    if (rest > 0) {
        const bool other = flips && (full & 1);
        if (other) swap_sets(c);
        r = get_graph(c, itrecord, rest, &ch);
        if (other) swap_sets(c);
    }
    return r;
}

extern "C" int fd23_sync(fd23_ctx *c) {
    if (!c) return fail(FD23_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return FD23_OK;
}

// ---------------------------------------------------------------------------
// outputs and state
// ---------------------------------------------------------------------------
extern "C" int fd23_get_seismograms(fd23_ctx *c, int shot, int comp, float *out, int nframes) {
    if (!c || !out || shot < 0 || shot >= c->d.nshots || comp < 0 || comp >= c->ncomp)
        return fail(FD23_EINVAL, "bad arguments");
    if (nframes < 0 || nframes > c->d.max_frames) return fail(FD23_EINVAL, "nframes %d > capacity", nframes);
    const int n = c->h_nrec[(size_t)shot * c->ncomp + comp];
    if (nframes == 0 || n == 0) return FD23_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    const long frame_stride = (long)c->d.nshots * c->ncomp * c->d.max_rec;
    const long off = ((long)shot * c->ncomp + comp) * c->d.max_rec;
    // copies rows of n values out of the frame-major buffer
    HIPCHK(hipMemcpy2D(out, (size_t)n * c->csize, (char *)c->seis + (size_t)off * c->csize,
                       (size_t)frame_stride * c->csize, (size_t)n * c->csize, nframes, hipMemcpyDeviceToHost));
    return FD23_OK;
}

static int internal_field(const fd23_ctx *c, int field) {
    if (is_psv(c)) {
        static const int map[5] = {0, 1, 2, 4, 3}; // U W Txx Txz Tzz -> U W Txx Tzz Txz
        return field >= 0 && field < 5 ? map[field] : -1;
    }
    return field >= 0 && field < 3 ? field : -1;
}

static float host_half_to_float(uint16_t h, int storage) {
    if (storage == FD23_STORE_BF16) {
        uint32_t u = (uint32_t)h << 16;
        float f;
        memcpy(&f, &u, 4);
        return f;
    }
    __half hv;
    memcpy(&hv, &h, 2);
    return __half2float(hv);
}
static uint16_t host_float_to_half(float v, int storage) {
    if (storage == FD23_STORE_BF16) {
        __hip_bfloat16 b = __float2bfloat16(v);
        uint16_t u;
        memcpy(&u, &b, 2);
        return u;
    }
    __half hv = __float2half(v);
    uint16_t u;
    memcpy(&u, &hv, 2);
    return u;
}

// copy one pitched array (n rows of w elements) between device and host,
// converting scaled storage to the compute type
static int xfer(fd23_ctx *c, void *dev, long pitch, int w, int h, void *host, float scale, bool to_host) {
    const int es = c->esize, cs = c->csize;
    std::vector<char> tmp((size_t)w * h * es);
    if (to_host) {
        HIPCHK(hipMemcpy2DAsync(tmp.data(), (size_t)w * es, dev, (size_t)pitch * es, (size_t)w * es, h,
                                hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (es == cs) {
            memcpy(host, tmp.data(), tmp.size());
        } else {
            const uint16_t *src = (const uint16_t *)tmp.data();
            float *dst = (float *)host;
            for (size_t i = 0; i < (size_t)w * h; i++) dst[i] = host_half_to_float(src[i], c->d.storage) * scale;
        }
    } else {
        if (es == cs) {
            memcpy(tmp.data(), host, tmp.size());
        } else {
            const float *src = (const float *)host;
            uint16_t *dst = (uint16_t *)tmp.data();
            for (size_t i = 0; i < (size_t)w * h; i++) dst[i] = host_float_to_half(src[i] / scale, c->d.storage);
        }
        HIPCHK(hipMemcpy2DAsync(dev, (size_t)pitch * es, tmp.data(), (size_t)w * es, (size_t)w * es, h,
                                hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return FD23_OK;
}

static int field_xfer(fd23_ctx *c, int shot, int field, void *host, bool to_host) {
    const int fi = internal_field(c, field);
    if (fi < 0 || shot < 0 || shot >= c->d.nshots || !host) return fail(FD23_EINVAL, "bad field/shot");
    float sc[16];
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(sc, c->scl, sizeof sc, hipMemcpyDeviceToHost));
    void *dev = (char *)c->fields[fi] + (size_t)shot * c->sstride * c->esize;
    return xfer(c, dev, c->px, c->d.nx, c->d.nz, host, sc[fi], to_host);
}

extern "C" int fd23_get_field(fd23_ctx *c, int shot, int field, void *out) {
    if (!c) return fail(FD23_EINVAL, "null context");
    return field_xfer(c, shot, field, out, true);
}
extern "C" int fd23_set_field(fd23_ctx *c, int shot, int field, const void *in) {
    if (!c) return fail(FD23_EINVAL, "null context");
    return field_xfer(c, shot, field, (void *)in, false);
}

extern "C" size_t fd23_state_bytes(const fd23_ctx *c) {
    if (!c) return 0;
    const size_t N = (size_t)c->d.nx * c->d.nz;
    const int half = c->npsi / 2;
    return ((size_t)c->nfield * N + (size_t)half * (c->xpsi + c->zpsi)) * c->csize;
}

static int state_xfer(fd23_ctx *c, int shot, void *buf, size_t bytes, bool to_host) {
    if (!c || !buf || shot < 0 || shot >= c->d.nshots) return fail(FD23_EINVAL, "bad arguments");
    if (bytes < fd23_state_bytes(c)) return fail(FD23_EINVAL, "state buffer too small");
    float sc[16];
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(sc, c->scl, sizeof sc, hipMemcpyDeviceToHost));
    char *p = (char *)buf;
    const size_t N = (size_t)c->d.nx * c->d.nz;
    for (int f = 0; f < c->nfield; f++) {
        const int fi = internal_field(c, f);
        void *dev = (char *)c->fields[fi] + (size_t)shot * c->sstride * c->esize;
        int r = xfer(c, dev, c->px, c->d.nx, c->d.nz, p, sc[fi], to_host);
        if (r) return r;
        p += N * c->csize;
    }
    const int half = c->npsi / 2;
    for (int q = 0; q < c->npsi; q++) {
        const long n = q < half ? c->xpsi : c->zpsi;
        if (n == 0) continue;
        void *dev = (char *)c->psi[q] + (size_t)shot * n * c->esize;
        int r = xfer(c, dev, n, (int)n, 1, p, sc[c->nfield + q], to_host);
        if (r) return r;
        p += (size_t)n * c->csize;
    }
    return FD23_OK;
}

extern "C" int fd23_get_state(fd23_ctx *c, int shot, void *buf, size_t bytes) {
    return state_xfer(c, shot, buf, bytes, true);
}
extern "C" int fd23_set_state(fd23_ctx *c, int shot, const void *buf, size_t bytes) {
    return state_xfer(c, shot, (void *)buf, bytes, false);
}

// ---------------------------------------------------------------------------
// tuning and measurement
// ---------------------------------------------------------------------------
extern "C" int fd23_set_tile(fd23_ctx *c, int tile) {
    if (!c || tile < 0 || tile >= kNumTiles) return fail(FD23_EINVAL, "bad tile %d", tile);
    if (tile_fused(tile) && !is_psv(c)) return fail(FD23_EINVAL, "the fused step is P-SV only");
    if (tile_block(tile) && (is_psv(c) || scaled_storage(c)))
        return fail(FD23_EINVAL, "the block step is SH with FP32 or FP64 storage only");
    if (tile_fused(tile) && kTiles[tile].fr == 0 && scaled_storage(c))
        return fail(FD23_EINVAL, "the tile-parallel fused step keeps FP32 or FP64 storage");
    if (c->tile != tile) {
        c->tile = tile;
        fd23_ctx_invalidate_graphs(c);
    }
    return FD23_OK;
}
extern "C" int fd23_get_tile(const fd23_ctx *c) { return c ? c->tile : -1; }

extern "C" double fd23_bytes_per_step(const fd23_ctx *c) {
    if (!c) return 0;
    const double S = c->d.nshots, N = (double)c->d.nx * c->d.nz;
    const double wf = c->esize, wm = c->csize;
    const double nxp = 2.0 * c->d.npml * c->d.nz;
    const double nzp = (c->d.topabsorb ? 2.0 : 1.0) * c->d.npml * c->d.nx;
    // the fused step reads the five fields once and writes them once
    if (is_psv(c) && use_fused(c)) return S * (10 * wf * N + 8 * wf * (nxp + nzp)) + 5 * wm * N;
    // the block step reads the state (V Tx Tz, four psi) and the material
    // once per launch of bk steps and writes the state once
    if (use_block(c)) return (S * (6 * wf * N + 4 * wf * (nxp + nzp)) + 3 * wm * N) / kTiles[c->tile].bk;
    if (is_psv(c)) return S * (15 * wf * N + 8 * wf * (nxp + nzp)) + 5 * wm * N;
    return S * (9 * wf * N + 4 * wf * (nxp + nzp)) + 3 * wm * N;
}

extern "C" int fd23_time_passes(fd23_ctx *c, int nsteps, double *stress_ms, double *velocity_ms) {
    int r = ready(c);
    if (r) return r;
    if (nsteps < 1 || nsteps > 4096) return fail(FD23_EINVAL, "nsteps must be in [1, 4096]");
    hipEvent_t e0, e1, e2;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventCreate(&e2));
    const bool sc = scaled_storage(c);
    const bool fused = use_fused(c), block = use_block(c);
    if (fused && ((r = ensure_alt(c)) || (r = build_fused_events(c)))) return r;
    if (block && (r = ensure_alt(c))) return r;
    if (block) {
        // nsteps block launches of bk steps back to back (no sources or
        // receivers): the first value is the time per step, the second 0
        const int K = kTiles[c->tile].bk;
        dispatch_block(c, 0, K, 1, false);
        hipEventRecord(e0, c->stream);
        for (int i = 0; i < nsteps; i++) dispatch_block(c, 0, K, 1, false);
        hipEventRecord(e1, c->stream);
        hipEventRecord(e2, c->stream);
    } else if (fused) {
        // nsteps fused steps back to back (no recording): the first value
        // is the step time, the second 0
        dispatch_fused(c, 0, 1, false);
        if (sc) {
            enqueue_commit(c, true);
            enqueue_commit(c, false);
        }
        hipEventRecord(e0, c->stream);
        for (int i = 0; i < nsteps; i++) dispatch_fused(c, 0, 1, false);
        hipEventRecord(e1, c->stream);
        hipEventRecord(e2, c->stream);
    } else {
        // nsteps stress passes back to back, then nsteps velocity passes:
        // the queue never drains and no event sits between two passes, so
        // each interval over nsteps is one kernel's duration plus its dispatch
        dispatch_pass(c, true, kNoEvents); // first pass outside the timing, then
        if (sc) enqueue_commit(c, true);   // the scales of its arrays are current
        dispatch_pass(c, false, kNoEvents);
        if (sc) enqueue_commit(c, false);
        hipEventRecord(e0, c->stream);
        for (int i = 0; i < nsteps; i++) dispatch_pass(c, true, kNoEvents);
        hipEventRecord(e1, c->stream);
        for (int i = 0; i < nsteps; i++) dispatch_pass(c, false, kNoEvents);
        hipEventRecord(e2, c->stream);
    }
    const hipError_t se = hipEventSynchronize(e2);
    float a = 0, b = 0;
    if (se == hipSuccess) {
        hipEventElapsedTime(&a, e0, e1);
        hipEventElapsedTime(&b, e1, e2);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipEventDestroy(e2);
    if (se != hipSuccess) return fail(FD23_EHIP, "%s", hipGetErrorString(se));
    const int per = block ? kTiles[c->tile].bk : 1; // steps per launch
    if (stress_ms) *stress_ms = a / nsteps / per;
    if (velocity_ms) *velocity_ms = b / nsteps / per;
    return check_launch();
}

// Tile autotuner: every candidate runs `steps` stress+velocity steps twice
// on the context's own arrays and the faster run counts; the state is saved
// first and restored afterwards so tuning leaves the simulation untouched.
// milliseconds per step of n graph-replayed steps (no injection, no
// recording) with the context's current shape and chain setting: the
// fastest of three replays after the capture; the graphs are dropped
// afterwards
static int time_chunk(fd23_ctx *c, int n, double *ms) {
    fd23_ctx_invalidate_graphs(c);
    c->timing = true;
    hipEvent_t a = nullptr, b = nullptr;
    int r = FD23_OK;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) r = fail(FD23_EHIP, "event");
    if (r == FD23_OK) r = launch_chunk(c, 1, n); // captures and runs once
    double best = 1e30;
    for (int rep = 0; rep < 3 && r == FD23_OK; rep++) {
        hipEventRecord(a, c->stream);
        r = launch_chunk(c, 1, n);
        hipEventRecord(b, c->stream);
        float t = 0;
        if (r == FD23_OK && (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&t, a, b) != hipSuccess))
            r = fail(FD23_EHIP, "chunk timing");
        best = std::min(best, (double)t);
    }
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    c->timing = false;
    fd23_ctx_invalidate_graphs(c);
    *ms = best / n;
    return r;
}

extern "C" int fd23_set_chains(fd23_ctx *c, int n) {
    if (!c) return fail(FD23_EINVAL, "null context");
    if (n < 0 || n > 4 || n == 3) return fail(FD23_EINVAL, "chains: 0, 1, 2 or 4");
    n = std::max(n, 1);
    if (n != c->chains) {
        c->chains = n;
        fd23_ctx_invalidate_graphs(c);
    }
    return FD23_OK;
}
extern "C" int fd23_get_chains(const fd23_ctx *c) { return c ? (use_chains(c) ? c->chains : 1) : -1; }

extern "C" int fd23_autotune(fd23_ctx *c, int steps, double *best_ms) {
    int r = ready(c);
    if (r) return r;
    if (steps < 1) steps = 4;
    HIPCHK(hipStreamSynchronize(c->stream));
    // snapshot fields and psi on the device
    std::vector<void *> saved;
    std::vector<size_t> sizes;
    auto snap = [&](void *p, size_t n) -> int {
        void *q = nullptr;
        if (hipMalloc(&q, n) != hipSuccess) return fail(FD23_ENOMEM, "autotune snapshot");
        hipMemcpyAsync(q, p, n, hipMemcpyDeviceToDevice, c->stream);
        saved.push_back(q);
        sizes.push_back(n);
        return FD23_OK;
    };
    const size_t S = c->d.nshots;
    for (int f = 0; f < c->nfield && !r; f++)
        r = snap(c->field_alloc[f], ((size_t)S * c->sstride + 2 * kGuard) * c->esize);
    const int half = c->npsi / 2;
    for (int q = 0; q < c->npsi && !r; q++)
        r = snap(c->psi[q], std::max<size_t>(1, S * (q < half ? c->xpsi : c->zpsi)) * c->esize);
    if (r) {
        for (auto p : saved) hipFree(p);
        return r;
    }
    // scales and running maxima too (FP16/BF16 storage)
    float scl_saved[2 * kNextScale];
    std::vector<unsigned> amax_saved((size_t)16 * kAmaxSpread);
    HIPCHK(hipMemcpy(scl_saved, c->scl, sizeof scl_saved, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(amax_saved.data(), c->amax, sizeof(unsigned) * amax_saved.size(), hipMemcpyDeviceToHost));
    const int orig = c->tile, orig_set = c->cur;
    int best = orig;
    double bt = 1e30;
    std::vector<std::pair<double, int>> timed; // pass time, shape
    for (int t = 0; t < kNumTiles; t++) {
        if (kTiles[t].spb > 1 && c->d.nshots < kTiles[t].spb) continue;
        // shapes that would run as another one: SH and a first-arrival
        // window run split shapes unsplit, the window and the box run fused
        // shapes as tile 0, and the window runs block shapes as tile 0
        if (kTiles[t].split && (!is_psv(c) || c->punch)) continue;
        if (tile_fused(t) && (!is_psv(c) || c->box || c->punch)) continue;
        if (tile_block(t) && (is_psv(c) || scaled_storage(c) || c->punch)) continue;
        if (tile_fused(t) && kTiles[t].fr == 0 && scaled_storage(c)) continue;
        c->tile = t;
        double a, b;
        fd23_time_passes(c, 1, &a, &b); // warm-up
        // the faster of two timings, so that one slow interval does not
        // decide between shapes a few percent apart
        double tt = 1e30;
        for (int rep = 0; rep < 2 && r == FD23_OK; rep++)
            if ((r = fd23_time_passes(c, steps, &a, &b)) == FD23_OK) tt = std::min(tt, a + b);
        if (r) break;
        timed.emplace_back(tt, t);
        if (tt < bt) {
            bt = tt;
            best = t;
        }
    }
    // When the best shape can run as chains (a two-pass shape with FP32 or
    // FP64 storage, several shot groups, no box or window), the two-pass
    // shapes within 10% of it that can run as chains run 64 graph-replayed
    // steps, the way fd23_run steps, with one, two and four chains; in
    // table order, an option replaces the pick only when 1% faster, so
    // shapes within noise of each other do not alternate from one tuning to
    // the next.  (Fused and block shapes read event lists that tuning does
    // not build: they are not timed this way.)
    c->chains = 1;
    if (c->cur != orig_set) swap_sets(c);
    c->tile = best;
    auto two_pass = [](int t) { return !tile_fused(t) && !tile_block(t); };
    if (r == FD23_OK && two_pass(best) && chains_possible(c)) {
        std::vector<int> cands;
        for (const auto &ct : timed) {
            if (ct.first > 1.10 * bt || !two_pass(ct.second)) continue;
            c->tile = ct.second;
            if (chains_possible(c)) cands.push_back(ct.second);
        }
        c->tile = best;
        std::sort(cands.begin(), cands.end());
        double top = 1e30;
        int top_t = best, top_ch = 1;
        for (int t : cands) {
            c->tile = t;
            for (int ch : {1, 2, 4}) {
                if (r != FD23_OK || (ch > 1 && !chains_possible(c, ch))) continue;
                c->chains = ch;
                double ms = 0;
                if ((r = time_chunk(c, 64, &ms)) == FD23_OK && (top >= 1e30 || ms < 0.99 * top)) {
                    top = ms;
                    top_t = t;
                    top_ch = ch;
                }
            }
        }
        best = top_t;
        c->chains = r == FD23_OK ? top_ch : 1;
    }
    // restore (the fused candidates may have left the other set current)
    if (c->cur != orig_set) swap_sets(c);
    size_t k = 0;
    for (int f = 0; f < c->nfield; f++, k++)
        hipMemcpyAsync(c->field_alloc[f], saved[k], sizes[k], hipMemcpyDeviceToDevice, c->stream);
    for (int q = 0; q < c->npsi; q++, k++)
        hipMemcpyAsync(c->psi[q], saved[k], sizes[k], hipMemcpyDeviceToDevice, c->stream);
    hipMemcpyAsync(c->scl, scl_saved, sizeof scl_saved, hipMemcpyHostToDevice, c->stream);
    hipMemcpyAsync(c->amax, amax_saved.data(), sizeof(unsigned) * amax_saved.size(), hipMemcpyHostToDevice, c->stream);
    hipStreamSynchronize(c->stream);
    for (auto p : saved) hipFree(p);
    c->tile = orig;
    if (r) return r;
    fd23_set_tile(c, best);
    if (best_ms) *best_ms = bt;
    return FD23_OK;
}

// ---------------------------------------------------------------------------
// device copy with 16-byte lanes and non-temporal stores, grid-stride over
// the buffer: the practical HBM read+write bandwidth the stencil passes are
// measured against, and the plain copy whose counters scale FETCH_SIZE and
// WRITE_SIZE
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_stream_copy(float4 *__restrict__ dst, const float4 *__restrict__ src,
                                                     size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float4 v = src[i];
        __builtin_nontemporal_store(v.x, &dst[i].x);
        __builtin_nontemporal_store(v.y, &dst[i].y);
        __builtin_nontemporal_store(v.z, &dst[i].z);
        __builtin_nontemporal_store(v.w, &dst[i].w);
    }
}

extern "C" int fd23_device_copy(void *dst, const void *src, size_t bytes, int grid, void *stream) {
    if (!dst || !src || bytes % 16 != 0 || grid < 1 || grid > (1 << 20))
        return fail(FD23_EINVAL, "device copy needs device buffers, a multiple of 16 bytes and 1..2^20 workgroups");
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, (hipStream_t)stream, (float4 *)dst,
                       (const float4 *)src, bytes / 16);
    return check_launch();
}

// ---------------------------------------------------------------------------
// The passes' memory pattern without their arithmetic: NR arrays read and
// NW written, nx x nz FP32 each, in the passes' 64 x 16 tiles (64 x 4
// threads, 4 rows each) and their XCD band order.  out_k = sum of the
// reads + k.  The stress pass reads 8 arrays and writes 3, the velocity
// pass reads 7 and writes 2.
// ---------------------------------------------------------------------------
struct StreamArrays {
    const float *r[8];
    float *w[3];
};
template <int NR, int NW>
__global__ __launch_bounds__(TX *BY) void k_stream_tiles(StreamArrays a, int nx, int ntx, int ntiles) {
    const int q = gridDim.x >> 3;
    const int L = (int)(blockIdx.x & 7) * q + (int)(blockIdx.x >> 3);
    if (L >= ntiles) return;
    const int tz = L / ntx, tx = L - tz * ntx;
    const int x = tx * TX + (int)threadIdx.x, z0 = tz * 16 + (int)threadIdx.y * 4;
    float acc[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const size_t i = (size_t)(z0 + r) * nx + x;
        float s = a.r[0][i];
#pragma unroll
        for (int k = 1; k < NR; ++k) s += a.r[k][i];
        acc[r] = s;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const size_t i = (size_t)(z0 + r) * nx + x;
#pragma unroll
        for (int k = 0; k < NW; ++k) a.w[k][i] = acc[r] + (float)k;
    }
}

template <int NR>
static void launch_stream_tiles(int nw, const StreamArrays &a, int nx, int ntx, int ntiles, hipStream_t s) {
    const dim3 grid(round_up(ntiles, 8)), block(TX, BY);
    if (nw == 1) hipLaunchKernelGGL((k_stream_tiles<NR, 1>), grid, block, 0, s, a, nx, ntx, ntiles);
    else if (nw == 2) hipLaunchKernelGGL((k_stream_tiles<NR, 2>), grid, block, 0, s, a, nx, ntx, ntiles);
    else hipLaunchKernelGGL((k_stream_tiles<NR, 3>), grid, block, 0, s, a, nx, ntx, ntiles);
}

extern "C" int fd23_device_stream_tiles(const float *const *reads, int nreads, float *const *writes, int nwrites,
                                        int nx, int nz, void *stream) {
    if (!reads || !writes || nreads < 1 || nreads > 8 || nwrites < 1 || nwrites > 3)
        return fail(FD23_EINVAL, "tiled stream needs 1..8 arrays read and 1..3 written");
    if (nx < TX || nz < 16 || nx % TX != 0 || nz % 16 != 0 || (long)(nx / TX) * (nz / 16) > (1L << 30))
        return fail(FD23_EINVAL, "tiled stream needs nx a multiple of 64 and nz a multiple of 16");
    StreamArrays a{};
    for (int k = 0; k < nreads; ++k) {
        if (!reads[k]) return fail(FD23_EINVAL, "tiled stream: null array");
        a.r[k] = reads[k];
    }
    for (int k = 0; k < nwrites; ++k) {
        if (!writes[k]) return fail(FD23_EINVAL, "tiled stream: null array");
        a.w[k] = writes[k];
    }
    const int ntx = nx / TX, ntiles = ntx * (nz / 16);
    const hipStream_t s = (hipStream_t)stream;
    switch (nreads) {
    case 1: launch_stream_tiles<1>(nwrites, a, nx, ntx, ntiles, s); break;
    case 2: launch_stream_tiles<2>(nwrites, a, nx, ntx, ntiles, s); break;
    case 3: launch_stream_tiles<3>(nwrites, a, nx, ntx, ntiles, s); break;
    case 4: launch_stream_tiles<4>(nwrites, a, nx, ntx, ntiles, s); break;
    case 5: launch_stream_tiles<5>(nwrites, a, nx, ntx, ntiles, s); break;
    case 6: launch_stream_tiles<6>(nwrites, a, nx, ntx, ntiles, s); break;
    case 7: launch_stream_tiles<7>(nwrites, a, nx, ntx, ntiles, s); break;
    default: launch_stream_tiles<8>(nwrites, a, nx, ntx, ntiles, s); break;
    }
    return check_launch();
}
