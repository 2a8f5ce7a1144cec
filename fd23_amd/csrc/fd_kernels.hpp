// Stress and velocity passes (P-SV and SH) for gfx950, with the C-PML
// memory-variable updates fused in, plus source injection and receiver
// gathering.  Semantics follow the reference's GPU kernels:
//   P-SV  PSV/clean_nobox/fd/cu_fd.cu:96-388
//   SH    SH/fd/cu_fd.cu:77-273
//
// Tiling: a workgroup is 64 (x, one wavefront per row) x 4 (waves) threads
// and owns a TX x TZ tile, TZ = 4*RZ; each thread updates RZ consecutive
// cells of its column.  The fields whose derivatives are needed are staged
// in LDS with their 4-cell halos (x halo only for the rows of the tile, z
// halo only for the 64 tile columns).  Every global load of a tile -- the
// staged fields, the arrays the pass updates, the material -- is issued
// before the first LDS barrier, so a workgroup waits on memory once.  A
// workgroup steps SPB shots of the same model with the material held in
// registers, so the material is read once per SPB shots.
//
// Grid: one dimension; workgroup b runs logical tile (b % 8)*q + b/8
// (q = padded grid / 8).  Workgroups are dispatched round-robin over the 8
// XCDs, so each XCD (own L2) sweeps one contiguous band of tiles and the
// halo rows/columns a tile shares with its neighbours stay in that L2.
//
// Each tile is classified once: tiles that touch no C-PML strip, no edge
// band of reduced order and not the free surface take a branch-free path
// with the full 8th-order coefficients; all others evaluate the
// reference's per-cell branch structure exactly.
//
// In the graph-captured loop the source injection is fused into the stress
// pass and the receiver gather into the velocity pass (per-tile lists built
// on the host); the stand-alone k_inject / k_record kernels serve the
// staged entry points.  Fusing is exact: the stress update of a cell reads
// only U and W, so adding the source to a cell's stress right after its own
// update equals adding it after the whole pass.
#pragma once
#include "fd_device.hpp"

#pragma clang fp contract(off)

namespace fd23 {

// tile classification shared by every pass
__device__ __forceinline__ bool tile_interior(const Grid &g, int x0, int z0, int TZ) {
    const int lo = max(g.npml, 4);
    const int ztop = g.top ? lo : 4;
    return x0 >= lo && x0 + TX <= g.nx - lo && z0 >= ztop && z0 + TZ <= g.nz - lo;
}

// x-psi index (cu_fd.cu:192,202) and z-psi index (:217,226)
__device__ __forceinline__ long xpsi_idx(int ix, int iz, int nx, int npml) {
    return ix < npml ? (long)ix + (long)iz * 2 * npml : (long)(npml + nx - 1 - ix) + (long)iz * 2 * npml;
}
__device__ __forceinline__ long zpsi_idx(int ix, int iz, int nx, int nz, int npml, bool topcell) {
    return topcell ? (long)iz * nx + ix : (long)(npml + nz - 1 - iz) * nx + ix;
}

// ---------------------------------------------------------------------------
// work mapping
// ---------------------------------------------------------------------------
// A launch covers one region of the tile grid: all tiles, the rectangle of
// interior tiles [ix0, ix1) x [iz0, iz1), or the frame of edge tiles around
// it.  Interior and edge tiles are launched apart so that each kernel is
// compiled for one path: the interior kernel then needs a third fewer
// registers than one holding both paths, and keeps more waves per SIMD.
enum { REGION_ALL = 0, REGION_INTERIOR = 1, REGION_EDGE = 2 };
struct Work {
    int ntx, ntz, ngroups; // tiles of the whole grid in x, in z; shot groups
    int ix0, ix1, iz0, iz1; // interior rectangle (tile indices)
    int ntiles;            // tiles of this launch's region
    int nwork;             // ntiles*ngroups (logical workgroups)
    int g0;                // first shot group of the launch
};
struct TileId {
    int tile, group, x0, z0; // tile: index in the whole grid (ntx per row)
};
template <int TZ, int REGION = REGION_ALL> __device__ __forceinline__ bool map_tile(const Work &w, TileId &t) {
    const int q = gridDim.x >> 3; // grid padded to a multiple of 8
    const int L = (int)(blockIdx.x & 7) * q + (int)(blockIdx.x >> 3);
    if (L >= w.nwork) return false;
    // tiles of one shot group are consecutive, so neighbouring tiles (which
    // read each other's halos) run together on one XCD and share its L2
    const int gl = L / w.ntiles;
    int k = L - gl * w.ntiles;
    t.group = w.g0 + gl;
    int tx, tz;
    if (REGION == REGION_INTERIOR) {
        const int wx = w.ix1 - w.ix0;
        tz = k / wx;
        tx = w.ix0 + (k - tz * wx);
        tz += w.iz0;
    } else if (REGION == REGION_EDGE) {
        // rows above the rectangle, the side pieces of its rows, rows below
        const int top = w.iz0 * w.ntx, side = w.ix0 + (w.ntx - w.ix1), mid = (w.iz1 - w.iz0) * side;
        if (k < top) {
            tz = k / w.ntx;
            tx = k - tz * w.ntx;
        } else if (k - top < mid) {
            k -= top;
            const int r = k / side, c = k - r * side;
            tz = w.iz0 + r;
            tx = c < w.ix0 ? c : w.ix1 + (c - w.ix0);
        } else {
            k -= top + mid;
            tz = w.iz1 + k / w.ntx;
            tx = k - (tz - w.iz1) * w.ntx;
        }
    } else {
        tz = k / w.ntx;
        tx = k - tz * w.ntx;
    }
    t.tile = tz * w.ntx + tx;
    t.x0 = tx * TX;
    t.z0 = tz * TZ;
    return true;
}

// ---------------------------------------------------------------------------
// fused events (graph loop): step = clock[0] + k
// ---------------------------------------------------------------------------
struct TapT {
    int cell;   // lz*TX + lx inside the tile
    int field;  // internal field index (P-SV U W Txx Tzz Txz, SH V Tx Tz)
    double m;   // moment component (already scaled)
    double q;   // multiplier: +-1, +-0.25 (P-SV) or +-0.5 (SH)
    int mode;   // 0: float product then add (P-SV), 1: double sum (SH)
};
struct RecT {
    int cell;   // lz*TX + lx inside the tile
    int comp;
    int out;    // index inside one seismogram frame
};
struct SrcInfo {
    int ntap, tap0; // taps [tap0, tap0+ntap) of the stand-alone tap table
    int lsrc;
    int clamp;      // 1: it >= lsrc uses stf[lsrc-1] ('p' / stype 13)
    long stf0;      // offset of this shot's samples in the stf table
    int t0;         // the shot's first step: sample i is injected at step
                    // t0 + i, frame f recorded at step t0 + f*itrecord
};
struct Events {
    const int *clock;     // nullptr: no fused events
    int k;                // step offset inside the captured chunk
    int itrecord;
    const int *tap_start; // CSR over (tile*nshots + shot)
    const TapT *taps;
    const SrcInfo *src;
    const double *stf;
    const int *rec_start; // CSR over (tile*nshots + shot)
    const RecT *recs;
    void *seis;           // [max_frames][frame_size] compute type
    int max_frames;
    long frame_size;
    // first-arrival window per (tile*nshots + shot): x < it < y when the
    // workgroup has work (a block may be open, or the shot has a tap or
    // receiver in the tile), z < it < w when every block is open
    const int4 *pwin;
    // shots 0..63 whose bit is set are idle slots, not stepped
    unsigned long long idle;
};

// Idle shot slot (a kernel argument, so the test waits on no memory):
// the graph loop skips such a shot's work.  Shots from 64 on are never
// idle.
__device__ __forceinline__ bool shot_idle(unsigned long long idle, int shot) {
    return shot < 64 && ((idle >> shot) & 1ull);
}

// first-arrival window (the graph loop applies it: the step index is known
// there only).  Bit r of rows_open is set when row r of the thread's RZ
// rows is updated this step; rows off the grid are clear.
__device__ __forceinline__ bool punch_on(const Grid &g, const Events &ev) {
    return g.pfirst != nullptr && ev.clock != nullptr && ev.pwin != nullptr;
}
template <int RZ>
__device__ __forceinline__ unsigned rows_open(const Grid &g, bool pun, int gx, int gz0, int it) {
    if (!pun) return ~0u;
    unsigned m = 0;
#pragma unroll
    for (int r = 0; r < RZ; r++)
        if (gx < g.nx && gz0 + r < g.nz && punch_open(g, gx, gz0 + r, it)) m |= 1u << r;
    return m;
}

// what a workgroup does under a window (uniform per workgroup), from the
// per-tile table (one scalar load per shot): nothing when no shot of its
// group has work in the tile; the interior path when every block is open;
// else the exact path with the per-row mask
enum { kTileIdle = 0, kTileMixed = 1, kTileOpen = 2 };
template <int SPB>
__device__ __forceinline__ int tile_window(const Grid &g, const Events &ev, const TileId &t, int it) {
    const int k0 = t.tile * g.nshots + t.group * SPB;
    const int ns = min(SPB, g.nshots - t.group * SPB);
    const int4 w0 = ev.pwin[k0];
    bool work = w0.x < it && it < w0.y;
    for (int s = 1; s < ns && !work; s++) {
        const int4 w = ev.pwin[k0 + s];
        work = w.x < it && it < w.y;
    }
    if (!work) return kTileIdle;
    return w0.z < it && it < w0.w ? kTileOpen : kTileMixed;
}

// source sample of a shot at step it, or false when inactive (it >= lsrc
// without clamping), PSV/clean_nobox/point_source/cu_point_source.cu:47,110-112
__device__ __forceinline__ bool src_sample(const SrcInfo &si, const double *stf, int it, double &sv) {
    if (si.lsrc <= 0) return false;
    int t = it - si.t0;
    if (t < 0) return false;
    if (t >= si.lsrc) {
        if (!si.clamp) return false;
        t = si.lsrc - 1;
    }
    sv = stf[si.stf0 + t];
    return true;
}

template <class C, class T> __device__ __forceinline__ C apply_tap(C v, const T &tp, double sv) {
    if (tp.mode == 0) {
        const C prod = C(sv) * C(tp.m);
        return v + prod * C(tp.q);
    }
    return C((double)v + tp.q * sv * tp.m);
}

// ---------------------------------------------------------------------------
// tile staging through registers: all loads first, LDS writes afterwards
//  ZH: the 4-row halos above/below (all TX columns)
//  XH: the 4-column halos left/right (tile rows only)
// LDS row r <-> global row z0-4+r; LDS column c <-> global column x0-4+c.
// ---------------------------------------------------------------------------
template <class S, int TZ, bool ZH, bool XH> struct Stage {
    using C = typename Store<S>::C;
    static constexpr int LW = TX + 2 * HALO;
    static constexpr int RB = ZH ? 0 : HALO, RE = ZH ? TZ + 2 * HALO : TZ + HALO;
    static constexpr int NR = (RE - RB) / BY;
    static constexpr int NE = XH ? TZ * 2 * HALO : 0;
    static constexpr int NH = XH ? (NE + TX * BY - 1) / (TX * BY) : 0;
    static_assert((RE - RB) % BY == 0, "tile rows must split evenly over the waves");
    C r[NR];
    C h[NH > 0 ? NH : 1];

    __device__ __forceinline__ void load(const S *__restrict__ f, long base, int px, int x0, int z0, float scl) {
        const int tx = threadIdx.x, ty = threadIdx.y;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const int row = RB + ty + i * BY;
            r[i] = Store<S>::ld(f, base + (long)(z0 - HALO + row) * px + x0 + tx, scl);
        }
        if (XH) {
            const int tid = ty * TX + tx;
#pragma unroll
            for (int i = 0; i < NH; i++) {
                const int e = tid + i * TX * BY;
                if (NE % (TX * BY) == 0 || e < NE) {
                    const int row = e >> 3, c = e & 7;
                    const int col = c < HALO ? c - HALO : TX + c - HALO;
                    h[i] = Store<S>::ld(f, base + (long)(z0 + row) * px + x0 + col, scl);
                }
            }
        }
    }
    // the same loads through a buffer resource whose origin is element
    // (row -HALO, column -HALO) of the shot: pxb = row pitch in bytes,
    // vcol = this lane's byte offset of column x0 + tx; ty is wave-uniform
    __device__ __forceinline__ void loadb(BufRsrc rs, unsigned pxb, unsigned vcol, int x0, int z0, int ty,
                                          float scl) {
        constexpr unsigned SZ = sizeof(S);
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const int row = RB + ty + i * BY; // global row z0 - HALO + row
            r[i] = Store<S>::ldb(rs, vcol, (unsigned)(z0 + row) * pxb, scl);
        }
        if (XH) {
            const int tid = ty * TX + (int)threadIdx.x;
#pragma unroll
            for (int i = 0; i < NH; i++) {
                const int e = tid + i * TX * BY;
                if (NE % (TX * BY) == 0 || e < NE) {
                    const int row = e >> 3, c = e & 7;
                    const int col = c < HALO ? c - HALO : TX + c - HALO;
                    const unsigned vo = (unsigned)(row + HALO) * pxb + (unsigned)(x0 + col + HALO) * SZ;
                    h[i] = Store<S>::ldb(rs, vo, (unsigned)z0 * pxb, scl);
                }
            }
        }
    }
    __device__ __forceinline__ void store(C *sm) const {
        const int tx = threadIdx.x, ty = threadIdx.y;
#pragma unroll
        for (int i = 0; i < NR; i++) sm[(RB + ty + i * BY) * LW + tx + HALO] = r[i];
        if (XH) {
            const int tid = ty * TX + tx;
#pragma unroll
            for (int i = 0; i < NH; i++) {
                const int e = tid + i * TX * BY;
                if (NE % (TX * BY) == 0 || e < NE) {
                    const int row = e >> 3, c = e & 7;
                    const int col = c < HALO ? c - HALO : TX + c - HALO;
                    sm[(row + HALO) * LW + col + HALO] = h[i];
                }
            }
        }
    }
};

// ===========================================================================
// P-SV stress pass (cu_fd.cu:96-245)
// ===========================================================================
template <class C> struct PsvStressMat {
    const C *LAM, *MU, *MUA; // pitch px
};
template <class S> struct PsvStressPsi {
    S *Ux, *Wx, *Uz, *Wz; // shot 0
};
template <class C> struct PsvStressProf {
    Prof<C> Ux, Wx, Uz, Wz;
};
// scale slots: fields U W Txx Tzz Txz = 0..4; psi Txx_x Txz_x U_x W_x
// Tzz_z Txz_z U_z W_z = 5..12
enum { SC_U = 0, SC_W, SC_TXX, SC_TZZ, SC_TXZ, SC_PTXX_X, SC_PTXZ_X, SC_PU_X, SC_PW_X,
       SC_PTZZ_Z, SC_PTXZ_Z, SC_PU_Z, SC_PW_Z, SC_PSV_N };
// scale buffer: current scales [0, kNextScale), next scales after them
constexpr int kNextScale = 16;
constexpr int kScaleExpMax = 100;

// ---------------------------------------------------------------------------
// per-cell P-SV updates, shared by the two-pass kernels and the fused step so
// that both run the same arithmetic (cu_fd.cu:183-243 and :334-386)
// ---------------------------------------------------------------------------
struct PsiScl4 {
    float r[4], w[4]; // read scales and inverse write scales
};

template <class C>
__device__ __forceinline__ void psv_stress_interior(C ux, C uz, C wx, C wz, C lm, C mm, C ma, C &a, C &b, C &cc) {
    const C gam = lm + C(2.0) * mm;
    a = mad(gam, ux, a); b = mad(lm, ux, b); cc = mad(ma, wx, cc);
    a = mad(lm, wz, a); b = mad(gam, wz, b); cc = mad(ma, uz, cc);
}

// exact path of one stress cell; psi read from rd and (when own) written to
// wr; sc: U_x W_x U_z W_z; mp: running max of U_x W_x W_z U_z.  PRE: the
// psi values were loaded ahead (q: U_x W_x W_z U_z of the cell) and rd is
// not read.
template <class S, class C, bool PRE = false>
__device__ __forceinline__ void psv_stress_exact(const Grid &g, int ix, int iz, int shot, C ux, C uz, C wx, C wz,
                                                 C lm, C mm, C ma, C &a, C &b, C &cc, const PsvStressPsi<S> &rd,
                                                 const PsvStressPsi<S> &wr, bool own,
                                                 const PsvStressProf<C> &pr, const PsiScl4 &sc, float (&mp)[4],
                                                 C q0 = C(0), C q1 = C(0), C q2 = C(0), C q3 = C(0)) {
    using ST = Store<S>;
    const int nx = g.nx, nz = g.nz, npml = g.npml;
    const C gam = lm + C(2.0) * mm;
    if (ix < npml || ix >= nx - npml) {
        const long pi = (long)shot * g.xpsi_stride + xpsi_idx(ix, iz, nx, npml);
        C p1 = mad(pr.Ux.b[ix], PRE ? q0 : ST::ld(rd.Ux, pi, sc.r[0]), pr.Ux.c[ix] * ux);
        if (own) ST::st(wr.Ux, pi, p1, sc.w[0]);
        const C e1 = mad(ux, pr.Ux.k[ix], p1);
        a = mad(gam, e1, a);
        b = mad(lm, e1, b);
        C p2 = mad(pr.Wx.b[ix], PRE ? q1 : ST::ld(rd.Wx, pi, sc.r[1]), pr.Wx.c[ix] * wx);
        if (own) ST::st(wr.Wx, pi, p2, sc.w[1]);
        cc = mad(ma, mad(wx, pr.Wx.k[ix], p2), cc);
        if (ST::scaled && own) {
            mp[0] = fmaxf(mp[0], fabsf((float)p1));
            mp[1] = fmaxf(mp[1], fabsf((float)p2));
        }
    } else if (iz != 0) {
        a = mad(gam, ux, a); b = mad(lm, ux, b); cc = mad(ma, wx, cc);
    }
    const bool topcell = iz < npml && g.top;
    if (topcell || iz >= nz - npml) {
        const long pi = (long)shot * g.zpsi_stride + zpsi_idx(ix, iz, nx, nz, npml, topcell);
        C p1 = mad(pr.Wz.b[iz], PRE ? q2 : ST::ld(rd.Wz, pi, sc.r[3]), pr.Wz.c[iz] * wz);
        if (own) ST::st(wr.Wz, pi, p1, sc.w[3]);
        const C e1 = mad(wz, pr.Wz.k[iz], p1);
        a = mad(lm, e1, a);
        b = mad(gam, e1, b);
        C p2 = mad(pr.Uz.b[iz], PRE ? q3 : ST::ld(rd.Uz, pi, sc.r[2]), pr.Uz.c[iz] * uz);
        if (own) ST::st(wr.Uz, pi, p2, sc.w[2]);
        cc = mad(ma, mad(uz, pr.Uz.k[iz], p2), cc);
        if (ST::scaled && own) {
            mp[2] = fmaxf(mp[2], fabsf((float)p1));
            mp[3] = fmaxf(mp[3], fabsf((float)p2));
        }
    } else if (iz == 0) {
        a = mad(C(2.0) * mm, ux, a);
        b = C(0);
        cc = C(0);
    } else {
        a = mad(lm, wz, a); b = mad(gam, wz, b); cc = mad(ma, uz, cc);
    }
}

template <class S, int RZ, int SPB, int REGION>
__global__ __launch_bounds__(TX *BY) void k_psv_stress(PsvFields<S> f, PsvStressMat<typename Store<S>::C> m,
                                                       PsvStressPsi<S> ps,
                                                       PsvStressProf<typename Store<S>::C> pr, Grid g, Work w,
                                                       Events ev, const float *__restrict__ scl,
                                                       unsigned *__restrict__ amax) {
    using C = typename Store<S>::C;
    using ST = Store<S>;
    constexpr int TZ = BY * RZ;
    constexpr int LW = TX + 2 * HALO;
    __shared__ C sU[(TZ + 2 * HALO) * LW];
    __shared__ C sW[(TZ + 2 * HALO) * LW];

    TileId t;
    if (!map_tile<TZ, REGION>(w, t)) return;
    // a wavefront is one row of 64 threads, so threadIdx.y is wave-uniform
    const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int x0 = t.x0, z0 = t.z0;
    const int nx = g.nx, nz = g.nz, px = g.px;
    const int gx = x0 + tx;
    const int lz0 = ty * RZ;
    bool interior = REGION == REGION_INTERIOR  ? true
                    : REGION == REGION_EDGE    ? false
                                               : tile_interior(g, x0, z0, TZ);
    const int it = ev.clock ? ev.clock[0] + ev.k : 0;
    // first-arrival window: the interior-only kernel is never launched with
    // one; a tile with a closed block takes the exact path, which masks rows
    const bool pun = REGION != REGION_INTERIOR && punch_on(g, ev); // uniform
    if (pun) {
        const int tw = tile_window<SPB>(g, ev, t, it);
        if (tw == kTileIdle) return;
        interior = interior && tw == kTileOpen;
    }
    const unsigned upd = rows_open<RZ>(g, pun, gx, z0 + lz0, it);

    // stored arrays are read in the current scales scl[0..15] and the
    // updated ones written in the next scales scl[16..31] (equal except
    // right after a scale change)
    float sc_u = 1.f, sc_w = 1.f, sc_xx = 1.f, sc_zz = 1.f, sc_xz = 1.f;
    float wi_xx = 1.f, wi_zz = 1.f, wi_xz = 1.f;
    if (ST::scaled) {
        sc_u = scl[SC_U]; sc_w = scl[SC_W];
        sc_xx = scl[SC_TXX]; sc_zz = scl[SC_TZZ]; sc_xz = scl[SC_TXZ];
        wi_xx = inv_pow2(scl[kNextScale + SC_TXX]); wi_zz = inv_pow2(scl[kNextScale + SC_TZZ]);
        wi_xz = inv_pow2(scl[kNextScale + SC_TXZ]);
    }
    float mx_xx = 0.f, mx_zz = 0.f, mx_xz = 0.f;
    float mp[4] = {0.f, 0.f, 0.f, 0.f}; // psi U_x W_x W_z U_z
    // psi read scales and inverse write scales (exact path): U_x W_x U_z W_z
    PsiScl4 psc{{1.f, 1.f, 1.f, 1.f}, {1.f, 1.f, 1.f, 1.f}};
    if (ST::scaled && !interior) {
        psc.r[0] = scl[SC_PU_X]; psc.r[1] = scl[SC_PW_X]; psc.r[2] = scl[SC_PU_Z]; psc.r[3] = scl[SC_PW_Z];
        psc.w[0] = inv_pow2(scl[kNextScale + SC_PU_X]); psc.w[1] = inv_pow2(scl[kNextScale + SC_PW_X]);
        psc.w[2] = inv_pow2(scl[kNextScale + SC_PU_Z]); psc.w[3] = inv_pow2(scl[kNextScale + SC_PW_Z]);
    }
    constexpr bool kPrePsi = REGION == REGION_EDGE;

    // buffer addressing: byte pitch of a row and this lane's column
    constexpr unsigned SZ = sizeof(S);
    const unsigned pxb = (unsigned)px * SZ, vcol = (unsigned)(gx + HALO) * SZ;
    const unsigned pxm = (unsigned)px * sizeof(C), vmat = (unsigned)gx * sizeof(C);

    // material for this thread's cells, reused across the SPB shots
    const BufRsrc rLAM = buf_rsrc(m.LAM), rMU = buf_rsrc(m.MU), rMUA = buf_rsrc(m.MUA);
    C lam[RZ], mu[RZ], mua[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        const int gz = z0 + lz0 + r;
        const bool live = gx < nx && gz < nz;
        const unsigned so = (unsigned)gz * pxm;
        lam[r] = live ? Store<C>::ldb(rLAM, vmat, so, 1.f) : C(0);
        mu[r] = live ? Store<C>::ldb(rMU, vmat, so, 1.f) : C(0);
        mua[r] = live ? Store<C>::ldb(rMUA, vmat, so, 1.f) : C(0);
    }

    const int sh0 = t.group * SPB;
    for (int s = 0; s < SPB; s++) {
        const int shot = sh0 + s;
        if (shot >= g.nshots) break;
        if (ev.clock && shot_idle(ev.idle, shot)) continue; // uniform per workgroup
        // buffer origin of this shot's arrays: element (row -HALO, column -HALO)
        const long org = (long)shot * g.sstride - (long)HALO * px - HALO;
        const BufRsrc bU = buf_rsrc(f.U + org), bW = buf_rsrc(f.W + org);
        const BufRsrc bXX = buf_rsrc(f.Txx + org), bZZ = buf_rsrc(f.Tzz + org), bXZ = buf_rsrc(f.Txz + org);
        Stage<S, TZ, true, true> rU, rW;
        rU.loadb(bU, pxb, vcol, x0, z0, ty, sc_u);
        rW.loadb(bW, pxb, vcol, x0, z0, ty, sc_w);
        C txx[RZ], tzz[RZ], txz[RZ];
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const unsigned so = (unsigned)(z0 + lz0 + r + HALO) * pxb;
            txx[r] = ST::ldb(bXX, vcol, so, sc_xx);
            tzz[r] = ST::ldb(bZZ, vcol, so, sc_zz);
            txz[r] = ST::ldb(bXZ, vcol, so, sc_xz);
        }
        // edge-only kernels: this shot's C-PML memory variables (U_x W_x
        // W_z U_z) are loaded with the fields, ahead of the barrier
        C pq[4][kPrePsi ? RZ : 1];
        if constexpr (kPrePsi) {
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int iz = z0 + lz0 + r;
                pq[0][r] = pq[1][r] = pq[2][r] = pq[3][r] = C(0);
                if (gx >= nx || iz >= nz || !(upd >> r & 1u)) continue;
                if (gx < g.npml || gx >= nx - g.npml) {
                    const long pi = (long)shot * g.xpsi_stride + xpsi_idx(gx, iz, nx, g.npml);
                    pq[0][r] = ST::ld(ps.Ux, pi, psc.r[0]);
                    pq[1][r] = ST::ld(ps.Wx, pi, psc.r[1]);
                }
                const bool topcell = iz < g.npml && g.top;
                if (topcell || iz >= nz - g.npml) {
                    const long pi = (long)shot * g.zpsi_stride + zpsi_idx(gx, iz, nx, nz, g.npml, topcell);
                    pq[2][r] = ST::ld(ps.Wz, pi, psc.r[3]);
                    pq[3][r] = ST::ld(ps.Uz, pi, psc.r[2]);
                }
            }
        }
        if (s > 0) __syncthreads();
        rU.store(sU);
        rW.store(sW);
        __syncthreads();

        if (interior) {
            const C c0 = C(1225. / 1024.), c1 = C(245. / 3072.), c2 = C(49. / 5120.), c3 = C(5. / 7168.);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int lz = lz0 + r;
                const C *pu = &sU[(lz + HALO) * LW + tx + HALO];
                const C *pw = &sW[(lz + HALO) * LW + tx + HALO];
                const C ux = dm(pu, 1, c0, c1, c2, c3);
                const C uz = dm(pu, LW, c0, c1, c2, c3);
                const C wx = dp(pw, 1, c0, c1, c2, c3);
                const C wz = dp(pw, LW, c0, c1, c2, c3);
                psv_stress_interior(ux, uz, wx, wz, lam[r], mu[r], mua[r], txx[r], tzz[r], txz[r]);
            }
        } else {
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int lz = lz0 + r;
                const int ix = gx, iz = z0 + lz;
                if (ix >= nx || iz >= nz || !(upd >> r & 1u)) continue;
                C c0, c1, c2, c3;
                Coef<C>::get(order_at(ix, iz, nx, nz), c0, c1, c2, c3);
                const C *pu = &sU[(lz + HALO) * LW + tx + HALO];
                const C *pw = &sW[(lz + HALO) * LW + tx + HALO];
                const C ux = dm(pu, 1, c0, c1, c2, c3);
                const C uz = dm(pu, LW, c0, c1, c2, c3);
                const C wx = dp(pw, 1, c0, c1, c2, c3);
                const C wz = dp(pw, LW, c0, c1, c2, c3);
                if constexpr (kPrePsi)
                    psv_stress_exact<S, C, true>(g, ix, iz, shot, ux, uz, wx, wz, lam[r], mu[r], mua[r], txx[r],
                                                 tzz[r], txz[r], ps, ps, true, pr, psc, mp, pq[0][r], pq[1][r],
                                                 pq[2][r], pq[3][r]);
                else
                    psv_stress_exact<S, C>(g, ix, iz, shot, ux, uz, wx, wz, lam[r], mu[r], mua[r], txx[r], tzz[r],
                                           txz[r], ps, ps, true, pr, psc, mp);
            }
        }
        // fused source injection: the owner of a tap cell adds it after the
        // cell's own update, taps in the reference's order
        if (ev.clock) {
            const int key = t.tile * g.nshots + shot;
            const int e0 = ev.tap_start[key], e1 = ev.tap_start[key + 1];
            double sv;
            if (e1 > e0 && src_sample(ev.src[shot], ev.stf, it, sv)) {
                for (int e = e0; e < e1; e++) {
                    const TapT tp = ev.taps[e];
                    const int lx = tp.cell % TX, lz = tp.cell / TX;
                    if (lx != tx || lz < lz0 || lz >= lz0 + RZ) continue;
#pragma unroll
                    for (int r = 0; r < RZ; r++) {
                        if (lz != lz0 + r) continue;
                        if (tp.field == 2) txx[r] = apply_tap(txx[r], tp, sv);
                        else if (tp.field == 3) tzz[r] = apply_tap(tzz[r], tp, sv);
                        else txz[r] = apply_tap(txz[r], tp, sv);
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int gz = z0 + lz0 + r;
            if (gx >= nx || gz >= nz) continue;
            const unsigned so = (unsigned)(gz + HALO) * pxb;
            ST::stb(bXX, vcol, so, txx[r], wi_xx);
            ST::stb(bZZ, vcol, so, tzz[r], wi_zz);
            ST::stb(bXZ, vcol, so, txz[r], wi_xz);
            if (ST::scaled && amax) { // the maxima are gathered on the step before a check
                mx_xx = fmaxf(mx_xx, fabsf((float)txx[r]));
                mx_zz = fmaxf(mx_zz, fabsf((float)tzz[r]));
                mx_xz = fmaxf(mx_xz, fabsf((float)txz[r]));
            }
        }
    }
    if (ST::scaled) {
        amax_update(amax, SC_TXX, mx_xx);
        amax_update(amax, SC_TZZ, mx_zz);
        amax_update(amax, SC_TXZ, mx_xz);
        amax_update(amax, SC_PU_X, mp[0]);
        amax_update(amax, SC_PW_X, mp[1]);
        amax_update(amax, SC_PW_Z, mp[2]);
        amax_update(amax, SC_PU_Z, mp[3]);
    }
}

// gather fused into a velocity pass: the new values of the tile are parked
// in LDS (the staging buffers are free once every wave has computed), then
// the tile's receivers read them.  Uniform per workgroup.
template <class C, int RZ, int NC>
__device__ __forceinline__ void fused_record(const Events &ev, int it, int key, C *park0, C *park1,
                                             const C (&val)[NC][RZ], int lz0, int t0) {
    const int e0 = ev.rec_start[key], e1 = ev.rec_start[key + 1];
    const int ts = it - t0; // step of the shot
    if (e1 <= e0 || ts < 0 || ts % ev.itrecord != 0) return;
    const int frame = ts / ev.itrecord;
    if (frame >= ev.max_frames) return;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        park0[(lz0 + r) * TX + threadIdx.x] = val[0][r];
        if (NC > 1) park1[(lz0 + r) * TX + threadIdx.x] = val[NC - 1][r];
    }
    __syncthreads();
    C *seis = (C *)ev.seis + (long)frame * ev.frame_size;
    for (int e = e0 + threadIdx.y * TX + threadIdx.x; e < e1; e += TX * BY) {
        const RecT rc = ev.recs[e];
        seis[rc.out] = rc.comp == 0 ? park0[rc.cell] : park1[rc.cell];
    }
}

// ===========================================================================
// P-SV velocity pass (cu_fd.cu:246-388)
// ===========================================================================
template <class C> struct PsvVelMat {
    const C *BU, *BW;
};
template <class S> struct PsvVelPsi {
    S *Txx_x, *Txz_x, *Txz_z, *Tzz_z;
};
template <class C> struct PsvVelProf {
    Prof<C> Txx_x, Txz_x, Txz_z, Tzz_z;
};

template <class C>
__device__ __forceinline__ void psv_velocity_interior(C dxTxx, C dzTzz, C dxTxz, C dzTxz, C bu, C bw, C &u, C &w) {
    u = mad(bu, dxTxx, u); w = mad(bw, dxTxz, w);
    u = mad(bu, dzTxz, u); w = mad(bw, dzTzz, w);
}

// exact path of one velocity cell; sc: Txx_x Txz_x Txz_z Tzz_z; mp: the same.
// PRE: the psi values were loaded ahead (q, in the same order).
template <class S, class C, bool PRE = false>
__device__ __forceinline__ void psv_velocity_exact(const Grid &g, int ix, int iz, int shot, C dxTxx, C dzTzz,
                                                   C dxTxz, C dzTxz, C b_u, C b_w, C &u, C &ww,
                                                   const PsvVelPsi<S> &rd, const PsvVelPsi<S> &wr, bool own,
                                                   const PsvVelProf<C> &pr, const PsiScl4 &sc, float (&mp)[4],
                                                   C q0 = C(0), C q1 = C(0), C q2 = C(0), C q3 = C(0)) {
    using ST = Store<S>;
    const int nx = g.nx, nz = g.nz, npml = g.npml;
    if (ix < npml || ix >= nx - npml) {
        const long pi = (long)shot * g.xpsi_stride + xpsi_idx(ix, iz, nx, npml);
        C p1 = mad(pr.Txx_x.b[ix], PRE ? q0 : ST::ld(rd.Txx_x, pi, sc.r[0]), pr.Txx_x.c[ix] * dxTxx);
        if (own) ST::st(wr.Txx_x, pi, p1, sc.w[0]);
        u = mad(b_u, mad(dxTxx, pr.Txx_x.k[ix], p1), u);
        C p2 = mad(pr.Txz_x.b[ix], PRE ? q1 : ST::ld(rd.Txz_x, pi, sc.r[1]), pr.Txz_x.c[ix] * dxTxz);
        if (own) ST::st(wr.Txz_x, pi, p2, sc.w[1]);
        ww = mad(b_w, mad(dxTxz, pr.Txz_x.k[ix], p2), ww);
        if (ST::scaled && own) {
            mp[0] = fmaxf(mp[0], fabsf((float)p1));
            mp[1] = fmaxf(mp[1], fabsf((float)p2));
        }
    } else if (iz != 0) {
        u = mad(b_u, dxTxx, u);
        ww = mad(b_w, dxTxz, ww);
    }
    const bool topcell = iz < npml && g.top;
    if (topcell || iz >= nz - npml) {
        const long pi = (long)shot * g.zpsi_stride + zpsi_idx(ix, iz, nx, nz, npml, topcell);
        C p1 = mad(pr.Txz_z.b[iz], PRE ? q2 : ST::ld(rd.Txz_z, pi, sc.r[2]), pr.Txz_z.c[iz] * dzTxz);
        C p2 = mad(pr.Tzz_z.b[iz], PRE ? q3 : ST::ld(rd.Tzz_z, pi, sc.r[3]), pr.Tzz_z.c[iz] * dzTzz);
        if (own) {
            ST::st(wr.Txz_z, pi, p1, sc.w[2]);
            ST::st(wr.Tzz_z, pi, p2, sc.w[3]);
        }
        u = mad(b_u, mad(dzTxz, pr.Txz_z.k[iz], p1), u);
        ww = mad(b_w, mad(dzTzz, pr.Tzz_z.k[iz], p2), ww);
        if (ST::scaled && own) {
            mp[2] = fmaxf(mp[2], fabsf((float)p1));
            mp[3] = fmaxf(mp[3], fabsf((float)p2));
        }
    } else if (iz == 0) {
        u = mad(b_u, dxTxx + dzTxz, u);
        ww = C(0);
    } else {
        u = mad(b_u, dzTxz, u);
        ww = mad(b_w, dzTzz, ww);
    }
}

template <class S, int RZ, int SPB, int REGION>
__global__ __launch_bounds__(TX *BY) void k_psv_velocity(PsvFields<S> f, PsvVelMat<typename Store<S>::C> m,
                                                         PsvVelPsi<S> ps,
                                                         PsvVelProf<typename Store<S>::C> pr, Grid g, Work w,
                                                         Events ev, const float *__restrict__ scl,
                                                         unsigned *__restrict__ amax) {
    using C = typename Store<S>::C;
    using ST = Store<S>;
    constexpr int TZ = BY * RZ;
    constexpr int LW = TX + 2 * HALO;
    __shared__ C sXX[(TZ + 2 * HALO) * LW];
    __shared__ C sZZ[(TZ + 2 * HALO) * LW];
    __shared__ C sXZ[(TZ + 2 * HALO) * LW];

    TileId t;
    if (!map_tile<TZ, REGION>(w, t)) return;
    // a wavefront is one row of 64 threads, so threadIdx.y is wave-uniform
    const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int x0 = t.x0, z0 = t.z0;
    const int nx = g.nx, nz = g.nz, px = g.px;
    const int gx = x0 + tx;
    const int lz0 = ty * RZ;
    bool interior = REGION == REGION_INTERIOR  ? true
                    : REGION == REGION_EDGE    ? false
                                               : tile_interior(g, x0, z0, TZ);
    const int it = ev.clock ? ev.clock[0] + ev.k : 0;
    // first-arrival window: the interior-only kernel is never launched with
    // one; a tile with a closed block takes the exact path, which masks rows
    const bool pun = REGION != REGION_INTERIOR && punch_on(g, ev); // uniform
    if (pun) {
        const int tw = tile_window<SPB>(g, ev, t, it);
        if (tw == kTileIdle) return;
        interior = interior && tw == kTileOpen;
    }
    const unsigned upd = rows_open<RZ>(g, pun, gx, z0 + lz0, it);

    float sc_u = 1.f, sc_w = 1.f, sc_xx = 1.f, sc_zz = 1.f, sc_xz = 1.f;
    float wi_u = 1.f, wi_w = 1.f;
    if (ST::scaled) {
        sc_u = scl[SC_U]; sc_w = scl[SC_W];
        sc_xx = scl[SC_TXX]; sc_zz = scl[SC_TZZ]; sc_xz = scl[SC_TXZ];
        wi_u = inv_pow2(scl[kNextScale + SC_U]); wi_w = inv_pow2(scl[kNextScale + SC_W]);
    }
    float mx_u = 0.f, mx_w = 0.f;
    float mp[4] = {0.f, 0.f, 0.f, 0.f}; // psi Txx_x Txz_x Txz_z Tzz_z
    PsiScl4 psc{{1.f, 1.f, 1.f, 1.f}, {1.f, 1.f, 1.f, 1.f}};
    if (ST::scaled && !interior) {
        psc.r[0] = scl[SC_PTXX_X]; psc.r[1] = scl[SC_PTXZ_X]; psc.r[2] = scl[SC_PTXZ_Z]; psc.r[3] = scl[SC_PTZZ_Z];
        psc.w[0] = inv_pow2(scl[kNextScale + SC_PTXX_X]); psc.w[1] = inv_pow2(scl[kNextScale + SC_PTXZ_X]);
        psc.w[2] = inv_pow2(scl[kNextScale + SC_PTXZ_Z]); psc.w[3] = inv_pow2(scl[kNextScale + SC_PTZZ_Z]);
    }
    constexpr bool kPrePsi = REGION == REGION_EDGE;

    // buffer addressing: byte pitch of a row and this lane's column
    constexpr unsigned SZ = sizeof(S);
    const unsigned pxb = (unsigned)px * SZ, vcol = (unsigned)(gx + HALO) * SZ;
    const unsigned pxm = (unsigned)px * sizeof(C), vmat = (unsigned)gx * sizeof(C);

    const BufRsrc rBU = buf_rsrc(m.BU), rBW = buf_rsrc(m.BW);
    C bu[RZ], bw[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        const int gz = z0 + lz0 + r;
        const bool live = gx < nx && gz < nz;
        const unsigned so = (unsigned)gz * pxm;
        bu[r] = live ? Store<C>::ldb(rBU, vmat, so, 1.f) : C(0);
        bw[r] = live ? Store<C>::ldb(rBW, vmat, so, 1.f) : C(0);
    }

    const int sh0 = t.group * SPB;
    for (int s = 0; s < SPB; s++) {
        const int shot = sh0 + s;
        if (shot >= g.nshots) break;
        // idle slots: skipped with one shot per workgroup (the test in a
        // two-shot loop costs these kernels registers); stepping one is harmless
        if (SPB == 1 && ev.clock && shot_idle(ev.idle, shot)) continue;
        // buffer origin of this shot's arrays: element (row -HALO, column -HALO)
        const long org = (long)shot * g.sstride - (long)HALO * px - HALO;
        const BufRsrc bU = buf_rsrc(f.U + org), bW = buf_rsrc(f.W + org);
        const BufRsrc bXX = buf_rsrc(f.Txx + org), bZZ = buf_rsrc(f.Tzz + org), bXZ = buf_rsrc(f.Txz + org);
        Stage<S, TZ, false, true> rXX;
        Stage<S, TZ, true, false> rZZ;
        Stage<S, TZ, true, true> rXZ;
        rXX.loadb(bXX, pxb, vcol, x0, z0, ty, sc_xx);
        rZZ.loadb(bZZ, pxb, vcol, x0, z0, ty, sc_zz);
        rXZ.loadb(bXZ, pxb, vcol, x0, z0, ty, sc_xz);
        C nv[2][RZ];
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const unsigned so = (unsigned)(z0 + lz0 + r + HALO) * pxb;
            nv[0][r] = ST::ldb(bU, vcol, so, sc_u);
            nv[1][r] = ST::ldb(bW, vcol, so, sc_w);
        }
        // edge-only kernels: this shot's C-PML memory variables (Txx_x
        // Txz_x Txz_z Tzz_z) are loaded with the fields, ahead of the barrier
        C pq[4][kPrePsi ? RZ : 1];
        if constexpr (kPrePsi) {
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int iz = z0 + lz0 + r;
                pq[0][r] = pq[1][r] = pq[2][r] = pq[3][r] = C(0);
                if (gx >= nx || iz >= nz || !(upd >> r & 1u)) continue;
                if (gx < g.npml || gx >= nx - g.npml) {
                    const long pi = (long)shot * g.xpsi_stride + xpsi_idx(gx, iz, nx, g.npml);
                    pq[0][r] = ST::ld(ps.Txx_x, pi, psc.r[0]);
                    pq[1][r] = ST::ld(ps.Txz_x, pi, psc.r[1]);
                }
                const bool topcell = iz < g.npml && g.top;
                if (topcell || iz >= nz - g.npml) {
                    const long pi = (long)shot * g.zpsi_stride + zpsi_idx(gx, iz, nx, nz, g.npml, topcell);
                    pq[2][r] = ST::ld(ps.Txz_z, pi, psc.r[2]);
                    pq[3][r] = ST::ld(ps.Tzz_z, pi, psc.r[3]);
                }
            }
        }
        if (s > 0) __syncthreads();
        rXX.store(sXX);
        rZZ.store(sZZ);
        rXZ.store(sXZ);
        __syncthreads();

        if (interior) {
            const C c0 = C(1225. / 1024.), c1 = C(245. / 3072.), c2 = C(49. / 5120.), c3 = C(5. / 7168.);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int lz = lz0 + r;
                const int o = (lz + HALO) * LW + tx + HALO;
                const C dxTxx = dp(&sXX[o], 1, c0, c1, c2, c3);
                const C dzTzz = dm(&sZZ[o], LW, c0, c1, c2, c3);
                const C dxTxz = dm(&sXZ[o], 1, c0, c1, c2, c3);
                const C dzTxz = dp(&sXZ[o], LW, c0, c1, c2, c3);
                psv_velocity_interior(dxTxx, dzTzz, dxTxz, dzTxz, bu[r], bw[r], nv[0][r], nv[1][r]);
            }
        } else {
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int lz = lz0 + r;
                const int ix = gx, iz = z0 + lz;
                if (ix >= nx || iz >= nz || !(upd >> r & 1u)) continue;
                C c0, c1, c2, c3;
                Coef<C>::get(order_at(ix, iz, nx, nz), c0, c1, c2, c3);
                const int o = (lz + HALO) * LW + tx + HALO;
                const C dxTxx = dp(&sXX[o], 1, c0, c1, c2, c3);
                const C dzTzz = dm(&sZZ[o], LW, c0, c1, c2, c3);
                const C dxTxz = dm(&sXZ[o], 1, c0, c1, c2, c3);
                const C dzTxz = dp(&sXZ[o], LW, c0, c1, c2, c3);
                if constexpr (kPrePsi)
                    psv_velocity_exact<S, C, true>(g, ix, iz, shot, dxTxx, dzTzz, dxTxz, dzTxz, bu[r], bw[r],
                                                   nv[0][r], nv[1][r], ps, ps, true, pr, psc, mp, pq[0][r],
                                                   pq[1][r], pq[2][r], pq[3][r]);
                else
                    psv_velocity_exact<S, C>(g, ix, iz, shot, dxTxx, dzTzz, dxTxz, dzTxz, bu[r], bw[r], nv[0][r],
                                             nv[1][r], ps, ps, true, pr, psc, mp);
            }
        }
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int gz = z0 + lz0 + r;
            if (gx >= nx || gz >= nz) continue;
            const unsigned so = (unsigned)(gz + HALO) * pxb;
            ST::stb(bU, vcol, so, nv[0][r], wi_u);
            ST::stb(bW, vcol, so, nv[1][r], wi_w);
            if (ST::scaled && amax) { // the maxima are gathered on the step before a check
                mx_u = fmaxf(mx_u, fabsf((float)nv[0][r]));
                mx_w = fmaxf(mx_w, fabsf((float)nv[1][r]));
            }
        }
        if (ev.clock) fused_record<C, RZ, 2>(ev, it, t.tile * g.nshots + shot, sXX, sZZ, nv, lz0, ev.src[shot].t0);
    }
    if (ST::scaled) {
        amax_update(amax, SC_U, mx_u);
        amax_update(amax, SC_W, mx_w);
        amax_update(amax, SC_PTXX_X, mp[0]);
        amax_update(amax, SC_PTXZ_X, mp[1]);
        amax_update(amax, SC_PTXZ_Z, mp[2]);
        amax_update(amax, SC_PTZZ_Z, mp[3]);
    }
}

// ===========================================================================
// SH stress pass (SH/fd/cu_fd.cu:77-171)
// scale slots: V Tx Tz = 0..2; psi Tx_x V_x Tz_z V_z = 3..6
// ===========================================================================
enum { SC_V = 0, SC_TX, SC_TZ, SC_PTX_X, SC_PV_X, SC_PTZ_Z, SC_PV_Z, SC_SH_N };

template <class C> struct ShStressMat {
    const C *MUX, *MUZ;
};
template <class S> struct ShPsi2 {
    S *x, *z; // (V_x, V_z) for stress, (Tx_x, Tz_z) for velocity
};
template <class C> struct ShProf2 {
    Prof<C> x, z;
};

template <class S, int RZ, int SPB>
__global__ __launch_bounds__(TX *BY) void k_sh_stress(ShFields<S> f, ShStressMat<typename Store<S>::C> m,
                                                      ShPsi2<S> ps, ShProf2<typename Store<S>::C> pr, Grid g,
                                                      Work w, Events ev, const float *__restrict__ scl,
                                                      unsigned *__restrict__ amax) {
    using C = typename Store<S>::C;
    using ST = Store<S>;
    constexpr int TZ = BY * RZ;
    constexpr int LW = TX + 2 * HALO;
    __shared__ C sV[(TZ + 2 * HALO) * LW];

    TileId t;
    if (!map_tile<TZ>(w, t)) return;
    // a wavefront is one row of 64 threads, so threadIdx.y is wave-uniform
    const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int x0 = t.x0, z0 = t.z0;
    const int nx = g.nx, nz = g.nz, px = g.px, npml = g.npml;
    const int gx = x0 + tx;
    const int lz0 = ty * RZ;
    bool interior = tile_interior(g, x0, z0, TZ);
    const int it = ev.clock ? ev.clock[0] + ev.k : 0;
    // first-arrival window: a tile with a closed block takes the exact
    // path, which masks rows
    const bool pun = punch_on(g, ev); // uniform
    if (pun) {
        const int tw = tile_window<SPB>(g, ev, t, it);
        if (tw == kTileIdle) return;
        interior = interior && tw == kTileOpen;
    }
    const unsigned upd = rows_open<RZ>(g, pun, gx, z0 + lz0, it);

    float sc_v = 1.f, sc_x = 1.f, sc_z = 1.f, wi_x = 1.f, wi_z = 1.f;
    if (ST::scaled) {
        sc_v = scl[SC_V]; sc_x = scl[SC_TX]; sc_z = scl[SC_TZ];
        wi_x = inv_pow2(scl[kNextScale + SC_TX]); wi_z = inv_pow2(scl[kNextScale + SC_TZ]);
    }
    float mx_x = 0.f, mx_z = 0.f;
    float mp[2] = {0.f, 0.f}; // psi V_x V_z

    // buffer addressing: byte pitch of a row and this lane's column
    constexpr unsigned SZ = sizeof(S);
    const unsigned pxb = (unsigned)px * SZ, vcol = (unsigned)(gx + HALO) * SZ;
    const unsigned pxm = (unsigned)px * sizeof(C), vmat = (unsigned)gx * sizeof(C);

    const BufRsrc rMUX = buf_rsrc(m.MUX), rMUZ = buf_rsrc(m.MUZ);
    C mux[RZ], muz[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        const int gz = z0 + lz0 + r;
        const bool live = gx < nx && gz < nz;
        const unsigned so = (unsigned)gz * pxm;
        mux[r] = live ? Store<C>::ldb(rMUX, vmat, so, 1.f) : C(0);
        muz[r] = live ? Store<C>::ldb(rMUZ, vmat, so, 1.f) : C(0);
    }

    const int sh0 = t.group * SPB;
    for (int s = 0; s < SPB; s++) {
        const int shot = sh0 + s;
        if (shot >= g.nshots) break;
        if (ev.clock && shot_idle(ev.idle, shot)) continue; // uniform per workgroup
        // buffer origin of this shot's arrays: element (row -HALO, column -HALO)
        const long org = (long)shot * g.sstride - (long)HALO * px - HALO;
        const BufRsrc bV = buf_rsrc(f.V + org), bX = buf_rsrc(f.Tx + org), bZ = buf_rsrc(f.Tz + org);
        Stage<S, TZ, true, true> rV;
        rV.loadb(bV, pxb, vcol, x0, z0, ty, sc_v);
        C vx[RZ], vz[RZ];
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const unsigned so = (unsigned)(z0 + lz0 + r + HALO) * pxb;
            vx[r] = ST::ldb(bX, vcol, so, sc_x);
            vz[r] = ST::ldb(bZ, vcol, so, sc_z);
        }
        if (s > 0) __syncthreads();
        rV.store(sV);
        __syncthreads();
        if (interior) {
            const C c0 = C(1225. / 1024.), c1 = C(245. / 3072.), c2 = C(49. / 5120.), c3 = C(5. / 7168.);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int lz = lz0 + r;
                const C *pv = &sV[(lz + HALO) * LW + tx + HALO];
                const C dxV = dp(pv, 1, c0, c1, c2, c3);
                const C dzV = dm(pv, LW, c0, c1, c2, c3);
                vx[r] = mad(mux[r], dxV, vx[r]);
                vz[r] = mad(muz[r], dzV, vz[r]);
            }
        } else {
            float sp_x = 1.f, sp_z = 1.f, wp_x = 1.f, wp_z = 1.f;
            if (ST::scaled) {
                sp_x = scl[SC_PV_X]; sp_z = scl[SC_PV_Z];
                wp_x = inv_pow2(scl[kNextScale + SC_PV_X]); wp_z = inv_pow2(scl[kNextScale + SC_PV_Z]);
            }
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int lz = lz0 + r;
                const int ix = gx, iz = z0 + lz;
                // early return of SH/fd/cu_fd.cu:140
                if (ix >= nx || iz >= nz || ix == 0 || ix >= nx - 1 || iz >= nz - 1 || !(upd >> r & 1u)) continue;
                C c0, c1, c2, c3;
                Coef<C>::get(order_at(ix, iz, nx, nz), c0, c1, c2, c3);
                const C *pv = &sV[(lz + HALO) * LW + tx + HALO];
                const C dxV = dp(pv, 1, c0, c1, c2, c3);
                const C dzV = dm(pv, LW, c0, c1, c2, c3);
                C tx_ = vx[r], tz_ = vz[r];
                if (ix < npml || ix >= nx - npml) {
                    const long pi = (long)shot * g.xpsi_stride + xpsi_idx(ix, iz, nx, npml);
                    C p = mad(pr.x.b[ix], ST::ld(ps.x, pi, sp_x), pr.x.c[ix] * dxV);
                    ST::st(ps.x, pi, p, wp_x);
                    tx_ = mad(mux[r], mad(dxV, pr.x.k[ix], p), tx_);
                    if (ST::scaled) mp[0] = fmaxf(mp[0], fabsf((float)p));
                } else {
                    tx_ = mad(mux[r], dxV, tx_);
                }
                const bool topcell = iz < npml && g.top;
                if (topcell || iz >= nz - npml) {
                    const long pi = (long)shot * g.zpsi_stride + zpsi_idx(ix, iz, nx, nz, npml, topcell);
                    C p = mad(pr.z.b[iz], ST::ld(ps.z, pi, sp_z), pr.z.c[iz] * dzV);
                    ST::st(ps.z, pi, p, wp_z);
                    tz_ = mad(muz[r], mad(dzV, pr.z.k[iz], p), tz_);
                    if (ST::scaled) mp[1] = fmaxf(mp[1], fabsf((float)p));
                } else {
                    tz_ = mad(muz[r], dzV, tz_);
                }
                vx[r] = tx_; vz[r] = tz_;
            }
        }
        // fused SH source, stress taps (Tx, Tz); a point force on V is added
        // by the velocity pass, which is the only pass that writes V
        if (ev.clock) {
            const int key = t.tile * g.nshots + shot;
            const int e0 = ev.tap_start[key], e1 = ev.tap_start[key + 1];
            double sv;
            if (e1 > e0 && src_sample(ev.src[shot], ev.stf, it, sv)) {
                for (int e = e0; e < e1; e++) {
                    const TapT tp = ev.taps[e];
                    const int lx = tp.cell % TX, lz = tp.cell / TX;
                    if (tp.field == 0 || lx != tx || lz < lz0 || lz >= lz0 + RZ) continue;
#pragma unroll
                    for (int r = 0; r < RZ; r++) {
                        if (lz != lz0 + r) continue;
                        if (tp.field == 1) vx[r] = apply_tap(vx[r], tp, sv);
                        else vz[r] = apply_tap(vz[r], tp, sv);
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int gz = z0 + lz0 + r;
            if (gx >= nx || gz >= nz) continue;
            const unsigned so = (unsigned)(gz + HALO) * pxb;
            ST::stb(bX, vcol, so, vx[r], wi_x);
            ST::stb(bZ, vcol, so, vz[r], wi_z);
            if (ST::scaled && amax) { // the maxima are gathered on the step before a check
                mx_x = fmaxf(mx_x, fabsf((float)vx[r]));
                mx_z = fmaxf(mx_z, fabsf((float)vz[r]));
            }
        }
    }
    if (ST::scaled) {
        amax_update(amax, SC_TX, mx_x);
        amax_update(amax, SC_TZ, mx_z);
        amax_update(amax, SC_PV_X, mp[0]);
        amax_update(amax, SC_PV_Z, mp[1]);
    }
}

// ===========================================================================
// SH velocity pass (SH/fd/cu_fd.cu:172-273); Tz row 0 is first set to
// -Tz row 1 (:226), so every read of Tz row 0 in this pass sees -Tz[1].
// ===========================================================================
template <class S, int RZ, int SPB>
__global__ __launch_bounds__(TX *BY) void k_sh_velocity(ShFields<S> f, const typename Store<S>::C *__restrict__ BV,
                                                        ShPsi2<S> ps, ShProf2<typename Store<S>::C> pr, Grid g,
                                                        Work w, Events ev, const float *__restrict__ scl,
                                                        unsigned *__restrict__ amax) {
    using C = typename Store<S>::C;
    using ST = Store<S>;
    constexpr int TZ = BY * RZ;
    constexpr int LW = TX + 2 * HALO;
    __shared__ C sX[(TZ + 2 * HALO) * LW];
    __shared__ C sZ[(TZ + 2 * HALO) * LW];

    TileId t;
    if (!map_tile<TZ>(w, t)) return;
    // a wavefront is one row of 64 threads, so threadIdx.y is wave-uniform
    const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int x0 = t.x0, z0 = t.z0;
    const int nx = g.nx, nz = g.nz, px = g.px, npml = g.npml;
    const int gx = x0 + tx;
    const int lz0 = ty * RZ;
    bool interior = tile_interior(g, x0, z0, TZ);
    const int it = ev.clock ? ev.clock[0] + ev.k : 0;
    // first-arrival window: a tile with a closed block takes the exact
    // path, which masks rows
    const bool pun = punch_on(g, ev); // uniform
    if (pun) {
        const int tw = tile_window<SPB>(g, ev, t, it);
        if (tw == kTileIdle) return;
        interior = interior && tw == kTileOpen;
    }
    const unsigned upd = rows_open<RZ>(g, pun, gx, z0 + lz0, it);
    // the surface imaging of column gx runs when its row-0 block is open
    const bool img_open = (rows_open<1>(g, pun, gx, 0, it) & 1u) != 0;

    float sc_v = 1.f, sc_x = 1.f, sc_z = 1.f, wi_v = 1.f;
    if (ST::scaled) {
        sc_v = scl[SC_V]; sc_x = scl[SC_TX]; sc_z = scl[SC_TZ];
        wi_v = inv_pow2(scl[kNextScale + SC_V]);
    }
    float mx_v = 0.f;
    float mp[2] = {0.f, 0.f}; // psi Tx_x Tz_z

    // buffer addressing: byte pitch of a row and this lane's column
    constexpr unsigned SZ = sizeof(S);
    const unsigned pxb = (unsigned)px * SZ, vcol = (unsigned)(gx + HALO) * SZ;
    const unsigned pxm = (unsigned)px * sizeof(C), vmat = (unsigned)gx * sizeof(C);

    const BufRsrc rBV = buf_rsrc(BV);
    C bv[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        const int gz = z0 + lz0 + r;
        const bool live = gx < nx && gz < nz;
        bv[r] = live ? Store<C>::ldb(rBV, vmat, (unsigned)gz * pxm, 1.f) : C(0);
    }

    const int sh0 = t.group * SPB;
    for (int s = 0; s < SPB; s++) {
        const int shot = sh0 + s;
        if (shot >= g.nshots) break;
        // idle slots: skipped with one shot per workgroup (the test in a
        // two-shot loop costs these kernels registers); stepping one is harmless
        if (SPB == 1 && ev.clock && shot_idle(ev.idle, shot)) continue;
        const long base = (long)shot * g.sstride;
        // buffer origin of this shot's arrays: element (row -HALO, column -HALO)
        const long org = base - (long)HALO * px - HALO;
        const BufRsrc bV = buf_rsrc(f.V + org), bX = buf_rsrc(f.Tx + org), bZ = buf_rsrc(f.Tz + org);
        Stage<S, TZ, false, true> rX;
        Stage<S, TZ, true, false> rZ;
        rX.loadb(bX, pxb, vcol, x0, z0, ty, sc_x);
        rZ.loadb(bZ, pxb, vcol, x0, z0, ty, sc_z);
        C nv[1][RZ];
#pragma unroll
        for (int r = 0; r < RZ; r++) nv[0][r] = ST::ldb(bV, vcol, (unsigned)(z0 + lz0 + r + HALO) * pxb, sc_v);
        // stress imaging at the surface: the LDS copy of row 0 (LDS row
        // HALO - z0) takes -Tz[row 1], and the tile owning row 0 writes it back
        const bool img = z0 - HALO <= 0 && z0 + TZ + HALO > 0;
        C img_v = C(0);
        if (img && ty == 0 && img_open) img_v = -ST::ld(f.Tz, base + (long)px + gx, sc_z);
        if (s > 0) __syncthreads();
        rX.store(sX);
        rZ.store(sZ);
        if (img) {
            __syncthreads();
            if (ty == 0 && img_open) {
                sZ[(HALO - z0) * LW + tx + HALO] = img_v;
                if (z0 == 0 && gx < nx) ST::st(f.Tz, base + gx, img_v, inv_pow2(sc_z));
            }
        }
        __syncthreads();
        // fused SH point force (cu_point_force): V += source before the
        // cell's velocity update, as the reference injects between passes
        if (ev.clock) {
            const int key = t.tile * g.nshots + shot;
            const int e0 = ev.tap_start[key], e1 = ev.tap_start[key + 1];
            double sv;
            if (e1 > e0 && src_sample(ev.src[shot], ev.stf, it, sv)) {
                for (int e = e0; e < e1; e++) {
                    const TapT tp = ev.taps[e];
                    const int lx = tp.cell % TX, lz = tp.cell / TX;
                    if (tp.field != 0 || lx != tx || lz < lz0 || lz >= lz0 + RZ) continue;
#pragma unroll
                    for (int r = 0; r < RZ; r++)
                        if (lz == lz0 + r) nv[0][r] = apply_tap(nv[0][r], tp, sv);
                }
            }
        }
        if (interior) {
            const C c0 = C(1225. / 1024.), c1 = C(245. / 3072.), c2 = C(49. / 5120.), c3 = C(5. / 7168.);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int lz = lz0 + r;
                const int o = (lz + HALO) * LW + tx + HALO;
                const C Tx_x = dm(&sX[o], 1, c0, c1, c2, c3);
                const C Tz_z = dp(&sZ[o], LW, c0, c1, c2, c3);
                C v = nv[0][r];
                v = mad(bv[r], Tx_x, v);
                v = mad(bv[r], Tz_z, v);
                nv[0][r] = v;
            }
        } else {
            float sp_x = 1.f, sp_z = 1.f, wp_x = 1.f, wp_z = 1.f;
            if (ST::scaled) {
                sp_x = scl[SC_PTX_X]; sp_z = scl[SC_PTZ_Z];
                wp_x = inv_pow2(scl[kNextScale + SC_PTX_X]); wp_z = inv_pow2(scl[kNextScale + SC_PTZ_Z]);
            }
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int lz = lz0 + r;
                const int ix = gx, iz = z0 + lz;
                if (ix >= nx || iz >= nz || !(upd >> r & 1u)) continue;
                C c0, c1, c2, c3;
                Coef<C>::get(order_at(ix, iz, nx, nz), c0, c1, c2, c3);
                const int o = (lz + HALO) * LW + tx + HALO;
                const C Tx_x = dm(&sX[o], 1, c0, c1, c2, c3);
                const C Tz_z = dp(&sZ[o], LW, c0, c1, c2, c3);
                C v = nv[0][r];
                if (ix < npml || ix >= nx - npml) {
                    const long pi = (long)shot * g.xpsi_stride + xpsi_idx(ix, iz, nx, npml);
                    C p = mad(pr.x.b[ix], ST::ld(ps.x, pi, sp_x), pr.x.c[ix] * Tx_x);
                    ST::st(ps.x, pi, p, wp_x);
                    v = mad(bv[r], mad(Tx_x, pr.x.k[ix], p), v);
                    if (ST::scaled) mp[0] = fmaxf(mp[0], fabsf((float)p));
                } else {
                    v = mad(bv[r], Tx_x, v);
                }
                const bool topcell = iz < npml && g.top;
                if (topcell || iz >= nz - npml) {
                    const long pi = (long)shot * g.zpsi_stride + zpsi_idx(ix, iz, nx, nz, npml, topcell);
                    C p = mad(pr.z.b[iz], ST::ld(ps.z, pi, sp_z), pr.z.c[iz] * Tz_z);
                    ST::st(ps.z, pi, p, wp_z);
                    v = mad(bv[r], mad(Tz_z, pr.z.k[iz], p), v);
                    if (ST::scaled) mp[1] = fmaxf(mp[1], fabsf((float)p));
                } else {
                    v = mad(bv[r], Tz_z, v);
                }
                nv[0][r] = v;
            }
        }
#pragma unroll
        for (int r = 0; r < RZ; r++) {
            const int gz = z0 + lz0 + r;
            if (gx >= nx || gz >= nz) continue;
            ST::stb(bV, vcol, (unsigned)(gz + HALO) * pxb, nv[0][r], wi_v);
            if (ST::scaled && amax) mx_v = fmaxf(mx_v, fabsf((float)nv[0][r]));
        }
        if (ev.clock) fused_record<C, RZ, 1>(ev, it, t.tile * g.nshots + shot, sX, sX, nv, lz0, ev.src[shot].t0);
    }
    if (ST::scaled) {
        amax_update(amax, SC_V, mx_v);
        amax_update(amax, SC_PTX_X, mp[0]);
        amax_update(amax, SC_PTZ_Z, mp[1]);
    }
}

// ===========================================================================
// Stand-alone source injection (staged entry point fd23_inject): one thread
// per shot walks its taps in the reference's order
// (PSV/clean_nobox/point_source/cu_point_source.cu:45-70, :115-158;
// SH/point_source/cu_point_source.cu:36-64, :100-158, :192-198).
// ===========================================================================
struct Tap {
    int field;  // index into the context's field-pointer table
    long off;   // element offset inside the shot (row*px + col)
    double m;
    double q;
    int mode;
};

template <class S>
__global__ void k_inject(S *const *fields, long sstride, const Tap *__restrict__ taps,
                         const SrcInfo *__restrict__ src, const double *__restrict__ stf, int nshots, int it,
                         const float *__restrict__ scl) {
    using ST = Store<S>;
    for (int shot = threadIdx.x; shot < nshots; shot += blockDim.x) {
        const SrcInfo si = src[shot];
        double sv;
        if (si.ntap == 0 || !src_sample(si, stf, it, sv)) continue;
        for (int i = 0; i < si.ntap; i++) {
            const Tap tp = taps[si.tap0 + i];
            S *fp = fields[tp.field];
            const float sc = ST::scaled ? scl[tp.field] : 1.f;
            const long q = (long)shot * sstride + tp.off;
            const TapT tt{0, tp.field, tp.m, tp.q, tp.mode};
            ST::st(fp, q, apply_tap(ST::ld(fp, q, sc), tt, sv), inv_pow2(sc));
        }
    }
}

// ===========================================================================
// Stand-alone receiver gather (staged entry point fd23_record),
// SH/record/cu_record.cu:136-144, into [frame][shot][comp][max_rec];
// offset -1 reads 0 (flat index outside the grid).
// ===========================================================================
template <class S>
__global__ void k_record(S *const *comps, int ncomp, long sstride, const long *__restrict__ roff,
                         const int *__restrict__ nrec, int max_rec, int nshots,
                         typename Store<S>::C *__restrict__ seis, int max_frames, int it, int itrecord,
                         const float *__restrict__ scl, const int *__restrict__ comp_field,
                         const SrcInfo *__restrict__ src) {
    using C = typename Store<S>::C;
    using ST = Store<S>;
    const long total = (long)nshots * ncomp * max_rec;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int r = (int)(i % max_rec);
        const int sc_ = (int)(i / max_rec);
        const int comp = sc_ % ncomp, shot = sc_ / ncomp;
        if (r >= nrec[shot * ncomp + comp]) continue;
        // frames count from the shot's first step
        const int ts = it - src[shot].t0;
        if (ts < 0 || ts % itrecord != 0 || ts / itrecord >= max_frames) continue;
        const int frame = ts / itrecord;
        const long o = roff[i];
        C v = C(0);
        if (o >= 0) {
            const float s = ST::scaled ? scl[comp_field[comp]] : 1.f;
            v = ST::ld(comps[comp], (long)shot * sstride + o, s);
        }
        seis[(long)frame * total + i] = v;
    }
}

// ===========================================================================
// Per-array power-of-two scales of FP16/BF16 storage.  Every few steps the
// control kernel turns each array's running max |value| into a new scale
// when the stored max has left [2^lo, 2^hi], aiming at 2^target, and the
// rescale kernel multiplies the stored data by the (exact) power-of-two
// ratio of old to new scale.
// ===========================================================================
__global__ void k_scale_control(float *scl, unsigned *amax, int n, int lo, int hi, int target) {
    __shared__ unsigned red[kAmaxSpread];
    const int t = threadIdx.x; // blockDim.x == kAmaxSpread
    for (int f = 0; f < n; f++) {
        red[t] = amax[f * kAmaxSpread + t];
        amax[f * kAmaxSpread + t] = 0u;
        __syncthreads();
        for (int o = kAmaxSpread / 2; o > 0; o >>= 1) {
            if (t < o) red[t] = max(red[t], red[t + o]);
            __syncthreads();
        }
        if (t == 0) {
            const float m = __uint_as_float(red[0]);
            const float cur = scl[f];
            float nw = cur;
            if (m > 0.f && isfinite(m)) {
                const float stored = m / cur;
                if (stored > ldexpf(1.f, hi) || (lo > -126 && stored < ldexpf(1.f, lo))) {
                    int e;
                    frexpf(m, &e); // m in [2^(e-1), 2^e)
                    // kept within [2^-100, 2^100] so that 1/scale stays a normal float
                    nw = ldexpf(1.f, max(-kScaleExpMax, min(kScaleExpMax, e - target)));
                }
            }
            scl[kNextScale + f] = nw;
        }
        __syncthreads();
    }
}

// the arrays a pass has rewritten now hold values in the next scales
__global__ void k_commit_scales(float *scl, unsigned mask) {
    const int f = threadIdx.x;
    if (f < kNextScale && ((mask >> f) & 1u)) scl[f] = scl[kNextScale + f];
}

// advance the step clock after a captured chunk of steps
__global__ void k_tick(int *clock, int steps) {
    if (threadIdx.x == 0) clock[0] += steps;
}

// rewrite a small device pointer table in stream order
struct PtrTable5 {
    void *p[5];
};
__global__ void k_set_ptrs(void **tab, PtrTable5 v, int n) {
    if (threadIdx.x < n) tab[threadIdx.x] = v.p[threadIdx.x];
}

// set the step clock in stream order (no host round trip)
__global__ void k_set_clock(int *clock, int it0) {
    if (threadIdx.x == 0) clock[0] = it0;
}

} // namespace fd23
