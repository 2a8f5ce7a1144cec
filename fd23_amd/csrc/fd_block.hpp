// SH block step: K time steps of the SH loop in one launch, for grids small
// enough that the two passes per step wait on dispatch and memory latency
// rather than on bandwidth.
//
// A workgroup (64 x NW threads, NW wavefronts) owns an OX x OZ tile of the
// grid and loads the region around it, 64 columns x NW*CZ rows: the tile
// plus a halo of H = 7*K cells on every side.  It then runs the K steps on
// the region in LDS and registers, recomputing the halo, and writes only
// the tile.  One SH step moves information 7 cells (the stress at x reads V
// at x-3..x+4, the velocity at x reads Tx at x-4..x+3, and the same in z),
// so after k steps the cells within 7k of the region's border are stale
// while the tile is exact.  Reads come from one buffer set and the tile is
// written to the other (ping-pong): neighbouring workgroups read the same
// halo cells in the launch.
//
// Thread (tx, ty) holds column tx and rows ty*CZ .. ty*CZ+CZ-1 of the
// region -- few rows per thread and many wavefronts per workgroup, so that
// each SIMD interleaves several wavefronts through the LDS latencies -- and
// keeps V, Tx, Tz, the four C-PML memory variables and the three material
// arrays of its cells in registers.  V, Tx and Tz are mirrored in LDS for
// the neighbours' derivatives; LDS cells outside the region hold 0.  Cells
// outside the grid hold 0 and are never updated, as the padding of the
// arrays in HBM.
//
// Every cell runs the operations of k_sh_stress / k_sh_velocity's exact
// path in the same order (SH/fd/cu_fd.cu:77-273), with the sources and the
// receivers of those kernels: results are bitwise equal to the two-pass
// loop.
#pragma once
#include "fd_kernels.hpp"

namespace fd23 {

template <class S> struct ShPsi4 {
    S *Tx_x, *V_x, *Tz_z, *V_z;
};
template <class C> struct ShProf4 {
    Prof<C> Tx_x, V_x, Tz_z, V_z;
};
template <class C> struct ShMat3 {
    const C *BV, *MUX, *MUZ;
};
struct BlockWork {
    int ntx, ntz;  // tiles in x and z
    int nwork;     // ntx*ntz*nshots (logical workgroups)
    int nsub;      // steps of this launch, 1..K
};

constexpr int kBlockReach = 7; // cells one SH step moves information
constexpr int kBlockTaps = 32;  // most source taps one (tile, shot) list may hold (lanes of a wavefront)

// tap e of the list whose entry i lane i of the wavefront holds (e uniform)
__device__ __forceinline__ TapT lane_tap(const TapT &t, int e) {
    auto rl64 = [e](double d) {
        const unsigned long long b = __builtin_bit_cast(unsigned long long, d);
        const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, e);
        const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), e);
        return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
    };
    TapT r;
    r.cell = __builtin_amdgcn_readlane(t.cell, e);
    r.field = __builtin_amdgcn_readlane(t.field, e);
    r.mode = __builtin_amdgcn_readlane(t.mode, e);
    r.m = rl64(t.m);
    r.q = rl64(t.q);
    return r;
}

template <int K, int CZ, int NW> struct BlockShape {
    static constexpr int H = kBlockReach * K; // halo
    static constexpr int RH = NW * CZ;         // region rows
    static constexpr int OX = TX - 2 * H;      // tile columns
    static constexpr int OZ = RH - 2 * H;      // tile rows
    static constexpr int LW = TX + 2 * HALO, LH = RH + 2 * HALO;
    static_assert(OX > 0 && OZ > 0, "halo larger than the region");
};

template <class S, int K, int CZ, int NW>
__global__ __launch_bounds__(TX *NW) void k_sh_block(ShFields<S> fa, ShFields<S> fb, ShMat3<typename Store<S>::C> m,
                                                     ShPsi4<S> pa, ShPsi4<S> pb, ShProf4<typename Store<S>::C> pr,
                                                     Grid g, BlockWork w, Events ev) {
    using C = typename Store<S>::C;
    static_assert(!Store<S>::scaled, "the block step keeps FP32 or FP64 storage");
    using B = BlockShape<K, CZ, NW>;
    constexpr int H = B::H, LW = B::LW, LH = B::LH;
    __shared__ C sV[LH * LW], sX[LH * LW], sZ[LH * LW];
    __shared__ C sZP[6][B::RH]; // z profiles of the region's rows: V_z b c k, Tz_z b c k

    // tiles of one shot are consecutive, and consecutive logical workgroups
    // run on one XCD, so neighbours read each other's halos through one L2
    const int q = gridDim.x >> 3;
    const int L = (int)(blockIdx.x & 7) * q + (int)(blockIdx.x >> 3);
    if (L >= w.nwork) return;
    const int ntiles = w.ntx * w.ntz;
    // idle shot slots are stepped too (their state is reset before reuse):
    // a test here would put a wait on the kernel arguments ahead of the
    // prologue's loads, and this kernel runs where that latency counts
    const int shot = L / ntiles;
    const int tile = L - shot * ntiles;
    const int tzi = tile / w.ntx, txi = tile - tzi * w.ntx;
    const int rx0 = txi * B::OX - H, rz0 = tzi * B::OZ - H; // region origin
    const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int nx = g.nx, nz = g.nz, px = g.px, npml = g.npml;
    const int gx = rx0 + tx;
    const int lz0 = ty * CZ;
    const bool colin = gx >= 0 && gx < nx;
    const bool xpml = colin && (gx < npml || gx >= nx - npml);
    const long base = (long)shot * g.sstride;

    const int tid = ty * TX + tx;
    const int key = tile * g.nshots + shot;

    // Every global load of the launch is issued before the first barrier,
    // in two dependent levels, so that the steps wait on no global memory.
    // Level 1: the heads of the (tile, shot) lists, the step, the source.
    int it0 = 0, e0 = 0, e1 = 0, r0 = 0, r1 = 0;
    SrcInfo si{};
    if (ev.clock) {
        it0 = ev.clock[0] + ev.k;
        e0 = ev.tap_start[key];
        e1 = ev.tap_start[key + 1];
        r0 = ev.rec_start[key];
        r1 = ev.rec_start[key + 1];
        si = ev.src[shot];
    }

    // The region's fields, material, C-PML memory and profiles, with no
    // branch around a load: coordinates are clamped into the grid and a
    // cell off the C-PML reads element 0 of a psi array; cells outside the
    // grid take 0 afterwards.
    const int cx = min(max(gx, 0), nx - 1);
    C xVb = pr.V_x.b[cx], xVc = pr.V_x.c[cx], xVk = pr.V_x.k[cx];
    C xTb = pr.Tx_x.b[cx], xTc = pr.Tx_x.c[cx], xTk = pr.Tx_x.k[cx];
    // z profiles of region row min(tid, RH-1), for the LDS table
    const int zrow = min(tid, B::RH - 1);
    const int zg = min(max(rz0 + zrow, 0), nz - 1);
    C zp[6] = {pr.V_z.b[zg], pr.V_z.c[zg], pr.V_z.k[zg], pr.Tz_z.b[zg], pr.Tz_z.c[zg], pr.Tz_z.k[zg]};

    C v[CZ], sx[CZ], sz[CZ], bv[CZ], mux[CZ], muz[CZ];
    C pTxx[CZ], pVx[CZ], pTzz[CZ], pVz[CZ];
#pragma unroll
    for (int r = 0; r < CZ; r++) {
        const int gz = rz0 + lz0 + r;
        const int cz = min(max(gz, 0), nz - 1);
        const long o = base + (long)cz * px + cx;
        const long mo = (long)cz * px + cx;
        v[r] = fa.V[o];
        sx[r] = fa.Tx[o];
        sz[r] = fa.Tz[o];
        bv[r] = m.BV[mo];
        mux[r] = m.MUX[mo];
        muz[r] = m.MUZ[mo];
        const bool zpml = (cz < npml && g.top) || cz >= nz - npml;
        const long xi = xpml ? (long)shot * g.xpsi_stride + xpsi_idx(cx, cz, nx, npml) : 0;
        const long zi = zpml ? (long)shot * g.zpsi_stride + zpsi_idx(cx, cz, nx, nz, npml, cz < npml && g.top) : 0;
        pTxx[r] = pa.Tx_x[xi];
        pVx[r] = pa.V_x[xi];
        pTzz[r] = pa.Tz_z[zi];
        pVz[r] = pa.V_z[zi];
    }

    // Level 2: the region's source taps (lane i of every wavefront holds list
    // entry min(i, n-1), read back with readlane; the lists end with one
    // spare entry, so an empty list reads that), the source samples of the
    // launch's steps, the thread's first receiver.  Nothing waits for them
    // before the first stress half needs the taps.
    const int ntap = e1 - e0;
    TapT tap{};
    double sval[K];
    bool son[K];
    RecT rec0{0, 0, 0};
    const int re0 = r0 + tid;
#pragma unroll
    for (int j = 0; j < K; j++) {
        son[j] = false;
        sval[j] = 0;
    }
    if (ev.clock) {
        tap = ev.taps[e0 + min(tx, max(ntap - 1, 0))]; // lane i: entry i of the list
        rec0 = ev.recs[min(re0, max(r1 - 1, r0))];
#pragma unroll
        for (int j = 0; j < K; j++) {
            // src_sample, with the index clamped instead of a branch
            const int t = it0 + j - si.t0;
            son[j] = ntap > 0 && j < w.nsub && si.lsrc > 0 && t >= 0 && (t < si.lsrc || si.clamp);
            sval[j] = ev.stf[son[j] ? si.stf0 + min(t, si.lsrc - 1) : 0];
        }
    }

    // cells outside the grid hold 0; off the C-PML the psi are unused
#pragma unroll
    for (int r = 0; r < CZ; r++) {
        const int gz = rz0 + lz0 + r;
        const bool in = colin && gz >= 0 && gz < nz;
        v[r] = in ? v[r] : C(0);
        sx[r] = in ? sx[r] : C(0);
        sz[r] = in ? sz[r] : C(0);
        bv[r] = in ? bv[r] : C(0);
        mux[r] = in ? mux[r] : C(0);
        muz[r] = in ? muz[r] : C(0);
    }

    // LDS: cells outside the region 0 (the rows above and below, the
    // columns left and right), the region's V, the z profiles
    for (int i = tid; i < LH * LW; i += TX * NW) {
        const int r = i / LW, col = i - r * LW;
        if (r < HALO || r >= LH - HALO || col < HALO || col >= LW - HALO) {
            sV[i] = C(0);
            sX[i] = C(0);
            sZ[i] = C(0);
        }
    }
#pragma unroll
    for (int r = 0; r < CZ; r++) sV[(lz0 + r + HALO) * LW + tx + HALO] = v[r];
    if (tid < B::RH) {
        const bool zin = rz0 + tid >= 0 && rz0 + tid < nz;
#pragma unroll
        for (int q = 0; q < 6; q++) sZP[q][tid] = zin ? zp[q] : C(0);
    }
    __syncthreads();

    // grid row 0 inside the region: the velocity half of each step first
    // sets Tz row 0 to -Tz row 1 (SH/fd/cu_fd.cu:226); uniform per workgroup
    const int lrow0 = -rz0;
    const bool img = lrow0 >= 0 && lrow0 + 1 < B::RH;

    // The reduced order of a cell is min(ox, oz) with ox = min(ix, nx-1-ix,
    // 4) of the column and oz the same of the row (order_at): the column's
    // coefficients are set once, a row at full order in z keeps them.
    const int ox = min(min(gx, nx - 1 - gx), 4);
    C cx0, cx1, cx2, cx3;
    Coef<C>::get(ox, cx0, cx1, cx2, cx3);
    const C k0 = C(1225. / 1024.), k1 = C(245. / 3072.), k2 = C(49. / 5120.), k3 = C(5. / 7168.);
    // a wavefront whose cells all run at full order with no C-PML, no grid
    // edge and no free surface takes the branch-free interior path of the
    // two-pass kernels (the test of tile_interior); uniform per wavefront
    const int lo = max(npml, 4), ztop = g.top ? lo : 4;
    const bool wint = rx0 >= lo && rx0 + TX <= nx - lo && rz0 + lz0 >= ztop && rz0 + lz0 + CZ <= nz - lo;

#pragma unroll
    for (int j = 0; j < K; j++) {
        if (j >= w.nsub) break;
        const int it = it0 + j;
        const int nt_ = son[j] ? ntap : 0; // taps active this step

        // ---- stress half: Tx, Tz from V (k_sh_stress)
        if (wint) {
#pragma unroll
            for (int r = 0; r < CZ; r++) {
                const C *pv = &sV[(lz0 + r + HALO) * LW + tx + HALO];
                const C dxV = dp(pv, 1, k0, k1, k2, k3);
                const C dzV = dm(pv, LW, k0, k1, k2, k3);
                sx[r] = mad(mux[r], dxV, sx[r]);
                sz[r] = mad(muz[r], dzV, sz[r]);
            }
        } else {
            // the exact path with selects in place of branches: a cell the
            // reference leaves alone (SH/fd/cu_fd.cu:140) keeps its values,
            // an x-psi only changes in the x C-PML; the z C-PML is a row
            // property, uniform over the wavefront
#pragma unroll
            for (int r = 0; r < CZ; r++) {
                const int lz = lz0 + r;
                const int iz = rz0 + lz;
                const bool upd = colin && iz >= 0 && iz < nz - 1 && gx != 0 && gx < nx - 1;
                const int oz = min(min(iz, nz - 1 - iz), 4);
                C c0 = cx0, c1 = cx1, c2 = cx2, c3 = cx3;
                if (oz < 4) Coef<C>::get(min(ox, oz), c0, c1, c2, c3);
                const C *pv = &sV[(lz + HALO) * LW + tx + HALO];
                const C dxV = dp(pv, 1, c0, c1, c2, c3);
                const C dzV = dm(pv, LW, c0, c1, c2, c3);
                const C p = mad(xVb, pVx[r], xVc * dxV);
                pVx[r] = xpml && upd ? p : pVx[r];
                const C tx_ = mad(mux[r], xpml ? mad(dxV, xVk, p) : dxV, sx[r]);
                C tz_;
                if ((iz < npml && g.top) || iz >= nz - npml) {
                    const C q = mad(sZP[0][lz], pVz[r], sZP[1][lz] * dzV);
                    pVz[r] = upd ? q : pVz[r];
                    tz_ = mad(muz[r], mad(dzV, sZP[2][lz], q), sz[r]);
                } else {
                    tz_ = mad(muz[r], dzV, sz[r]);
                }
                sx[r] = upd ? tx_ : sx[r];
                sz[r] = upd ? tz_ : sz[r];
            }
        }
        // stress taps of the source (Tx, Tz), after the cell's update
        for (int e = 0; e < nt_; e++) {
            const TapT tp = lane_tap(tap, e);
            const int lx = tp.cell % TX, lz = tp.cell / TX;
            if (tp.field == 0 || lx != tx || lz < lz0 || lz >= lz0 + CZ) continue;
#pragma unroll
            for (int r = 0; r < CZ; r++) {
                if (lz != lz0 + r) continue;
                if (tp.field == 1) sx[r] = apply_tap(sx[r], tp, sval[j]);
                else sz[r] = apply_tap(sz[r], tp, sval[j]);
            }
        }
#pragma unroll
        for (int r = 0; r < CZ; r++) {
            sX[(lz0 + r + HALO) * LW + tx + HALO] = sx[r];
            sZ[(lz0 + r + HALO) * LW + tx + HALO] = sz[r];
        }
        __syncthreads();
        if (img) {
            // the owner of row 0 in each grid column takes -Tz[row 1]
            if (colin && lrow0 >= lz0 && lrow0 < lz0 + CZ) {
                const C im = -sZ[(lrow0 + 1 + HALO) * LW + tx + HALO];
#pragma unroll
                for (int r = 0; r < CZ; r++)
                    if (lz0 + r == lrow0) sz[r] = im;
                sZ[(lrow0 + HALO) * LW + tx + HALO] = im;
            }
            __syncthreads();
        }

        // point force on V, before the cell's velocity update
        for (int e = 0; e < nt_; e++) {
            const TapT tp = lane_tap(tap, e);
            const int lx = tp.cell % TX, lz = tp.cell / TX;
            if (tp.field != 0 || lx != tx || lz < lz0 || lz >= lz0 + CZ) continue;
#pragma unroll
            for (int r = 0; r < CZ; r++)
                if (lz == lz0 + r) v[r] = apply_tap(v[r], tp, sval[j]);
        }

        // ---- velocity half: V from Tx, Tz (k_sh_velocity)
        if (wint) {
#pragma unroll
            for (int r = 0; r < CZ; r++) {
                const int o = (lz0 + r + HALO) * LW + tx + HALO;
                const C Tx_x = dm(&sX[o], 1, k0, k1, k2, k3);
                const C Tz_z = dp(&sZ[o], LW, k0, k1, k2, k3);
                C vv = v[r];
                vv = mad(bv[r], Tx_x, vv);
                vv = mad(bv[r], Tz_z, vv);
                v[r] = vv;
            }
        } else {
#pragma unroll
            for (int r = 0; r < CZ; r++) {
                const int lz = lz0 + r;
                const int iz = rz0 + lz;
                const bool upd = colin && iz >= 0 && iz < nz;
                const int oz = min(min(iz, nz - 1 - iz), 4);
                C c0 = cx0, c1 = cx1, c2 = cx2, c3 = cx3;
                if (oz < 4) Coef<C>::get(min(ox, oz), c0, c1, c2, c3);
                const int o = (lz + HALO) * LW + tx + HALO;
                const C Tx_x = dm(&sX[o], 1, c0, c1, c2, c3);
                const C Tz_z = dp(&sZ[o], LW, c0, c1, c2, c3);
                const C p = mad(xTb, pTxx[r], xTc * Tx_x);
                pTxx[r] = xpml && upd ? p : pTxx[r];
                C vv = mad(bv[r], xpml ? mad(Tx_x, xTk, p) : Tx_x, v[r]);
                if ((iz < npml && g.top) || iz >= nz - npml) {
                    const C q = mad(sZP[3][lz], pTzz[r], sZP[4][lz] * Tz_z);
                    pTzz[r] = upd ? q : pTzz[r];
                    vv = mad(bv[r], mad(Tz_z, sZP[5][lz], q), vv);
                } else {
                    vv = mad(bv[r], Tz_z, vv);
                }
                v[r] = upd ? vv : v[r];
            }
        }
        // V of the region is read again by the next stress half only, after
        // the barrier below; the stress half's reads of sV ended at the
        // barrier above
#pragma unroll
        for (int r = 0; r < CZ; r++) sV[(lz0 + r + HALO) * LW + tx + HALO] = v[r];

        __syncthreads();
        // receivers of the tile (the lists hold tile cells only), one per
        // thread from the V just written to LDS; the next writes of sV come
        // after the next stress half's barrier
        // frames count from the shot's first step
        const int ts = it - si.t0;
        if (ev.clock && ts >= 0 && ts % ev.itrecord == 0 && ts / ev.itrecord < ev.max_frames) {
            C *seis = (C *)ev.seis + (long)(ts / ev.itrecord) * ev.frame_size;
            for (int e = re0; e < r1; e += TX * NW) {
                const RecT rc = e == re0 ? rec0 : ev.recs[e];
                const int lx = rc.cell % TX, lz = rc.cell / TX;
                seis[rc.out] = sV[(lz + HALO) * LW + lx + HALO];
            }
        }
    }

    // the tile's cells to the other buffer set
    const bool xown = tx >= H && tx < H + B::OX;
#pragma unroll
    for (int r = 0; r < CZ; r++) {
        const int lz = lz0 + r;
        const int gz = rz0 + lz;
        if (!xown || lz < H || lz >= H + B::OZ || !colin || gz < 0 || gz >= nz) continue;
        const long o = base + (long)gz * px + gx;
        fb.V[o] = v[r];
        fb.Tx[o] = sx[r];
        fb.Tz[o] = sz[r];
        if (xpml) {
            const long pi = (long)shot * g.xpsi_stride + xpsi_idx(gx, gz, nx, npml);
            pb.Tx_x[pi] = pTxx[r];
            pb.V_x[pi] = pVx[r];
        }
        const bool topcell = gz < npml && g.top;
        if (topcell || gz >= nz - npml) {
            const long pi = (long)shot * g.zpsi_stride + zpsi_idx(gx, gz, nx, nz, npml, topcell);
            pb.Tz_z[pi] = pTzz[r];
            pb.V_z[pi] = pVz[r];
        }
    }
}

} // namespace fd23
