// Fused P-SV step: the stress and the velocity update of one time step in a
// single pass, reading the fields of step it from one buffer set and
// writing those of step it+1 to the other (ping-pong), so no ordering
// between workgroups is needed.
//
// One wavefront per workgroup marches down a strip of the grid:
//  - lane l holds column xc = x0 - 4 + l; lanes 4..59 own columns
//    x0 .. x0+55 (56 per strip), lanes 0-3 and 60-63 recompute the
//    neighbours' stresses that the owned velocities need;
//  - the work item covers velocity rows [zv0, zv0+H) of one shot; in blocks
//    of R rows it computes the stresses of rows rs0..rs0+R-1 and then the
//    velocities of the rows 4 above, which need new stresses down to rs0+R-1;
//  - z-derivatives come from per-lane register queues of rows, x-derivatives
//    from LDS rows;
//  - all loads of the next block are issued before the current block is
//    computed, so each wavefront keeps R rows of every array in flight.
// Every cell runs the same per-cell code as the two-pass kernels
// (psv_stress_* / psv_velocity_*), so results are bitwise equal to them.
#pragma once
#include "fd_kernels.hpp"

namespace fd23 {

constexpr int FW = 64;             // lanes = stress columns of a strip
constexpr int FO = FW - 2 * HALO;  // owned columns of a strip

struct FusedWork {
    int nstrip, nseg, H, nwork;    // work item = (segment, strip, shot), shot fastest
};
struct TapF {                      // a source tap seen by one strip
    int row, lane, field, mode;
    double m, q;
};
struct RecF {                      // a receiver inside one (segment, strip)
    int row, lane, comp, out;
};
struct FusedEvents {
    const int *clock;              // step = clock[0] + k
    int k, itrecord;
    const int *tap_start;          // CSR over strip*nshots + shot
    const TapF *taps;
    const SrcInfo *src;
    const double *stf;
    const int *rec_start;          // CSR over (seg*nstrip + strip)*nshots + shot
    const RecF *recs;              // sorted by row within a list
    void *seis;
    int max_frames;
    long frame_size;
    unsigned long long idle;       // idle shot slots 0..63, not stepped
};
template <class C> struct PsvMat5 {
    const C *BU, *BW, *MU, *MUA, *LAM; // pitch px
};

// 8-tap differences on register queues, the same operation order as dm/dp
template <class C>
__device__ __forceinline__ C dm8(C fm4, C fm3, C fm2, C fm1, C f0, C f1, C f2, C f3, C c0, C c1, C c2, C c3) {
    C r = c0 * (f0 - fm1);
    r = mad(-c1, f1 - fm2, r);
    r = mad(c2, f2 - fm3, r);
    return mad(-c3, f3 - fm4, r);
}
template <class C>
__device__ __forceinline__ C dp8(C fm3, C fm2, C fm1, C f0, C f1, C f2, C f3, C f4, C c0, C c1, C c2, C c3) {
    C r = c0 * (f1 - f0);
    r = mad(-c1, f2 - fm1, r);
    r = mad(c2, f3 - fm2, r);
    return mad(-c3, f4 - fm3, r);
}

template <class S, int H, int R>
__global__ __launch_bounds__(FW) void k_psv_fused(PsvFields<S> fa, PsvFields<S> fb, PsvMat5<typename Store<S>::C> m,
                                                  PsvStressPsi<S> spa, PsvStressPsi<S> spb, PsvVelPsi<S> vpa,
                                                  PsvVelPsi<S> vpb, PsvStressProf<typename Store<S>::C> spr,
                                                  PsvVelProf<typename Store<S>::C> vpr, Grid g, FusedWork w,
                                                  FusedEvents ev, const float *__restrict__ scl,
                                                  unsigned *__restrict__ amax) {
    using C = typename Store<S>::C;
    using ST = Store<S>;
    static_assert(H % R == 0, "segment height must be a multiple of the row block");
    constexpr int LWF = FW + 2 * HALO;
    __shared__ C lU[R][LWF], lW[R][LWF], lXX[R][FW], lXZ[R][FW], pU[FW], pW[FW];

    const int qd = gridDim.x >> 3; // grid padded to a multiple of 8
    const int L = (int)(blockIdx.x & 7) * qd + (int)(blockIdx.x >> 3);
    if (L >= w.nwork) return;
    const int S_ = g.nshots;
    const int shot = L % S_;
    if (shot_idle(ev.idle, shot)) return; // uniform per workgroup
    const int rest = L / S_;
    const int strip = rest % w.nstrip;
    const int seg = rest / w.nstrip;
    const int lane = threadIdx.x;
    const int nx = g.nx, nz = g.nz, px = g.px;
    const int x0 = strip * FO;
    const int xc = x0 - HALO + lane;
    const bool xin = xc >= 0 && xc < nx;
    const bool own_x = lane >= HALO && lane < FW - HALO && xin;
    const int zv0 = seg * H, zv1 = min(zv0 + H, nz);
    const long base = (long)shot * g.sstride;
    const int lo = max(g.npml, 4);
    const int ztop = g.top ? lo : 4;
    // strip columns x0-4 .. x0+59 all at full order with no C-PML
    const bool xint = x0 - HALO >= lo && x0 + FO + HALO <= nx - lo;
    const int it = ev.clock[0] + ev.k;
    double sv = 0.0;
    const SrcInfo si = ev.src[shot];
    const bool src_on = src_sample(si, ev.stf, it, sv);
    const int t0 = ev.tap_start[strip * S_ + shot], t1 = ev.tap_start[strip * S_ + shot + 1];
    const int rkey = (seg * w.nstrip + strip) * S_ + shot;
    int r0 = ev.rec_start[rkey];
    const int r1 = ev.rec_start[rkey + 1];
    const int ts = it - si.t0; // step of the shot: frames count from its first step
    const bool rec_step = ts >= 0 && (ts % ev.itrecord) == 0 && ts / ev.itrecord < ev.max_frames;

    // scales: fields U W Txx Tzz Txz read (current) and written (next)
    float sc_u = 1.f, sc_w = 1.f, sc_xx = 1.f, sc_zz = 1.f, sc_xz = 1.f;
    float wi_u = 1.f, wi_w = 1.f, wi_xx = 1.f, wi_zz = 1.f, wi_xz = 1.f;
    PsiScl4 ssc{{1.f, 1.f, 1.f, 1.f}, {1.f, 1.f, 1.f, 1.f}}, vsc{{1.f, 1.f, 1.f, 1.f}, {1.f, 1.f, 1.f, 1.f}};
    if (ST::scaled) {
        sc_u = scl[SC_U]; sc_w = scl[SC_W]; sc_xx = scl[SC_TXX]; sc_zz = scl[SC_TZZ]; sc_xz = scl[SC_TXZ];
        wi_u = inv_pow2(scl[kNextScale + SC_U]); wi_w = inv_pow2(scl[kNextScale + SC_W]);
        wi_xx = inv_pow2(scl[kNextScale + SC_TXX]); wi_zz = inv_pow2(scl[kNextScale + SC_TZZ]);
        wi_xz = inv_pow2(scl[kNextScale + SC_TXZ]);
        const int so[4] = {SC_PU_X, SC_PW_X, SC_PU_Z, SC_PW_Z}, vo[4] = {SC_PTXX_X, SC_PTXZ_X, SC_PTXZ_Z, SC_PTZZ_Z};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            ssc.r[i] = scl[so[i]]; ssc.w[i] = inv_pow2(scl[kNextScale + so[i]]);
            vsc.r[i] = scl[vo[i]]; vsc.w[i] = inv_pow2(scl[kNextScale + vo[i]]);
        }
    }
    float mx[5] = {0.f, 0.f, 0.f, 0.f, 0.f}; // U W Txx Tzz Txz
    float smp[4] = {0.f, 0.f, 0.f, 0.f}, vmp[4] = {0.f, 0.f, 0.f, 0.f};

    // Loads are unconditional, so that all loads of a block are in flight at
    // once: a row outside the domain is clamped to the nearest zero pad row
    // (fields and material both have HALO zero rows above row 0 and more
    // below row nz-1), so it reads zero with no test.  Columns left of the
    // domain read zero pads; columns right of it may read the next row, but
    // only lanes outside the domain use them and their results are dropped.
    auto ldrow = [&](const S *f, int row, int col, float s) -> C {
        const int rc = min(max(row, -HALO), nz + HALO);
        return ST::ld(f, base + (long)rc * px + col, s);
    };
    auto ldmat = [&](const C *f, int row) -> C {
        const int rc = min(max(row, -HALO), nz + HALO);
        return f[(long)rc * px + xc];
    };
    // the 8 extra LDS columns: lanes 0-3 -> x0-8..x0-5, lanes 4-7 -> x0+60..x0+63
    // (the other lanes load their own column and drop it)
    const int xcol = lane < 4 ? x0 - 2 * HALO + lane : lane < 8 ? x0 + FO + HALO + (lane - 4) : xc;

    // Row block b: stresses of rows rs0 .. rs0+R-1 (rs0 = zv0-4+R*b) and
    // velocities of rows rs0-4 .. rs0+R-5.  Queues, lowest row first:
    //   qU  old U    rows rs0-4 .. rs0+R+2     qW  old W    rows rs0-4 .. rs0+R+3
    //   qZZ new Tzz  rows rs0-8 .. rs0+R-1     qXZ new Txz  rows rs0-8 .. rs0+R-1
    //   qXX new Txx  rows rs0-4 .. rs0+R-1
    C qU[R + 7], qW[R + 8], qZZ[R + 8], qXZ[R + 8], qXX[R + 4];
    int rs0 = zv0 - HALO;
#pragma unroll
    for (int i = 0; i < R + 7; i++) qU[i] = ldrow(fa.U, rs0 - 4 + i, xc, sc_u);
#pragma unroll
    for (int i = 0; i < R + 8; i++) qW[i] = ldrow(fa.W, rs0 - 4 + i, xc, sc_w);
#pragma unroll
    for (int i = 0; i < R + 8; i++) qZZ[i] = qXZ[i] = C(0);
#pragma unroll
    for (int i = 0; i < R + 4; i++) qXX[i] = C(0);

    // operands of the current block
    C hU[R], hW[R], oxx[R], ozz[R], oxz[R], lam[R], mu[R], mua[R], bu[R], bw[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
        const int rs = rs0 + k, rv = rs - HALO;
        hU[k] = ldrow(fa.U, rs, xcol, sc_u);
        hW[k] = ldrow(fa.W, rs, xcol, sc_w);
        oxx[k] = ldrow(fa.Txx, rs, xc, sc_xx);
        ozz[k] = ldrow(fa.Tzz, rs, xc, sc_zz);
        oxz[k] = ldrow(fa.Txz, rs, xc, sc_xz);
        lam[k] = ldmat(m.LAM, rs);
        mu[k] = ldmat(m.MU, rs);
        mua[k] = ldmat(m.MUA, rs);
        bu[k] = ldmat(m.BU, rv);
        bw[k] = ldmat(m.BW, rv);
    }

    constexpr int NB = (H + 2 * HALO) / R;
#pragma unroll 1
    for (int b = 0; b < NB; b++, rs0 += R) {
        const int rv0 = rs0 - HALO;
        // ---- prefetch the next block's operands and queue rows
        C nU[R], nW[R], hU_n[R], hW_n[R], oxx_n[R], ozz_n[R], oxz_n[R], lam_n[R], mu_n[R], mua_n[R], bu_n[R], bw_n[R];
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int rs = rs0 + R + k, rv = rs - HALO;
            nU[k] = ldrow(fa.U, rs + 3, xc, sc_u);
            nW[k] = ldrow(fa.W, rs + 4, xc, sc_w);
            hU_n[k] = ldrow(fa.U, rs, xcol, sc_u);
            hW_n[k] = ldrow(fa.W, rs, xcol, sc_w);
            oxx_n[k] = ldrow(fa.Txx, rs, xc, sc_xx);
            ozz_n[k] = ldrow(fa.Tzz, rs, xc, sc_zz);
            oxz_n[k] = ldrow(fa.Txz, rs, xc, sc_xz);
            lam_n[k] = ldmat(m.LAM, rs);
            mu_n[k] = ldmat(m.MU, rs);
            mua_n[k] = ldmat(m.MUA, rs);
            bu_n[k] = ldmat(m.BU, rv);
            bw_n[k] = ldmat(m.BW, rv);
        }

        // ---- stage the U and W rows rs0..rs0+R-1 for the x-derivatives
#pragma unroll
        for (int k = 0; k < R; k++) {
            lU[k][lane + HALO] = qU[k + 4];
            lW[k][lane + HALO] = qW[k + 4];
            if (lane < 8) {
                const int o = lane < 4 ? lane : FW + lane;
                lU[k][o] = hU[k];
                lW[k][o] = hW[k];
            }
        }
        __syncthreads();

        // ---- stresses of rows rs0 .. rs0+R-1
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int rs = rs0 + k;
            C a = oxx[k], bb = ozz[k], cc = oxz[k];
            if (rs >= 0 && rs < nz && xin) {
                if (xint && rs >= ztop && rs < nz - lo) {
                    const C c0 = C(1225. / 1024.), c1 = C(245. / 3072.), c2 = C(49. / 5120.), c3 = C(5. / 7168.);
                    const C ux = dm(&lU[k][lane + HALO], 1, c0, c1, c2, c3);
                    const C wx = dp(&lW[k][lane + HALO], 1, c0, c1, c2, c3);
                    const C uz = dm8(qU[k], qU[k + 1], qU[k + 2], qU[k + 3], qU[k + 4], qU[k + 5], qU[k + 6], qU[k + 7],
                                     c0, c1, c2, c3);
                    const C wz = dp8(qW[k + 1], qW[k + 2], qW[k + 3], qW[k + 4], qW[k + 5], qW[k + 6], qW[k + 7],
                                     qW[k + 8], c0, c1, c2, c3);
                    psv_stress_interior(ux, uz, wx, wz, lam[k], mu[k], mua[k], a, bb, cc);
                } else {
                    C c0, c1, c2, c3;
                    Coef<C>::get(order_at(xc, rs, nx, nz), c0, c1, c2, c3);
                    const C ux = dm(&lU[k][lane + HALO], 1, c0, c1, c2, c3);
                    const C wx = dp(&lW[k][lane + HALO], 1, c0, c1, c2, c3);
                    const C uz = dm8(qU[k], qU[k + 1], qU[k + 2], qU[k + 3], qU[k + 4], qU[k + 5], qU[k + 6], qU[k + 7],
                                     c0, c1, c2, c3);
                    const C wz = dp8(qW[k + 1], qW[k + 2], qW[k + 3], qW[k + 4], qW[k + 5], qW[k + 6], qW[k + 7],
                                     qW[k + 8], c0, c1, c2, c3);
                    const bool own = own_x && rs >= zv0 && rs < zv1;
                    psv_stress_exact<S, C>(g, xc, rs, shot, ux, uz, wx, wz, lam[k], mu[k], mua[k], a, bb, cc, spa, spb,
                                           own, spr, ssc, smp);
                }
                // source taps after the cell's update, in the reference's order
                if (src_on) {
                    for (int e = t0; e < t1; e++) {
                        const TapF tp = ev.taps[e];
                        if (tp.row != rs || tp.lane != lane) continue;
                        if (tp.field == 2) a = apply_tap(a, tp, sv);
                        else if (tp.field == 3) bb = apply_tap(bb, tp, sv);
                        else cc = apply_tap(cc, tp, sv);
                    }
                }
                if (own_x && rs >= zv0 && rs < zv1) {
                    const long q = base + (long)rs * px + xc;
                    ST::st(fb.Txx, q, a, wi_xx);
                    ST::st(fb.Tzz, q, bb, wi_zz);
                    ST::st(fb.Txz, q, cc, wi_xz);
                    if (ST::scaled) {
                        mx[2] = fmaxf(mx[2], fabsf((float)a));
                        mx[3] = fmaxf(mx[3], fabsf((float)bb));
                        mx[4] = fmaxf(mx[4], fabsf((float)cc));
                    }
                }
            } else {
                a = bb = cc = C(0); // outside the domain the new stresses are zero
            }
            qXX[4 + k] = a;
            qZZ[8 + k] = bb;
            qXZ[8 + k] = cc;
        }

        // ---- stage the new Txx and Txz rows rv0..rv0+R-1 for the x-derivatives
#pragma unroll
        for (int k = 0; k < R; k++) {
            lXX[k][lane] = qXX[k];
            lXZ[k][lane] = qXZ[k + 4];
        }
        __syncthreads();

        // ---- velocities of rows rv0 .. rv0+R-1
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int rv = rv0 + k;
            if (rv < zv0 || rv >= zv1) continue;
            C u = qU[k], ww = qW[k];
            if (own_x) {
                if (xint && rv >= ztop && rv < nz - lo) {
                    const C c0 = C(1225. / 1024.), c1 = C(245. / 3072.), c2 = C(49. / 5120.), c3 = C(5. / 7168.);
                    const C dxTxx = dp(&lXX[k][lane], 1, c0, c1, c2, c3);
                    const C dxTxz = dm(&lXZ[k][lane], 1, c0, c1, c2, c3);
                    const C dzTzz = dm8(qZZ[k], qZZ[k + 1], qZZ[k + 2], qZZ[k + 3], qZZ[k + 4], qZZ[k + 5], qZZ[k + 6],
                                        qZZ[k + 7], c0, c1, c2, c3);
                    const C dzTxz = dp8(qXZ[k + 1], qXZ[k + 2], qXZ[k + 3], qXZ[k + 4], qXZ[k + 5], qXZ[k + 6],
                                        qXZ[k + 7], qXZ[k + 8], c0, c1, c2, c3);
                    psv_velocity_interior(dxTxx, dzTzz, dxTxz, dzTxz, bu[k], bw[k], u, ww);
                } else {
                    C c0, c1, c2, c3;
                    Coef<C>::get(order_at(xc, rv, nx, nz), c0, c1, c2, c3);
                    const C dxTxx = dp(&lXX[k][lane], 1, c0, c1, c2, c3);
                    const C dxTxz = dm(&lXZ[k][lane], 1, c0, c1, c2, c3);
                    const C dzTzz = dm8(qZZ[k], qZZ[k + 1], qZZ[k + 2], qZZ[k + 3], qZZ[k + 4], qZZ[k + 5], qZZ[k + 6],
                                        qZZ[k + 7], c0, c1, c2, c3);
                    const C dzTxz = dp8(qXZ[k + 1], qXZ[k + 2], qXZ[k + 3], qXZ[k + 4], qXZ[k + 5], qXZ[k + 6],
                                        qXZ[k + 7], qXZ[k + 8], c0, c1, c2, c3);
                    psv_velocity_exact<S, C>(g, xc, rv, shot, dxTxx, dzTzz, dxTxz, dzTxz, bu[k], bw[k], u, ww, vpa,
                                             vpb, true, vpr, vsc, vmp);
                }
                const long q = base + (long)rv * px + xc;
                ST::st(fb.U, q, u, wi_u);
                ST::st(fb.W, q, ww, wi_w);
                if (ST::scaled) {
                    mx[0] = fmaxf(mx[0], fabsf((float)u));
                    mx[1] = fmaxf(mx[1], fabsf((float)ww));
                }
            }
            // receivers of row rv: park the row in LDS, then gather
            if (rec_step && r0 < r1 && ev.recs[r0].row == rv) {
                pU[lane] = u;
                pW[lane] = ww;
                __syncthreads();
                int r = r0;
                while (r < r1 && ev.recs[r].row == rv) r++;
                C *seis = (C *)ev.seis + (long)(ts / ev.itrecord) * ev.frame_size;
                for (int e = r0 + lane; e < r; e += FW) {
                    const RecF rc = ev.recs[e];
                    seis[rc.out] = rc.comp == 0 ? pU[rc.lane] : pW[rc.lane];
                }
                r0 = r;
                __syncthreads();
            }
        }

        // ---- advance the queues by R rows and take the prefetched operands
#pragma unroll
        for (int i = 0; i < 7; i++) qU[i] = qU[i + R];
#pragma unroll
        for (int i = 0; i < 8; i++) qW[i] = qW[i + R];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            qZZ[i] = qZZ[i + R];
            qXZ[i] = qXZ[i + R];
        }
#pragma unroll
        for (int i = 0; i < 4; i++) qXX[i] = qXX[i + R];
#pragma unroll
        for (int k = 0; k < R; k++) {
            qU[7 + k] = nU[k];
            qW[8 + k] = nW[k];
            hU[k] = hU_n[k]; hW[k] = hW_n[k];
            oxx[k] = oxx_n[k]; ozz[k] = ozz_n[k]; oxz[k] = oxz_n[k];
            lam[k] = lam_n[k]; mu[k] = mu_n[k]; mua[k] = mua_n[k];
            bu[k] = bu_n[k]; bw[k] = bw_n[k];
        }
    }
    if (ST::scaled) {
        amax_update(amax, SC_U, mx[0]);
        amax_update(amax, SC_W, mx[1]);
        amax_update(amax, SC_TXX, mx[2]);
        amax_update(amax, SC_TZZ, mx[3]);
        amax_update(amax, SC_TXZ, mx[4]);
        amax_update(amax, SC_PU_X, smp[0]);
        amax_update(amax, SC_PW_X, smp[1]);
        amax_update(amax, SC_PW_Z, smp[2]);
        amax_update(amax, SC_PU_Z, smp[3]);
        amax_update(amax, SC_PTXX_X, vmp[0]);
        amax_update(amax, SC_PTXZ_X, vmp[1]);
        amax_update(amax, SC_PTXZ_Z, vmp[2]);
        amax_update(amax, SC_PTZZ_Z, vmp[3]);
    }
}


// ---------------------------------------------------------------------------
// Tile-parallel fused P-SV step (FP32/FP64 storage).  The same strips of
// 64 stress columns (56 owned) and segments of OZ = 4*CZ rows as
// k_psv_fused, but a workgroup of four wavefronts takes one (segment,
// strip, shot) as a 2D tile with every global load issued before its first
// barrier, as the two-pass kernels do, instead of marching one wavefront
// down the segment:
//  - U and W of rows z0-8 .. z0+OZ+7 and columns x0-8 .. x0+63 are staged
//    in LDS;
//  - wavefront ty computes the new stresses of rows z0-4+ty*(CZ+2) ..
//    (CZ+2 rows, 64 lanes: the owned cells and the 4-cell ring the owned
//    velocities read), stores the owned ones, and puts all of them in the
//    same LDS space once every wavefront has read U and W;
//  - it then computes the velocities of rows z0+ty*CZ .. (CZ rows, lanes
//    4-59) from LDS.
// Per-cell arithmetic is k_psv_fused's (psv_stress_* / psv_velocity_*), so
// results are bitwise equal to the two-pass kernels.  Tiles (strips x
// segments, one shot group per shot) are mapped by map_tile as the
// two-pass kernels': REGION_INTERIOR compiles the interior path only, for
// the rectangle of tiles whose stress cells are all interior, REGION_EDGE
// the exact path for the frame around it.
// ---------------------------------------------------------------------------
template <class S, int CZ, int REGION>
__global__ __launch_bounds__(FW * 4) void k_psv_fused_tile(PsvFields<S> fa, PsvFields<S> fb,
                                                           PsvMat5<typename Store<S>::C> m, PsvStressPsi<S> spa,
                                                           PsvStressPsi<S> spb, PsvVelPsi<S> vpa, PsvVelPsi<S> vpb,
                                                           PsvStressProf<typename Store<S>::C> spr,
                                                           PsvVelProf<typename Store<S>::C> vpr, Grid g, Work w,
                                                           FusedEvents ev) {
    using C = typename Store<S>::C;
    using ST = Store<S>;
    static_assert(!ST::scaled, "the tile-parallel fused step keeps FP32 or FP64 storage");
    constexpr int NW = 4;
    constexpr int OZ = NW * CZ;             // owned rows
    constexpr int SR = CZ + 2;              // stress rows per wavefront
    constexpr int LWU = FW + 2 * HALO;      // staged U/W columns
    constexpr int NU = (OZ + 4 * HALO) * LWU;
    constexpr int NT = (OZ + 2 * HALO) * FW;
    constexpr int NL = 2 * NU > 3 * NT ? 2 * NU : 3 * NT;
    __shared__ C lds[NL];
    C *sU = lds, *sW = lds + NU;                            // rows z0-8 .., columns x0-8 ..
    C *sXX = lds, *sZZ = lds + NT, *sXZ = lds + 2 * NT;     // rows z0-4 .., lanes

    TileId t;
    if (!map_tile<OZ, REGION>(w, t)) return;
    const int S_ = g.nshots;
    const int shot = t.group;
    if (shot_idle(ev.idle, shot)) return; // uniform per workgroup
    const int strip = t.tile % w.ntx;
    const int seg = t.tile / w.ntx;
    const int lane = threadIdx.x, ty = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int tid = ty * FW + lane;
    const int nx = g.nx, nz = g.nz, px = g.px;
    const int x0 = strip * FO;
    const int xc = x0 - HALO + lane;
    const bool xin = xc >= 0 && xc < nx;
    const bool own_x = lane >= HALO && lane < FW - HALO && xin;
    const int z0 = seg * OZ;
    const long base = (long)shot * g.sstride;
    const int lo = max(g.npml, 4);
    const int ztop = g.top ? lo : 4;
    const bool xint = x0 - HALO >= lo && x0 + FO + HALO <= nx - lo;

    // rows outside the domain read the zero pad rows (clamped), columns left
    // of it the pad columns; columns right of it may read the next row, used
    // only by lanes outside the domain, whose results are dropped
    auto rowc = [&](int row) { return min(max(row, -HALO), nz + HALO); };
    const int xcol = lane < 4 ? x0 - 2 * HALO + lane : x0 + FO + HALO + (lane - 4); // lanes 0-7: the 8 extra columns

    // ---- every global load of the tile
    constexpr int UR = (OZ + 4 * HALO) / NW; // staged rows per wavefront
    C u_[UR], w_[UR], ue[UR], we[UR];
#pragma unroll
    for (int i = 0; i < UR; i++) {
        const long r = base + (long)rowc(z0 - 2 * HALO + ty * UR + i) * px;
        u_[i] = ST::ld(fa.U, r + xc, 1.f);
        w_[i] = ST::ld(fa.W, r + xc, 1.f);
        ue[i] = lane < 8 ? ST::ld(fa.U, r + xcol, 1.f) : C(0);
        we[i] = lane < 8 ? ST::ld(fa.W, r + xcol, 1.f) : C(0);
    }
    C oxx[SR], ozz[SR], oxz[SR], lam[SR], mu[SR], mua[SR];
#pragma unroll
    for (int i = 0; i < SR; i++) {
        const int rs = z0 - HALO + ty * SR + i;
        const long r = (long)rowc(rs) * px + xc;
        oxx[i] = ST::ld(fa.Txx, base + r, 1.f);
        ozz[i] = ST::ld(fa.Tzz, base + r, 1.f);
        oxz[i] = ST::ld(fa.Txz, base + r, 1.f);
        lam[i] = m.LAM[r];
        mu[i] = m.MU[r];
        mua[i] = m.MUA[r];
    }
    C bu[CZ], bw[CZ];
#pragma unroll
    for (int i = 0; i < CZ; i++) {
        const long r = (long)rowc(z0 + ty * CZ + i) * px + xc;
        bu[i] = m.BU[r];
        bw[i] = m.BW[r];
    }
    const int it = ev.clock[0] + ev.k;
    double sv = 0.0;
    const SrcInfo si = ev.src[shot];
    const bool src_on = src_sample(si, ev.stf, it, sv);
    const int t0 = ev.tap_start[strip * S_ + shot], t1 = ev.tap_start[strip * S_ + shot + 1];
    const int rkey = (seg * w.ntx + strip) * S_ + shot;
    const int r0 = ev.rec_start[rkey], r1 = ev.rec_start[rkey + 1];
    const int ts = it - si.t0; // step of the shot: frames count from its first step
    const bool rec_step = ts >= 0 && (ts % ev.itrecord) == 0 && ts / ev.itrecord < ev.max_frames;

#pragma unroll
    for (int i = 0; i < UR; i++) {
        const int row = (ty * UR + i) * LWU;
        sU[row + lane + HALO] = u_[i];
        sW[row + lane + HALO] = w_[i];
        if (lane < 8) {
            const int o = lane < 4 ? lane : FW + lane;
            sU[row + o] = ue[i];
            sW[row + o] = we[i];
        }
    }
    __syncthreads();

    const PsiScl4 one{{1.f, 1.f, 1.f, 1.f}, {1.f, 1.f, 1.f, 1.f}};
    float mp[4] = {0.f, 0.f, 0.f, 0.f};
    const C k0 = C(1225. / 1024.), k1 = C(245. / 3072.), k2 = C(49. / 5120.), k3 = C(5. / 7168.);

    // ---- new stresses of rows z0-4+ty*SR .. (all 64 lanes)
#pragma unroll
    for (int i = 0; i < SR; i++) {
        const int rs = z0 - HALO + ty * SR + i;
        const int ru = rs - z0 + 2 * HALO; // staged row
        const C *pu = &sU[ru * LWU + lane + HALO], *pw = &sW[ru * LWU + lane + HALO];
        C a = oxx[i], bb = ozz[i], cc = oxz[i];
        if (REGION == REGION_INTERIOR || (rs >= 0 && rs < nz && xin)) {
            if (REGION == REGION_INTERIOR || (REGION == REGION_ALL && xint && rs >= ztop && rs < nz - lo)) {
                const C ux = dm(pu, 1, k0, k1, k2, k3), wx = dp(pw, 1, k0, k1, k2, k3);
                const C uz = dm(pu, LWU, k0, k1, k2, k3), wz = dp(pw, LWU, k0, k1, k2, k3);
                psv_stress_interior(ux, uz, wx, wz, lam[i], mu[i], mua[i], a, bb, cc);
            } else {
                C c0, c1, c2, c3;
                Coef<C>::get(order_at(xc, rs, nx, nz), c0, c1, c2, c3);
                const C ux = dm(pu, 1, c0, c1, c2, c3), wx = dp(pw, 1, c0, c1, c2, c3);
                const C uz = dm(pu, LWU, c0, c1, c2, c3), wz = dp(pw, LWU, c0, c1, c2, c3);
                const bool own = own_x && rs >= z0 && rs < z0 + OZ;
                psv_stress_exact<S, C>(g, xc, rs, shot, ux, uz, wx, wz, lam[i], mu[i], mua[i], a, bb, cc, spa, spb,
                                       own, spr, one, mp);
            }
            // source taps after the cell's update, in the reference's order
            if (src_on) {
                for (int e = t0; e < t1; e++) {
                    const TapF tp = ev.taps[e];
                    if (tp.row != rs || tp.lane != lane) continue;
                    if (tp.field == 2) a = apply_tap(a, tp, sv);
                    else if (tp.field == 3) bb = apply_tap(bb, tp, sv);
                    else cc = apply_tap(cc, tp, sv);
                }
            }
            if (own_x && rs >= z0 && rs < z0 + OZ) {
                const long q = base + (long)rs * px + xc;
                ST::st(fb.Txx, q, a, 1.f);
                ST::st(fb.Tzz, q, bb, 1.f);
                ST::st(fb.Txz, q, cc, 1.f);
            }
        } else {
            a = bb = cc = C(0); // outside the domain the new stresses are zero
        }
        oxx[i] = a;
        ozz[i] = bb;
        oxz[i] = cc;
    }
    // the owned cells' U and W, read before the LDS space takes the stresses
    C u[CZ], ww[CZ];
#pragma unroll
    for (int i = 0; i < CZ; i++) {
        const int ru = ty * CZ + i + 2 * HALO;
        u[i] = sU[ru * LWU + lane + HALO];
        ww[i] = sW[ru * LWU + lane + HALO];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SR; i++) {
        const int rt = (ty * SR + i) * FW + lane;
        sXX[rt] = oxx[i];
        sZZ[rt] = ozz[i];
        sXZ[rt] = oxz[i];
    }
    __syncthreads();

    // ---- velocities of rows z0+ty*CZ .. (lanes 4-59)
#pragma unroll
    for (int i = 0; i < CZ; i++) {
        const int rv = z0 + ty * CZ + i;
        if (!own_x || rv >= nz) continue;
        const int o = (ty * CZ + i + HALO) * FW + lane;
        if (REGION == REGION_INTERIOR || (REGION == REGION_ALL && xint && rv >= ztop && rv < nz - lo)) {
            const C dxTxx = dp(&sXX[o], 1, k0, k1, k2, k3), dxTxz = dm(&sXZ[o], 1, k0, k1, k2, k3);
            const C dzTzz = dm(&sZZ[o], FW, k0, k1, k2, k3), dzTxz = dp(&sXZ[o], FW, k0, k1, k2, k3);
            psv_velocity_interior(dxTxx, dzTzz, dxTxz, dzTxz, bu[i], bw[i], u[i], ww[i]);
        } else {
            C c0, c1, c2, c3;
            Coef<C>::get(order_at(xc, rv, nx, nz), c0, c1, c2, c3);
            const C dxTxx = dp(&sXX[o], 1, c0, c1, c2, c3), dxTxz = dm(&sXZ[o], 1, c0, c1, c2, c3);
            const C dzTzz = dm(&sZZ[o], FW, c0, c1, c2, c3), dzTxz = dp(&sXZ[o], FW, c0, c1, c2, c3);
            psv_velocity_exact<S, C>(g, xc, rv, shot, dxTxx, dzTzz, dxTxz, dzTxz, bu[i], bw[i], u[i], ww[i], vpa,
                                     vpb, true, vpr, one, mp);
        }
        const long q = base + (long)rv * px + xc;
        ST::st(fb.U, q, u[i], 1.f);
        ST::st(fb.W, q, ww[i], 1.f);
    }

    // ---- receivers of the tile: park U and W in LDS, one entry per thread
    if (rec_step && r1 > r0) {
        __syncthreads();
        C *pU = lds, *pW = lds + OZ * FW;
#pragma unroll
        for (int i = 0; i < CZ; i++) {
            pU[(ty * CZ + i) * FW + lane] = u[i];
            pW[(ty * CZ + i) * FW + lane] = ww[i];
        }
        __syncthreads();
        C *seis = (C *)ev.seis + (long)(ts / ev.itrecord) * ev.frame_size;
        for (int e = r0 + tid; e < r1; e += FW * NW) {
            const RecF rc = ev.recs[e];
            const int c = (rc.row - z0) * FW + rc.lane;
            seis[rc.out] = rc.comp == 0 ? pU[c] : pW[c];
        }
    }
}

} // namespace fd23
