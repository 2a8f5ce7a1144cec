// Fused P-SV step: the stress and the velocity update of one time step in a
// single pass, reading the fields of step it from one buffer set and
// writing those of step it+1 to the other (ping-pong), so no ordering
// between workgroups is needed.
//
// One wavefront per workgroup marches down a strip of the grid:
//  - lane l holds column xc = x0 - 4 + l; lanes 4..59 own columns
//    x0 .. x0+55 (56 per strip), lanes 0-3 and 60-63 recompute the
//    neighbours' stresses that the owned velocities need;
//  - the work item covers velocity rows [zv0, zv0+H) of one shot; at row
//    step j it computes the stresses of row rs = zv0-4+j and the velocities
//    of row rv = rs-4, which need new stresses down to rs;
//  - z-derivatives come from per-lane register queues (old U rows rs-4..rs+3,
//    old W rows rs-4..rs+4, new Tzz rows rs-8..rs, new Txz rows rs-7..rs,
//    new Txx rows rs-4..rs), x-derivatives from one LDS row per field;
//  - the next row's loads are issued before the current row is computed.
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

template <class S, int H>
__global__ __launch_bounds__(FW) void k_psv_fused(PsvFields<S> fa, PsvFields<S> fb, PsvMat5<typename Store<S>::C> m,
                                                  PsvStressPsi<S> spa, PsvStressPsi<S> spb, PsvVelPsi<S> vpa,
                                                  PsvVelPsi<S> vpb, PsvStressProf<typename Store<S>::C> spr,
                                                  PsvVelProf<typename Store<S>::C> vpr, Grid g, FusedWork w,
                                                  FusedEvents ev, const float *__restrict__ scl,
                                                  unsigned *__restrict__ amax) {
    using C = typename Store<S>::C;
    using ST = Store<S>;
    constexpr int LWF = FW + 2 * HALO;
    __shared__ C lU[LWF], lW[LWF], lXX[FW], lXZ[FW], pU[FW], pW[FW];

    const int qd = gridDim.x >> 3; // grid padded to a multiple of 8
    const int L = (int)(blockIdx.x & 7) * qd + (int)(blockIdx.x >> 3);
    if (L >= w.nwork) return;
    const int S_ = g.nshots;
    const int shot = L % S_;
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
    // strip columns x0-4 .. x0+59 all at full order with no C-PML
    const bool xint = x0 - HALO >= lo && x0 + FO + HALO <= nx - lo;
    const int it = ev.clock[0] + ev.k;
    double sv = 0.0;
    const bool src_on = src_sample(ev.src[shot], ev.stf, it, sv);
    const int t0 = ev.tap_start[strip * S_ + shot], t1 = ev.tap_start[strip * S_ + shot + 1];
    const int rkey = (seg * w.nstrip + strip) * S_ + shot;
    int r0 = ev.rec_start[rkey];
    const int r1 = ev.rec_start[rkey + 1];
    const bool rec_step = (it % ev.itrecord) == 0 && it / ev.itrecord < ev.max_frames;

    // scales: fields U W Txx Tzz Txz read (current) and written (next)
    float sc_u = 1.f, sc_w = 1.f, sc_xx = 1.f, sc_zz = 1.f, sc_xz = 1.f;
    float wi_u = 1.f, wi_w = 1.f, wi_xx = 1.f, wi_zz = 1.f, wi_xz = 1.f;
    PsiScl4 ssc{{1.f, 1.f, 1.f, 1.f}, {1.f, 1.f, 1.f, 1.f}}, vsc{{1.f, 1.f, 1.f, 1.f}, {1.f, 1.f, 1.f, 1.f}};
    if (ST::scaled) {
        sc_u = scl[SC_U]; sc_w = scl[SC_W]; sc_xx = scl[SC_TXX]; sc_zz = scl[SC_TZZ]; sc_xz = scl[SC_TXZ];
        wi_u = 1.f / scl[kNextScale + SC_U]; wi_w = 1.f / scl[kNextScale + SC_W];
        wi_xx = 1.f / scl[kNextScale + SC_TXX]; wi_zz = 1.f / scl[kNextScale + SC_TZZ];
        wi_xz = 1.f / scl[kNextScale + SC_TXZ];
        const int so[4] = {SC_PU_X, SC_PW_X, SC_PU_Z, SC_PW_Z}, vo[4] = {SC_PTXX_X, SC_PTXZ_X, SC_PTXZ_Z, SC_PTZZ_Z};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            ssc.r[i] = scl[so[i]]; ssc.w[i] = 1.f / scl[kNextScale + so[i]];
            vsc.r[i] = scl[vo[i]]; vsc.w[i] = 1.f / scl[kNextScale + vo[i]];
        }
    }
    float mx[5] = {0.f, 0.f, 0.f, 0.f, 0.f}; // U W Txx Tzz Txz
    float smp[4] = {0.f, 0.f, 0.f, 0.f}, vmp[4] = {0.f, 0.f, 0.f, 0.f};

    // row loads (zero outside the domain rows; columns outside are pads)
    auto ldrow = [&](const S *f, int row, int col, float s) -> C {
        return (row >= 0 && row < nz) ? ST::ld(f, base + (long)row * px + col, s) : C(0);
    };
    auto ldmat = [&](const C *f, int row) -> C {
        return (row >= 0 && row < nz && xin) ? f[(long)row * px + xc] : C(0);
    };
    // the 8 extra LDS columns: lanes 0-3 -> x0-8..x0-5, lanes 4-7 -> x0+60..x0+63
    const int xcol = lane < 4 ? x0 - 2 * HALO + lane : x0 + FO + HALO + (lane - 4);

    C qU[8], qW[9], qZZ[9], qXZ[8], qXX[5];
#pragma unroll
    for (int i = 0; i < 8; i++) qU[i] = ldrow(fa.U, zv0 - 8 + i, xc, sc_u);
#pragma unroll
    for (int i = 0; i < 9; i++) qW[i] = ldrow(fa.W, zv0 - 8 + i, xc, sc_w);
#pragma unroll
    for (int i = 0; i < 9; i++) qZZ[i] = C(0);
#pragma unroll
    for (int i = 0; i < 8; i++) qXZ[i] = C(0);
#pragma unroll
    for (int i = 0; i < 5; i++) qXX[i] = C(0);

    // operands of row step j (prefetched one step ahead)
    int rs = zv0 - HALO;
    C hU = lane < 8 ? ldrow(fa.U, rs, xcol, sc_u) : C(0);
    C hW = lane < 8 ? ldrow(fa.W, rs, xcol, sc_w) : C(0);
    C oxx = ldrow(fa.Txx, rs, xc, sc_xx), ozz = ldrow(fa.Tzz, rs, xc, sc_zz), oxz = ldrow(fa.Txz, rs, xc, sc_xz);
    C lam = ldmat(m.LAM, rs), mu = ldmat(m.MU, rs), mua = ldmat(m.MUA, rs);
    C bu = ldmat(m.BU, rs - HALO), bw = ldmat(m.BW, rs - HALO);
    C nU = ldrow(fa.U, rs + HALO, xc, sc_u), nW = ldrow(fa.W, rs + HALO + 1, xc, sc_w);

    for (int j = 0; j < H + 2 * HALO; j++, rs++) {
        const int rv = rs - HALO;
        // ---- prefetch the next row step's operands
        const int rn = rs + 1;
        const C hU_n = lane < 8 ? ldrow(fa.U, rn, xcol, sc_u) : C(0);
        const C hW_n = lane < 8 ? ldrow(fa.W, rn, xcol, sc_w) : C(0);
        const C oxx_n = ldrow(fa.Txx, rn, xc, sc_xx), ozz_n = ldrow(fa.Tzz, rn, xc, sc_zz);
        const C oxz_n = ldrow(fa.Txz, rn, xc, sc_xz);
        const C lam_n = ldmat(m.LAM, rn), mu_n = ldmat(m.MU, rn), mua_n = ldmat(m.MUA, rn);
        const C bu_n = ldmat(m.BU, rn - HALO), bw_n = ldmat(m.BW, rn - HALO);
        const C nU_n = ldrow(fa.U, rn + HALO, xc, sc_u), nW_n = ldrow(fa.W, rn + HALO + 1, xc, sc_w);

        // ---- stage the U and W rows rs for the x-derivatives
        lU[lane + HALO] = qU[4];
        lW[lane + HALO] = qW[4];
        if (lane < 8) {
            const int o = lane < 4 ? lane : FW + lane;
            lU[o] = hU;
            lW[o] = hW;
        }
        __syncthreads();

        // ---- stresses of row rs
        C a = oxx, b = ozz, cc = oxz;
        if (rs >= 0 && rs < nz && xin) {
            const bool rint = rs >= (g.top ? lo : 4) && rs < nz - lo;
            if (xint && rint) {
                const C c0 = C(1225. / 1024.), c1 = C(245. / 3072.), c2 = C(49. / 5120.), c3 = C(5. / 7168.);
                const C ux = dm(&lU[lane + HALO], 1, c0, c1, c2, c3);
                const C wx = dp(&lW[lane + HALO], 1, c0, c1, c2, c3);
                const C uz = dm8(qU[0], qU[1], qU[2], qU[3], qU[4], qU[5], qU[6], qU[7], c0, c1, c2, c3);
                const C wz = dp8(qW[1], qW[2], qW[3], qW[4], qW[5], qW[6], qW[7], qW[8], c0, c1, c2, c3);
                psv_stress_interior(ux, uz, wx, wz, lam, mu, mua, a, b, cc);
            } else {
                C c0, c1, c2, c3;
                Coef<C>::get(order_at(xc, rs, nx, nz), c0, c1, c2, c3);
                const C ux = dm(&lU[lane + HALO], 1, c0, c1, c2, c3);
                const C wx = dp(&lW[lane + HALO], 1, c0, c1, c2, c3);
                const C uz = dm8(qU[0], qU[1], qU[2], qU[3], qU[4], qU[5], qU[6], qU[7], c0, c1, c2, c3);
                const C wz = dp8(qW[1], qW[2], qW[3], qW[4], qW[5], qW[6], qW[7], qW[8], c0, c1, c2, c3);
                const bool own = own_x && rs >= zv0 && rs < zv1;
                psv_stress_exact<S, C>(g, xc, rs, shot, ux, uz, wx, wz, lam, mu, mua, a, b, cc, spa, spb, own,
                                       spr, ssc, smp);
            }
            // source taps after the cell's update, in the reference's order
            if (src_on) {
                for (int e = t0; e < t1; e++) {
                    const TapF tp = ev.taps[e];
                    if (tp.row != rs || tp.lane != lane) continue;
                    if (tp.field == 2) a = apply_tap(a, tp, sv);
                    else if (tp.field == 3) b = apply_tap(b, tp, sv);
                    else cc = apply_tap(cc, tp, sv);
                }
            }
            if (own_x && rs >= zv0 && rs < zv1) {
                const long q = base + (long)rs * px + xc;
                ST::st(fb.Txx, q, a, wi_xx);
                ST::st(fb.Tzz, q, b, wi_zz);
                ST::st(fb.Txz, q, cc, wi_xz);
                if (ST::scaled) {
                    mx[2] = fmaxf(mx[2], fabsf((float)a));
                    mx[3] = fmaxf(mx[3], fabsf((float)b));
                    mx[4] = fmaxf(mx[4], fabsf((float)cc));
                }
            }
        } else {
            a = b = cc = C(0); // outside the domain the new stresses are zero
        }
#pragma unroll
        for (int i = 0; i < 8; i++) qZZ[i] = qZZ[i + 1];
        qZZ[8] = b;
#pragma unroll
        for (int i = 0; i < 7; i++) qXZ[i] = qXZ[i + 1];
        qXZ[7] = cc;
#pragma unroll
        for (int i = 0; i < 4; i++) qXX[i] = qXX[i + 1];
        qXX[4] = a;

        // ---- stage the new Txx and Txz rows rv for the x-derivatives
        lXX[lane] = qXX[0];
        lXZ[lane] = qXZ[3];
        __syncthreads();

        // ---- velocities of row rv
        if (rv >= zv0 && rv < zv1) {
            C u = qU[0], ww = qW[0];
            if (own_x) {
                const bool rint = rv >= (g.top ? lo : 4) && rv < nz - lo;
                if (xint && rint) {
                    const C c0 = C(1225. / 1024.), c1 = C(245. / 3072.), c2 = C(49. / 5120.), c3 = C(5. / 7168.);
                    const C dxTxx = dp(&lXX[lane], 1, c0, c1, c2, c3);
                    const C dxTxz = dm(&lXZ[lane], 1, c0, c1, c2, c3);
                    const C dzTzz = dm8(qZZ[0], qZZ[1], qZZ[2], qZZ[3], qZZ[4], qZZ[5], qZZ[6], qZZ[7], c0, c1, c2, c3);
                    const C dzTxz = dp8(qXZ[0], qXZ[1], qXZ[2], qXZ[3], qXZ[4], qXZ[5], qXZ[6], qXZ[7], c0, c1, c2, c3);
                    psv_velocity_interior(dxTxx, dzTzz, dxTxz, dzTxz, bu, bw, u, ww);
                } else {
                    C c0, c1, c2, c3;
                    Coef<C>::get(order_at(xc, rv, nx, nz), c0, c1, c2, c3);
                    const C dxTxx = dp(&lXX[lane], 1, c0, c1, c2, c3);
                    const C dxTxz = dm(&lXZ[lane], 1, c0, c1, c2, c3);
                    const C dzTzz = dm8(qZZ[0], qZZ[1], qZZ[2], qZZ[3], qZZ[4], qZZ[5], qZZ[6], qZZ[7], c0, c1, c2, c3);
                    const C dzTxz = dp8(qXZ[0], qXZ[1], qXZ[2], qXZ[3], qXZ[4], qXZ[5], qXZ[6], qXZ[7], c0, c1, c2, c3);
                    psv_velocity_exact<S, C>(g, xc, rv, shot, dxTxx, dzTzz, dxTxz, dzTxz, bu, bw, u, ww, vpa, vpb,
                                             true, vpr, vsc, vmp);
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
                C *seis = (C *)ev.seis + (long)(it / ev.itrecord) * ev.frame_size;
                for (int e = r0 + lane; e < r; e += FW) {
                    const RecF rc = ev.recs[e];
                    seis[rc.out] = rc.comp == 0 ? pU[rc.lane] : pW[rc.lane];
                }
                r0 = r;
            }
        }

        // ---- advance the queues and the prefetched operands
#pragma unroll
        for (int i = 0; i < 7; i++) qU[i] = qU[i + 1];
        qU[7] = nU;
#pragma unroll
        for (int i = 0; i < 8; i++) qW[i] = qW[i + 1];
        qW[8] = nW;
        hU = hU_n; hW = hW_n;
        oxx = oxx_n; ozz = ozz_n; oxz = oxz_n;
        lam = lam_n; mu = mu_n; mua = mua_n;
        bu = bu_n; bw = bw_n;
        nU = nU_n; nW = nW_n;
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

} // namespace fd23
