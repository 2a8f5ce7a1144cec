// Stress and velocity passes (P-SV and SH) for gfx950, with the C-PML
// memory-variable updates fused in, plus source injection and receiver
// gathering.  Semantics follow the reference's GPU kernels:
//   P-SV  PSV/clean_nobox/fd/cu_fd.cu:96-388
//   SH    SH/fd/cu_fd.cu:77-273
//
// Tiling: a workgroup is 64 (x, one wavefront per row) x 4 (waves) threads
// and owns a TX x TZ tile, TZ = 4*RZ; each thread updates RZ consecutive
// cells of its column.  The fields whose derivatives are needed are staged
// once per tile in LDS with their 4-cell halos (x halo only for the rows of
// the tile, z halo only for the 64 tile columns).  A workgroup steps SPB
// shots of the same model with the material held in registers, so the
// material is read once per SPB shots.
//
// Each tile is classified once: tiles that touch no C-PML strip, no edge
// band of reduced order and not the free surface take a branch-free path
// with the full 8th-order coefficients; all others evaluate the
// reference's per-cell branch structure exactly.
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

// Cooperative LDS staging of one field tile.
//  ZH: stage the 4-row halos above/below (all TX columns)
//  XH: stage the 4-column halos left/right (tile rows only)
// LDS row r <-> global row z0-4+r; LDS column c <-> global column x0-4+c.
template <class S, int TZ, bool ZH, bool XH>
__device__ __forceinline__ void stage(typename Store<S>::C *sm, const S *__restrict__ f, long base,
                                      int px, int x0, int z0, float scl) {
    constexpr int LW = TX + 2 * HALO;
    const int tx = threadIdx.x, ty = threadIdx.y;
    constexpr int rbeg = ZH ? 0 : HALO, rend = ZH ? TZ + 2 * HALO : TZ + HALO;
    static_assert((rend - rbeg) % BY == 0, "tile rows must split evenly over the waves");
#pragma unroll
    for (int i = 0; i < (rend - rbeg) / BY; i++) {
        const int r = rbeg + ty + i * BY;
        sm[r * LW + tx + HALO] = Store<S>::ld(f, base + (long)(z0 - HALO + r) * px + x0 + tx, scl);
    }
    if (XH) {
        const int tid = ty * TX + tx;
        constexpr int NE = TZ * 2 * HALO;
#pragma unroll
        for (int i = 0; i < (NE + TX * BY - 1) / (TX * BY); i++) {
            const int e = tid + i * TX * BY;
            if (NE % (TX * BY) != 0 && e >= NE) break;
            const int r = e >> 3, c = e & 7;
            const int col = c < HALO ? c - HALO : TX + c - HALO;
            sm[(r + HALO) * LW + col + HALO] =
                Store<S>::ld(f, base + (long)(z0 + r) * px + x0 + col, scl);
        }
    }
}

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

template <class S, int RZ, int SPB>
__global__ __launch_bounds__(TX *BY) void k_psv_stress(PsvFields<S> f, PsvStressMat<typename Store<S>::C> m,
                                                       PsvStressPsi<S> ps,
                                                       PsvStressProf<typename Store<S>::C> pr, Grid g,
                                                       const float *__restrict__ scl,
                                                       unsigned *__restrict__ amax) {
    using C = typename Store<S>::C;
    using ST = Store<S>;
    constexpr int TZ = BY * RZ;
    constexpr int LW = TX + 2 * HALO;
    __shared__ C sU[(TZ + 2 * HALO) * LW];
    __shared__ C sW[(TZ + 2 * HALO) * LW];

    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x0 = blockIdx.x * TX, z0 = blockIdx.y * TZ;
    const int nx = g.nx, nz = g.nz, px = g.px, npml = g.npml;
    const int gx = x0 + tx;
    const int lz0 = ty * RZ;
    const bool interior = tile_interior(g, x0, z0, TZ);

    float sc_u = 1.f, sc_w = 1.f, sc_xx = 1.f, sc_zz = 1.f, sc_xz = 1.f;
    if (ST::scaled) {
        sc_u = scl[SC_U]; sc_w = scl[SC_W];
        sc_xx = scl[SC_TXX]; sc_zz = scl[SC_TZZ]; sc_xz = scl[SC_TXZ];
    }
    float mx_xx = 0.f, mx_zz = 0.f, mx_xz = 0.f;

    // material for this thread's cells, reused across the SPB shots
    C lam[RZ], mu[RZ], mua[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        const int gz = z0 + lz0 + r;
        const bool live = gx < nx && gz < nz;
        const long q = (long)gz * px + gx;
        lam[r] = live ? m.LAM[q] : C(0);
        mu[r] = live ? m.MU[q] : C(0);
        mua[r] = live ? m.MUA[q] : C(0);
    }

    const int sh0 = blockIdx.z * SPB;
    for (int s = 0; s < SPB; s++) {
        const int shot = sh0 + s;
        if (shot >= g.nshots) break;
        const long base = (long)shot * g.sstride;
        if (s > 0) __syncthreads();
        stage<S, TZ, true, true>(sU, f.U, base, px, x0, z0, sc_u);
        stage<S, TZ, true, true>(sW, f.W, base, px, x0, z0, sc_w);
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
                const C gam = lam[r] + C(2.0) * mu[r];
                const long q = base + (long)(z0 + lz) * px + gx;
                C txx = ST::ld(f.Txx, q, sc_xx), tzz = ST::ld(f.Tzz, q, sc_zz), txz = ST::ld(f.Txz, q, sc_xz);
                txx = mad(gam, ux, txx); tzz = mad(lam[r], ux, tzz); txz = mad(mua[r], wx, txz);
                txx = mad(lam[r], wz, txx); tzz = mad(gam, wz, tzz); txz = mad(mua[r], uz, txz);
                ST::st(f.Txx, q, txx, 1.f / sc_xx);
                ST::st(f.Tzz, q, tzz, 1.f / sc_zz);
                ST::st(f.Txz, q, txz, 1.f / sc_xz);
                if (ST::scaled) {
                    mx_xx = fmaxf(mx_xx, fabsf((float)txx));
                    mx_zz = fmaxf(mx_zz, fabsf((float)tzz));
                    mx_xz = fmaxf(mx_xz, fabsf((float)txz));
                }
            }
        } else {
            float sp_ux = 1.f, sp_wx = 1.f, sp_uz = 1.f, sp_wz = 1.f;
            if (ST::scaled) {
                sp_ux = scl[SC_PU_X]; sp_wx = scl[SC_PW_X]; sp_uz = scl[SC_PU_Z]; sp_wz = scl[SC_PW_Z];
            }
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int lz = lz0 + r;
                const int ix = gx, iz = z0 + lz;
                if (ix >= nx || iz >= nz) continue;
                C c0, c1, c2, c3;
                Coef<C>::get(order_at(ix, iz, nx, nz), c0, c1, c2, c3);
                const C *pu = &sU[(lz + HALO) * LW + tx + HALO];
                const C *pw = &sW[(lz + HALO) * LW + tx + HALO];
                const C ux = dm(pu, 1, c0, c1, c2, c3);
                const C uz = dm(pu, LW, c0, c1, c2, c3);
                const C wx = dp(pw, 1, c0, c1, c2, c3);
                const C wz = dp(pw, LW, c0, c1, c2, c3);
                const C lm = lam[r], mm = mu[r], ma = mua[r];
                const C gam = lm + C(2.0) * mm;
                const long q = base + (long)iz * px + ix;
                C txx = ST::ld(f.Txx, q, sc_xx), tzz = ST::ld(f.Tzz, q, sc_zz), txz = ST::ld(f.Txz, q, sc_xz);
                if (ix < npml || ix >= nx - npml) {
                    const long pi = (long)shot * g.xpsi_stride + xpsi_idx(ix, iz, nx, npml);
                    C p1 = mad(pr.Ux.b[ix], ST::ld(ps.Ux, pi, sp_ux), pr.Ux.c[ix] * ux);
                    ST::st(ps.Ux, pi, p1, 1.f / sp_ux);
                    const C e1 = mad(ux, pr.Ux.k[ix], p1);
                    txx = mad(gam, e1, txx);
                    tzz = mad(lm, e1, tzz);
                    C p2 = mad(pr.Wx.b[ix], ST::ld(ps.Wx, pi, sp_wx), pr.Wx.c[ix] * wx);
                    ST::st(ps.Wx, pi, p2, 1.f / sp_wx);
                    txz = mad(ma, mad(wx, pr.Wx.k[ix], p2), txz);
                    if (ST::scaled) {
                        atomicMax(&amax[SC_PU_X], __float_as_uint(fabsf((float)p1)));
                        atomicMax(&amax[SC_PW_X], __float_as_uint(fabsf((float)p2)));
                    }
                } else if (iz != 0) {
                    txx = mad(gam, ux, txx); tzz = mad(lm, ux, tzz); txz = mad(ma, wx, txz);
                }
                const bool topcell = iz < npml && g.top;
                if (topcell || iz >= nz - npml) {
                    const long pi = (long)shot * g.zpsi_stride + zpsi_idx(ix, iz, nx, nz, npml, topcell);
                    C p1 = mad(pr.Wz.b[iz], ST::ld(ps.Wz, pi, sp_wz), pr.Wz.c[iz] * wz);
                    ST::st(ps.Wz, pi, p1, 1.f / sp_wz);
                    const C e1 = mad(wz, pr.Wz.k[iz], p1);
                    txx = mad(lm, e1, txx);
                    tzz = mad(gam, e1, tzz);
                    C p2 = mad(pr.Uz.b[iz], ST::ld(ps.Uz, pi, sp_uz), pr.Uz.c[iz] * uz);
                    ST::st(ps.Uz, pi, p2, 1.f / sp_uz);
                    txz = mad(ma, mad(uz, pr.Uz.k[iz], p2), txz);
                    if (ST::scaled) {
                        atomicMax(&amax[SC_PW_Z], __float_as_uint(fabsf((float)p1)));
                        atomicMax(&amax[SC_PU_Z], __float_as_uint(fabsf((float)p2)));
                    }
                } else if (iz == 0) {
                    txx = mad(C(2.0) * mm, ux, txx);
                    tzz = C(0);
                    txz = C(0);
                } else {
                    txx = mad(lm, wz, txx); tzz = mad(gam, wz, tzz); txz = mad(ma, uz, txz);
                }
                ST::st(f.Txx, q, txx, 1.f / sc_xx);
                ST::st(f.Tzz, q, tzz, 1.f / sc_zz);
                ST::st(f.Txz, q, txz, 1.f / sc_xz);
                if (ST::scaled) {
                    mx_xx = fmaxf(mx_xx, fabsf((float)txx));
                    mx_zz = fmaxf(mx_zz, fabsf((float)tzz));
                    mx_xz = fmaxf(mx_xz, fabsf((float)txz));
                }
            }
        }
    }
    if (ST::scaled) {
        amax_update(&amax[SC_TXX], mx_xx);
        amax_update(&amax[SC_TZZ], mx_zz);
        amax_update(&amax[SC_TXZ], mx_xz);
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

template <class S, int RZ, int SPB>
__global__ __launch_bounds__(TX *BY) void k_psv_velocity(PsvFields<S> f, PsvVelMat<typename Store<S>::C> m,
                                                         PsvVelPsi<S> ps,
                                                         PsvVelProf<typename Store<S>::C> pr, Grid g,
                                                         const float *__restrict__ scl,
                                                         unsigned *__restrict__ amax) {
    using C = typename Store<S>::C;
    using ST = Store<S>;
    constexpr int TZ = BY * RZ;
    constexpr int LW = TX + 2 * HALO;
    __shared__ C sXX[(TZ + 2 * HALO) * LW];
    __shared__ C sZZ[(TZ + 2 * HALO) * LW];
    __shared__ C sXZ[(TZ + 2 * HALO) * LW];

    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x0 = blockIdx.x * TX, z0 = blockIdx.y * TZ;
    const int nx = g.nx, nz = g.nz, px = g.px, npml = g.npml;
    const int gx = x0 + tx;
    const int lz0 = ty * RZ;
    const bool interior = tile_interior(g, x0, z0, TZ);

    float sc_u = 1.f, sc_w = 1.f, sc_xx = 1.f, sc_zz = 1.f, sc_xz = 1.f;
    if (ST::scaled) {
        sc_u = scl[SC_U]; sc_w = scl[SC_W];
        sc_xx = scl[SC_TXX]; sc_zz = scl[SC_TZZ]; sc_xz = scl[SC_TXZ];
    }
    float mx_u = 0.f, mx_w = 0.f;

    C bu[RZ], bw[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        const int gz = z0 + lz0 + r;
        const bool live = gx < nx && gz < nz;
        const long q = (long)gz * px + gx;
        bu[r] = live ? m.BU[q] : C(0);
        bw[r] = live ? m.BW[q] : C(0);
    }

    const int sh0 = blockIdx.z * SPB;
    for (int s = 0; s < SPB; s++) {
        const int shot = sh0 + s;
        if (shot >= g.nshots) break;
        const long base = (long)shot * g.sstride;
        if (s > 0) __syncthreads();
        stage<S, TZ, false, true>(sXX, f.Txx, base, px, x0, z0, sc_xx);
        stage<S, TZ, true, false>(sZZ, f.Tzz, base, px, x0, z0, sc_zz);
        stage<S, TZ, true, true>(sXZ, f.Txz, base, px, x0, z0, sc_xz);
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
                const long q = base + (long)(z0 + lz) * px + gx;
                C u = ST::ld(f.U, q, sc_u), w = ST::ld(f.W, q, sc_w);
                u = mad(bu[r], dxTxx, u); w = mad(bw[r], dxTxz, w);
                u = mad(bu[r], dzTxz, u); w = mad(bw[r], dzTzz, w);
                ST::st(f.U, q, u, 1.f / sc_u);
                ST::st(f.W, q, w, 1.f / sc_w);
                if (ST::scaled) {
                    mx_u = fmaxf(mx_u, fabsf((float)u));
                    mx_w = fmaxf(mx_w, fabsf((float)w));
                }
            }
        } else {
            float sp_xx = 1.f, sp_xzx = 1.f, sp_xzz = 1.f, sp_zz = 1.f;
            if (ST::scaled) {
                sp_xx = scl[SC_PTXX_X]; sp_xzx = scl[SC_PTXZ_X]; sp_xzz = scl[SC_PTXZ_Z]; sp_zz = scl[SC_PTZZ_Z];
            }
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int lz = lz0 + r;
                const int ix = gx, iz = z0 + lz;
                if (ix >= nx || iz >= nz) continue;
                C c0, c1, c2, c3;
                Coef<C>::get(order_at(ix, iz, nx, nz), c0, c1, c2, c3);
                const int o = (lz + HALO) * LW + tx + HALO;
                const C dxTxx = dp(&sXX[o], 1, c0, c1, c2, c3);
                const C dzTzz = dm(&sZZ[o], LW, c0, c1, c2, c3);
                const C dxTxz = dm(&sXZ[o], 1, c0, c1, c2, c3);
                const C dzTxz = dp(&sXZ[o], LW, c0, c1, c2, c3);
                const C b_u = bu[r], b_w = bw[r];
                const long q = base + (long)iz * px + ix;
                C u = ST::ld(f.U, q, sc_u), w = ST::ld(f.W, q, sc_w);
                if (ix < npml || ix >= nx - npml) {
                    const long pi = (long)shot * g.xpsi_stride + xpsi_idx(ix, iz, nx, npml);
                    C p1 = mad(pr.Txx_x.b[ix], ST::ld(ps.Txx_x, pi, sp_xx), pr.Txx_x.c[ix] * dxTxx);
                    ST::st(ps.Txx_x, pi, p1, 1.f / sp_xx);
                    u = mad(b_u, mad(dxTxx, pr.Txx_x.k[ix], p1), u);
                    C p2 = mad(pr.Txz_x.b[ix], ST::ld(ps.Txz_x, pi, sp_xzx), pr.Txz_x.c[ix] * dxTxz);
                    ST::st(ps.Txz_x, pi, p2, 1.f / sp_xzx);
                    w = mad(b_w, mad(dxTxz, pr.Txz_x.k[ix], p2), w);
                    if (ST::scaled) {
                        atomicMax(&amax[SC_PTXX_X], __float_as_uint(fabsf((float)p1)));
                        atomicMax(&amax[SC_PTXZ_X], __float_as_uint(fabsf((float)p2)));
                    }
                } else if (iz != 0) {
                    u = mad(b_u, dxTxx, u);
                    w = mad(b_w, dxTxz, w);
                }
                const bool topcell = iz < npml && g.top;
                if (topcell || iz >= nz - npml) {
                    const long pi = (long)shot * g.zpsi_stride + zpsi_idx(ix, iz, nx, nz, npml, topcell);
                    C p1 = mad(pr.Txz_z.b[iz], ST::ld(ps.Txz_z, pi, sp_xzz), pr.Txz_z.c[iz] * dzTxz);
                    C p2 = mad(pr.Tzz_z.b[iz], ST::ld(ps.Tzz_z, pi, sp_zz), pr.Tzz_z.c[iz] * dzTzz);
                    ST::st(ps.Txz_z, pi, p1, 1.f / sp_xzz);
                    ST::st(ps.Tzz_z, pi, p2, 1.f / sp_zz);
                    u = mad(b_u, mad(dzTxz, pr.Txz_z.k[iz], p1), u);
                    w = mad(b_w, mad(dzTzz, pr.Tzz_z.k[iz], p2), w);
                    if (ST::scaled) {
                        atomicMax(&amax[SC_PTXZ_Z], __float_as_uint(fabsf((float)p1)));
                        atomicMax(&amax[SC_PTZZ_Z], __float_as_uint(fabsf((float)p2)));
                    }
                } else if (iz == 0) {
                    u = mad(b_u, dxTxx + dzTxz, u);
                    w = C(0);
                } else {
                    u = mad(b_u, dzTxz, u);
                    w = mad(b_w, dzTzz, w);
                }
                ST::st(f.U, q, u, 1.f / sc_u);
                ST::st(f.W, q, w, 1.f / sc_w);
                if (ST::scaled) {
                    mx_u = fmaxf(mx_u, fabsf((float)u));
                    mx_w = fmaxf(mx_w, fabsf((float)w));
                }
            }
        }
    }
    if (ST::scaled) {
        amax_update(&amax[SC_U], mx_u);
        amax_update(&amax[SC_W], mx_w);
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
                                                      ShPsi2<S> ps, ShProf2<typename Store<S>::C> pr,
                                                      Grid g, const float *__restrict__ scl,
                                                      unsigned *__restrict__ amax) {
    using C = typename Store<S>::C;
    using ST = Store<S>;
    constexpr int TZ = BY * RZ;
    constexpr int LW = TX + 2 * HALO;
    __shared__ C sV[(TZ + 2 * HALO) * LW];

    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x0 = blockIdx.x * TX, z0 = blockIdx.y * TZ;
    const int nx = g.nx, nz = g.nz, px = g.px, npml = g.npml;
    const int gx = x0 + tx;
    const int lz0 = ty * RZ;
    const bool interior = tile_interior(g, x0, z0, TZ);

    float sc_v = 1.f, sc_x = 1.f, sc_z = 1.f;
    if (ST::scaled) { sc_v = scl[SC_V]; sc_x = scl[SC_TX]; sc_z = scl[SC_TZ]; }
    float mx_x = 0.f, mx_z = 0.f;

    C mux[RZ], muz[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        const int gz = z0 + lz0 + r;
        const bool live = gx < nx && gz < nz;
        const long q = (long)gz * px + gx;
        mux[r] = live ? m.MUX[q] : C(0);
        muz[r] = live ? m.MUZ[q] : C(0);
    }

    const int sh0 = blockIdx.z * SPB;
    for (int s = 0; s < SPB; s++) {
        const int shot = sh0 + s;
        if (shot >= g.nshots) break;
        const long base = (long)shot * g.sstride;
        if (s > 0) __syncthreads();
        stage<S, TZ, true, true>(sV, f.V, base, px, x0, z0, sc_v);
        __syncthreads();
        if (interior) {
            const C c0 = C(1225. / 1024.), c1 = C(245. / 3072.), c2 = C(49. / 5120.), c3 = C(5. / 7168.);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int lz = lz0 + r;
                const C *pv = &sV[(lz + HALO) * LW + tx + HALO];
                const C dxV = dp(pv, 1, c0, c1, c2, c3);
                const C dzV = dm(pv, LW, c0, c1, c2, c3);
                const long q = base + (long)(z0 + lz) * px + gx;
                C tx_ = mad(mux[r], dxV, ST::ld(f.Tx, q, sc_x));
                C tz_ = mad(muz[r], dzV, ST::ld(f.Tz, q, sc_z));
                ST::st(f.Tx, q, tx_, 1.f / sc_x);
                ST::st(f.Tz, q, tz_, 1.f / sc_z);
                if (ST::scaled) {
                    mx_x = fmaxf(mx_x, fabsf((float)tx_));
                    mx_z = fmaxf(mx_z, fabsf((float)tz_));
                }
            }
        } else {
            float sp_x = 1.f, sp_z = 1.f;
            if (ST::scaled) { sp_x = scl[SC_PV_X]; sp_z = scl[SC_PV_Z]; }
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int lz = lz0 + r;
                const int ix = gx, iz = z0 + lz;
                // early return of SH/fd/cu_fd.cu:140
                if (ix >= nx || iz >= nz || ix == 0 || ix >= nx - 1 || iz >= nz - 1) continue;
                C c0, c1, c2, c3;
                Coef<C>::get(order_at(ix, iz, nx, nz), c0, c1, c2, c3);
                const C *pv = &sV[(lz + HALO) * LW + tx + HALO];
                const C dxV = dp(pv, 1, c0, c1, c2, c3);
                const C dzV = dm(pv, LW, c0, c1, c2, c3);
                const long q = base + (long)iz * px + ix;
                C tx_ = ST::ld(f.Tx, q, sc_x), tz_ = ST::ld(f.Tz, q, sc_z);
                if (ix < npml || ix >= nx - npml) {
                    const long pi = (long)shot * g.xpsi_stride + xpsi_idx(ix, iz, nx, npml);
                    C p = mad(pr.x.b[ix], ST::ld(ps.x, pi, sp_x), pr.x.c[ix] * dxV);
                    ST::st(ps.x, pi, p, 1.f / sp_x);
                    tx_ = mad(mux[r], mad(dxV, pr.x.k[ix], p), tx_);
                    if (ST::scaled) atomicMax(&amax[SC_PV_X], __float_as_uint(fabsf((float)p)));
                } else {
                    tx_ = mad(mux[r], dxV, tx_);
                }
                const bool topcell = iz < npml && g.top;
                if (topcell || iz >= nz - npml) {
                    const long pi = (long)shot * g.zpsi_stride + zpsi_idx(ix, iz, nx, nz, npml, topcell);
                    C p = mad(pr.z.b[iz], ST::ld(ps.z, pi, sp_z), pr.z.c[iz] * dzV);
                    ST::st(ps.z, pi, p, 1.f / sp_z);
                    tz_ = mad(muz[r], mad(dzV, pr.z.k[iz], p), tz_);
                    if (ST::scaled) atomicMax(&amax[SC_PV_Z], __float_as_uint(fabsf((float)p)));
                } else {
                    tz_ = mad(muz[r], dzV, tz_);
                }
                ST::st(f.Tx, q, tx_, 1.f / sc_x);
                ST::st(f.Tz, q, tz_, 1.f / sc_z);
                if (ST::scaled) {
                    mx_x = fmaxf(mx_x, fabsf((float)tx_));
                    mx_z = fmaxf(mx_z, fabsf((float)tz_));
                }
            }
        }
    }
    if (ST::scaled) {
        amax_update(&amax[SC_TX], mx_x);
        amax_update(&amax[SC_TZ], mx_z);
    }
}

// ===========================================================================
// SH velocity pass (SH/fd/cu_fd.cu:172-273); Tz row 0 is first set to
// -Tz row 1 (:226), so every read of Tz row 0 in this pass sees -Tz[1].
// ===========================================================================
template <class S, int RZ, int SPB>
__global__ __launch_bounds__(TX *BY) void k_sh_velocity(ShFields<S> f, const typename Store<S>::C *__restrict__ BV,
                                                        ShPsi2<S> ps, ShProf2<typename Store<S>::C> pr,
                                                        Grid g, const float *__restrict__ scl,
                                                        unsigned *__restrict__ amax) {
    using C = typename Store<S>::C;
    using ST = Store<S>;
    constexpr int TZ = BY * RZ;
    constexpr int LW = TX + 2 * HALO;
    __shared__ C sX[(TZ + 2 * HALO) * LW];
    __shared__ C sZ[(TZ + 2 * HALO) * LW];

    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x0 = blockIdx.x * TX, z0 = blockIdx.y * TZ;
    const int nx = g.nx, nz = g.nz, px = g.px, npml = g.npml;
    const int gx = x0 + tx;
    const int lz0 = ty * RZ;
    const bool interior = tile_interior(g, x0, z0, TZ);

    float sc_v = 1.f, sc_x = 1.f, sc_z = 1.f;
    if (ST::scaled) { sc_v = scl[SC_V]; sc_x = scl[SC_TX]; sc_z = scl[SC_TZ]; }
    float mx_v = 0.f;

    C bv[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        const int gz = z0 + lz0 + r;
        const bool live = gx < nx && gz < nz;
        bv[r] = live ? BV[(long)gz * px + gx] : C(0);
    }

    const int sh0 = blockIdx.z * SPB;
    for (int s = 0; s < SPB; s++) {
        const int shot = sh0 + s;
        if (shot >= g.nshots) break;
        const long base = (long)shot * g.sstride;
        if (s > 0) __syncthreads();
        stage<S, TZ, false, true>(sX, f.Tx, base, px, x0, z0, sc_x);
        stage<S, TZ, true, false>(sZ, f.Tz, base, px, x0, z0, sc_z);
        // stress imaging at the surface: the LDS copy of row 0 (LDS row
        // HALO - z0) takes -Tz[row 1], and its owner writes it back.
        if (z0 - HALO <= 0 && z0 + TZ + HALO > 0) {
            __syncthreads();
            if (ty == 0) {
                const int r0 = HALO - z0; // LDS row of global row 0
                const C v = -ST::ld(f.Tz, base + (long)px + gx, sc_z);
                sZ[r0 * LW + tx + HALO] = v;
                if (z0 == 0 && gx < nx) ST::st(f.Tz, base + gx, v, 1.f / sc_z);
            }
        }
        __syncthreads();
        if (interior) {
            const C c0 = C(1225. / 1024.), c1 = C(245. / 3072.), c2 = C(49. / 5120.), c3 = C(5. / 7168.);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int lz = lz0 + r;
                const int o = (lz + HALO) * LW + tx + HALO;
                const C Tx_x = dm(&sX[o], 1, c0, c1, c2, c3);
                const C Tz_z = dp(&sZ[o], LW, c0, c1, c2, c3);
                const long q = base + (long)(z0 + lz) * px + gx;
                C v = ST::ld(f.V, q, sc_v);
                v = mad(bv[r], Tx_x, v);
                v = mad(bv[r], Tz_z, v);
                ST::st(f.V, q, v, 1.f / sc_v);
                if (ST::scaled) mx_v = fmaxf(mx_v, fabsf((float)v));
            }
        } else {
            float sp_x = 1.f, sp_z = 1.f;
            if (ST::scaled) { sp_x = scl[SC_PTX_X]; sp_z = scl[SC_PTZ_Z]; }
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int lz = lz0 + r;
                const int ix = gx, iz = z0 + lz;
                if (ix >= nx || iz >= nz) continue;
                C c0, c1, c2, c3;
                Coef<C>::get(order_at(ix, iz, nx, nz), c0, c1, c2, c3);
                const int o = (lz + HALO) * LW + tx + HALO;
                const C Tx_x = dm(&sX[o], 1, c0, c1, c2, c3);
                const C Tz_z = dp(&sZ[o], LW, c0, c1, c2, c3);
                const long q = base + (long)iz * px + ix;
                C v = ST::ld(f.V, q, sc_v);
                if (ix < npml || ix >= nx - npml) {
                    const long pi = (long)shot * g.xpsi_stride + xpsi_idx(ix, iz, nx, npml);
                    C p = mad(pr.x.b[ix], ST::ld(ps.x, pi, sp_x), pr.x.c[ix] * Tx_x);
                    ST::st(ps.x, pi, p, 1.f / sp_x);
                    v = mad(bv[r], mad(Tx_x, pr.x.k[ix], p), v);
                    if (ST::scaled) atomicMax(&amax[SC_PTX_X], __float_as_uint(fabsf((float)p)));
                } else {
                    v = mad(bv[r], Tx_x, v);
                }
                const bool topcell = iz < npml && g.top;
                if (topcell || iz >= nz - npml) {
                    const long pi = (long)shot * g.zpsi_stride + zpsi_idx(ix, iz, nx, nz, npml, topcell);
                    C p = mad(pr.z.b[iz], ST::ld(ps.z, pi, sp_z), pr.z.c[iz] * Tz_z);
                    ST::st(ps.z, pi, p, 1.f / sp_z);
                    v = mad(bv[r], mad(Tz_z, pr.z.k[iz], p), v);
                    if (ST::scaled) atomicMax(&amax[SC_PTZ_Z], __float_as_uint(fabsf((float)p)));
                } else {
                    v = mad(bv[r], Tz_z, v);
                }
                ST::st(f.V, q, v, 1.f / sc_v);
                if (ST::scaled) mx_v = fmaxf(mx_v, fabsf((float)v));
            }
        }
    }
    if (ST::scaled) amax_update(&amax[SC_V], mx_v);
}

// ===========================================================================
// Source injection (one workgroup): one thread per shot walks its taps in the
// reference's order (PSV/clean_nobox/point_source/cu_point_source.cu:45-70,
// :115-158; SH/point_source/cu_point_source.cu:36-64, :100-158, :192-198).
// ===========================================================================
struct Tap {
    int field;  // index into the context's field-pointer table
    long off;   // element offset inside the shot (row*px + col)
    double m;   // moment component (already scaled)
    double q;   // multiplier: +-1, +-0.25 (P-SV) or +-0.5 (SH)
    int mode;   // 0: float product then add (P-SV), 1: double sum (SH)
};
struct SrcInfo {
    int ntap, tap0; // taps [tap0, tap0+ntap) of the tap table
    int lsrc;
    int clamp;      // 1: it >= lsrc uses stf[lsrc-1] ('p' / stype 13)
    long stf0;      // offset of this shot's samples in the stf table
};

template <class S>
__global__ void k_inject(S *const *fields, int nfields_per_shot, long sstride, const Tap *__restrict__ taps,
                         const SrcInfo *__restrict__ src, const double *__restrict__ stf, int nshots,
                         int it_arg, int *clock, const float *__restrict__ scl) {
    using C = typename Store<S>::C;
    using ST = Store<S>;
    int it = it_arg;
    if (clock) it = clock[0];
    for (int shot = threadIdx.x; shot < nshots; shot += blockDim.x) {
        const SrcInfo si = src[shot];
        int t = it;
        bool act = si.ntap > 0 && si.lsrc > 0;
        if (t >= si.lsrc) {
            if (si.clamp) t = si.lsrc - 1;
            else act = false;
        }
        if (act) {
            const double sv = stf[si.stf0 + t];
            for (int i = 0; i < si.ntap; i++) {
                const Tap tp = taps[si.tap0 + i];
                S *fp = fields[tp.field];
                const float sc = ST::scaled ? scl[tp.field] : 1.f;
                const long q = (long)shot * sstride + tp.off;
                C v = ST::ld(fp, q, sc);
                if (tp.mode == 0) {
                    const C prod = C(sv) * C(tp.m);
                    v += prod * C(tp.q);
                } else {
                    v = C((double)v + tp.q * sv * tp.m);
                }
                ST::st(fp, q, v, 1.f / sc);
            }
        }
    }
    (void)nfields_per_shot;
    if (clock) {
        __syncthreads();
        if (threadIdx.x == 0) {
            clock[1] = it;
            clock[0] = it + 1;
        }
    }
}

// ===========================================================================
// Receiver gather (SH/record/cu_record.cu:136-144) into the on-device
// seismogram buffer [frame][shot][comp][max_rec]; offset -1 reads 0.
// ===========================================================================
template <class S>
__global__ void k_record(S *const *comps, int ncomp, long sstride, const long *__restrict__ roff,
                         const int *__restrict__ nrec, int max_rec, int nshots,
                         typename Store<S>::C *__restrict__ seis, int max_frames, int it_arg,
                         const int *clock, int itrecord, const float *__restrict__ scl,
                         const int *__restrict__ comp_field) {
    using C = typename Store<S>::C;
    using ST = Store<S>;
    int it = it_arg;
    if (clock) it = clock[1];
    if (it % itrecord != 0) return;
    const int frame = it / itrecord;
    if (frame >= max_frames) return;
    const long total = (long)nshots * ncomp * max_rec;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int r = (int)(i % max_rec);
        const int sc_ = (int)(i / max_rec);
        const int comp = sc_ % ncomp, shot = sc_ / ncomp;
        if (r >= nrec[shot * ncomp + comp]) continue;
        const long o = roff[i];
        C v = C(0);
        if (o >= 0) {
            const float s = ST::scaled ? scl[comp_field[comp]] : 1.f;
            v = ST::ld(comps[comp], (long)shot * sstride + o, s);
        }
        seis[(long)frame * total + i] = v;
    }
}

} // namespace fd23
