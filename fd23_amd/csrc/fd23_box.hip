// Hybrid box coupling on the device, generalised from the reference's
// square, file-driven box (PSV/psv_clean_double/box/box.cpp:10-210,
// box/cu_box.cu:6-78, fd/cu_fd.cu:93-207, main.cpp:153-200) to a bx*bz
// rectangle driven by an analytic plane wave.
//
// Geometry (box-local cell (i,j), global (x0+i, z0+j)); d = distance to the
// box edge, min(i, j, bx-1-i, bz-1-j):
//   sub ring  d in [0,3]: r = a - s   (the reference's s0..s1 / s6..s7 band)
//   add ring  d in [4,7]: a = r + s   (the reference's s2..s3 / s4..s5 band)
// a = global field, r = residual grid (bx*bz, zero outside, no C-PML, full
// 8th order at every cell, homogeneous box medium), s = drive field.
// Per step (the reference's order): add(U,W); stress; sub(U,W); stress_r;
// add(Txx,Txz,Tzz); velocity (+record); sub(Txx,Txz,Tzz); velocity_r.
//
// Drive: s = -(incident plane wave), so the residual grid carries the total
// field inside the box and the global grid the scattered field outside it
// (the total-field / scattered-field arrangement; with s = +incident the
// ring arithmetic is the reference's point-source box unchanged).  The
// plane wave travels upward at angle th from the vertical, unit direction
// n = (sin th, -cos th); P polarisation n, SV polarisation p = (cos th,
// sin th).  Particle velocity v = A g(tau) pol, stress from the strain of the
// displacement, tau = t - (n.x - n.x_first)/c - t0, g = Ricker(f0).
// Velocities are sampled at t = (it+1/2) dt, stresses at t = (it+1) dt, at
// each field's staggered position (U at x+1/2, W at z-1/2, Txz at both).
// The reference's half-size copy of the drive (cu_box.cu:39-43, which
// copies sizeof(int)*bw*bw bytes of double arrays) is not reproduced.
//
// File drive (fd23_box_configure_frames, the reference's own box): a
// bw x bw square around (xs, zs) whose drive comes from per-step
// displacement frames of (bw+8)^2 cells (U then W), as BOX::setboxsource
// (box.cpp:120-197) converts them: velocities (u[it+1] - u[it])/dt and
// stresses C grad u[it+1] / dt with the box material, the derivative order
// reduced toward the box centre.  All frames are uploaded once and a
// one-workgroup kernel converts frames it and it+1 at the start of each
// step, so the graph loop needs no host work per step.  The box material
// follows MATERIAL::init_for_box_from_full (material.cpp:39-99): a cell of
// the global material is forced to the homogeneous source medium only
// when one of its five values differs from it by more than 1e-4
// relative; the residual grid reads the global material at its cells.
#include "fd23.h"
#include "fd_device.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#pragma clang fp contract(off)

using namespace fd23;

struct fd23_ctx;
hipStream_t fd23_ctx_stream(fd23_ctx *c);
void *fd23_ctx_field(fd23_ctx *c, int internal);
void *fd23_ctx_material(fd23_ctx *c, int i);
int fd23_ctx_px(fd23_ctx *c);
long fd23_ctx_sstride(fd23_ctx *c);
void *&fd23_ctx_box(fd23_ctx *c);
const fd23_desc &fd23_ctx_desc(fd23_ctx *c);
int *fd23_ctx_clock(fd23_ctx *c);
void fd23_ctx_invalidate_graphs(fd23_ctx *c);
extern "C" void fd23_set_error_(const char *msg);
void fd23_box_free(fd23_ctx *c);

namespace {

struct BoxPar {
    int x0, z0, bx, bz;
    int px;          // global row pitch
    long sstride;    // global shot stride
    int nshots;
    double h, dt;
    double nx_, nz_; // propagation direction
    double qx, qz;   // polarisation
    double c;        // phase velocity
    double amp, f0, t0, nref;
    double sxx, szz, sxz; // stress = s** * A g(tau)
    double sign;     // s = sign * incident
};

__device__ __forceinline__ double ricker(double f0, double tau) {
    const double a = 3.14159265358979323846 * f0 * tau;
    const double a2 = a * a;
    return (1.0 - 2.0 * a2) * exp(-a2);
}

// drive value of one field at global (x, z) in cells (may be half-integer)
__device__ __forceinline__ double drive(const BoxPar &p, int field, double xc, double zc, double t) {
    const double x = xc * p.h, z = zc * p.h;
    const double tau = t - (p.nx_ * x + p.nz_ * z - p.nref) / p.c - p.t0;
    const double g = p.sign * p.amp * ricker(p.f0, tau);
    switch (field) {
    case 0: return g * p.qx;      // U
    case 1: return g * p.qz;      // W
    case 2: return g * p.sxx;     // Txx
    case 3: return g * p.szz;     // Tzz
    default: return g * p.sxz;    // Txz
    }
}
// staggered position offsets of internal fields U W Txx Tzz Txz
__device__ __forceinline__ void offsets(int field, double &ox, double &oz) {
    ox = (field == 0 || field == 4) ? 0.5 : 0.0;
    oz = (field == 1 || field == 4) ? -0.5 : 0.0;
}

__device__ __forceinline__ int ring_depth(int i, int j, int bx, int bz) {
    return min(min(i, j), min(bx - 1 - i, bz - 1 - j));
}

// add (a = r + s) or sub (r = a - s) over one ring for fields [f0, f1)
template <class C>
__global__ void k_box_ring(BoxPar p, C *const *glob, C *resid, int f0, int f1, bool add, const int *clock, int k,
                           int it_arg, double toff, const double *stab) {
    const int it = clock ? clock[0] + k : it_arg;
    const double t = (it + toff) * p.dt;
    const long nb = (long)p.bx * p.bz;
    const int dlo = add ? 4 : 0, dhi = add ? 7 : 3;
    const long total = nb * (f1 - f0) * p.nshots;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const long cell = e % nb;
        const int fi = f0 + (int)((e / nb) % (f1 - f0));
        const int shot = (int)(e / (nb * (f1 - f0)));
        const int i = (int)(cell % p.bx), j = (int)(cell / p.bx);
        const int d = ring_depth(i, j, p.bx, p.bz);
        if (d < dlo || d > dhi) continue;
        C s;
        if (stab) { // file drive: this step's converted frames
            s = C(stab[(long)fi * nb + cell]);
        } else {
            double ox, oz;
            offsets(fi, ox, oz);
            s = C(drive(p, fi, p.x0 + i + ox, p.z0 + j + oz, t));
        }
        C *a = glob[fi] + (long)shot * p.sstride + (long)(p.z0 + j) * p.px + (p.x0 + i);
        C *r = resid + ((long)shot * 5 + fi) * nb + cell;
        if (add)
            *a = *r + s;
        else
            *r = *a - s;
    }
}

template <class C> __device__ __forceinline__ C rat(const C *f, int bx, int bz, int i, int j) {
    return (i < 0 || i >= bx || j < 0 || j >= bz) ? C(0) : f[(long)j * bx + i];
}

// residual grid passes (PSV/psv_clean_double/fd/cu_fd.cu:93-207): full
// order, zero outside the box, homogeneous (lam, mu, b) already scaled
template <class C>
__global__ void k_box_stress_r(BoxPar p, C *resid, const C *LAMg, const C *MUg, const C *MUAg) {
    const long nb = (long)p.bx * p.bz;
    const C c0 = C(1225. / 1024.), c1 = C(245. / 3072.), c2 = C(49. / 5120.), c3 = C(5. / 7168.);
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nb * p.nshots; e += (long)gridDim.x * blockDim.x) {
        const int shot = (int)(e / nb);
        const long cell = e % nb;
        const int i = (int)(cell % p.bx), j = (int)(cell / p.bx);
        C *R = resid + (long)shot * 5 * nb;
        const C *U = R, *W = R + nb;
        C *Txx = R + 2 * nb, *Tzz = R + 3 * nb, *Txz = R + 4 * nb;
        const int bx = p.bx, bz = p.bz;
        auto dmx = [&](const C *f) {
            C r = c0 * (rat(f, bx, bz, i, j) - rat(f, bx, bz, i - 1, j));
            r = mad(-c1, rat(f, bx, bz, i + 1, j) - rat(f, bx, bz, i - 2, j), r);
            r = mad(c2, rat(f, bx, bz, i + 2, j) - rat(f, bx, bz, i - 3, j), r);
            return mad(-c3, rat(f, bx, bz, i + 3, j) - rat(f, bx, bz, i - 4, j), r);
        };
        auto dmz = [&](const C *f) {
            C r = c0 * (rat(f, bx, bz, i, j) - rat(f, bx, bz, i, j - 1));
            r = mad(-c1, rat(f, bx, bz, i, j + 1) - rat(f, bx, bz, i, j - 2), r);
            r = mad(c2, rat(f, bx, bz, i, j + 2) - rat(f, bx, bz, i, j - 3), r);
            return mad(-c3, rat(f, bx, bz, i, j + 3) - rat(f, bx, bz, i, j - 4), r);
        };
        auto dpx = [&](const C *f) {
            C r = c0 * (rat(f, bx, bz, i + 1, j) - rat(f, bx, bz, i, j));
            r = mad(-c1, rat(f, bx, bz, i + 2, j) - rat(f, bx, bz, i - 1, j), r);
            r = mad(c2, rat(f, bx, bz, i + 3, j) - rat(f, bx, bz, i - 2, j), r);
            return mad(-c3, rat(f, bx, bz, i + 4, j) - rat(f, bx, bz, i - 3, j), r);
        };
        auto dpz = [&](const C *f) {
            C r = c0 * (rat(f, bx, bz, i, j + 1) - rat(f, bx, bz, i, j));
            r = mad(-c1, rat(f, bx, bz, i, j + 2) - rat(f, bx, bz, i, j - 1), r);
            r = mad(c2, rat(f, bx, bz, i, j + 3) - rat(f, bx, bz, i, j - 2), r);
            return mad(-c3, rat(f, bx, bz, i, j + 4) - rat(f, bx, bz, i, j - 3), r);
        };
        const C ux = dmx(U), uz = dmz(U), wx = dpx(W), wz = dpz(W);
        const long qg = (long)(p.z0 + j) * p.px + p.x0 + i;
        const C lam = LAMg[qg], mu = MUg[qg], mua = MUAg[qg];
        const C gam = lam + C(2.0) * mu;
        const long q = (long)j * bx + i;
        Txx[q] = Txx[q] + (gam * ux + lam * wz);
        Tzz[q] = Tzz[q] + (lam * ux + gam * wz);
        Txz[q] = Txz[q] + mua * (wx + uz);
    }
}

template <class C> __global__ void k_box_velocity_r(BoxPar p, C *resid, const C *BUg, const C *BWg) {
    const long nb = (long)p.bx * p.bz;
    const C c0 = C(1225. / 1024.), c1 = C(245. / 3072.), c2 = C(49. / 5120.), c3 = C(5. / 7168.);
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nb * p.nshots; e += (long)gridDim.x * blockDim.x) {
        const int shot = (int)(e / nb);
        const long cell = e % nb;
        const int i = (int)(cell % p.bx), j = (int)(cell / p.bx);
        C *R = resid + (long)shot * 5 * nb;
        C *U = R, *W = R + nb;
        const C *Txx = R + 2 * nb, *Tzz = R + 3 * nb, *Txz = R + 4 * nb;
        const int bx = p.bx, bz = p.bz;
        auto dmx = [&](const C *f) {
            C r = c0 * (rat(f, bx, bz, i, j) - rat(f, bx, bz, i - 1, j));
            r = mad(-c1, rat(f, bx, bz, i + 1, j) - rat(f, bx, bz, i - 2, j), r);
            r = mad(c2, rat(f, bx, bz, i + 2, j) - rat(f, bx, bz, i - 3, j), r);
            return mad(-c3, rat(f, bx, bz, i + 3, j) - rat(f, bx, bz, i - 4, j), r);
        };
        auto dmz = [&](const C *f) {
            C r = c0 * (rat(f, bx, bz, i, j) - rat(f, bx, bz, i, j - 1));
            r = mad(-c1, rat(f, bx, bz, i, j + 1) - rat(f, bx, bz, i, j - 2), r);
            r = mad(c2, rat(f, bx, bz, i, j + 2) - rat(f, bx, bz, i, j - 3), r);
            return mad(-c3, rat(f, bx, bz, i, j + 3) - rat(f, bx, bz, i, j - 4), r);
        };
        auto dpx = [&](const C *f) {
            C r = c0 * (rat(f, bx, bz, i + 1, j) - rat(f, bx, bz, i, j));
            r = mad(-c1, rat(f, bx, bz, i + 2, j) - rat(f, bx, bz, i - 1, j), r);
            r = mad(c2, rat(f, bx, bz, i + 3, j) - rat(f, bx, bz, i - 2, j), r);
            return mad(-c3, rat(f, bx, bz, i + 4, j) - rat(f, bx, bz, i - 3, j), r);
        };
        auto dpz = [&](const C *f) {
            C r = c0 * (rat(f, bx, bz, i, j + 1) - rat(f, bx, bz, i, j));
            r = mad(-c1, rat(f, bx, bz, i, j + 2) - rat(f, bx, bz, i, j - 1), r);
            r = mad(c2, rat(f, bx, bz, i, j + 3) - rat(f, bx, bz, i, j - 2), r);
            return mad(-c3, rat(f, bx, bz, i, j + 4) - rat(f, bx, bz, i, j - 3), r);
        };
        const C dxTxx = dpx(Txx), dzTzz = dmz(Tzz), dxTxz = dmx(Txz), dzTxz = dpz(Txz);
        const long qg = (long)(p.z0 + j) * p.px + p.x0 + i;
        const C bu = BUg[qg], bw = BWg[qg];
        const long q = (long)j * bx + i;
        U[q] = U[q] + bu * (dxTxx + dzTxz);
        W[q] = W[q] + bw * (dxTxz + dzTzz);
    }
}

// force the box medium into the global material (the reference's
// init_for_box_from_full, PSV/psv_clean_double/material/material.cpp:39-99)
template <class C>
__global__ void k_box_material(BoxPar p, C *BU, C *BW, C *MU, C *MUA, C *LAM, C b, C mu, C lam) {
    const long nb = (long)p.bx * p.bz;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nb; e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e % p.bx), j = (int)(e / p.bx);
        const long q = (long)(p.z0 + j) * p.px + p.x0 + i;
        BU[q] = b;
        BW[q] = b;
        MU[q] = mu;
        MUA[q] = mu;
        LAM[q] = lam;
    }
}

// MATERIAL::init_for_box_from_full (material.cpp:58-97): a box cell keeps
// the global material unless one of its five values differs from the
// homogeneous source medium by more than eps relative; then all five are
// set to it.  Comparisons in double, as the FP64 tree.
template <class C>
__global__ void k_box_material_force(BoxPar p, C *BU, C *BW, C *MU, C *MUA, C *LAM, double b, double mu,
                                     double lam, int *nforced) {
    const long nb = (long)p.bx * p.bz;
    const double eps = 1e-4;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nb; e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e % p.bx), j = (int)(e / p.bx);
        const long q = (long)(p.z0 + j) * p.px + p.x0 + i;
        if (fabs(b - (double)BU[q]) > b * eps || fabs(b - (double)BW[q]) > b * eps ||
            fabs(mu - (double)MU[q]) > mu * eps || fabs(mu - (double)MUA[q]) > mu * eps ||
            fabs(lam - (double)LAM[q]) > lam * eps) {
            BU[q] = C(b);
            BW[q] = C(b);
            MU[q] = C(mu);
            MUA[q] = C(mu);
            LAM[q] = C(lam);
            atomicAdd(nforced, 1);
        }
    }
}

// BOX::setboxsource (box.cpp:142-196) for step it on the device: frames
// it and it+1 (each U then W, (bw+8)^2, x fastest) to the drive of the
// five fields, s[field][j*bw + i] in the internal order U W Txx Tzz Txz,
// with the box material at the box cells.  Same double arithmetic and
// order of operations as the host code, no contraction.
template <class C>
__global__ void k_box_frames_drive(BoxPar p, const double *frames, int nframes, const int *clock, int k, int it_arg,
                                   const C *LAM, const C *MU, const C *MUA, double *s) {
    const int it = clock ? clock[0] + k : it_arg;
    const int bw = p.bx, fw = bw + 8;
    const long fsz = (long)fw * fw;
    // the host checks every run's frame range; the clamp only keeps a
    // step beyond it inside the table
    const int f1 = min(max(it + 1, 1), nframes - 1), f0 = f1 - 1;
    const double *uo = frames + (long)f0 * 2 * fsz, *wo = uo + fsz;
    const double *un = frames + (long)f1 * 2 * fsz, *wn = un + fsz;
    const long nb = (long)bw * bw;
    const int mid = (bw + 1) / 2;
    for (long e = threadIdx.x; e < nb; e += blockDim.x) {
        const int tx = (int)(e % bw) + 4, ty = (int)(e / bw) + 4;
        int ord = min(4, max(abs(ty - 3 - mid) - 1, abs(tx - 3 - mid) - 1));
        ord = max(ord, 0);
        double c0, c1, c2, c3;
        Coef<double>::get(ord, c0, c1, c2, c3);
        auto U = [&](int x, int z) { return un[(long)z * fw + x]; };
        auto W = [&](int x, int z) { return wn[(long)z * fw + x]; };
        double ux = c0 * (U(tx, ty) - U(tx - 1, ty));
        ux = ux - c1 * (U(tx + 1, ty) - U(tx - 2, ty));
        ux = ux + c2 * (U(tx + 2, ty) - U(tx - 3, ty));
        ux = ux - c3 * (U(tx + 3, ty) - U(tx - 4, ty));
        double uz = c0 * (U(tx, ty) - U(tx, ty - 1));
        uz = uz - c1 * (U(tx, ty + 1) - U(tx, ty - 2));
        uz = uz + c2 * (U(tx, ty + 2) - U(tx, ty - 3));
        uz = uz - c3 * (U(tx, ty + 3) - U(tx, ty - 4));
        double wx = c0 * (W(tx + 1, ty) - W(tx, ty));
        wx = wx - c1 * (W(tx + 2, ty) - W(tx - 1, ty));
        wx = wx + c2 * (W(tx + 3, ty) - W(tx - 2, ty));
        wx = wx - c3 * (W(tx + 4, ty) - W(tx - 3, ty));
        double wz = c0 * (W(tx, ty + 1) - W(tx, ty));
        wz = wz - c1 * (W(tx, ty + 2) - W(tx, ty - 1));
        wz = wz + c2 * (W(tx, ty + 3) - W(tx, ty - 2));
        wz = wz - c3 * (W(tx, ty + 4) - W(tx, ty - 3));
        const long qf = (long)ty * fw + tx;
        const long qg = (long)(p.z0 + ty - 4) * p.px + p.x0 + tx - 4;
        const double lam = (double)LAM[qg], mu = (double)MU[qg], mua = (double)MUA[qg];
        const double gam = lam + 2.0 * mu;
        s[e] = (un[qf] - uo[qf]) / p.dt;
        s[nb + e] = (wn[qf] - wo[qf]) / p.dt;
        s[2 * nb + e] = (gam * ux + lam * wz) / p.dt;
        s[3 * nb + e] = (lam * ux + gam * wz) / p.dt;
        s[4 * nb + e] = (mua * (wx + uz)) / p.dt;
    }
}

struct Box {
    BoxPar p;
    void *resid = nullptr; // [shot][5][bz][bx], compute type
    void **gtab = nullptr; // device table of the 5 global field pointers
    double lam, mu, b;     // scaled box medium
    bool f64 = false;
    int staged_it = 0;     // step index of the staged entry points
    // file drive: all frames on the device and this step's drive
    double *frames = nullptr;
    int nframes = 0;
    double *sdrive = nullptr; // [5][bw*bw]
    int nforced = 0;          // box cells whose material was forced
};

Box *get(fd23_ctx *c) { return (Box *)fd23_ctx_box(c); }

template <class C> int launch_ring(fd23_ctx *c, Box *bx, int f0, int f1, bool add, int k, double toff) {
    const long total = (long)bx->p.bx * bx->p.bz * (f1 - f0) * bx->p.nshots;
    const int nblk = (int)std::min<long>((total + 255) / 256, 2048);
    hipLaunchKernelGGL((k_box_ring<C>), dim3(nblk), dim3(256), 0, fd23_ctx_stream(c), bx->p, (C *const *)bx->gtab,
                       (C *)bx->resid, f0, f1, add, k >= 0 ? fd23_ctx_clock(c) : nullptr, k < 0 ? 0 : k,
                       bx->staged_it, toff, bx->sdrive);
    return FD23_OK;
}

// the file drive of this step, ahead of its first ring
template <class C> void launch_frames_drive(fd23_ctx *c, Box *bx, int k) {
    hipLaunchKernelGGL((k_box_frames_drive<C>), dim3(1), dim3(1024), 0, fd23_ctx_stream(c), bx->p, bx->frames,
                       bx->nframes, k >= 0 ? fd23_ctx_clock(c) : nullptr, k < 0 ? 0 : k, bx->staged_it,
                       (const C *)fd23_ctx_material(c, 4), (const C *)fd23_ctx_material(c, 2),
                       (const C *)fd23_ctx_material(c, 3), bx->sdrive);
}

} // namespace

// the graph loop reads the step from the device clock; the staged entry
// points (k < 0) count steps here, and fd23_run leaves the count at its end
void fd23_box_set_step(fd23_ctx *c, int it) {
    Box *b = get(c);
    if (b) b->staged_it = it;
}

int fd23_box_launch_pre_stress(fd23_ctx *c, int k) {
    Box *b = get(c);
    if (!b) return FD23_OK;
    if (b->frames) {
        if (k < 0 && b->staged_it + 1 >= b->nframes) {
            fd23_set_error_("the box source frames end before this step");
            return FD23_EINVAL;
        }
        if (b->f64)
            launch_frames_drive<double>(c, b, k);
        else
            launch_frames_drive<float>(c, b, k);
    }
    return b->f64 ? launch_ring<double>(c, b, 0, 2, true, k, 0.5) : launch_ring<float>(c, b, 0, 2, true, k, 0.5);
}

// frames a run of steps [it0, it0+nt) needs: it0+nt+1
int fd23_box_check_run(fd23_ctx *c, int it0, int nt) {
    Box *b = get(c);
    if (!b || !b->frames || nt == 0) return FD23_OK;
    if ((long)it0 + nt + 1 > b->nframes) {
        char msg[160];
        snprintf(msg, sizeof msg, "the box source has %d frames; steps %d..%d need %ld", b->nframes, it0,
                 it0 + nt - 1, (long)it0 + nt + 1);
        fd23_set_error_(msg);
        return FD23_EINVAL;
    }
    return FD23_OK;
}

int fd23_box_launch_post_stress(fd23_ctx *c, int k) {
    Box *b = get(c);
    if (!b) return FD23_OK;
    const long nb = (long)b->p.bx * b->p.bz * b->p.nshots;
    const int nblk = (int)std::min<long>((nb + 255) / 256, 2048);
    if (b->f64) {
        launch_ring<double>(c, b, 0, 2, false, k, 0.5);
        hipLaunchKernelGGL((k_box_stress_r<double>), dim3(nblk), dim3(256), 0, fd23_ctx_stream(c), b->p,
                           (double *)b->resid, (const double *)fd23_ctx_material(c, 4),
                           (const double *)fd23_ctx_material(c, 2), (const double *)fd23_ctx_material(c, 3));
        launch_ring<double>(c, b, 2, 5, true, k, 1.0);
    } else {
        launch_ring<float>(c, b, 0, 2, false, k, 0.5);
        hipLaunchKernelGGL((k_box_stress_r<float>), dim3(nblk), dim3(256), 0, fd23_ctx_stream(c), b->p,
                           (float *)b->resid, (const float *)fd23_ctx_material(c, 4),
                           (const float *)fd23_ctx_material(c, 2), (const float *)fd23_ctx_material(c, 3));
        launch_ring<float>(c, b, 2, 5, true, k, 1.0);
    }
    return FD23_OK;
}

int fd23_box_launch_post_velocity(fd23_ctx *c, int k) {
    Box *b = get(c);
    if (!b) return FD23_OK;
    const long nb = (long)b->p.bx * b->p.bz * b->p.nshots;
    const int nblk = (int)std::min<long>((nb + 255) / 256, 2048);
    if (b->f64) {
        launch_ring<double>(c, b, 2, 5, false, k, 1.0);
        hipLaunchKernelGGL((k_box_velocity_r<double>), dim3(nblk), dim3(256), 0, fd23_ctx_stream(c), b->p,
                           (double *)b->resid, (const double *)fd23_ctx_material(c, 0),
                           (const double *)fd23_ctx_material(c, 1));
    } else {
        launch_ring<float>(c, b, 2, 5, false, k, 1.0);
        hipLaunchKernelGGL((k_box_velocity_r<float>), dim3(nblk), dim3(256), 0, fd23_ctx_stream(c), b->p,
                           (float *)b->resid, (const float *)fd23_ctx_material(c, 0),
                           (const float *)fd23_ctx_material(c, 1));
    }
    if (k < 0) b->staged_it++;
    return FD23_OK;
}

extern "C" int fd23_box_configure(fd23_ctx *c, int shot, int x0, int z0, int bx, int bz, int incident,
                                  double angle_deg, double f0, double t0, double amp, double vp, double vs,
                                  double rho, double h, double dt) {
    char msg[256];
    if (!c) {
        fd23_set_error_("null context");
        return FD23_EINVAL;
    }
    const fd23_desc &d = fd23_ctx_desc(c);
    if (d.wave != FD23_WAVE_PSV || (d.storage != FD23_STORE_F32 && d.storage != FD23_STORE_F64)) {
        fd23_set_error_("box coupling needs a P-SV context with FP32 or FP64 storage");
        return FD23_EINVAL;
    }
    if (shot != -1 && shot != 0) {
        fd23_set_error_("the box applies to every shot of a context: pass shot -1");
        return FD23_EINVAL;
    }
    // the reference rejects boxes too close to the edge (param.cpp:179-182);
    // here the rings and their stencils must stay inside the grid
    if (bx < 16 || bz < 16 || x0 < 4 || z0 < 4 || x0 + bx > d.nx - 4 || z0 + bz > d.nz - 4 || incident < 0 ||
        incident > 1 || !(vp > 0) || !(rho > 0) || vs < 0 || !(h > 0) || !(dt > 0) || !(f0 > 0)) {
        snprintf(msg, sizeof msg, "bad box: x0=%d z0=%d bx=%d bz=%d in %dx%d", x0, z0, bx, bz, d.nx, d.nz);
        fd23_set_error_(msg);
        return FD23_EINVAL;
    }
    fd23_box_free(c);
    Box *b = new Box();
    BoxPar &p = b->p;
    p.x0 = x0;
    p.z0 = z0;
    p.bx = bx;
    p.bz = bz;
    p.px = fd23_ctx_px(c);
    p.sstride = fd23_ctx_sstride(c);
    p.nshots = d.nshots;
    p.h = h;
    p.dt = dt;
    const double th = angle_deg * 3.14159265358979323846 / 180.0;
    p.nx_ = std::sin(th);
    p.nz_ = -std::cos(th);
    const double mu = rho * vs * vs, lam = rho * vp * vp - 2.0 * mu;
    if (incident == 0) {
        p.qx = p.nx_;
        p.qz = p.nz_;
        p.c = vp;
        p.sxx = -(lam + 2.0 * mu * p.nx_ * p.nx_) / vp;
        p.szz = -(lam + 2.0 * mu * p.nz_ * p.nz_) / vp;
        p.sxz = -(2.0 * mu * p.nx_ * p.nz_) / vp;
    } else {
        if (!(vs > 0)) {
            delete b;
            fd23_set_error_("SV incidence needs vs > 0");
            return FD23_EINVAL;
        }
        p.qx = std::cos(th);
        p.qz = std::sin(th);
        p.c = vs;
        p.sxx = -2.0 * mu * p.qx * p.nx_ / vs;
        p.szz = -2.0 * mu * p.qz * p.nz_ / vs;
        p.sxz = -mu * (p.qx * p.nz_ + p.qz * p.nx_) / vs;
    }
    p.amp = amp;
    p.f0 = f0;
    p.t0 = t0;
    // the wavefront reaches the box first at the corner minimising n.x
    double nref = 1e300;
    for (int cx = 0; cx < 2; cx++)
        for (int cz = 0; cz < 2; cz++)
            nref = std::min(nref, p.nx_ * (x0 + cx * (bx - 1)) * h + p.nz_ * (z0 + cz * (bz - 1)) * h);
    p.nref = nref - 1.0 * h; // half-cell staggering margin
    p.sign = -1.0;
    b->f64 = d.storage == FD23_STORE_F64;
    const double fac = dt / h;
    b->lam = lam * fac;
    b->mu = mu * fac;
    b->b = fac / rho;
    const size_t cs = b->f64 ? 8 : 4;
    const size_t bytes = (size_t)d.nshots * 5 * bx * bz * cs;
    if (hipMalloc(&b->resid, bytes) != hipSuccess || hipMalloc((void **)&b->gtab, 5 * sizeof(void *)) != hipSuccess) {
        fd23_set_error_("box allocation failed");
        if (b->resid) (void)hipFree(b->resid);
        delete b;
        return FD23_ENOMEM;
    }
    hipStream_t st = fd23_ctx_stream(c);
    void *tab[5];
    for (int f = 0; f < 5; f++) tab[f] = fd23_ctx_field(c, f);
    (void)hipMemsetAsync(b->resid, 0, bytes, st);
    (void)hipMemcpyAsync(b->gtab, tab, sizeof tab, hipMemcpyHostToDevice, st);
    // homogeneous box medium in the global material (BU BW MU MUA LAM)
    const int nblk = (int)std::min<long>(((long)bx * bz + 255) / 256, 2048);
    if (b->f64)
        hipLaunchKernelGGL((k_box_material<double>), dim3(nblk), dim3(256), 0, st, p,
                           (double *)fd23_ctx_material(c, 0), (double *)fd23_ctx_material(c, 1),
                           (double *)fd23_ctx_material(c, 2), (double *)fd23_ctx_material(c, 3),
                           (double *)fd23_ctx_material(c, 4), b->b, b->mu, b->lam);
    else
        hipLaunchKernelGGL((k_box_material<float>), dim3(nblk), dim3(256), 0, st, p,
                           (float *)fd23_ctx_material(c, 0), (float *)fd23_ctx_material(c, 1),
                           (float *)fd23_ctx_material(c, 2), (float *)fd23_ctx_material(c, 3),
                           (float *)fd23_ctx_material(c, 4), (float)b->b, (float)b->mu, (float)b->lam);
    if (hipStreamSynchronize(st) != hipSuccess) {
        fd23_set_error_("box set-up failed");
        (void)hipFree(b->resid);
        (void)hipFree(b->gtab);
        delete b;
        return FD23_EHIP;
    }
    fd23_ctx_box(c) = b;
    fd23_ctx_invalidate_graphs(c);
    return FD23_OK;
}

// The reference's box (BOX::init_box, box.cpp:10-119): a bw x bw square
// whose cell (i, j) is global (xs + idxstart + i, zs + idxstart + j),
// idxstart = (bw+1)/2 + 1 - bw, driven by nframes displacement frames
// (frame f: U then W, (bw+8)^2 doubles, x fastest: the sourcefile layout
// read at box.cpp:103-105 and 124-137); step it uses frames it and it+1.
// The source medium (srcpvel, srcsvel, srcden) is forced into the global
// material where it differs (material.cpp:39-99).
extern "C" int fd23_box_configure_frames(fd23_ctx *c, int xs, int zs, int bw, const double *frames, int nframes,
                                         double srcpvel, double srcsvel, double srcden, double h, double dt) {
    char msg[256];
    if (!c || !frames) {
        fd23_set_error_("null argument");
        return FD23_EINVAL;
    }
    const fd23_desc &d = fd23_ctx_desc(c);
    if (d.wave != FD23_WAVE_PSV || (d.storage != FD23_STORE_F32 && d.storage != FD23_STORE_F64)) {
        fd23_set_error_("box coupling needs a P-SV context with FP32 or FP64 storage");
        return FD23_EINVAL;
    }
    const int idxstart = (bw + 1) / 2 + 1 - bw;
    const int x0 = xs + idxstart, z0 = zs + idxstart;
    if (bw < 16 || bw > 1024 || nframes < 2 || x0 < 4 || z0 < 4 || x0 + bw > d.nx - 4 || z0 + bw > d.nz - 4 ||
        !(srcden > 0) || !(srcpvel > 0) || srcsvel < 0 || !(h > 0) || !(dt > 0)) {
        snprintf(msg, sizeof msg, "bad box: xs=%d zs=%d bw=%d nframes=%d in %dx%d", xs, zs, bw, nframes, d.nx,
                 d.nz);
        fd23_set_error_(msg);
        return FD23_EINVAL;
    }
    fd23_box_free(c);
    Box *b = new Box();
    BoxPar &p = b->p;
    p.x0 = x0;
    p.z0 = z0;
    p.bx = bw;
    p.bz = bw;
    p.px = fd23_ctx_px(c);
    p.sstride = fd23_ctx_sstride(c);
    p.nshots = d.nshots;
    p.h = h;
    p.dt = dt;
    b->f64 = d.storage == FD23_STORE_F64;
    b->nframes = nframes;
    // the source medium scaled as material.cpp:52-58 forms it
    double mu = srcsvel * srcsvel * srcden;
    double lam = (srcpvel * srcpvel * srcden - 2.0 * mu);
    double bb = 1.0 / srcden;
    mu = mu * dt / h;
    lam = lam * dt / h;
    bb = bb * dt / h;
    b->lam = lam;
    b->mu = mu;
    b->b = bb;
    const size_t cs = b->f64 ? 8 : 4;
    const size_t rbytes = (size_t)d.nshots * 5 * bw * bw * cs;
    const size_t fbytes = (size_t)nframes * 2 * (bw + 8) * (bw + 8) * sizeof(double);
    const size_t sbytes = (size_t)5 * bw * bw * sizeof(double);
    int *d_nforced = nullptr;
    if (hipMalloc(&b->resid, rbytes) != hipSuccess || hipMalloc((void **)&b->gtab, 5 * sizeof(void *)) != hipSuccess ||
        hipMalloc((void **)&b->frames, fbytes) != hipSuccess ||
        hipMalloc((void **)&b->sdrive, sbytes) != hipSuccess ||
        hipMalloc((void **)&d_nforced, sizeof(int)) != hipSuccess) {
        fd23_set_error_("box allocation failed");
        if (b->resid) (void)hipFree(b->resid);
        if (b->gtab) (void)hipFree(b->gtab);
        if (b->frames) (void)hipFree(b->frames);
        if (b->sdrive) (void)hipFree(b->sdrive);
        if (d_nforced) (void)hipFree(d_nforced);
        delete b;
        return FD23_ENOMEM;
    }
    hipStream_t st = fd23_ctx_stream(c);
    void *tab[5];
    for (int f = 0; f < 5; f++) tab[f] = fd23_ctx_field(c, f);
    (void)hipMemsetAsync(b->resid, 0, rbytes, st);
    (void)hipMemsetAsync(b->sdrive, 0, sbytes, st);
    (void)hipMemsetAsync(d_nforced, 0, sizeof(int), st);
    (void)hipMemcpyAsync(b->gtab, tab, sizeof tab, hipMemcpyHostToDevice, st);
    hipError_t e = hipMemcpyAsync(b->frames, frames, fbytes, hipMemcpyHostToDevice, st);
    const int nblk = (int)std::min<long>(((long)bw * bw + 255) / 256, 2048);
    if (b->f64)
        hipLaunchKernelGGL((k_box_material_force<double>), dim3(nblk), dim3(256), 0, st, p,
                           (double *)fd23_ctx_material(c, 0), (double *)fd23_ctx_material(c, 1),
                           (double *)fd23_ctx_material(c, 2), (double *)fd23_ctx_material(c, 3),
                           (double *)fd23_ctx_material(c, 4), bb, mu, lam, d_nforced);
    else
        hipLaunchKernelGGL((k_box_material_force<float>), dim3(nblk), dim3(256), 0, st, p,
                           (float *)fd23_ctx_material(c, 0), (float *)fd23_ctx_material(c, 1),
                           (float *)fd23_ctx_material(c, 2), (float *)fd23_ctx_material(c, 3),
                           (float *)fd23_ctx_material(c, 4), bb, mu, lam, d_nforced);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&b->nforced, d_nforced, sizeof(int), hipMemcpyDeviceToHost, st);
    const hipError_t es = hipStreamSynchronize(st);
    if (e == hipSuccess) e = es;
    (void)hipFree(d_nforced);
    if (e != hipSuccess) {
        snprintf(msg, sizeof msg, "box set-up failed: %s", hipGetErrorString(e));
        fd23_set_error_(msg);
        (void)hipFree(b->resid);
        (void)hipFree(b->gtab);
        (void)hipFree(b->frames);
        (void)hipFree(b->sdrive);
        delete b;
        return FD23_EHIP;
    }
    fd23_ctx_box(c) = b;
    fd23_ctx_invalidate_graphs(c);
    return FD23_OK;
}

// box cells whose global material the last fd23_box_configure_frames set
// to the source medium (the reference prints each, material.cpp:77-85)
extern "C" int fd23_box_forced_cells(fd23_ctx *c) {
    Box *b = get(c);
    return b ? b->nforced : 0;
}

// residual-grid fields of one shot (U W Txx Tzz Txz, bz*bx each, compute type)
extern "C" int fd23_box_get_residual(fd23_ctx *c, int shot, void *out) {
    Box *b = get(c);
    if (!b || !out || shot < 0 || shot >= b->p.nshots) {
        fd23_set_error_("no box or bad shot");
        return FD23_EINVAL;
    }
    const size_t cs = b->f64 ? 8 : 4;
    const size_t n = (size_t)5 * b->p.bx * b->p.bz;
    hipStream_t st = fd23_ctx_stream(c);
    if (hipMemcpyAsync(out, (char *)b->resid + shot * n * cs, n * cs, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        fd23_set_error_("box read-back failed");
        return FD23_EHIP;
    }
    return FD23_OK;
}

void fd23_box_free(fd23_ctx *c) {
    Box *b = get(c);
    if (!b) return;
    if (b->resid) (void)hipFree(b->resid);
    if (b->gtab) (void)hipFree(b->gtab);
    if (b->frames) (void)hipFree(b->frames);
    if (b->sdrive) (void)hipFree(b->sdrive);
    delete b;
    fd23_ctx_box(c) = nullptr;
}
