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
#include "fd23.h"
#include "fd_device.hpp"

#include <hip/hip_runtime.h>

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
                           int it_arg, double toff) {
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
        double ox, oz;
        offsets(fi, ox, oz);
        const C s = C(drive(p, fi, p.x0 + i + ox, p.z0 + j + oz, t));
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
__global__ void k_box_stress_r(BoxPar p, C *resid, C lam, C mu, C mua) {
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
        const C gam = lam + C(2.0) * mu;
        const long q = (long)j * bx + i;
        Txx[q] = Txx[q] + (gam * ux + lam * wz);
        Tzz[q] = Tzz[q] + (lam * ux + gam * wz);
        Txz[q] = Txz[q] + mua * (wx + uz);
    }
}

template <class C> __global__ void k_box_velocity_r(BoxPar p, C *resid, C bu, C bw) {
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

struct Box {
    BoxPar p;
    void *resid = nullptr; // [shot][5][bz][bx], compute type
    void **gtab = nullptr; // device table of the 5 global field pointers
    double lam, mu, b;     // scaled box medium
    bool f64 = false;
    int staged_it = 0;     // step index of the staged entry points
};

Box *get(fd23_ctx *c) { return (Box *)fd23_ctx_box(c); }

template <class C> int launch_ring(fd23_ctx *c, Box *bx, int f0, int f1, bool add, int k, double toff) {
    const long total = (long)bx->p.bx * bx->p.bz * (f1 - f0) * bx->p.nshots;
    const int nblk = (int)std::min<long>((total + 255) / 256, 2048);
    hipLaunchKernelGGL((k_box_ring<C>), dim3(nblk), dim3(256), 0, fd23_ctx_stream(c), bx->p, (C *const *)bx->gtab,
                       (C *)bx->resid, f0, f1, add, k >= 0 ? fd23_ctx_clock(c) : nullptr, k < 0 ? 0 : k,
                       bx->staged_it, toff);
    return FD23_OK;
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
    return b->f64 ? launch_ring<double>(c, b, 0, 2, true, k, 0.5) : launch_ring<float>(c, b, 0, 2, true, k, 0.5);
}

int fd23_box_launch_post_stress(fd23_ctx *c, int k) {
    Box *b = get(c);
    if (!b) return FD23_OK;
    const long nb = (long)b->p.bx * b->p.bz * b->p.nshots;
    const int nblk = (int)std::min<long>((nb + 255) / 256, 2048);
    if (b->f64) {
        launch_ring<double>(c, b, 0, 2, false, k, 0.5);
        hipLaunchKernelGGL((k_box_stress_r<double>), dim3(nblk), dim3(256), 0, fd23_ctx_stream(c), b->p,
                           (double *)b->resid, b->lam, b->mu, b->mu);
        launch_ring<double>(c, b, 2, 5, true, k, 1.0);
    } else {
        launch_ring<float>(c, b, 0, 2, false, k, 0.5);
        hipLaunchKernelGGL((k_box_stress_r<float>), dim3(nblk), dim3(256), 0, fd23_ctx_stream(c), b->p,
                           (float *)b->resid, (float)b->lam, (float)b->mu, (float)b->mu);
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
                           (double *)b->resid, b->b, b->b);
    } else {
        launch_ring<float>(c, b, 2, 5, false, k, 1.0);
        hipLaunchKernelGGL((k_box_velocity_r<float>), dim3(nblk), dim3(256), 0, fd23_ctx_stream(c), b->p,
                           (float *)b->resid, (float)b->b, (float)b->b);
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
    delete b;
    fd23_ctx_box(c) = nullptr;
}
