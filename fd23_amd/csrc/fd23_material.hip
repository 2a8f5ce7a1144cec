// Material precompute on the device: the context's material arrays from
// vp, vs and den, the same values as the host precompute
// (host_precompute.cpp, checked against the reference's own host code),
// computed by one pointwise kernel instead of a host pass over the grid.
//
// P-SV, PSV/clean_nobox/material/material.cpp:35-240: Lame parameters per
// cell (T arithmetic, differences in double), the water treatment of the
// 5-cell left/right/bottom edges (vs/vp < 0.01: mu = vp^2/3 rho), buoyancy
// BU toward +x (wrapping across rows) and BW toward -z (the last row uses
// its own density), MUA the harmonic mean of four mu with the reference's
// repeated m2m3m4 term in the denominator, row 0 as the Mittet free surface
// (ix in [1, nx-2]: BU = 4/(rho+rho'), LAM = 0, MU halved), then every
// array times dt/h.  SH, SH/material/material.cpp:76-130: BV = 1/rho and
// harmonic means of mu toward +x (wrapping) and -z (row 0 uses itself).
// The stability number needs the largest vp (vs for SH), and the FP16/BF16
// scale estimate the largest buoyancy: both are reduced on the device.
#include "fd23.h"
#include "fd_device.hpp"
#include "mdl.hpp"
#include "keys.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#pragma clang fp contract(off)

using namespace fd23;

struct fd23_ctx;
hipStream_t fd23_ctx_stream(fd23_ctx *c);
void *fd23_ctx_material(fd23_ctx *c, int i);
int fd23_ctx_px(fd23_ctx *c);
const fd23_desc &fd23_ctx_desc(fd23_ctx *c);
void fd23_ctx_material_ready(fd23_ctx *c, double bmax);
extern "C" void fd23_set_error_(const char *msg);

namespace {

int merr(int code, const std::string &m) {
    fd23_set_error_(m.c_str());
    return code;
}
#define MCHK(x)                                                                                      \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return merr(FD23_EHIP, std::string(#x ": ") + hipGetErrorString(e_)); \
    } while (0)

// order-preserving map of a double onto unsigned 64-bit integers, so that
// atomicMax on the keys is a max on the values
__device__ __forceinline__ unsigned long long order_key(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
double order_value(unsigned long long k) {
    const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double v;
    memcpy(&v, &b, sizeof v);
    return v;
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = v;
}

// workgroup-wide max of two values (256 threads), then one atomic per
// workgroup and value: the grid is a few thousand workgroups that stride
// over the cells, so the two addresses see a few thousand atomics
constexpr int kMatBlock = 256, kMatGrid = 2048;
__device__ __forceinline__ void block_max_to(unsigned long long *dst, double a, double b) {
    __shared__ double sa[kMatBlock / 64], sb[kMatBlock / 64];
    for (int o = 32; o > 0; o >>= 1) {
        a = fmax(a, __shfl_xor(a, o));
        b = fmax(b, __shfl_xor(b, o));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sa[w] = a;
        sb[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kMatBlock / 64; i++) {
            a = fmax(a, sa[i]);
            b = fmax(b, sb[i]);
        }
        atomicMax(&dst[0], order_key(a));
        atomicMax(&dst[1], order_key(b));
    }
}

template <class T>
__device__ __forceinline__ void lame(const T *vp, const T *vs, const T *den, long q, bool edge, T &mu, T &lam) {
    mu = vs[q] * vs[q] * den[q];
    lam = (T)(vp[q] * vp[q] * den[q] - 2.0 * mu);
    if (edge && vs[q] / vp[q] < 0.01) { // water at the model edges
        mu = (T)(vp[q] * vp[q] / 3.0 * den[q]);
        lam = (T)(vp[q] * vp[q] * den[q] - 2.0 * mu);
    }
}

template <class T>
__global__ __launch_bounds__(kMatBlock) void k_psv_material(const T *__restrict__ vp, const T *__restrict__ vs,
                                                            const T *__restrict__ den, int nx, int nz, int px,
                                                            T fac, T *BU, T *BW, T *MU, T *MUA, T *LAM,
                                                            unsigned long long *maxk) {
    const long N = (long)nx * nz;
    double vmax = -1e300, bmax = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < N; q += (long)gridDim.x * blockDim.x) {
        const int iz = (int)(q / nx), ix = (int)(q - (long)iz * nx);
        auto edge = [&](int x, int z) { return x < 5 || x >= nx - 5 || z >= nz - 5; };
        T mu, lam;
        lame(vp, vs, den, q, edge(ix, iz), mu, lam);
        T bu = 0, bw = 0, mua = 0;
        if (iz >= 1) {
            bu = (T)(2.0 / ((double)den[q] + den[q + 1 < N ? q + 1 : q]));
            bw = (T)(2.0 / ((double)den[q] + den[q + nx < N ? q - nx : q]));
            if (ix + 1 < nx && iz + 1 < nz) {
                T m2, m3, m4, l_;
                lame(vp, vs, den, q + 1, edge(ix + 1, iz), m2, l_);
                lame(vp, vs, den, q - nx, edge(ix, iz - 1), m3, l_);
                lame(vp, vs, den, q + 1 - nx, edge(ix + 1, iz - 1), m4, l_);
                const double d1 = mu, d2 = m2, d3 = m3, d4 = m4;
                const double dd = fmax(d2 * d3 * d4 + d1 * d3 * d4 + d1 * d2 * d4 + d2 * d3 * d4, 1E-10);
                mua = (T)(4.0 * d1 * d2 * d3 * d4 / dd);
            } else {
                mua = mu;
            }
        } else if (ix >= 1 && ix < nx - 1) { // free surface row
            bu = (T)(4.0 / ((double)den[ix] + den[ix + 1]));
            lam = 0;
            mu = (T)(0.5 * mu);
        }
        const long o = (long)iz * px + ix;
        BU[o] = bu * fac;
        BW[o] = bw * fac;
        MU[o] = mu * fac;
        MUA[o] = mua * fac;
        LAM[o] = lam * fac;
        vmax = fmax(vmax, (double)vp[q]);
        bmax = fmax(bmax, fmax(fabs((double)(bu * fac)), fabs((double)(bw * fac))));
    }
    block_max_to(maxk, vmax, bmax);
}

__global__ __launch_bounds__(kMatBlock) void k_sh_material(const float *__restrict__ vs,
                                                           const float *__restrict__ den, int nx, int nz, int px,
                                                           float fac, float *BV, float *MUX, float *MUZ,
                                                           unsigned long long *maxk) {
    const long N = (long)nx * nz;
    double vmax = -1e300, bmax = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < N; q += (long)gridDim.x * blockDim.x) {
        const int iz = (int)(q / nx), ix = (int)(q - (long)iz * nx);
        auto mu_of = [&](long k) -> double { return vs[k] * vs[k] * den[k]; };
        const double m1 = mu_of(q);
        double m2 = q + 1 < N ? mu_of(q + 1) : m1;
        const float mux = (float)(2.0 * m1 * m2 / fmax(m1 + m2, 1E-10));
        m2 = q >= nx ? mu_of(q - nx) : m1;
        const float muz = (float)(2.0 * m1 * m2 / fmax(m1 + m2, 1E-10));
        const float bv = (float)(1.0 / den[q]) * fac;
        const long o = (long)iz * px + ix;
        BV[o] = bv;
        MUX[o] = mux * fac;
        MUZ[o] = muz * fac;
        vmax = fmax(vmax, (double)vs[q]);
        bmax = fmax(bmax, fabs((double)bv));
    }
    block_max_to(maxk, vmax, bmax);
}

// crossing-number test of SH/model/model.cpp:437-469 in the host
// rasteriser's double arithmetic (host_precompute.cpp: inside)
__device__ int inside_poly(double xp, double yp, const float *vx, const float *vz, int n) {
    int nleft = 0;
    double x2 = vx[n - 1], y2 = vz[n - 1];
    for (int j = 0; j < n; j++) {
        const double x1 = x2, y1 = y2;
        x2 = vx[j];
        y2 = vz[j];
        if (fmin(y1, y2) >= yp) continue;
        if (fmax(y1, y2) < yp) continue;
        if (y1 == y2) {
            if (xp > fmax(x1, x2)) continue;
            if (xp >= fmin(x1, x2)) return 1;
            nleft++;
            continue;
        }
        const double xi = x1 + (yp - y1) * (x2 - x1) / (y2 - y1);
        if (xi == xp) return 1;
        if (xi > xp) nleft++;
    }
    return nleft % 2;
}

struct MdlDev {
    const float *vx, *vz, *vp, *vs, *den;
    const int *first, *count, *ix1, *ix2, *iz1, *iz2;
    int np;
};

// every cell takes the last polygon that contains it (polygons tested from
// the last down, the first hit wins), else the background
template <class T>
__global__ void k_mdl_raster(MdlDev m, int nx, int nz, double hh, T *vp, T *vs, T *den) {
    const long N = (long)nx * nz;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N) return;
    const int iz = (int)(q / nx), ix = (int)(q - (long)iz * nx);
    int hit = 0;
    for (int p = m.np - 1; p >= 1 && !hit; p--) {
        if (m.count[p] == 0 || ix < m.ix1[p] || ix > m.ix2[p] || iz < m.iz1[p] || iz > m.iz2[p]) continue;
        if (inside_poly(hh * ix, hh * iz, m.vx + m.first[p], m.vz + m.first[p], m.count[p])) hit = p;
    }
    vp[q] = (T)m.vp[hit];
    vs[q] = (T)m.vs[hit];
    den[q] = (T)m.den[hit];
}

const double kCoefSum = 1225. / 1024. + 245. / 3072. + 49. / 5120. + 5. / 7168.;

} // namespace

// The material kernels on device inputs `in` (vp vs den, or vs den for
// SH, nz*nx each) with the two 64-bit reduction keys at `keys`.  The keys
// are updated by 64-bit atomics, which fault the device on an address
// that is not 8-byte aligned (HSA_STATUS_ERROR_MEMORY_APERTURE_VIOLATION
// in k_psv_material): a staging layout that placed them after three
// arrays of an odd cell count did that, so the address is checked here
// before any launch.  Waits for the stream; the caller frees the buffers.
static int material_from_device(fd23_ctx *c, const char *in, unsigned long long *keys, double h, double dt,
                                double *stab) {
    if (!fd23_keys_aligned(keys))
        return merr(FD23_EINVAL, "set_model: reduction keys not 8-byte aligned (64-bit atomics need it)");
    const fd23_desc &d = fd23_ctx_desc(c);
    const bool psv = d.wave == FD23_WAVE_PSV;
    const bool f64 = psv && d.storage == FD23_STORE_F64;
    const size_t N = (size_t)d.nx * d.nz;
    hipStream_t st = fd23_ctx_stream(c);
    // 0 is below every key (the key of -NaN): the maxima start empty
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, keys, 2, 0ull);
    const int block = kMatBlock;
    const int grid = (int)std::min<size_t>((N + block - 1) / block, kMatGrid);
    const int px = fd23_ctx_px(c);
    if (psv && f64) {
        const double *v = (const double *)in;
        hipLaunchKernelGGL(k_psv_material<double>, dim3(grid), dim3(block), 0, st, v, v + N, v + 2 * N, d.nx,
                           d.nz, px, dt / h, (double *)fd23_ctx_material(c, 0), (double *)fd23_ctx_material(c, 1),
                           (double *)fd23_ctx_material(c, 2), (double *)fd23_ctx_material(c, 3),
                           (double *)fd23_ctx_material(c, 4), keys);
    } else if (psv) {
        const float *v = (const float *)in;
        const float ff = (float)dt / (float)h;
        hipLaunchKernelGGL(k_psv_material<float>, dim3(grid), dim3(block), 0, st, v, v + N, v + 2 * N, d.nx, d.nz,
                           px, ff, (float *)fd23_ctx_material(c, 0), (float *)fd23_ctx_material(c, 1),
                           (float *)fd23_ctx_material(c, 2), (float *)fd23_ctx_material(c, 3),
                           (float *)fd23_ctx_material(c, 4), keys);
    } else {
        const float *v = (const float *)in;
        const float ff = (float)dt / (float)h;
        hipLaunchKernelGGL(k_sh_material, dim3(grid), dim3(block), 0, st, v, v + N, d.nx, d.nz, px, ff,
                           (float *)fd23_ctx_material(c, 0), (float *)fd23_ctx_material(c, 1),
                           (float *)fd23_ctx_material(c, 2), keys);
    }
    unsigned long long hk[2] = {0, 0};
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(hk, keys, sizeof hk, hipMemcpyDeviceToHost, st);
    const hipError_t es_ = hipStreamSynchronize(st);
    if (e == hipSuccess) e = es_;
    if (e != hipSuccess) return merr(FD23_EHIP, std::string("set_model: ") + hipGetErrorString(e));
    const double vmax = std::max(order_value(hk[0]), -99999.0), bmax = order_value(hk[1]);
    fd23_ctx_material_ready(c, bmax);
    // stability number as the host precompute forms it
    const double s = psv && f64 ? vmax * dt * kCoefSum * std::sqrt(2.0) / h
                                : vmax * (double)(float)dt * kCoefSum * std::sqrt(2.0) / (double)(float)h;
    if (stab) *stab = s;
    if (s > 1) return merr(FD23_ESTAB, "model stability number " + std::to_string(s) + " > 1");
    return FD23_OK;
}

// material_from_device on a staging allocation (keys at buf, inputs at
// buf + kIn), which it frees
static int material_from_staging(fd23_ctx *c, char *buf, size_t kIn, double h, double dt, double *stab) {
    const int r = material_from_device(c, buf + kIn, (unsigned long long *)buf, h, dt, stab);
    hipFree(buf);
    return r;
}

static int check_model_ctx(fd23_ctx *c) {
    const fd23_desc &d = fd23_ctx_desc(c);
    if (d.nx < 2 || d.nz < 2) return merr(FD23_EINVAL, "set_model: grid too small");
    if (d.wave != FD23_WAVE_PSV && d.storage == FD23_STORE_F64)
        return merr(FD23_EINVAL, "set_model: SH material is computed in float; pass it with fd23_set_material");
    return FD23_OK;
}

// the staging allocation: the two 64-bit reduction keys first (8-byte
// aligned, as 64-bit atomics need), then the inputs from byte kIn
static const size_t kIn = 256;

// vp, vs, den: nz*nx host arrays (double for an FP64 P-SV context, float
// otherwise; SH reads vs and den only).
extern "C" int fd23_set_model(fd23_ctx *c, const void *vp, const void *vs, const void *den, double h, double dt,
                              double *stab) {
    if (!c || !vs || !den) return merr(FD23_EINVAL, "set_model: null argument");
    const fd23_desc &d = fd23_ctx_desc(c);
    const bool psv = d.wave == FD23_WAVE_PSV;
    if (psv && !vp) return merr(FD23_EINVAL, "set_model: null vp");
    int r = check_model_ctx(c);
    if (r) return r;
    const bool f64 = psv && d.storage == FD23_STORE_F64;
    const size_t N = (size_t)d.nx * d.nz, es = f64 ? 8 : 4;
    const int nin = psv ? 3 : 2;
    hipStream_t st = fd23_ctx_stream(c);
    char *buf = nullptr;
    MCHK(hipMalloc((void **)&buf, kIn + nin * N * es));
    const void *src[3] = {psv ? vp : vs, psv ? vs : den, den};
    for (int i = 0; i < nin; i++) {
        hipError_t e = hipMemcpyAsync(buf + kIn + i * N * es, src[i], N * es, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) {
            hipStreamSynchronize(st);
            hipFree(buf);
            return merr(FD23_EHIP, std::string("set_model copy: ") + hipGetErrorString(e));
        }
    }
    return material_from_staging(c, buf, kIn, h, dt, stab);
}

// ---------------------------------------------------------------------------
// No-allocation variant of fd23_set_model, for callers that cannot allocate
// device memory.  Its scratch is one file-scope device array per GPU: the
// two 64-bit reduction keys in the first 256 bytes, then the staged model
// arrays (vp vs den, or vs den for SH), at most kNoAllocStage bytes.  The
// host name `scratch` is a stand-in for the device array, never a device
// address: each device's address comes from hipGetSymbolAddress with that
// device current, once, and only that pointer is passed to copies and
// kernels.  Every call on a GPU shares that GPU's array, so calls that
// overlap on one GPU race; the keys are cleared by a kernel at each call.
// ---------------------------------------------------------------------------
static constexpr size_t kNoAllocKeys = 256;
static constexpr size_t kNoAllocStage = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocKeys + kNoAllocStage];

static std::mutex g_scratch_mu;
static std::vector<char *> g_scratch_dev; // per device ordinal

static int scratch_on_device(int dev, char **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_dev.size() <= dev) g_scratch_dev.resize(dev + 1, nullptr);
    if (!g_scratch_dev[dev]) {
        int cur = -1;
        MCHK(hipGetDevice(&cur));
        if (cur != dev) MCHK(hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) MCHK(hipSetDevice(cur));
        if (e != hipSuccess || !p)
            return merr(FD23_EHIP, std::string("set_model_noalloc: scratch address: ") + hipGetErrorString(e));
        g_scratch_dev[dev] = (char *)p;
    }
    *out = g_scratch_dev[dev];
    return FD23_OK;
}

extern "C" size_t fd23_set_model_noalloc_capacity(void) { return kNoAllocStage; }

extern "C" int fd23_set_model_noalloc(fd23_ctx *c, const void *vp, const void *vs, const void *den, double h,
                                      double dt, double *stab) {
    if (!c || !vs || !den) return merr(FD23_EINVAL, "set_model_noalloc: null argument");
    const fd23_desc &d = fd23_ctx_desc(c);
    const bool psv = d.wave == FD23_WAVE_PSV;
    if (psv && !vp) return merr(FD23_EINVAL, "set_model_noalloc: null vp");
    int r = check_model_ctx(c);
    if (r) return r;
    const bool f64 = psv && d.storage == FD23_STORE_F64;
    const size_t N = (size_t)d.nx * d.nz, es = f64 ? 8 : 4;
    const int nin = psv ? 3 : 2;
    if (N * es * nin > kNoAllocStage)
        return merr(FD23_EINVAL, "set_model_noalloc: the model needs " + std::to_string(N * es * nin) +
                                     " bytes of staging, above the fixed scratch of " +
                                     std::to_string(kNoAllocStage) + " bytes (use fd23_set_model)");
    hipStream_t st = fd23_ctx_stream(c);
    int dev = -1;
    MCHK(hipStreamGetDevice(st, &dev));
    char *buf = nullptr;
    if ((r = scratch_on_device(dev, &buf))) return r;
    const void *src[3] = {psv ? vp : vs, psv ? vs : den, den};
    for (int i = 0; i < nin; i++)
        MCHK(hipMemcpyAsync(buf + kNoAllocKeys + i * N * es, src[i], N * es, hipMemcpyHostToDevice, st));
    return material_from_device(c, buf + kNoAllocKeys, (unsigned long long *)buf, h, dt, stab);
}

// a .mdl model rasterised on the device, then its material; raster files
// (model.vp/.vs/.den) are read on the host and go through fd23_set_model
extern "C" int fd23_set_model_file(fd23_ctx *c, const char *model, double h, double dt, double *stab) {
    if (!c || !model) return merr(FD23_EINVAL, "set_model_file: null argument");
    int r = check_model_ctx(c);
    if (r) return r;
    const fd23_desc &d = fd23_ctx_desc(c);
    const bool psv = d.wave == FD23_WAVE_PSV;
    const bool f64 = psv && d.storage == FD23_STORE_F64;
    const size_t N = (size_t)d.nx * d.nz;
    const std::string nm(model);
    if (nm.size() < 4 || nm.compare(nm.size() - 4, 4, ".mdl") != 0) {
        std::vector<float> v(N), s_(N), rho(N);
        if ((r = fd23_model_read(model, d.nx, d.nz, h, v.data(), s_.data(), rho.data()))) return r;
        if (!f64) return fd23_set_model(c, v.data(), s_.data(), rho.data(), h, dt, stab);
        std::vector<double> a(v.begin(), v.end()), b(s_.begin(), s_.end()), e(rho.begin(), rho.end());
        return fd23_set_model(c, a.data(), b.data(), e.data(), h, dt, stab);
    }
    MdlFlat m;
    if ((r = fd23_mdl_flatten(model, d.nx, d.nz, h, m))) return r;
    const int np = (int)m.vp.size();
    // polygon tables: 2 vertex arrays, 3 property arrays, 6 int arrays
    std::vector<char> tab;
    auto put = [&](const void *p, size_t bytes) {
        const size_t o = tab.size();
        tab.resize(o + ((bytes + 255) & ~(size_t)255));
        if (bytes) memcpy(tab.data() + o, p, bytes);
        return o;
    };
    const size_t o_vx = put(m.vx.data(), m.vx.size() * 4), o_vz = put(m.vz.data(), m.vz.size() * 4);
    const size_t o_vp = put(m.vp.data(), np * 4), o_vs = put(m.vs.data(), np * 4), o_den = put(m.den.data(), np * 4);
    const size_t o_f = put(m.first.data(), np * 4), o_c = put(m.count.data(), np * 4);
    const size_t o_x1 = put(m.ix1.data(), np * 4), o_x2 = put(m.ix2.data(), np * 4);
    const size_t o_z1 = put(m.iz1.data(), np * 4), o_z2 = put(m.iz2.data(), np * 4);
    const size_t es = f64 ? 8 : 4;
    const size_t o_tab = kIn + 3 * N * es; // after the keys and the three rasterised arrays
    hipStream_t st = fd23_ctx_stream(c);
    char *buf = nullptr;
    MCHK(hipMalloc((void **)&buf, o_tab + tab.size()));
    hipError_t e = hipMemcpyAsync(buf + o_tab, tab.data(), tab.size(), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) {
        hipStreamSynchronize(st);
        hipFree(buf);
        return merr(FD23_EHIP, std::string("set_model_file copy: ") + hipGetErrorString(e));
    }
    const char *t = buf + o_tab;
    MdlDev md{(const float *)(t + o_vx), (const float *)(t + o_vz), (const float *)(t + o_vp),
              (const float *)(t + o_vs), (const float *)(t + o_den), (const int *)(t + o_f),
              (const int *)(t + o_c), (const int *)(t + o_x1), (const int *)(t + o_x2),
              (const int *)(t + o_z1), (const int *)(t + o_z2), np};
    // rasterised vp vs den at buf + kIn (SH: vs den there, vp after them,
    // the layout material_from_device reads)
    char *a0 = buf + kIn, *a1 = a0 + N * es, *a2 = a1 + N * es;
    char *ovp = psv ? a0 : a2, *ovs = psv ? a1 : a0, *oden = psv ? a2 : a1;
    const int block = 256, grid = (int)((N + block - 1) / block);
    const double hh = (float)h;
    if (f64)
        hipLaunchKernelGGL(k_mdl_raster<double>, dim3(grid), dim3(block), 0, st, md, d.nx, d.nz, hh, (double *)ovp,
                           (double *)ovs, (double *)oden);
    else
        hipLaunchKernelGGL(k_mdl_raster<float>, dim3(grid), dim3(block), 0, st, md, d.nx, d.nz, hh, (float *)ovp,
                           (float *)ovs, (float *)oden);
    return material_from_staging(c, buf, kIn, h, dt, stab);
}
