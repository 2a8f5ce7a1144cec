// Device-side building blocks for the fd23 staggered-grid loop on gfx950.
//
// Layout (one array per field, all shots of a context back to back):
//   shot s, row z, column x  ->  base + s*sstride + z*px + x
// with px = round_up(nx + 4, 64) and 4 zero rows above row 0 and at least 4
// below row nz-1.  Columns nx..px-1 and the extra rows are never written,
// so every tap outside [0,nx) x [0,nz) reads 0 with no bounds test -- the
// same "zero outside the domain" the reference gets from its zero-filled
// LDS halo (PSV/clean_nobox/fd/cu_fd.cu:72-91).  Column -1 of row z is
// column px-1 of row z-1, which is padding.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace fd23 {

constexpr int TX = 64;  // tile width = one wavefront across x
constexpr int BY = 4;   // wavefronts per workgroup (threadIdx.y)
constexpr int HALO = 4; // 8th-order staggered stencil reach

// ---------------------------------------------------------------------------
// storage traits: S is the element type in HBM, C the arithmetic type.
// Scaled types hold value/scale; scale is a power of two per array.
// ---------------------------------------------------------------------------
// Raw buffer resource over one array, for loads and stores addressed by a
// per-lane byte offset plus a wave-uniform byte offset that stays in a
// scalar register: an access then needs no vector address arithmetic.
// Word 3 is gfx9's (32-bit data format, no swizzle, no thread-id add); the
// range is 2 GiB, far above one shot's array, so no access is clipped.
using BufRsrc = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ BufRsrc buf_rsrc(const void *origin) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(origin), (short)0, 0x7fffffff, 0x00020000);
}

// ld/st: element i of an array; ldb/stb: the same through a buffer
// resource at byte offsets vo (per lane) + so (wave-uniform)
template <class S> struct Store;
template <> struct Store<float> {
    using C = float;
    static constexpr bool scaled = false;
    __device__ __forceinline__ static float ld(const float *p, long i, float) { return p[i]; }
    __device__ __forceinline__ static void st(float *p, long i, float v, float) { p[i] = v; }
    __device__ __forceinline__ static float ldb(BufRsrc r, unsigned vo, unsigned so, float) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
    }
    __device__ __forceinline__ static void stb(BufRsrc r, unsigned vo, unsigned so, float v, float) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, vo, so, 0);
    }
};
template <> struct Store<double> {
    using C = double;
    static constexpr bool scaled = false;
    __device__ __forceinline__ static double ld(const double *p, long i, float) { return p[i]; }
    __device__ __forceinline__ static void st(double *p, long i, double v, float) { p[i] = v; }
    __device__ __forceinline__ static double ldb(BufRsrc r, unsigned vo, unsigned so, float) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
    }
    __device__ __forceinline__ static void stb(BufRsrc r, unsigned vo, unsigned so, double v, float) {
        using U2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, vo, so, 0);
    }
};
template <> struct Store<__half> {
    using C = float;
    static constexpr bool scaled = true;
    __device__ __forceinline__ static float ld(const __half *p, long i, float s) {
        return __half2float(p[i]) * s;
    }
    __device__ __forceinline__ static void st(__half *p, long i, float v, float inv) {
        p[i] = __float2half_rn(v * inv);
    }
    __device__ __forceinline__ static float ldb(BufRsrc r, unsigned vo, unsigned so, float s) {
        const unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(r, vo, so, 0);
        return __half2float(__builtin_bit_cast(__half, b)) * s;
    }
    __device__ __forceinline__ static void stb(BufRsrc r, unsigned vo, unsigned so, float v, float inv) {
        const __half h = __float2half_rn(v * inv);
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, h), r, vo, so, 0);
    }
};
template <> struct Store<__hip_bfloat16> {
    using C = float;
    static constexpr bool scaled = true;
    __device__ __forceinline__ static float ld(const __hip_bfloat16 *p, long i, float s) {
        return __bfloat162float(p[i]) * s;
    }
    __device__ __forceinline__ static void st(__hip_bfloat16 *p, long i, float v, float inv) {
        p[i] = __float2bfloat16(v * inv);
    }
    __device__ __forceinline__ static float ldb(BufRsrc r, unsigned vo, unsigned so, float s) {
        const unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(r, vo, so, 0);
        return __bfloat162float(__builtin_bit_cast(__hip_bfloat16, b)) * s;
    }
    __device__ __forceinline__ static void stb(BufRsrc r, unsigned vo, unsigned so, float v, float inv) {
        const __hip_bfloat16 h = __float2bfloat16(v * inv);
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, h), r, vo, so, 0);
    }
};

// 8th-order coefficients by reduced order, PSV/clean_nobox/param/const.h:13-18
template <class C> struct Coef {
    __device__ __forceinline__ static void get(int ord, C &c0, C &c1, C &c2, C &c3) {
        c0 = ord == 4 ? C(1225. / 1024.) : ord == 3 ? C(75. / 64.) : ord == 2 ? C(9. / 8.) : C(1.0);
        c1 = ord == 4 ? C(245. / 3072.) : ord == 3 ? C(25. / 384.) : ord == 2 ? C(1. / 24.) : C(0.0);
        c2 = ord == 4 ? C(49. / 5120.) : ord == 3 ? C(3. / 640.) : C(0.0);
        c3 = ord == 4 ? C(5. / 7168.) : C(0.0);
    }
};

// reduced order of the reference (cu_fd.cu:165-169): the same scalar for
// both derivative directions, so the top rows run at low order in x too.
__device__ __forceinline__ int order_at(int ix, int iz, int nx, int nz) {
    int o = min(ix, iz);
    o = min(o, nx - 1 - ix);
    o = min(o, nz - 1 - iz);
    return min(o, 4);
}

// Every multiply-add in the kernels is an explicit fused mad and implicit
// contraction is off (fd_kernels.hpp), so a cell's result does not depend on
// how a tile shape unrolls: all tile shapes and batch sizes are bitwise
// identical.
__device__ __forceinline__ float mad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double mad(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Derivative from a strided LDS window: "m" form centres between taps i-1
// and i (UDIFF ux, DXDZTXZ dxTxz, DZTZZ), "p" form between i and i+1
// (WDIFF wx, DXTXX, DZTXZ).  Same association as the reference macros,
// ((c0 d0 - c1 d1) + c2 d2) - c3 d3, each step one fused mad.
template <class C>
__device__ __forceinline__ C dm(const C *f, int s, C c0, C c1, C c2, C c3) {
    C r = c0 * (f[0] - f[-s]);
    r = mad(-c1, f[s] - f[-2 * s], r);
    r = mad(c2, f[2 * s] - f[-3 * s], r);
    return mad(-c3, f[3 * s] - f[-4 * s], r);
}
template <class C>
__device__ __forceinline__ C dp(const C *f, int s, C c0, C c1, C c2, C c3) {
    C r = c0 * (f[s] - f[0]);
    r = mad(-c1, f[2 * s] - f[-s], r);
    r = mad(c2, f[3 * s] - f[-2 * s], r);
    return mad(-c3, f[4 * s] - f[-3 * s], r);
}

// 1/s for a power of two s = 2^e, |e| <= 126: the exponent negated, exact
// (the host only accepts power-of-two scales)
__device__ __forceinline__ float inv_pow2(float s) { return __uint_as_float(0x7f000000u - __float_as_uint(s)); }

// Largest of a non-negative value over the 64 lanes of the wavefront, in
// every lane.  Non-negative floats order like their bit patterns as
// integers.  DPP steps: swap lanes in pairs and in halves of each quad,
// mirror each half-row and each row (every lane then holds its row's max),
// then broadcast lane 15 into rows 1 and 3 and lane 31 into rows 2 and 3, so
// that lane 63 holds the wave's max.
__device__ __forceinline__ float wave_max_nonneg(float v) {
    int x = __float_as_int(v);
    x = max(x, __builtin_amdgcn_update_dpp(x, x, 0xb1, 0xf, 0xf, false));  // quad_perm [1,0,3,2]
    x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x4e, 0xf, 0xf, false));  // quad_perm [2,3,0,1]
    x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x141, 0xf, 0xf, false)); // row_half_mirror
    x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x140, 0xf, 0xf, false)); // row_mirror
    x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x142, 0xa, 0xf, false)); // row_bcast:15
    x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x143, 0xc, 0xf, false)); // row_bcast:31
    return __int_as_float(__builtin_amdgcn_readlane(x, 63));
}

// Running max of |v| per array (scaled storage): reduced across the 64
// lanes of the wavefront, then one atomicMax on the float bits.  Each
// array's maximum is spread over kAmaxSpread words picked by the workgroup
// index, so that the atomics of a pass do not all queue on one address; the
// scale check reduces the words.  amax is null on steps that no scale check
// follows: the maxima are only gathered on the step before each check.
constexpr int kAmaxSpread = 256;
__device__ __forceinline__ void amax_update(unsigned *amax, int slot, float v) {
    if (!amax) return;
    v = wave_max_nonneg(fabsf(v));
    if ((threadIdx.x & 63) == 0 && v > 0.f)
        atomicMax(&amax[slot * kAmaxSpread + (blockIdx.x & (kAmaxSpread - 1))], __float_as_uint(v));
}

// ---------------------------------------------------------------------------
// argument blocks
// ---------------------------------------------------------------------------
struct Grid {
    int nx, nz, px;     // domain and row pitch (elements)
    long sstride;       // elements between shots in a field array
    int npml, top;      // C-PML width, top absorbing flag
    int nshots;
    long xpsi_stride;   // elements between shots in an x-psi array (2*npml*nz)
    long zpsi_stride;   // ... in a z-psi array (2*npml*nx)
    // first-arrival window (null: every cell is updated every step): per
    // block of 32 x 16 cells, a cell is updated at step it only when
    // pfirst < it < plast of its block (SH/fd/cu_fd.cu:18-25)
    const int *pfirst, *plast;
    int pnxb;           // blocks per row, nx / 32
};

constexpr int kPunchBX = FD23_PUNCH_BX, kPunchBZ = FD23_PUNCH_BZ; // block of the window, SH/param/const.h:20-21
// ix < nx and iz < nz (nx and nz are multiples of the block)
__device__ __forceinline__ bool punch_open(const Grid &g, int ix, int iz, int it) {
    const int b = (int)((unsigned)iz / kPunchBZ) * g.pnxb + (int)((unsigned)ix / kPunchBX);
    return g.pfirst[b] < it && it < g.plast[b];
}

template <class S> struct PsvFields {
    S *U, *W, *Txx, *Tzz, *Txz; // point at shot 0, row 0, column 0
};
template <class S> struct ShFields {
    S *V, *Tx, *Tz;
};
// C-PML profile triple (b, c, k=1/kappa)
template <class C> struct Prof {
    const C *b, *c, *k;
};

} // namespace fd23
