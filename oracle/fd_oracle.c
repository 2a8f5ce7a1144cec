/*
 * fd23 oracle: a CPU restatement of the reference's staggered-grid
 * time-stepping loop (GPU-kernel semantics) in FP32 and FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  It is imported by tests/, by
 * __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline
 * leg; the product library never links, loads or calls it.
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -shared).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define REAL float
#define SFX _f32
#include "fd_oracle_impl.h"
#undef REAL
#undef SFX

#define REAL double
#define SFX _f64
#include "fd_oracle_impl.h"
#undef REAL
#undef SFX

/* Staggered-position Ricker wavelet, amplitude 1 at t0 (not in the
 * reference; the synthetic workloads use it). */
void or_ricker(int n, double dt, double f0, double t0, double *out)
{
    const double PI_ = 3.14159265358979323846;
    for (int i = 0; i < n; i++) {
        double a = PI_ * f0 * (i * dt - t0);
        a = a * a;
        out[i] = (1.0 - 2.0 * a) * exp(-a);
    }
}

/* Reduced-precision storage emulation (the FP16/BF16 error floor): round
 * the n doubles of a in place as they would be stored with one
 * power-of-two scale s = 2^(floor(log2 max|a|) - 8), which puts the largest
 * magnitude in [2^8, 2^9): fmt 1 = FP16 (11 significant bits, round to
 * nearest even, subnormals below 2^-14 spaced 2^-24), fmt 2 = BF16 (the
 * value first rounded to float, then to 8 significant bits, nearest even,
 * as a float -> bfloat16 conversion), fmt 0 = float.  Not in the
 * reference; the same rounding as tests/helpers.py round_storage, checked
 * against it. */
static double or_round_f16(double y)
{
    if (y == 0.0 || !isfinite(y)) return y;
    int e;
    frexp(y, &e);            /* |y| in [2^(e-1), 2^e) */
    int ee = e - 1;
    if (ee < -14) ee = -14;  /* subnormal range: fixed spacing */
    const double q = ldexp(1.0, ee - 10);
    return nearbyint(y / q) * q;
}

static double or_round_bf16(double y)
{
    float f = (float)y;
    unsigned int b;
    memcpy(&b, &f, 4);
    b = (unsigned int)(((unsigned long long)b + 0x7FFFu + ((b >> 16) & 1u)) & 0xFFFF0000u);
    memcpy(&f, &b, 4);
    return (double)f;
}

/* Block-scaled int16 (fmt 3): each block of 64 consecutive values shares
 * one power-of-two quantum q putting the block's largest magnitude in
 * [2^14, 2^15) q; values round to the nearest multiple of q (ties to
 * even) within |v| <= 32767 q.  A block whose largest magnitude is below
 * 1e-280 (beneath any exponent range such a format would carry) is
 * flushed to zero. */
static void or_round_block16(double *a, long n)
{
#pragma omp parallel for
    for (long b = 0; b < n; b += 64) {
        const long e_ = b + 64 < n ? b + 64 : n;
        double m = 0.0;
        for (long i = b; i < e_; i++) {
            const double v = fabs(a[i]);
            if (v > m) m = v;
        }
        if (m == 0.0) continue;
        if (m < 1e-280) {
            for (long i = b; i < e_; i++) a[i] = 0.0;
            continue;
        }
        int e;
        frexp(m, &e); /* m in [2^(e-1), 2^e) */
        const double q = ldexp(1.0, e - 15);
        for (long i = b; i < e_; i++) {
            double r = nearbyint(a[i] / q);
            if (r > 32767.0) r = 32767.0;
            if (r < -32767.0) r = -32767.0;
            a[i] = r * q;
        }
    }
}

void or_round_storage(double *a, long n, int fmt)
{
    if (fmt == 3) {
        or_round_block16(a, n);
        return;
    }
    double m = 0.0;
#pragma omp parallel for reduction(max : m)
    for (long i = 0; i < n; i++) {
        const double v = fabs(a[i]);
        if (v > m) m = v;
    }
    if (m == 0.0) return;
    if (fmt == 0) {
#pragma omp parallel for
        for (long i = 0; i < n; i++) a[i] = (double)(float)a[i];
        return;
    }
    int e;
    frexp(m, &e);
    const double s = ldexp(1.0, e - 1 - 8), inv = ldexp(1.0, -(e - 1 - 8));
#pragma omp parallel for
    for (long i = 0; i < n; i++) {
        const double y = a[i] * inv;
        a[i] = (fmt == 1 ? or_round_f16(y) : or_round_bf16(y)) * s;
    }
}

/* number of OpenMP threads the passes use (for the CPU baseline report) */
#ifdef _OPENMP
#include <omp.h>
int or_num_threads(void) { return omp_get_max_threads(); }
void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
#else
int or_num_threads(void) { return 1; }
void or_set_threads(int n) { (void)n; }
#endif
