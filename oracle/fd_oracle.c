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

/* number of OpenMP threads the passes use (for the CPU baseline report) */
#ifdef _OPENMP
#include <omp.h>
int or_num_threads(void) { return omp_get_max_threads(); }
void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
#else
int or_num_threads(void) { return 1; }
void or_set_threads(int n) { (void)n; }
#endif
