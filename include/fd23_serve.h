/* fd23_serve.h - request-serving engine over the fd23 time-stepping loop.
 *
 * The reference runs one solve per process (PSV/clean_nobox/main.cpp:
 * 24-170, SH/main.cpp:20-200); it has no serving layer, so nothing here
 * replaces a reference entry point.  The engine keeps many solves resident
 * on one GPU at once:
 *
 *  - admission by free HBM: a request (or a batch of compatible requests)
 *    is admitted when its device bytes (fd23_desc_device_bytes plus the
 *    box residual grid) fit in the budget that running solves leave free
 *    and a stream slot is idle; requests that do not fit wait, smaller
 *    ones behind them may pass;
 *  - batching: queued requests with the same non-zero model_key and the
 *    same grid, time step, wavelet, moment tensor and box are merged into
 *    one context (one launch steps all their shots, reusing the material);
 *    their numbers of steps may differ;
 *  - joining: a request that arrives while a compatible context runs
 *    takes that context's idle shot slots (left by requests that have
 *    finished) at the next chunk boundary, its shots starting from rest at
 *    that step (fd23_set_shot_slot); contexts with a box take no joins;
 *  - continuous scheduling: every resident context advances in chunks of
 *    chunk_steps graph-replayed steps on its own HIP stream, at most two
 *    chunks queued at a time and each ending where a request ends, so
 *    requests that arrive while others run are admitted at the next chunk
 *    boundary and their work overlaps on the device; a request's
 *    seismograms are copied out as soon as its last step completes, and a
 *    context is retired, its memory returned, when its last request ends.
 *
 * All device work runs on the engine's own thread; submit/wait/result are
 * safe to call from any thread.  Requests with FP16/BF16 storage and a box
 * run at FP32 storage (the box ring works on FP32/FP64 fields).
 */
#ifndef FD23_SERVE_H
#define FD23_SERVE_H

#include <stddef.h>

#include "fd23.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fd23_server fd23_server;

typedef struct fd23_server_desc {
    size_t hbm_budget;   /* device bytes the engine may hold; 0 = 90% of free */
    int max_active;      /* resident contexts (one stream each), 1..64       */
    int chunk_steps;     /* steps per scheduling chunk (>= 1)                */
    int max_batch_shots; /* shots merged into one context (>= 1)             */
    int autotune;        /* 1 = tune the tile once per (grid, shots) shape   */
    int spare_shots;     /* idle shot slots a new context keeps for requests
                          * that join it while it runs (requests with a
                          * non-zero model_key and no box; a context holds
                          * at most 64 shots with spares); 0 = none         */
} fd23_server_desc;

/* One solve.  All pointers are read during fd23_server_submit only (the
 * engine keeps copies).  Receivers: nrec per component, the same list for
 * every shot of the request. */
typedef struct fd23_request {
    int wave, storage;          /* FD23_WAVE_*, FD23_STORE_*                */
    int nx, nz, npml, topabsorb;
    int nshots, nt, itrecord;
    double dx, dt;
    const void *const *material; /* as fd23_set_material (f64 for F64)     */
    const double *cpml;          /* fd23_cpml_pack layout                  */
    int src_kind;                /* FD23_SRC_*, or -1 for no point source  */
    const int *xs, *zs;          /* nshots source cells                    */
    const double *stf;           /* lsrc samples, shared by all shots      */
    int lsrc;
    const double *M;             /* 3 moment-tensor factors                */
    int nrec;
    const int *rx, *rz;
    int box;                     /* 1 = plane-wave box (P-SV)              */
    int box_x0, box_z0, box_bx, box_bz, box_incident;
    double box_angle, box_f0, box_t0, box_amp, box_vp, box_vs, box_rho;
    unsigned long long model_key; /* equal non-zero keys: same material     */
    /* vp, vs, den (nx*nz, double for F64 P-SV, else float; SH reads vs and
     * den) in place of material (which is then NULL): the engine computes
     * the material on the device at admission (fd23_set_model) with dx, dt */
    const void *const *model;
} fd23_request;

#define FD23_REQ_QUEUED 0
#define FD23_REQ_RUNNING 1
#define FD23_REQ_DONE 2
#define FD23_REQ_FAILED 3

typedef struct fd23_server_stats {
    long long submitted, completed, failed, queued, running;
    long long launches;       /* contexts created                          */
    long long merged;         /* requests that shared a context            */
    long long peak_active;    /* most contexts resident at once            */
    size_t bytes_in_use, budget;
    double cell_updates;      /* completed shots * nx * nz * nt            */
    long long joined;         /* requests that entered a running context   */
} fd23_server_stats;

int fd23_server_create(const fd23_server_desc *desc, fd23_server **out);
/* stops the engine after the work in flight; queued requests fail       */
int fd23_server_destroy(fd23_server *srv);
int fd23_server_submit(fd23_server *srv, const fd23_request *req, long long *id);
/* waits up to timeout_s (<0: forever); *state = FD23_REQ_*               */
int fd23_server_wait(fd23_server *srv, long long id, double timeout_s, int *state);
/* bytes of the result: nshots * ncomp * ceil(nt/itrecord) * nrec values
 * of the compute type (float, double for F64)                           */
size_t fd23_server_result_bytes(fd23_server *srv, long long id);
/* copies the seismograms, layout [shot][comp][frame][rec], and releases
 * the request; latency_s = submit to completion, queue_s = submit to
 * admission.  A failed request returns its error code.                  */
int fd23_server_result(fd23_server *srv, long long id, void *out, size_t bytes, double *latency_s,
                       double *queue_s);
int fd23_server_stats_get(fd23_server *srv, fd23_server_stats *out);

#ifdef __cplusplus
}
#endif
#endif /* FD23_SERVE_H */
