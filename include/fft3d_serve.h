/* Request server for the MI355X 3D FFT: continuous batching under a device-memory budget.
 *
 * No counterpart in the reference (it has no serving path; its nearest mechanisms are the
 * plan-owned stream ring and per-stream graph execs of bench/gpu/impl_4/L17_dmma.cu:8-16,
 * 984-1026).  Requests arrive at any time from any thread; a scheduler thread inside the
 * server
 *   * admits queued requests in arrival order while the in+out bytes of everything admitted
 *     and unfinished stay within the budget (outputs the caller does not provide are
 *     sub-allocated from a device arena of that size);
 *   * packs admitted requests of the fused geometries (6^3, 8^3, 12^3, 17^3, 8x8x16) and
 *     one precision, up to max_items at a time, into ONE mixed persistent launch, replayed
 *     through a hipGraph captured once per launch slot (the kernel reads its items and task
 *     list from a device descriptor, refreshed from pinned memory before each replay);
 *   * runs the other geometries (24^3, 36^3, 24x24x48) through cached plans;
 *   * keeps up to `streams` launches in flight, so requests admitted while others run
 *     overlap them on the device, and retires each launch when its event completes.
 * The plain modes serve the same requests with one plan launch each, in arrival order:
 * FFT3D_SRV_PLAIN_SYNC waits for each launch before the next; FFT3D_SRV_PLAIN_ASYNC keeps up
 * to `streams` of them in flight on the same streams, without packing.
 *
 * Times are seconds on the host's steady clock.  Latency of a request = completion time
 * observed by the scheduler - its arrival time. */
#ifndef FFT3D_SERVE_H
#define FFT3D_SERVE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fft3d_srv fft3d_srv;

/* fft3d_srv_config.batched */
enum {
    FFT3D_SRV_PLAIN_SYNC = 0,  /* one launch per request, each waited for before the next */
    FFT3D_SRV_BATCHED = 1,     /* continuous batching into mixed launches */
    FFT3D_SRV_PLAIN_ASYNC = 2  /* one launch per request, up to `streams` in flight */
};

typedef struct fft3d_srv_config {
    size_t budget_bytes; /* admission cap on in+out bytes of unfinished requests; also the
                            size of the output arena */
    int streams;         /* launches kept in flight (1..16) */
    int batched;         /* FFT3D_SRV_BATCHED, FFT3D_SRV_PLAIN_SYNC or FFT3D_SRV_PLAIN_ASYNC */
    int max_items;       /* requests per mixed launch (1..16) */
    double window_us;    /* batched mode: hold a partial mixed batch open for up to this long
                            after its oldest request was admitted, so later arrivals can join
                            (0: launch as soon as a stream is free; at most 1e6) */
} fft3d_srv_config;

typedef struct fft3d_srv_request {
    int n0, n1, n2;
    int batch;      /* 1..512 volumes */
    int direction;  /* FFT3D_MX_FORWARD / FFT3D_MX_BACKWARD */
    int precision;  /* FFT3D_MX_C128 / FFT3D_MX_C64 */
    const void *in; /* device input, batch volumes, resident until the request completes */
    void *out;      /* device output, or NULL: the server places it in its arena and frees
                       it at completion (for load tests that discard results) */
} fft3d_srv_request;

typedef struct fft3d_srv_stats {
    long submitted, completed;
    long launches, mixed_launches, plan_launches;
    double mean_items_per_mixed_launch;
    size_t peak_inflight_bytes;
    double busy_seconds; /* scheduler wall time with at least one launch in flight */
} fft3d_srv_stats;

/* Plans every supported geometry/precision/direction up front (plan creation includes the
 * create-time race) and captures the launch graphs.  NULL on failure. */
fft3d_srv *fft3d_srv_create(const fft3d_srv_config *cfg);

/* Queue a request that arrived at `arrival` (seconds, steady clock; <= 0 means now).
 * Returns its id (>= 0) or a negative FFT3D_MX_E* code. */
long fft3d_srv_submit(fft3d_srv *s, const fft3d_srv_request *r, double arrival);

/* Block until request `id` completes; its latency in seconds goes to *latency.  0 or error. */
int fft3d_srv_wait(fft3d_srv *s, long id, double *latency);

/* Open-loop replay: request i is submitted at t0 + arrival[i] (t0 = call time); returns
 * when all have completed, with each latency in latency[i].  0 or error. */
int fft3d_srv_run_trace(fft3d_srv *s, const fft3d_srv_request *reqs, const double *arrival,
                        int n, double *latency);

/* Steady-clock seconds, the time base of arrival and latency. */
double fft3d_srv_now(void);

int fft3d_srv_get_stats(const fft3d_srv *s, fft3d_srv_stats *st);

/* Waits for every submitted request, then frees everything. */
void fft3d_srv_destroy(fft3d_srv *s);

#ifdef __cplusplus
}
#endif

#endif /* FFT3D_SERVE_H */
