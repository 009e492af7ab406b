/*
 * sonic_amd_serve.h -- a request loop over the batch (sonic_amd.h,
 * sf_batch_*): solves arrive at any time from any thread, are admitted into
 * the device arena as its memory and problem slots free up (continuous
 * batching), advance together in the persistent batch kernels with their
 * state resident on the device between steps, and are handed back when
 * their step count is reached.  Each result is the same as the problem
 * solved alone.
 *
 * No reference counterpart: the reference solves one case per process
 * (app/application/SF_application.cpp:111-118).
 */
#ifndef SONIC_AMD_SERVE_H
#define SONIC_AMD_SERVE_H

#include "sonic_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sf_server sf_server;

typedef struct {
    long long submitted;     /* requests accepted by sf_server_submit          */
    long long completed;     /* results ready                                  */
    long long failed;        /* requests that ended with an error              */
    long long queued;        /* waiting for admission                          */
    long long running;       /* admitted, not finished                         */
    double p50_ms;           /* latency (submit -> result ready), completed    */
    double p99_ms;
    double mean_ms;
    double max_ms;
    double solves_per_s;     /* completed / (last completion - first submit)   */
    double cell_updates;     /* sum of nodes x steps over completed requests   */
    double busy_s;           /* time the loop spent advancing the batch        */
    double admit_s;          /* time spent admitting (state upload included)   */
    double retire_s;         /* time spent handing results back (download)     */
    long long ticks;         /* batch advances                                 */
    int peak_items;          /* most problems resident at once                 */
    size_t peak_nodes;       /* most arena nodes in use at once                */
    size_t arena_nodes;      /* arena capacity in padded nodes                 */
} sf_serve_stats;

/* A server with an arena of device_bytes of device memory (160 bytes per
 * padded node, sf_batch_item_nodes), at most max_items resident problems,
 * and at most steps_per_tick steps between admission points.  Starts its
 * loop thread. */
int sf_server_create(int device, double gamma, size_t device_bytes, int max_items,
                     int steps_per_tick, sf_server **out);
/* Stops the loop (requests not finished are dropped) and frees everything. */
int sf_server_destroy(sf_server *server);
/* Queue a solve of `steps` SSPRK3 steps; the state (soa[v*N + j*nx + i]) is
 * copied.  SF_E_MEMORY when the problem could never fit the arena. */
int sf_server_submit(sf_server *server, const sf_batch_item *item, const double *soa,
                     long long steps, long long *ticket);
/* *status: 0 pending, 1 done, -1 failed (the message is then returned by
 * sf_server_result). */
int sf_server_poll(sf_server *server, long long ticket, int *status);
/* Wait up to timeout_s seconds for the ticket; *status as for poll. */
int sf_server_wait(sf_server *server, long long ticket, double timeout_s, int *status);
/* Copy the result out and release the ticket.  time, steps and latency_ms
 * may be NULL. */
int sf_server_result(sf_server *server, long long ticket, double *soa, double *time,
                     long long *steps, double *latency_ms);
/* Wait until every submitted request has finished (or timeout_s passed;
 * then SF_E_STATE). */
int sf_server_drain(sf_server *server, double timeout_s);
int sf_server_stats(sf_server *server, sf_serve_stats *stats);
/* Launch shape of the batch (as sf_batch_set_tiles / _occupancy /
 * _schedule); applied before the loop's next pass.  The default is
 * xlen 255, ylen 32, 16 workgroups per CU, one launch per variant. */
int sf_server_set_shape(sf_server *server, int xlen, int ylen, int blocks_per_cu, int fused);
/* How each pass launches its steps, as sf_batch_step's use_graph: 0 direct
 * launches, 1 one replayed hipGraph per step, 2 (the default) a graph when
 * the pass is long enough to pay for capturing it.  Results do not depend
 * on it. */
int sf_server_set_launch(sf_server *server, int use_graph);
/* Forget completed latencies and counters (the queue is kept). */
int sf_server_reset_stats(sf_server *server);

#ifdef __cplusplus
}
#endif
#endif
