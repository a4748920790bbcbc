/* Serving loop over the LC2 engine (lc2e.h): many client threads submit candidate pose
 * sets for resident pairs. One worker thread merges the queued requests of a pair into a
 * single fused launch and keeps launches of different pairs in flight on the device at
 * once, each on its own stream (a launch slot), while their work-groups fit the device
 * together; a launch that fills the GPU by itself runs alone. Requests that arrive while
 * others are on the GPU join the next launch of their pair as a slot and room on the
 * device free up (continuous batching).
 *
 * No counterpart in the reference: it evaluates one pose at a time from one optimizer
 * (optimizers.cu:130-177, :650-728) and has no request path. Its closest component, the
 * OpenIGTLink pose server (interpolator.cu:359-388), streams results and does not batch.
 *
 * Admission is by bytes. Volumes and pairs are created through lc2e.h against the engine's
 * HBM budget (LC2E_ERR_BUDGET when they do not fit), and so are the slots' staging buffers.
 * If a merged launch does not fit the budget, its requests are evaluated one by one.
 * Pairs are served in the order of their oldest queued request, and a pair's requests are
 * launched in the order they arrived; a pair has at most one launch in flight. Each request's values are those of lc2e_eval_batch on its own poses:
 * they do not depend on what it was merged with, and its status is its own.
 * While a server runs, the pairs it serves must not be evaluated through lc2e.h (the
 * engine is not re-entrant); other pairs and volumes may be created and destroyed. A pair
 * is destroyed only after its requests are collected, and the server only after every
 * lc2s_wait has returned. Status codes and lc2e_last_error() are those of lc2e.h. */
#ifndef LC2SERVE_H
#define LC2SERVE_H

#include "lc2e.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lc2s_server lc2s_server;

#define LC2S_MAX_SLOTS 8

typedef struct lc2s_counters {
    long long launches;  /* fused launches, including refused merged ones */
    long long requests;  /* completed requests */
    long long poses;     /* evaluated poses */
    long long queued;    /* requests waiting for a launch now */
    int max_in_flight;   /* most launches on the device at once so far */
} lc2s_counters;

/* max_poses: largest merged launch (poses); slots: launches in flight at once
 * (1..LC2S_MAX_SLOTS); window_us: how long a pair's oldest request waits for more requests
 * of its pair before it is launched (0: launch as soon as a slot is free). */
int lc2s_create(lc2e_engine* e, int max_poses, int slots, double window_us, lc2s_server** out);
/* Completes every accepted request, then stops the worker. */
void lc2s_destroy(lc2s_server* s);

/* Queue n poses (n * mats_per_pose 3x4 matrices, copied) for pair p; *ticket identifies
 * the request. */
int lc2s_submit(lc2s_server* s, lc2e_pair* p, int n, const float* mats, long long* ticket);
/* Wait up to timeout_ms (< 0: no limit) for a request; on completion copies its n LC2
 * values (and, if not NULL, sum of measures and of weights) and forgets the ticket.
 * LC2E_ERR_INVALID for an unknown ticket, LC2E_ERR_UNSUPPORTED on time-out (the request
 * stays queued). A launch that fails reports its status and lc2e_last_error() text to
 * every request merged into it, on the thread that waits. */
int lc2s_wait(lc2s_server* s, long long ticket, double timeout_ms, double* lc2,
              double* sum_measure, double* sum_w);

/* Counters since creation. */
int lc2s_stats(lc2s_server* s, lc2s_counters* out);

#ifdef __cplusplus
}
#endif

#endif /* LC2SERVE_H */
