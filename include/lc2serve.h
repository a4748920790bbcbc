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
 * HBM budget (LC2E_ERR_BUDGET when they do not fit), and so are the slots' matrix buffers
 * (results land in mapped pinned host memory).
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
    long long graph_launches; /* launches replayed from a captured HIP graph */
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

/* Graph replay of launches (off by default): each slot runs a (pair, poses) launch eagerly
 * the first time, captures it (matrix upload and kernels) as a HIP graph the
 * second time and replays the graph after that, until the pair's buffers or launch shape
 * change. Results are identical either way. Off by default because on MI355X with ROCm 7.2
 * a replayed launch took 5-13 us longer than an eager one (measured at 1 to 64 poses). */
int lc2s_use_graphs(lc2s_server* s, int enable);

/* ---- Whole registrations as requests ----------------------------------------------
 * A registrar runs registrations submitted by clients: each one is a pair built from two
 * resident volumes, an optimizer run over a pose vector (an optional DIRECT-L stage, then
 * BOBYQA; the objective is -log(LC2), as the reference's Optimizers* minimise,
 * optimizers.cu:650-728), and the pair's destruction. Every candidate set the optimizer
 * asks for goes to the serving loop above as one request, so evaluations of concurrent
 * registrations are batched and overlapped like any other requests. A registration's
 * result does not depend on what runs beside it.
 *
 * Admission: registrations start in submission order while fewer than max_sessions run
 * and their pair fits the engine's HBM budget (lc2e_pair_create). A registration whose
 * pair does not fit waits until a running one finishes and frees its pair; if it does not
 * fit with nothing else running, it fails with LC2E_ERR_BUDGET.
 *
 * No counterpart in the reference, which runs one registration per process
 * (main.cu:728-827). */
typedef struct lc2s_registrar lc2s_registrar;

typedef struct lc2s_registration {
    lc2e_volume* fixed;       /* resident volumes; they must outlive the registration */
    lc2e_volume* moving;
    float resampled_spacing[3];
    int high_res_is_moving, center_overlaid, patch_radius, dim;  /* as lc2e_pair_create */
    /* Tracked frames (a 2D-to-3D frame stack with one matrix per frame): n_frames > 0
     * gives each frame's VirtualToMoving 3x4 at the identity pose (n_frames x 12 floats,
     * copied); the pose is then a rigid 6-vector (TX, TY, TZ mm, RX, RY, RZ deg) A about
     * frames_centre and frame k evaluates A * F_k, composed in FP64 and rounded once.
     * n_frames = 0: the pose is the pair's parameter vector (lc2e_pose_to_matrix). */
    int n_frames;
    const float* frames;
    float frames_centre[3];
    int n_params;                          /* pose length */
    double x0[12], lower[12], upper[12];   /* start and absolute bounds */
    int direct_maxeval;   /* > 0: a DIRECT-L stage over [lower, upper] first */
    double xtol_rel;      /* BOBYQA: relative x tolerance */
    int maxeval;          /* BOBYQA: evaluation limit (0: none) */
} lc2s_registration;

typedef struct lc2s_registration_result {
    double x[12];         /* optimum pose */
    double lc2;           /* LC2 there */
    int evaluations;      /* LC2 evaluations (DIRECT + BOBYQA) */
    int status;           /* LC2E_OK or a negative status */
    double admit_ms;      /* from submission to admission */
    double run_ms;        /* from admission to result */
} lc2s_registration_result;

/* A registrar over a server (which must outlive it) and its engine; max_sessions >= 1
 * registrations run at once, each on its own host thread. */
int lc2s_registrar_create(lc2s_server* s, lc2e_engine* e, int max_sessions, lc2s_registrar** out);
/* Completes every accepted registration, then stops. */
void lc2s_registrar_destroy(lc2s_registrar* r);
int lc2s_register_submit(lc2s_registrar* r, const lc2s_registration* reg, long long* ticket);
/* Wait up to timeout_ms (< 0: no limit); LC2E_ERR_UNSUPPORTED on time-out, LC2E_ERR_INVALID
 * for an unknown ticket. Returns the registration's status. */
int lc2s_register_wait(lc2s_registrar* r, long long ticket, double timeout_ms,
                       lc2s_registration_result* out);

typedef struct lc2s_registrar_counters {
    long long submitted, completed, failed;
    int running, max_running;
    long long budget_waits;   /* admissions deferred because the pair did not fit */
} lc2s_registrar_counters;
int lc2s_registrar_stats(lc2s_registrar* r, lc2s_registrar_counters* out);

#ifdef __cplusplus
}
#endif

#endif /* LC2SERVE_H */
