/*
 * hos_serve.h -- continuous-batching request loop for many HOS simulations
 * on one GPU (host C++ over hos_gpu.h).  No counterpart in the reference,
 * which runs one simulation per MPI job (Csource/2dpar/main_2d_par.c).
 *
 * Requests (grid, order, depth, dt, number of RK steps, initial state) are
 * queued by hos_server_submit.  The scheduler groups them into pools, one
 * hos_ctx per (Nx, Ny, nlevs, Lx, Ly, g), with one slot per member:
 *   - admission: a queued request enters a free slot of its pool while the
 *     pool has no steps in flight; a pool is created when its device
 *     footprint (hos_footprint) fits the remaining HBM budget (empty pools
 *     are released first when it does not), otherwise the request waits
 *     until memory is released;
 *   - each pool runs on its own stream and is launched on its own: an idle
 *     pool with active members gets k = min(steps_per_launch, the fewest
 *     steps any of its members still needs) steps enqueued at once
 *     (hos_step_async), and the scheduler polls the pools (hos_query)
 *     instead of waiting for each, so a slow shape holds up no other pool;
 *   - when a pool's steps complete, a member that has done its steps is
 *     copied out, its slot is refilled by the pool's last active member
 *     (hos_copy_member), queued requests are admitted into the free slots,
 *     and the pool is launched again at once;
 *   - a pool whose launch or steps fail marks its requests HOS_REQ_FAILED
 *     (their steps are not counted as done) and is released.
 * Each member's arithmetic does not depend on the other members, so a
 * result is bit-identical whether its request ran alone or in a mix.
 *
 * All functions are thread-safe; hos_server_start runs the scheduler on a
 * background thread, or the caller drives it with hos_server_step.
 */
#ifndef HOS_SERVE_H
#define HOS_SERVE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hos_server hos_server;

typedef struct hos_server_params {
    double hbm_budget;  /* bytes of device memory the pools may use          */
    int precision;      /* HOS_FP32 or HOS_FP64                              */
    int use_graph;      /* replay one captured step per pool (1)             */
    int max_slots;      /* members per pool; 0 = by grid size (2..64)        */
    int compensated;    /* FP32: Kahan order sums                            */
    int steps_per_launch; /* most steps enqueued per pool launch; 0 = 8       */
} hos_server_params;

typedef struct hos_request {
    int Nx, Ny, nlevs;
    double Lx, Ly, g;
    double depth;  /* <= 0: deep water */
    double dt;
    int nsteps;    /* RK4 steps */
} hos_request;

/* request states */
#define HOS_REQ_QUEUED 0
#define HOS_REQ_RUNNING 1
#define HOS_REQ_DONE 2
#define HOS_REQ_FAILED 3

int hos_server_create(const hos_server_params *p, hos_server **out);
int hos_server_destroy(hos_server *s);

/* Queue a request; u0 is the initial state [2][Ny/2+1][Nx] complex double
 * (copied).  Returns HOS_E_INVALID at once for a request that can never fit
 * the budget or has an invalid shape. */
int hos_server_submit(hos_server *s, const hos_request *r, const double *u0, long long *id);

/* HOS_REQ_* of a request, or a negative error for an unknown id. */
int hos_server_status(hos_server *s, long long id);

/* Copy a finished request's state into u (may be NULL) and its record into
 * info[6] = {submit, admit, done (seconds since the server started), steps,
 * H, volume of the last step}; then forget the request. */
int hos_server_result(hos_server *s, long long id, double *u, double *info);

/* One scheduler iteration: collect the pools whose steps have completed,
 * retire their finished members, admit queued requests, and launch every idle
 * pool that has active members.  Does not wait for steps in flight.  *active
 * receives the number of running members. */
int hos_server_step(hos_server *s, int *active);

/* Run the scheduler on a background thread / stop it. */
int hos_server_start(hos_server *s);
int hos_server_stop(hos_server *s);

/* Block until no request is queued or running, or timeout_s passes
 * (returns HOS_E_STATE on timeout). */
int hos_server_drain(hos_server *s, double timeout_s);

/* Statistics: out[0..11] = {step p50 ms, step p99 ms, request latency p50 s,
 * request p99 s, completed requests, grid-point RK4 steps, busy seconds,
 * pools created, peak bytes in pools, peak running members, queue-wait p50 s,
 * queue-wait p99 s}. */
int hos_server_stats(hos_server *s, double *out, int n);

/* Failure injection for tests of the error path: the next n pool launches
 * fail on the host (nothing is enqueued), as a failed hos_step_async would. */
int hos_server_fail_launches(hos_server *s, int n);

#ifdef __cplusplus
}
#endif
#endif
