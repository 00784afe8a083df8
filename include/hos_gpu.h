/*
 * hos_gpu.h -- C ABI of the MI355X (gfx950) HOS time-step library.
 *
 * Two layers:
 *
 *  1. The native API (hos_*): a context holds one or more solver instances
 *     ("members") of one grid shape, their device-resident state and the
 *     precomputed FFT tables.  Every entry point returns 0 on success and a
 *     negative HOS_E* code on failure; hos_last_error() gives the message.
 *     All array arguments are HOST pointers in the reference's layouts:
 *       spectral field   double _Complex [Ny/2+1][Nx]   (index Nx*j + i,
 *                        FFT_TRANSPOSE 1, Csource/2dpar/euler_2d_par.h:22)
 *       state u          [2][Ny/2+1][Nx]  (eta-hat, then phi-hat at offset
 *                        Nx*(Ny/2+1), Csource/2dpar/main_2d_par.c:150-158)
 *       real field       double [Nx][Ny+2] (padded rows,
 *                        Csource/2dpar/fft_routines_2d_par.c:24-34)
 *
 *  2. include/hos_ref.h: the reference's own function names and signatures
 *     (Setup_TimeScheme, rhs_hos_setup, sol_update_RK, rhs_hos, Zvel, ...),
 *     implemented on top of layer 1 so that a C driver written against
 *     Csource/2dpar links against this library instead.
 *
 * Thread safety: a context must not be used from two host threads at once.
 */
#ifndef HOS_GPU_H
#define HOS_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define HOS_OK 0
#define HOS_E_INVALID (-1)   /* bad argument or configuration */
#define HOS_E_DEVICE (-2)    /* HIP runtime error */
#define HOS_E_NOMEM (-3)     /* device allocation failed */
#define HOS_E_STATE (-4)     /* call not valid in the current state */

/* precision of the FFTs, spectral multiplies and products.  The RK state,
 * the stage combination, the linear terms (-g eta, W^(1)) and the
 * diagnostics are always FP64. */
#define HOS_FP64 0
#define HOS_FP32 1

typedef struct hos_params {
    int Nx, Ny;        /* grid (powers of two, 16..2048)                      */
    int nlevs;         /* NLevs of Csource/2dpar/euler_2d_par.h:24; M = nlevs+1 */
    double Lx, Ly;     /* domain; Kx0 = 2pi/Lx (hdf5_routines_2d_par.c:98-99) */
    double g;
    double depth;      /* <= 0: deep water (the reference's live Dz);
                          > 0: finite depth h, d^n/dz^n -> k^n tanh(kh)^(n odd) */
    int rampflg;       /* Dommermuth ramp, hosm_2d_par.c:211-229 */
    double Tramp;
    int scheme;        /* Setup_TimeScheme flag, time_schemes_par.c:127-314:
                          1 Dormand-Prince (4th-order weights), 2 RK4,
                          3 Euler, 4 midpoint, 5 Heun */
    int precision;     /* HOS_FP64 or HOS_FP32 */
    int kx_quirk;      /* 1: reference kx table (i <= Nx/2+1 -> +i Kx0,
                          operators_par.c:55); 0: signed table */
    int mean_level;    /* 1: hold eta-hat(0,0) at its value when the state was
                          set (mean-level correction); 0: reference behaviour */
    int compensated;   /* FP32 only: Kahan-compensated order sums */
} hos_params;

/* Defaults: deep water, RK4, FP64, reference kx table, no ramp, g = 9.81. */
void hos_params_default(hos_params *p);

typedef struct hos_ctx hos_ctx;

/* Create a context with nmembers instances of the same shape on the current
 * HIP device.  Replaces fftw_mpi_local_size_2d_transposed + the FFTW plans
 * (main_2d_par.c:143,189-190) and rhs_hos_setup / Setup_TimeScheme
 * allocations (hosm_2d_par.c:51-97, time_schemes_par.c:157-163). */
int hos_create(const hos_params *p, int nmembers, hos_ctx **out);
int hos_destroy(hos_ctx *c);

/* Bytes of device memory a context of this shape would allocate. */
size_t hos_footprint(const hos_params *p, int nmembers);

/* Copy one member's state in / out (host [2][Ny/2+1][Nx] complex double).
 * hos_set_state applies no dealiasing: callers pass a masked state, as
 * main_2d_par.c:258-262 does (fft_2d then Dealias). */
int hos_set_state(hos_ctx *c, int member, const double *u);
int hos_get_state(hos_ctx *c, int member, double *u);

/* Per-member time t (double), advanced by dt every step, as
 * time_schemes_par.c:119. */
int hos_set_time(hos_ctx *c, int member, double t);
int hos_get_time(hos_ctx *c, int member, double *t);

/* Advance every member by nsteps sol_update_RK steps of size dt
 * (time_schemes_par.c:28-121); stage time is t_n for every stage (:70).
 * The stage-0 Hamiltonian of each step is recorded (hos_get_energy). */
int hos_step(hos_ctx *c, int nsteps, double dt);

/* Same, with a per-member dt array (length nmembers). */
int hos_step_dt(hos_ctx *c, int nsteps, const double *dt);

/* Batching and slot pools (ensembles, the serving loop):
 *  hos_set_active(c, n)   only members 0..n-1 are stepped (0 <= n <= nmembers);
 *  hos_get_active(c)      the current count;
 *  hos_set_member(c, m, depth, dt, t)   per-member depth (<= 0: deep), time
 *                         step and time;
 *  hos_copy_member(c, src, dst)   copy a member's state and parameters to
 *                         another slot (slot compaction);
 *  hos_step_async(c, n)   enqueue n steps of the active members with the
 *                         per-member dt already set, and return at once;
 *  hos_wait(c)            wait for them and record the step timings.
 * Contexts use their own HIP stream, so the steps of several contexts
 * enqueued with hos_step_async run concurrently on the device.  Until
 * hos_wait, the calls that write the state, the time, the launch shapes or
 * the context's scratch (hos_step*, hos_set_state, hos_set_time,
 * hos_set_member, hos_set_active, hos_copy_member, hos_set_tile,
 * hos_autotune, hos_rhs, hos_zvel, hos_fft_2d, hos_ifft_2d, hos_mult,
 * hos_spectral_op, hos_hamiltonian) return HOS_E_STATE and change nothing;
 * the getters wait for the steps in flight. */
int hos_set_active(hos_ctx *c, int n);
int hos_get_active(hos_ctx *c);
int hos_set_member(hos_ctx *c, int member, double depth, double dt, double t);
int hos_copy_member(hos_ctx *c, int src, int dst);
int hos_step_async(hos_ctx *c, int nsteps);
int hos_wait(hos_ctx *c);
/* Non-blocking form of hos_wait: 1 while the steps enqueued by
 * hos_step_async are still running; 0 once none is in flight (the steps that
 * had been enqueued are then recorded exactly as hos_wait records them);
 * < 0 on error. */
int hos_query(hos_ctx *c);

/* Stage-0 Hamiltonian of the last step, hosm_2d_par.c:462-523 (the
 * reference's H = mean(g eta^2 + phi eta_t), no 1/2), and volume = mean eta =
 * Re eta-hat(0,0), both FP64. */
int hos_get_energy(hos_ctx *c, int member, double *H, double *volume);

/* One RHS evaluation (rhs_hos, hosm_2d_par.c:138-235) of member 0 on a host
 * state u at time t; result into host hrhs [2][Ny/2+1][Nx].  Does not change
 * the context's state. */
int hos_rhs(hos_ctx *c, const double *u, double *hrhs, double t);

/* Allocation-free form of hos_rhs.  It has no counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 *
 * One RHS evaluation (rhs_hos, hosm_2d_par.c:138-235, by the same kernels as
 * hos_rhs) of the DEVICE state d_u [2][Ny/2+1][Nx] complex double at time t
 * into the DEVICE array d_rhs (same layout; it must not overlap d_u), for the
 * configuration *p (scheme and mean_level are not used).  The launches are
 * enqueued on `stream` (a hipStream_t; NULL: the null stream) and not waited
 * for, and no host-device copy is made.  Every other buffer (x-intermediates,
 * physical fields, FFT twiddles, wavenumber and tanh tables, the time) lives
 * in one static device array of HOS_NOALLOC_SCRATCH_BYTES per GPU, which each
 * GPU holds until the process ends.  A configuration that needs more returns
 * HOS_E_INVALID (FP64 fits 256^2 at every M <= 8 and 512^2 at M <= 4;
 * FP32 512^2 at every M <= 8).  Nothing is allocated: no hipMalloc, hipFree or
 * allocating library routine.  All calls on one GPU share that GPU's array,
 * which keeps the last call's contents (every kernel writes what a later
 * kernel of the same call reads): calls in one stream are ordered, calls
 * that overlap on one GPU (other streams, other host threads) race. */
#define HOS_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int hos_rhs_noalloc(const hos_params *p, const double *d_u, double *d_rhs, double t, void *stream);

/* W expansion of a host state (Zvel, hosm_2d_par.c:240-445): W^(M), W^(M-2),
 * (W^2)^(M), (W^2)^(M-2), each host [Ny/2+1][Nx] complex.  Any pointer may be
 * NULL. */
int hos_zvel(hos_ctx *c, const double *u, double *wM, double *wM2, double *w2M, double *w2M2);

/* fft_2d / ifft_2d of fft_routines_2d_par.c:21-103 on host arrays:
 * real [Nx][Ny+2] padded <-> spectral [Ny/2+1][Nx], r2c normalised by
 * 1/(Nx*Ny), c2r unnormalised.  c2r uses the real part of the ky=0 and
 * ky=Ny/2 columns after the x inverse (c2r of the Hermitian projection). */
int hos_fft_2d(hos_ctx *c, const double *u_padded, double *hu);
int hos_ifft_2d(hos_ctx *c, const double *hu, double *u_padded);

/* The reference's spectral operators on host fields [Ny/2+1][Nx], run on the
 * GPU with the context's wavenumber table and mask:
 *   HOS_OP_DX  i kx a            (Dx, operators_par.c:20-75)
 *   HOS_OP_DY  i ky a            (Dy, :85-120)
 *   HOS_OP_DZ  |k| a, or |k| tanh(|k| h) a at finite depth (Dz, :130-196)
 *   HOS_OP_SUM a + b             (Sum, :247-279)
 *   HOS_OP_DEALIAS  mask of a    (Dealias, :283-325)
 * b is used by HOS_OP_SUM only; out may alias a or b. */
#define HOS_OP_DX 1
#define HOS_OP_DY 2
#define HOS_OP_DZ 3
#define HOS_OP_SUM 4
#define HOS_OP_DEALIAS 5
int hos_spectral_op(hos_ctx *c, int op, const double *a, const double *b, double *out);

/* Mult (operators_par.c:210-235): pseudo-spectral product of a and b; out may
 * alias either input. */
int hos_mult(hos_ctx *c, const double *a, const double *b, double *out);

/* Hamiltonian (hosm_2d_par.c:462-523) of host fields, FP64, on the GPU. */
int hos_hamiltonian(hos_ctx *c, const double *heta, const double *heta_t, const double *hphi, double *H);

/* Device synchronisation and per-step timing of the last hos_step call
 * (milliseconds per step from HIP events). */
int hos_sync(hos_ctx *c);
int hos_last_step_ms(hos_ctx *c, double *ms);

/* Per-step durations (ms, HIP events between steps) of the last hos_step call;
 * returns the number of steps recorded and copies up to n of them. */
int hos_step_times(hos_ctx *c, double *ms, int n);

/* Kernel kinds for the per-kernel timing below. */
#define HOS_K_XINV 0        /* x inverse pass (spectral -> x-intermediate)        */
#define HOS_K_YC2R 1        /* y c2r pass (x-intermediate -> physical)            */
#define HOS_K_YR2C 2        /* y r2c pass with a one-field prologue               */
#define HOS_K_YR2C_PAIR 3   /* y r2c pass with the two-field RHS prologue         */
#define HOS_K_XFWDINV 4     /* x forward + d^j/dz^j + x inverse (one W level)     */
#define HOS_K_XFWD_RHS 5    /* x forward pass with the RHS / RK epilogue          */
#define HOS_K_YC2R_PHI 6    /* y c2r of a W level with the next level's y r2c     */
#define HOS_K_YDIAG 7       /* diagonal pass: y c2r of a W diagonal, W^(L), y r2c of phi^(L+1) */
#define HOS_K_NKINDS 8

/* 1: record a HIP event pair around every stream-ordered kernel launch of
 * later hos_step calls (not graph replays) and accumulate, per kernel kind,
 * the launch count, the summed duration and the algorithmic bytes (the
 * minimum HBM bytes the launch must read and write).  Resets the totals. */
int hos_profile(hos_ctx *c, int enable);
int hos_kernel_stats(hos_ctx *c, int kind, double *ms_total, long *launches, double *bytes_total);

/* Capture the launches of one step into a hipGraph and replay it for every
 * later hos_step of the same dt (1) or launch stream-ordered (0). */
int hos_use_graph(hos_ctx *c, int enable);

/* Launch-shape overrides chosen by the autotuner.  Kinds 0-3: tx = x-rows
 * per y-pass workgroup (4, 8 or 16 within the LDS limit; 0 = built-in
 * default) for one kind of y pass or for all of them.  Kind HOS_TILE_XPASS:
 * the x inverse passes run rows per workgroup (1) or one row per 64-lane
 * wave (2, FP32 form with Nx = 512 or 1024); 0 = built-in default.  Kind
 * HOS_TILE_GROUPS: the active members step as one launch chain (1) or as G =
 * 2..4 nearly equal ranges, each its own chain of launches (and graph) on its
 * own stream, forked from and joined to the context stream every step (with
 * the RK schemes 2-5, and only while every range keeps two members or more);
 * 0 = built-in default. */
#define HOS_TILE_C2R 0   /* k_yc2r: several fields per launch        */
#define HOS_TILE_R2C 1   /* k_yr2c: one field per launch              */
#define HOS_TILE_PAIR 2  /* k_yr2c_pair: the two RHS fields            */
#define HOS_TILE_ALL 3
#define HOS_TILE_XPASS 4 /* k_xinv, k_xfwdinv: workgroup rows or wave rows */
#define HOS_TILE_GROUPS 5 /* member groups: 1 chain, or 2..4 on their own streams */
/* HOS_TILE_BAND: 1 off, 2 on (0: default, on).  On: when every active
 * member's state is zero on the modes the dealiasing mask removes
 * (operators_par.c:309-322; hos_set_state checks), the RK steps of the
 * schemes 2-5 skip those modes: the level-0 x inverse transforms rows
 * ky < my only and loads the kept columns only, the y passes that read its
 * outputs read those rows only, the RHS pass writes them only, and the RK
 * epilogue touches the kept modes only.  The skipped modes are exact zeros,
 * and a masked state stays masked, so results are unchanged. */
#define HOS_TILE_BAND 6
/* HOS_TILE_PLAN: the radix plan of the RHS pass's row FFTs (0: default, 1):
 * 1 the built-in rule (1024 points: 16 8 8), 2 radices of at most 8
 * (8 8 4 4), 3 radix 4 (4 4 4 4 4), 4 the largest radix at every stage
 * (16 16 4); 2-4 need Ny >= 512.  Plans round differently; hos_autotune
 * tries them all. */
#define HOS_TILE_PLAN 7
/* HOS_TILE_FUSE: 1 off, 2 on (0: default, on).  On: the y c2r launch of a
 * W level's outputs also does the next level's y r2c in the workgroups of
 * the one output that r2c reads (d phi^(l)/dz, from LDS), using the c2r
 * tile; same results, bit for bit, with one launch and one field read fewer
 * per level from the second level on. */
#define HOS_TILE_FUSE 8
/* HOS_TILE_SKEW: 1 off, 2 on (0: default, on).  With two member groups
 * (HOS_TILE_GROUPS), the groups' launch chains run free over the steps of
 * one call (joined at its end, not after every step), and group 1 starts
 * when group 0's first RHS has done its W level M-2: its memory-bound level
 * passes then overlap group 0's issue-bound RHS pass.  Members are
 * independent, so results are the same bits; hos_step_times reports group
 * 0's steps.  Three or four groups join after every step either way.
 * hos_autotune leaves it as set. */
#define HOS_TILE_SKEW 9
/* HOS_TILE_DIAG: 1 off, 2 on (0: default, on).  On: the W expansion runs as
 * diagonal passes.  Pass L (L = 1..M-1) does the y c2r of the fields with
 * order + derivative = L+1, d^k phi^(L+1-k)/dz^k (k = 1..L), straight from
 * the x-intermediates, forms phi^(L+1) and W^(L) from them (the same fields
 * make both, hosm_2d_par.c:302-357), writes W^(L) to the physical grid and
 * does the y r2c of phi^(L+1); the RHS pass then reads W^(1)..W^(M-1)
 * instead of the M(M-1)/2 derivative fields.  No derivative field reaches
 * the physical grid.  Results agree with the mode off to rounding. */
#define HOS_TILE_DIAG 10
/* HOS_TILE_XBLOCK: 1 off, 2 on (0: default, on).  On, in the diagonal form
 * with row-per-wave x passes (HOS_TILE_DIAG and HOS_TILE_XPASS both 2, M >= 2),
 * the x-intermediates between the x and y passes are stored in blocks of 4
 * rows inside which the rows of each column pair lie side by side: the y
 * passes, which read and write one column pair of every row, then move 64
 * contiguous bytes per block instead of 16 bytes per row, and the x passes
 * store and load a block through LDS as one contiguous run.  Only where the
 * values lie changes: results are the same bits.  hos_autotune leaves it as
 * set. */
#define HOS_TILE_XBLOCK 11
int hos_set_tile(hos_ctx *c, int kind, int tx);
/* current setting of one kind (0-2: tx; HOS_TILE_XPASS: 1 or 2; HOS_TILE_GROUPS: 1..4;
 * HOS_TILE_BAND, HOS_TILE_FUSE, HOS_TILE_SKEW, HOS_TILE_DIAG, HOS_TILE_XBLOCK: 1 or 2;
 * HOS_TILE_PLAN: 1..4),
 * < 0 on error */
int hos_get_tile(hos_ctx *c, int kind);

/* Autotuner: time nsteps RK steps of the active members for every valid
 * (c2r, r2c, pair) tile combination, then with the best tiles the other
 * x-pass mode, then the other W-expansion form (HOS_TILE_DIAG), then every
 * FFT radix plan of the RHS pass (Ny >= 512), then
 * every member-group count (in calls of at least 8 steps, over which two
 * offset groups, HOS_TILE_SKEW, show their gain); keep the fastest and store it in an
 * in-process cache keyed by (Nx, Ny, M, precision, active members, scheme
 * kind, band); later calls with the same key apply the cached choice at once.
 * The members' state, time, dt and energies are restored afterwards.
 * tiles_out[3] (may be NULL) receives the y tiles; hos_get_tile gives the
 * other choices. */
int hos_autotune(hos_ctx *c, int nsteps, int *tiles_out);

const char *hos_last_error(void);

/* Plain device copy (k_copy_c64) of `bytes`, timed with HIP events over reps
 * launches: the practical copy bandwidth in GB/s (bytes read + written), and
 * the copy kernel of known size that calibrates the HBM byte counters. */
int hos_copy_bandwidth(size_t bytes, int reps, double *gbps);

/* Introspection used by tests and the benchmark. */
int hos_device_name(char *buf, int len);
int hos_num_members(hos_ctx *c);
/* Algorithmic HBM bytes per grid point per step of this configuration's
 * dominant kernels (FFT passes plus FP64 state traffic). */
double hos_bytes_per_point_step(const hos_params *p);

#ifdef __cplusplus
}
#endif
#endif /* HOS_GPU_H */
