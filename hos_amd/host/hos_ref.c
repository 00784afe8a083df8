/*
 * hos_ref.c -- the reference's call boundary (include/hos_ref.h) on top of the
 * native GPU API (include/hos_gpu.h).  Host C; every numerical operation runs
 * on the GPU through libhos_gpu.so.  Built twice: with hos_real = double
 * (libhos_ref.so, the reference's USE_DOUBLES build) and, with
 * -DHOS_REF_FLOAT, hos_real = float (libhos_ref_f32.so), where the arrays
 * are converted to and from FP64 at the boundary (din / dout / dback).
 */
#include "../../include/hos_ref.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "../../include/hos_gpu.h"

/* euler_2d_par.h:40-71 and fft_routines_2d_par.h:28-33 */
int Nx, Ny, runsubid, saveflg, rampflg;
hos_real Lx, Ly, Kx0, Ky0, g, T, dtsave, Tramp;
ptrdiff_t alloc_local, local_Nx, local_Nyhpo, local_N, local_0_start, local_1_start;
int mpi_rank = 0, mpi_size = 1;

int hos_nlevs = 2;
#ifdef HOS_REF_FLOAT
int hos_precision = HOS_FP32; /* the float build computes in the FP32 form */
#else
int hos_precision = HOS_FP64;
#endif
double hos_depth = 0.0;
int hos_kx_quirk = 1;
int hos_runtime_log = 1;

static hos_ctx *ctx;
static int scheme_flg = 4;      /* the driver's default, main_2d_par.c:252 */
static hos_cplx *reg_u;         /* host state whose mirror lives on the GPU */
static int host_stale, dev_stale;
static double gpu_time = NAN;      /* time the GPU member holds */

static void die(const char *where)
{
    fprintf(stderr, "hos: %s failed: %s\n", where, hos_last_error());
    exit(EXIT_FAILURE);
}

/* FP64 views of boundary arrays for the native API.  hos_real = double: the
 * array itself.  hos_real = float: a copy in one of the boundary's FP64
 * buffers (slot k), converted back by dback for outputs. */
#define NSLOT 6
static double *slot_buf[NSLOT];
static size_t slot_len[NSLOT];

static double *slot(int k, size_t n)
{
    if (slot_len[k] < n) {
        free(slot_buf[k]);
        slot_buf[k] = malloc(n * sizeof(double));
        if (!slot_buf[k]) {
            fprintf(stderr, "hos: out of host memory\n");
            exit(EXIT_FAILURE);
        }
        slot_len[k] = n;
    }
    return slot_buf[k];
}

/* n = number of hos_real values */
static const double *din(int k, const void *a, size_t n)
{
    if (!a) return NULL;
    if (sizeof(hos_real) == sizeof(double)) return (const double *)a;
    double *d = slot(k, n);
    const hos_real *f = (const hos_real *)a;
    for (size_t i = 0; i < n; i++) d[i] = f[i];
    return d;
}
static double *dout(int k, void *a, size_t n)
{
    if (!a) return NULL;
    if (sizeof(hos_real) == sizeof(double)) return (double *)a;
    return slot(k, n);
}
static void dback(void *a, const double *d, size_t n)
{
    if (!a || sizeof(hos_real) == sizeof(double)) return;
    hos_real *f = (hos_real *)a;
    for (size_t i = 0; i < n; i++) f[i] = (hos_real)d[i];
}
#define CHECK(call)                                                                                \
    do {                                                                                           \
        if ((call) != HOS_OK) die(#call);                                                          \
    } while (0)

ptrdiff_t hos_local_size_2d_transposed(ptrdiff_t n0, ptrdiff_t n1, ptrdiff_t *local_n0, ptrdiff_t *local_0_start_,
                                       ptrdiff_t *local_n1, ptrdiff_t *local_1_start_)
{
    if (local_n0) *local_n0 = n0;
    if (local_0_start_) *local_0_start_ = 0;
    if (local_n1) *local_n1 = n1;
    if (local_1_start_) *local_1_start_ = 0;
    local_Nx = n0;
    local_0_start = 0;
    local_Nyhpo = n1;
    local_1_start = 0;
    local_N = n0 * n1;
    alloc_local = n0 * n1;
    return alloc_local;
}

static void teardown(void)
{
    if (ctx) hos_destroy(ctx);
    ctx = NULL;
    reg_u = NULL;
    host_stale = dev_stale = 0;
    gpu_time = NAN;
}

static void ensure_ctx(void)
{
    if (ctx) return;
    if (alloc_local == 0) hos_local_size_2d_transposed(Nx, Ny / 2 + 1, NULL, NULL, NULL, NULL);
    if (alloc_local != (ptrdiff_t)Nx * (Ny / 2 + 1)) {
        fprintf(stderr, "hos: alloc_local %td does not match Nx*(Ny/2+1) for %dx%d\n", alloc_local, Nx, Ny);
        exit(EXIT_FAILURE);
    }
    hos_params p;
    hos_params_default(&p);
    p.Nx = Nx;
    p.Ny = Ny;
    p.nlevs = hos_nlevs;
    p.Lx = Lx;
    p.Ly = Ly;
    p.g = g;
    p.depth = hos_depth;
    p.rampflg = rampflg;
    p.Tramp = Tramp > 0 ? Tramp : 1.0;
    p.scheme = scheme_flg;
    p.precision = hos_precision;
    p.kx_quirk = hos_kx_quirk;
    CHECK(hos_create(&p, 1, &ctx));
}

struct hos_ctx *hos_ref_context(void) { return ctx; }

static int in_state(const void *p)
{
    if (!reg_u || !p) return 0;
    const char *a = (const char *)reg_u, *b = a + sizeof(hos_cplx) * 2 * (size_t)alloc_local;
    return (const char *)p >= a && (const char *)p < b;
}

void hos_ref_sync(void)
{
    if (ctx && reg_u && host_stale) {
        const size_t n = 4 * (size_t)alloc_local;
        double *d = dout(0, reg_u, n);
        CHECK(hos_get_state(ctx, 0, d));
        dback(reg_u, d, n);
        host_stale = 0;
    }
}

void hos_ref_host_modified(void) { dev_stale = 1; }

/* before a call reads host memory p, or writes host memory inside the state */
static void touch(const void *p)
{
    if (in_state(p)) hos_ref_sync();
}
static void wrote(const void *p)
{
    if (in_state(p)) dev_stale = 1;
}

/* ---------------------------------------------------------------- L4 */
void Setup_TimeScheme(int flg)
{
    if (flg < 1 || flg > 5) {
        fprintf(stderr, "hos: Setup_TimeScheme(%d): unknown scheme\n", flg);
        exit(EXIT_FAILURE);
    }
    if (ctx && flg != scheme_flg) {
        hos_ref_sync();
        teardown();
    }
    scheme_flg = flg;
}

void sol_update_RK(hos_cplx *u, hos_real *t, hos_real dt, char *dtflag)
{
    (void)dtflag;
    ensure_ctx();
    if (u != reg_u || dev_stale) {
        if (reg_u && u != reg_u) hos_ref_sync();
        CHECK(hos_set_state(ctx, 0, din(0, u, 4 * (size_t)alloc_local)));
        reg_u = u;
        dev_stale = host_stale = 0;
    }
    if (!(gpu_time == *t)) {
        CHECK(hos_set_time(ctx, 0, *t));
        gpu_time = *t;
    }
    CHECK(hos_step(ctx, 1, dt));
    host_stale = 1;
    if (hos_runtime_log) {
        double H = 0.0;
        CHECK(hos_get_energy(ctx, 0, &H, NULL));
        FILE *f = fopen("runtime.dat", "a+");
        if (!f) {
            perror("Unable to open runtime datafile");
            exit(EXIT_FAILURE);
        }
        fprintf(f, "%14.10f\t%.16g\n", *t, H);
        fclose(f);
    }
    *t = *t + dt;
    gpu_time = *t;
}

/* ---------------------------------------------------------------- L3 */
void rhs_hos_setup(void)
{
    /* a state still on the GPU is copied back before its context goes */
    hos_ref_sync();
    teardown();
    ensure_ctx();
}

void rhs_hos_clean(void)
{
    hos_ref_sync();
    teardown();
}

void rhs_hos(hos_cplx *hrhs, hos_cplx *hu, hos_real t)
{
    const size_t n = 4 * (size_t)alloc_local;
    ensure_ctx();
    touch(hu);
    touch(hrhs);
    const double *u = din(0, hu, n);
    double *r = dout(1, hrhs, n);
    CHECK(hos_rhs(ctx, u, r, t));
    dback(hrhs, r, n);
    wrote(hrhs);
}

void Zvel(hos_cplx *hu, hos_cplx *hZvelM, hos_cplx *hZvelM2, hos_cplx *hZvel2M, hos_cplx *hZvel2M2, hos_real t)
{
    (void)t;
    hos_cplx *outs[4] = {hZvelM, hZvelM2, hZvel2M, hZvel2M2};
    const size_t n = 2 * (size_t)alloc_local;
    double *d[4];
    ensure_ctx();
    touch(hu);
    for (int q = 0; q < 4; q++) touch(outs[q]);
    const double *u = din(0, hu, 2 * n);
    for (int q = 0; q < 4; q++) d[q] = dout(1 + q, outs[q], n);
    CHECK(hos_zvel(ctx, u, d[0], d[1], d[2], d[3]));
    for (int q = 0; q < 4; q++) {
        dback(outs[q], d[q], n);
        wrote(outs[q]);
    }
}

void ZvelLinear(const hos_cplx *hu, hos_cplx *out)
{
    const size_t n = 2 * (size_t)alloc_local;
    ensure_ctx();
    touch(hu);
    touch(out);
    const double *a = din(0, hu + alloc_local, n);
    double *o = dout(1, out, n);
    CHECK(hos_spectral_op(ctx, HOS_OP_DZ, a, NULL, o));
    dback(out, o, n);
    wrote(out);
}

hos_real Hamiltonian(const hos_cplx *heta, const hos_cplx *heta_t, const hos_cplx *hphi)
{
    const size_t n = 2 * (size_t)alloc_local;
    double H = 0.0;
    ensure_ctx();
    touch(heta);
    touch(heta_t);
    touch(hphi);
    CHECK(hos_hamiltonian(ctx, din(0, heta, n), din(1, heta_t, n), din(2, hphi, n), &H));
    return (hos_real)H;
}

hos_real RampFun(const hos_real t) { return 1 - exp(-t * t / Tramp / Tramp); }

/* ---------------------------------------------------------------- L2 */
static void unary(int op, const hos_cplx *a, const hos_cplx *b, hos_cplx *out, const char *name)
{
    const size_t n = 2 * (size_t)alloc_local;
    ensure_ctx();
    touch(a);
    touch(b);
    touch(out);
    const double *da = din(0, a, n), *db = din(1, b, n);
    double *o = dout(2, out, n);
    if (hos_spectral_op(ctx, op, da, db, o) != HOS_OK) die(name);
    dback(out, o, n);
    wrote(out);
}

void Dx(hos_cplx *hu, hos_cplx *hu_x) { unary(HOS_OP_DX, hu, NULL, hu_x, "Dx"); }
void Dy(hos_cplx *hu, hos_cplx *hu_y) { unary(HOS_OP_DY, hu, NULL, hu_y, "Dy"); }
void Dz(const hos_cplx *hu, hos_cplx *hu_z) { unary(HOS_OP_DZ, hu, NULL, hu_z, "Dz"); }
void Sum(hos_cplx *a, hos_cplx *b, hos_cplx *s) { unary(HOS_OP_SUM, a, b, s, "Sum"); }
void Dealias(hos_cplx *hu) { unary(HOS_OP_DEALIAS, hu, NULL, hu, "Dealias"); }

void Mult(hos_cplx *hu1, hos_cplx *hu2, hos_cplx *hprod)
{
    const size_t n = 2 * (size_t)alloc_local;
    ensure_ctx();
    touch(hu1);
    touch(hu2);
    touch(hprod);
    const double *a = din(0, hu1, n), *b = din(1, hu2, n);
    double *o = dout(2, hprod, n);
    CHECK(hos_mult(ctx, a, b, o));
    dback(hprod, o, n);
    wrote(hprod);
}

/* ---------------------------------------------------------------- L1 */
void fft_2d(hos_real *u, hos_cplx *hu, hos_plan plan)
{
    (void)plan;
    const size_t nr = (size_t)Nx * (Ny + 2), n = 2 * (size_t)alloc_local;
    ensure_ctx();
    touch(hu);
    const double *a = din(0, u, nr);
    double *o = dout(1, hu, n);
    CHECK(hos_fft_2d(ctx, a, o));
    dback(hu, o, n);
    wrote(hu);
}

void ifft_2d(hos_cplx *hu, hos_real *u, hos_plan plan)
{
    (void)plan;
    const size_t nr = (size_t)Nx * (Ny + 2), n = 2 * (size_t)alloc_local;
    ensure_ctx();
    touch(hu);
    const double *a = din(0, hu, n);
    double *o = dout(1, u, nr);
    CHECK(hos_ifft_2d(ctx, a, o));
    dback(u, o, nr);
}
