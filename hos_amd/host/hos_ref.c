/*
 * hos_ref.c -- the reference's call boundary (include/hos_ref.h) on top of the
 * native GPU API (include/hos_gpu.h).  Host C; every numerical operation runs
 * on the GPU through libhos_gpu.so.
 */
#include "../../include/hos_ref.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "../../include/hos_gpu.h"

/* euler_2d_par.h:40-71 and fft_routines_2d_par.h:28-33 */
int Nx, Ny, runsubid, saveflg, rampflg;
double Lx, Ly, Kx0, Ky0, g, T, dtsave, Tramp;
ptrdiff_t alloc_local, local_Nx, local_Nyhpo, local_N, local_0_start, local_1_start;
int mpi_rank = 0, mpi_size = 1;

int hos_nlevs = 2;
int hos_precision = HOS_FP64;
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
        CHECK(hos_get_state(ctx, 0, (double *)reg_u));
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

void sol_update_RK(hos_cplx *u, double *t, double dt, char *dtflag)
{
    (void)dtflag;
    ensure_ctx();
    if (u != reg_u || dev_stale) {
        if (reg_u && u != reg_u) hos_ref_sync();
        CHECK(hos_set_state(ctx, 0, (const double *)u));
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

void rhs_hos(hos_cplx *hrhs, hos_cplx *hu, double t)
{
    ensure_ctx();
    touch(hu);
    touch(hrhs);
    CHECK(hos_rhs(ctx, (const double *)hu, (double *)hrhs, t));
    wrote(hrhs);
}

void Zvel(hos_cplx *hu, hos_cplx *hZvelM, hos_cplx *hZvelM2, hos_cplx *hZvel2M, hos_cplx *hZvel2M2, double t)
{
    (void)t;
    hos_cplx *outs[4] = {hZvelM, hZvelM2, hZvel2M, hZvel2M2};
    ensure_ctx();
    touch(hu);
    for (int q = 0; q < 4; q++) touch(outs[q]);
    CHECK(hos_zvel(ctx, (const double *)hu, (double *)hZvelM, (double *)hZvelM2, (double *)hZvel2M,
                   (double *)hZvel2M2));
    for (int q = 0; q < 4; q++) wrote(outs[q]);
}

void ZvelLinear(const hos_cplx *hu, hos_cplx *out)
{
    ensure_ctx();
    touch(hu);
    touch(out);
    CHECK(hos_spectral_op(ctx, HOS_OP_DZ, (const double *)(hu + alloc_local), NULL, (double *)out));
    wrote(out);
}

double Hamiltonian(const hos_cplx *heta, const hos_cplx *heta_t, const hos_cplx *hphi)
{
    double H = 0.0;
    ensure_ctx();
    touch(heta);
    touch(heta_t);
    touch(hphi);
    CHECK(hos_hamiltonian(ctx, (const double *)heta, (const double *)heta_t, (const double *)hphi, &H));
    return H;
}

double RampFun(const double t) { return 1 - exp(-t * t / Tramp / Tramp); }

/* ---------------------------------------------------------------- L2 */
static void unary(int op, const hos_cplx *a, const hos_cplx *b, hos_cplx *out, const char *name)
{
    ensure_ctx();
    touch(a);
    touch(b);
    touch(out);
    if (hos_spectral_op(ctx, op, (const double *)a, (const double *)b, (double *)out) != HOS_OK) die(name);
    wrote(out);
}

void Dx(hos_cplx *hu, hos_cplx *hu_x) { unary(HOS_OP_DX, hu, NULL, hu_x, "Dx"); }
void Dy(hos_cplx *hu, hos_cplx *hu_y) { unary(HOS_OP_DY, hu, NULL, hu_y, "Dy"); }
void Dz(const hos_cplx *hu, hos_cplx *hu_z) { unary(HOS_OP_DZ, hu, NULL, hu_z, "Dz"); }
void Sum(hos_cplx *a, hos_cplx *b, hos_cplx *s) { unary(HOS_OP_SUM, a, b, s, "Sum"); }
void Dealias(hos_cplx *hu) { unary(HOS_OP_DEALIAS, hu, NULL, hu, "Dealias"); }

void Mult(hos_cplx *hu1, hos_cplx *hu2, hos_cplx *hprod)
{
    ensure_ctx();
    touch(hu1);
    touch(hu2);
    touch(hprod);
    CHECK(hos_mult(ctx, (const double *)hu1, (const double *)hu2, (double *)hprod));
    wrote(hprod);
}

/* ---------------------------------------------------------------- L1 */
void fft_2d(double *u, hos_cplx *hu, hos_plan plan)
{
    (void)plan;
    ensure_ctx();
    touch(hu);
    CHECK(hos_fft_2d(ctx, u, (double *)hu));
    wrote(hu);
}

void ifft_2d(hos_cplx *hu, double *u, hos_plan plan)
{
    (void)plan;
    ensure_ctx();
    touch(hu);
    CHECK(hos_ifft_2d(ctx, (const double *)hu, u));
}
