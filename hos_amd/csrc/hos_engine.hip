// hos_engine.hip -- host orchestration and the native C ABI (include/hos_gpu.h).
//
// One context = one grid shape, nm members, one HIP stream.  A step runs
// nstages RHS evaluations; each RHS is the level schedule
//
//   level 0   k_xinv  (eta, eta_x, eta_y, phi_x, phi_y, d^j phi/dz^j j=1..M)
//             k_yc2r  -> physical fields
//   level l   k_yr2c  (prologue phi^(l+1))             l = 1..M-1
//             k_xfwdinv (d^j/dz^j phi^(l+1), j=1..M-l)
//             k_yc2r  -> physical fields
//   final     k_yr2c_pair (prologue: nonlinear eta_t, phi_t)
//             k_xfwd_rhs  (+ W^(1), -g eta, ramp, dealias, RK stage update)
//
// which is the reference's rhs_hos + Zvel (hosm_2d_par.c:138-445) with each
// Mult's r2c/c2r round trip removed: every product is formed on the real grid
// and only fields that need a spectral operator (d^j/dz^j) are transformed.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <array>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/hos_gpu.h"
#include "hos_ctx.hpp"
#include "hos_misc_kernels.hpp"

using namespace hos;

// ---------------------------------------------------------------------------
static thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char b[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(b, sizeof b, fmt, ap);
    va_end(ap);
    g_err = b;
    return code;
}

static int ilog2(int n)
{
    int l = 0;
    while ((1 << l) < n) ++l;
    return (1 << l) == n ? l : -1;
}


static void harvest_prof(hos_ctx *c)
{
    for (auto &r : c->recs) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            c->kms[r.kind] += ms;
            c->kbytes[r.kind] += r.bytes;
            c->kn[r.kind] += 1;
        }
    }
    c->recs.clear();
    c->evused = 0;
}

extern "C" void hos_params_default(hos_params *p)
{
    std::memset(p, 0, sizeof *p);
    p->Nx = p->Ny = 256;
    p->nlevs = 2;
    p->Lx = p->Ly = 2 * M_PI;
    p->g = 9.81;
    p->depth = 0.0;
    p->rampflg = 0;
    p->Tramp = 1.0;
    p->scheme = 2;
    p->precision = HOS_FP64;
    p->kx_quirk = 1;
    p->mean_level = 0;
    p->compensated = 1;
}

extern "C" const char *hos_last_error(void) { return g_err.c_str(); }

static int setup_scheme(hos_ctx *c, int flg)
{
    std::memset(c->a, 0, sizeof c->a);
    std::memset(c->b, 0, sizeof c->b);
    std::memset(c->c, 0, sizeof c->c);
    switch (flg) {
    case 1:  // Dormand-Prince, 4th-order weights (time_schemes_par.c:181-238)
        c->S = 7;
        c->a[1] = 1.0 / 5; c->a[2] = 3.0 / 10; c->a[3] = 4.0 / 5; c->a[4] = 8.0 / 9; c->a[5] = 1; c->a[6] = 1;
        c->b[1][0] = 1.0 / 5;
        c->b[2][0] = 3.0 / 40; c->b[2][1] = 9.0 / 40;
        c->b[3][0] = 44.0 / 45; c->b[3][1] = -56.0 / 15; c->b[3][2] = 32.0 / 9;
        c->b[4][0] = 19372.0 / 6561; c->b[4][1] = -25360.0 / 2187; c->b[4][2] = 64448.0 / 6561;
        c->b[4][3] = -212.0 / 729;
        c->b[5][0] = 9017.0 / 3168; c->b[5][1] = -355.0 / 33; c->b[5][2] = 46732.0 / 5247;
        c->b[5][3] = 49.0 / 176; c->b[5][4] = -5103.0 / 18656;
        c->b[6][0] = 35.0 / 384; c->b[6][2] = 500.0 / 1113; c->b[6][3] = 125.0 / 192;
        c->b[6][4] = -2187.0 / 6784; c->b[6][5] = 11.0 / 84;
        c->c[0] = 1951.0 / 21600; c->c[2] = 22642.0 / 50085; c->c[3] = 451.0 / 720;
        c->c[4] = -12231.0 / 42400; c->c[5] = 649.0 / 6300; c->c[6] = 1.0 / 60;
        break;
    case 2:  // RK4 (:241-264)
        c->S = 4;
        c->a[1] = 0.5; c->a[2] = 0.5; c->a[3] = 1.0;
        c->b[1][0] = 0.5; c->b[2][1] = 0.5; c->b[3][2] = 1.0;
        c->c[0] = 1.0 / 6; c->c[1] = 1.0 / 3; c->c[2] = 1.0 / 3; c->c[3] = 1.0 / 6;
        break;
    case 3: c->S = 1; c->c[0] = 1.0; break;                                              // Euler
    case 4: c->S = 2; c->a[1] = 0.5; c->b[1][0] = 0.5; c->c[0] = 0; c->c[1] = 1; break; // midpoint
    case 5: c->S = 2; c->a[1] = 1; c->b[1][0] = 1; c->c[0] = 0.5; c->c[1] = 0.5; break; // Heun
    default: return fail(HOS_E_INVALID, "unknown time scheme %d", flg);
    }
    c->subdiag = true;
    for (int s = 0; s < c->S; ++s)
        for (int q = 0; q < s - 1; ++q)
            if (c->b[s][q] != 0.0) c->subdiag = false;
    return HOS_OK;
}

// y-pass tiles: TX x-rows per workgroup, TX/2 complex rows of Ny in LDS
// (tile_lds_max).  Wider tiles write longer transposed segments; narrower ones
// give more workgroups when a launch carries few fields.
static bool tx_fits(bool f32, int logny, int tx)
{
    // LDS rows in the Lds<T, N> layout: N + N/16 elements plus 128 bytes each
    const int bytes = f32 ? 8 : 16, n = 1 << logny;
    return (tx == 4 || tx == 8 || tx == 16) && (tx / 2) * ((n + n / 16) * bytes + 128) <= tile_lds_max(tx);
}

static int default_tx(bool f32, int logny, int kind)
{
    // c2r launches carry several fields, r2c ones one (or one pair)
    int tx = (kind == HOS_TILE_C2R) ? 8 : 4;
    while (tx > 4 && !tx_fits(f32, logny, tx)) tx /= 2;
    return tx;
}

// x inverse passes one row per wave by default where available and the
// batch is large enough to fill the chip with waves (measured: one 1024^2
// member runs faster with rows per workgroup)
static int default_xwave(const hos_ctx *c) { return x_wave_ok(c) && c->nm >= 4 ? 1 : 0; }

// slots per active count in the graph cache (group counts 1..MAXG)
constexpr int GSLOTS = hos_ctx::MAXG * (hos_ctx::MAXG + 1) / 2;

// two member groups on their own streams by default from 4 members (measured
// at 8 x 1024^2: faster than one chain; the latency-bound phases of one chain
// overlap the other's memory-bound passes)
static int default_groups(const hos_ctx *c) { return c->nm >= 4 ? 2 : 1; }

// the W levels as diagonal passes (k_ydiag) by default: they move about a
// quarter less HBM traffic per RHS (8 x 1024^2, M = 5, FP32: step 2.23 ->
// 1.97 ms; FP64 plain 0.89 -> 0.80 ms); hos_autotune tries both
static int default_diag(const hos_ctx *) { return 1; }

static int validate(const hos_params *p, int nm)
{
    if (!p) return fail(HOS_E_INVALID, "null params");
    const int lx = ilog2(p->Nx), ly = ilog2(p->Ny);
    if (lx < LOGN_MIN || lx > LOGN_MAX || ly < LOGN_MIN || ly > LOGN_MAX)
        return fail(HOS_E_INVALID, "grid %dx%d: Nx and Ny must be powers of two in [%d, %d]", p->Nx, p->Ny,
                    1 << LOGN_MIN, 1 << LOGN_MAX);
    if (p->nlevs < 0 || p->nlevs + 1 > MAXM) return fail(HOS_E_INVALID, "nlevs %d out of range [0, %d]", p->nlevs, MAXM - 1);
    if (!(p->Lx > 0) || !(p->Ly > 0)) return fail(HOS_E_INVALID, "domain lengths must be positive");
    if (p->precision != HOS_FP64 && p->precision != HOS_FP32) return fail(HOS_E_INVALID, "bad precision");
    if (p->rampflg == 1 && !(p->Tramp > 0)) return fail(HOS_E_INVALID, "Tramp must be positive");
    if (nm < 1) return fail(HOS_E_INVALID, "nmembers must be >= 1");
    return HOS_OK;
}

static void layout(hos_ctx *c)
{
    const hos_params &p = c->p;
    c->L = p.nlevs;
    c->M = p.nlevs + 1;
    c->Nk = p.Ny / 2 + 1;
    c->lognx = ilog2(p.Nx);
    c->logny = ilog2(p.Ny);
    c->mx = (int)std::floor(0.5 * p.Nx / (1 + 0.5 * p.nlevs));  // operators_par.c:290-291
    c->my = (int)std::floor(0.5 * p.Ny / (1 + 0.5 * p.nlevs));
    c->f32 = p.precision == HOS_FP32;
    c->fs = (long)c->Nk * p.Nx;
    // the x-intermediates' rows 16 elements longer than Nx: the y passes read
    // and write them as column tiles, and a power-of-two row stride (8 KB at
    // Nx = 1024) put every row of a tile on the same HBM channels (transposed
    // tile reads 3.33 -> 3.58 TB/s; y c2r 41.7 -> 39.1 us, FP64 plain step
    // 1.126 -> 1.114 ms, the two-group FP32 batch unchanged)
    c->ip = p.Nx + 16;
    c->fsI = (long)xi_rows(c->Nk) * c->ip;  // whole blocks of rows (xi_at)
    // FP32: one spare row per physical field.  Fields a power of two apart (4 MB
    // at 1024^2) put the same point of every field a pass reads on the same
    // HBM channels; one row off, the level r2c passes went from 33 to 31 us
    // (8 x 1024^2 step 2.99 -> 2.97 ms).  The FP64 plain step measured 0.7%
    // slower with the row, so FP64 fields stay packed
    c->fsP = (long)p.Nx * p.Ny + (c->f32 ? p.Ny : 0);
    // x-intermediate slots: all level-0 outputs (5 + M, the Zvel path), or the
    // working slots 0 .. max(2, M)-1 plus the M + 4 fields the RHS pass
    // transforms itself
    c->nI = std::max(5 + c->M, std::max(2, c->M) + c->M + 4);
    HosIdx &ix = c->ix;
    std::memset(&ix, 0, sizeof ix);
    ix.M = c->M;
    ix.eta = 0; ix.etax = 1; ix.etay = 2; ix.phix = 3; ix.phiy = 4;
    int f = 5;
    for (int a = 1; a <= c->M; ++a)
        for (int j = 1; j <= c->M - a + 1; ++j) ix.d[a][j] = f++;
    // W^(a) of the diagonal passes in the physical fields of d phi^(a)/dz,
    // which those passes leave unused
    for (int a = 1; a < c->M; ++a) ix.w[a] = ix.d[a][1];
    c->nP = f;
    double fact = 1.0;
    ix.invfact[0] = 1.0;
    for (int k = 1; k <= MAXM; ++k) {
        fact *= k;
        ix.invfact[k] = 1.0 / fact;
    }
}

// the x-intermediate slots of the diagonal passes (HOS_TILE_DIAG): slots 0
// (phi^(L+1)) and 1 (eta), M(M-1)/2 diagonal slots and the M + 4 fields the
// RHS pass transforms itself.  A context holds them whether or not the mode
// is on, so that it can be switched without reallocating.
static void diag_slots(hos_ctx *c)
{
    c->nI = std::max(c->nI, Eng<float>::diag_base(c->M) + c->M + 4);
}

static size_t footprint(const hos_ctx &c)
{
    const size_t rb = c.f32 ? 4 : 8;
    size_t b = 0;
    b += 3 * (size_t)c.nm * 2 * c.fs * 16;         // U, A, Y
    b += (size_t)c.nK * c.nm * 2 * c.fs * 16;      // K
    b += (size_t)c.nm * c.nI * c.fsI * 2 * rb;     // I
    b += (size_t)c.nm * c.nP * c.fsP * rb;         // P
    b += (size_t)c.fsP * 8;                        // staging
    b += (size_t)c.nm * c.fs * 8;                  // tanh(|k| h) tables
    return b;
}

extern "C" size_t hos_footprint(const hos_params *p, int nmembers)
{
    if (validate(p, nmembers) != HOS_OK) return 0;
    hos_ctx c;
    c.p = *p;
    c.nm = nmembers;
    layout(&c);
    diag_slots(&c);
    setup_scheme(&c, p->scheme);
    c.nK = c.subdiag ? 0 : c.S;
    size_t b = footprint(c);
    b += (size_t)2 * 2 * c.fs * 16;  // K slots for hos_rhs / hos_zvel
    return b;
}

template <class T>
static int upload_twiddles(int n, void **dst)
{
    std::vector<cx<T>> h(n);
    const long double PI = 3.14159265358979323846264338327950288L;
    for (int k = 0; k < n; ++k) {
        long double ang = -2.0L * PI * (long double)k / (long double)n;
        h[k] = {T(cosl(ang)), T(sinl(ang))};
    }
    HIP_TRY(hipMalloc(dst, sizeof(cx<T>) * n));
    HIP_TRY(hipMemcpy(*dst, h.data(), sizeof(cx<T>) * n, hipMemcpyHostToDevice));
    return HOS_OK;
}

static void drop_graphs(hos_ctx *c);
static int th_table(hos_ctx *c, int m);

static void release(hos_ctx *c)
{
    drop_graphs(c);
    void *ptrs[] = {c->U, c->A, c->Y, c->K, c->I, c->P, c->twx, c->twy, c->kx, c->t, c->dt, c->H, c->vol,
                    c->hpart, c->mean0, c->depth, c->th, c->hostbuf};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->step_ev) (void)hipEventDestroy(e);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    for (int g = 0; g < hos_ctx::MAXG; ++g) {
        if (c->gst[g]) (void)hipStreamDestroy(c->gst[g]);
        if (c->gjoin[g]) (void)hipEventDestroy(c->gjoin[g]);
    }
    if (c->gfork) (void)hipEventDestroy(c->gfork);
    if (c->gskew) (void)hipEventDestroy(c->gskew);
}


extern "C" int hos_create(const hos_params *p, int nmembers, hos_ctx **out)
{
    if (!out) return fail(HOS_E_INVALID, "null output pointer");
    *out = nullptr;
    int rc = validate(p, nmembers);
    if (rc) return rc;
    hos_ctx *c = new hos_ctx();
    c->p = *p;
    c->nm = nmembers;
    layout(c);
    diag_slots(c);
    if ((rc = setup_scheme(c, p->scheme))) {
        delete c;
        return rc;
    }
    c->nK = c->subdiag ? 2 : (c->S > 2 ? c->S : 2);
    for (int k = HOS_TILE_C2R; k <= HOS_TILE_PAIR; ++k) c->tx[k] = default_tx(c->f32, c->logny, k);
    c->xwave = default_xwave(c);
    c->ngroups = default_groups(c);
    c->diag = default_diag(c);
    const int nm = nmembers;
    auto bail = [&](int code) {
        release(c);
        delete c;
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(HOS_E_DEVICE, "hipStreamCreate failed"));
    for (int g = 0; g < hos_ctx::MAXG; ++g) {
        if (hipStreamCreateWithFlags(&c->gst[g], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->gjoin[g], hipEventDisableTiming) != hipSuccess)
            return bail(fail(HOS_E_DEVICE, "group stream set-up failed"));
    }
    if (hipEventCreateWithFlags(&c->gfork, hipEventDisableTiming) != hipSuccess)
        return bail(fail(HOS_E_DEVICE, "group event set-up failed"));
    (void)hipEventCreate(&c->ev0);
    (void)hipEventCreate(&c->ev1);
    const size_t rb = c->f32 ? 4 : 8;
    const size_t st = (size_t)nm * 2 * c->fs * sizeof(cx<double>);
    const size_t kslots = (size_t)(c->subdiag ? 2 : c->S * nm);
    const size_t kbytes = (kslots < 2 ? 2 : kslots) * 2 * c->fs * sizeof(cx<double>);
    const int nparts = x_nparts(c);
    struct {
        void **p;
        size_t n;
    } allocs[] = {{(void **)&c->U, st},
                  {(void **)&c->A, st},
                  {(void **)&c->Y, st},
                  {(void **)&c->K, kbytes},
                  {&c->I, (size_t)nm * c->nI * c->fsI * 2 * rb},
                  {&c->P, (size_t)nm * c->nP * c->fsP * rb},
                  {(void **)&c->kx, sizeof(double) * p->Nx},
                  {(void **)&c->t, sizeof(double) * nm},
                  {(void **)&c->dt, sizeof(double) * nm},
                  {(void **)&c->H, sizeof(double) * nm},
                  {(void **)&c->vol, sizeof(double) * nm},
                  {(void **)&c->mean0, sizeof(double) * nm},
                  {(void **)&c->depth, sizeof(double) * nm},
                  {(void **)&c->th, sizeof(double) * nm * c->fs},
                  {(void **)&c->hpart, sizeof(double) * nm * nparts},
                  {(void **)&c->hostbuf, sizeof(double) * c->fsP}};
    for (auto &al : allocs) {
        if (hipMalloc(al.p, al.n) != hipSuccess)
            return bail(fail(HOS_E_NOMEM, "hipMalloc of %zu bytes failed", al.n));
        (void)hipMemset(*al.p, 0, al.n);
    }
    // the clears above run on the null stream; the context stream does not wait for it
    if (hipDeviceSynchronize() != hipSuccess) return bail(fail(HOS_E_DEVICE, "device sync failed"));
    // kx table, operators_par.c:55-63 (reference) or signed
    std::vector<double> kx(p->Nx);
    const double Kx0 = 2 * M_PI / p->Lx;
    for (int i = 0; i < p->Nx; ++i) {
        if (p->kx_quirk)
            kx[i] = (i <= p->Nx / 2 + 1) ? i * Kx0 : -(p->Nx - i) * Kx0;
        else
            kx[i] = (i <= p->Nx / 2) ? i * Kx0 : -(p->Nx - i) * Kx0;
    }
    if (hipMemcpy(c->kx, kx.data(), sizeof(double) * p->Nx, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(HOS_E_DEVICE, "kx upload failed"));
    rc = c->f32 ? upload_twiddles<float>(p->Nx, &c->twx) : upload_twiddles<double>(p->Nx, &c->twx);
    if (!rc) rc = c->f32 ? upload_twiddles<float>(p->Ny, &c->twy) : upload_twiddles<double>(p->Ny, &c->twy);
    if (rc) return bail(rc);
    c->host_t.assign(nm, 0.0);
    c->host_dt.assign(nm, 0.0);
    c->masked.assign(nm, 1);  // zero states
    c->active = nm;
    {
        std::vector<double> dep(nm, p->depth);
        if (hipMemcpy(c->depth, dep.data(), sizeof(double) * nm, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(HOS_E_DEVICE, "depth upload failed"));
        for (int m = 0; m < nm; ++m)
            if ((rc = th_table(c, m))) return bail(rc);
    }
    c->gexec_by_active.assign(GSLOTS * (nm + 1), nullptr);
    c->graph_by_active.assign(GSLOTS * (nm + 1), nullptr);
    if (hipDeviceSynchronize() != hipSuccess) return bail(fail(HOS_E_DEVICE, "device sync failed"));
    *out = c;
    return HOS_OK;
}

extern "C" int hos_destroy(hos_ctx *c)
{
    if (!c) return HOS_OK;
    (void)hipStreamSynchronize(c->stream);
    release(c);
    delete c;
    return HOS_OK;
}

extern "C" int hos_num_members(hos_ctx *c) { return c ? c->nm : 0; }

// ---------------------------------------------------------------------------

static RkArgs rk_base(hos_ctx *c)
{
    RkArgs rk{};
    rk.g = c->p.g;
    rk.rampflg = c->p.rampflg;
    rk.tramp = c->p.Tramp;
    rk.t = c->t;
    rk.dt = c->dt;
    rk.U = c->U;
    rk.A = c->A;
    rk.Yn = c->Y;
    rk.K = c->K;
    rk.sK = 2 * c->fs;
    rk.hpart = c->hpart;
    rk.vol = c->vol;
    rk.mean_level = c->p.mean_level;
    rk.mean0 = c->mean0;
    rk.nstages = c->S;
    return rk;
}

// groups this step runs in: the setting, while every group keeps at least
// two members; the generic Butcher path and the per-kernel profile run as one
// chain
static int groups_now(const hos_ctx *c)
{
    if (!c->subdiag || c->prof_on) return 1;
    int ng = c->ngroups;
    while (ng > 1 && c->active < 2 * ng) --ng;
    return ng;
}
// graph slot of (active count, group count, group): groups 1..MAXG take
// 1 + 2 + .. + MAXG slots per active count
static size_t graph_slot(const hos_ctx *c, int ng, int g)
{
    return GSLOTS * (size_t)c->active + (size_t)(ng * (ng - 1) / 2 + g);
}

// members [m0, m0+n) of group g of ng: nearly equal ranges in order
static void group_range(const hos_ctx *c, int g, int ng, int *m0, int *n)
{
    *m0 = (int)((long)c->active * g / ng);
    *n = (int)((long)c->active * (g + 1) / ng) - *m0;
}

// the RK steps skip the masked modes when every active member's state is
// masked (hos_set_state checks), with the band on and a sub-diagonal scheme
static bool band_now(const hos_ctx *c)
{
    if (!c->band_on || !c->subdiag) return false;
    for (int m = 0; m < c->active; ++m)
        if (!c->masked[m]) return false;
    return true;
}

// one RK step of members m0 .. m0+n-1, launched on stream st
template <class T>
static int launch_step(hos_ctx *c, int m0, int n, hipStream_t st, hipEvent_t mark, int mark_at)
{
    Eng<T> e{c, n, m0, st};
    e.inflight = c->active;  // the other member groups step concurrently
    e.band = band_now(c);
    const long s2 = 2 * c->fs;  // member stride of the FP64 [2][Nk][Nx] arrays
    int rc = HOS_OK, nparts = x_nparts(c);
    for (int s = 0; s < c->S && !rc; ++s) {
        RkArgs rk = rk_base(c);
        rk.stage = s;
        rk.c_s = c->c[s];
        rk.stream = e.many_rows();
        rk.band = e.band;
        nparts = e.rhs_parts(rk);  // the same at every stage
        rk.t += m0;
        rk.dt += m0;
        rk.U += m0 * s2;
        rk.A += m0 * s2;
        rk.Yn += m0 * s2;
        rk.hpart += (long)m0 * nparts;
        rk.vol += m0;
        rk.mean0 += m0;
        e.mark = s == 0 ? mark : nullptr;
        e.mark_at = mark_at;
        if (c->subdiag) {
            rk.mode = EPI_RK_SUBDIAG;
            rk.b_next = (s + 1 < c->S) ? c->b[s + 1][s] : 0.0;
            // stage 0 reads the state itself, later stages the stage input Y
            rc = e.rhs((s == 0 ? c->U : c->Y) + m0 * s2, rk);
        } else {
            // generic Butcher (one chain, m0 = 0): K_s stored, then
            // Y_{s+1} = u + dt sum_p b[s+1][p] K_p
            rk.mode = EPI_STORE_K;
            rk.K = c->K + (long)s * c->nm * s2;
            rc = e.rhs(c->Y, rk);
            CombArgs ca{};
            const bool last = (s == c->S - 1);
            ca.np = s + 1;
            for (int q = 0; q <= s; ++q) ca.coef[q] = last ? c->c[q] : c->b[s + 1][q];
            dim3 grid((unsigned)std::min<long>((s2 + NT - 1) / NT, 2048), n);
            if (!rc) {
                hipLaunchKernelGGL(k_combine, grid, dim3(NT), 0, e.S(), c->U, c->K, (long)c->nm * s2, s2,
                                   last ? c->U : c->Y, s2, c->dt, ca,
                                   (last && c->p.mean_level) ? (const double *)c->mean0 : nullptr);
                if (last)
                    hipLaunchKernelGGL(k_copy_c64, dim3(2048), dim3(NT), 0, e.S(), c->U, c->Y, (long)n * s2);
            }
        }
    }
    if (!rc)
        hipLaunchKernelGGL(k_end_step, dim3(n), dim3(NT), 0, e.S(), c->hpart + (long)m0 * nparts, nparts,
                           c->H + m0, c->t + m0, c->dt + m0, 1);
    return rc;
}

static int launch_group(hos_ctx *c, int g, int ng, hipStream_t st, hipEvent_t mark = nullptr, int mark_at = -1)
{
    int m0 = 0, n = 0;
    group_range(c, g, ng, &m0, &n);
    return c->f32 ? launch_step<float>(c, m0, n, st, mark, mark_at) : launch_step<double>(c, m0, n, st, mark, mark_at);
}

// one step, stream-ordered: one chain on the context stream, or two groups
// forked from it and joined back
static int step_once(hos_ctx *c)
{
    const int ng = groups_now(c);
    if (ng == 1) return launch_group(c, 0, 1, c->stream);
    HIP_TRY(hipEventRecord(c->gfork, c->stream));
    for (int g = 0; g < ng; ++g) {
        HIP_TRY(hipStreamWaitEvent(c->gst[g], c->gfork, 0));
        int rc = launch_group(c, g, ng, c->gst[g]);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(c->gjoin[g], c->gst[g]));
        HIP_TRY(hipStreamWaitEvent(c->stream, c->gjoin[g], 0));
    }
    return HOS_OK;
}


extern "C" int hos_use_graph(hos_ctx *c, int enable)
{
    if (!c) return fail(HOS_E_INVALID, "null context");
    c->use_graph = enable ? 1 : 0;
    return HOS_OK;
}

static void drop_graphs(hos_ctx *c)
{
    for (size_t i = 0; i < c->gexec_by_active.size(); ++i) {
        if (c->gexec_by_active[i]) (void)hipGraphExecDestroy(c->gexec_by_active[i]);
        if (c->graph_by_active[i]) (void)hipGraphDestroy(c->graph_by_active[i]);
        c->gexec_by_active[i] = nullptr;
        c->graph_by_active[i] = nullptr;
    }
}

// one captured step per (active-member count, group): grid sizes are baked in.
// Each group's chain is its own graph, launched on the group's stream, so the
// two chains run concurrently.
static int build_graph(hos_ctx *c)
{
    const int ng = groups_now(c);
    for (int g = 0; g < ng; ++g) {
        const size_t slot = graph_slot(c, ng, g);
        if (c->gexec_by_active[slot]) continue;
        hipStream_t st = ng == 1 ? c->stream : c->gst[g];
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        c->capturing = true;
        int rc = launch_group(c, g, ng, st);
        c->capturing = false;
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(st, &gr);
        if (rc) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        if (e != hipSuccess) return fail(HOS_E_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
        c->graph_by_active[slot] = gr;
        HIP_TRY(hipGraphInstantiate(&c->gexec_by_active[slot], gr, nullptr, nullptr, 0));
    }
    return HOS_OK;
}

// one step by graph replay: one graph on the context stream, or the groups'
// graphs forked from it and joined back
static int replay_step(hos_ctx *c)
{
    const int ng = groups_now(c);
    if (ng == 1) {
        HIP_TRY(hipGraphLaunch(c->gexec_by_active[graph_slot(c, 1, 0)], c->stream));
        return HOS_OK;
    }
    HIP_TRY(hipEventRecord(c->gfork, c->stream));
    for (int g = 0; g < ng; ++g) {
        HIP_TRY(hipStreamWaitEvent(c->gst[g], c->gfork, 0));
        HIP_TRY(hipGraphLaunch(c->gexec_by_active[graph_slot(c, ng, g)], c->gst[g]));
        HIP_TRY(hipEventRecord(c->gjoin[g], c->gst[g]));
        HIP_TRY(hipStreamWaitEvent(c->stream, c->gjoin[g], 0));
    }
    return HOS_OK;
}

// enqueue nsteps steps of the active members with the dt already on the device
static int step_enqueue(hos_ctx *c, int nsteps)
{
    if (c->async_pending) return fail(HOS_E_STATE, "a previous hos_step_async has not been waited for");
    if (nsteps < 0) return fail(HOS_E_INVALID, "nsteps < 0");
    {
        // the captured graphs bake in the band's launch shapes
        const int b = band_now(c) ? 1 : 0;
        if (b != c->graph_band) {
            drop_graphs(c);
            c->graph_band = b;
        }
    }
    if (c->use_graph && c->active > 0) {
        int rc = build_graph(c);
        if (rc) return rc;
    }
    while ((int)c->step_ev.size() < nsteps + 1) {
        hipEvent_t e = nullptr;
        HIP_TRY(hipEventCreate(&e));
        c->step_ev.push_back(e);
    }
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    HIP_TRY(hipEventRecord(c->step_ev[0], c->stream));
    const int ng = c->active > 0 ? groups_now(c) : 1;
    const bool skew = ng == 2 && c->skew && nsteps > 0;
    if (skew) {
        // The two member groups run free over the nsteps (no join per step):
        // their members are independent.  Group 1 starts when group 0's
        // first RHS has done its W level M-2, so while group 0 is in its
        // RHS pass (bound by instruction issue) group 1 is in its level
        // passes (bound by memory), and the offset holds over the steps
        // (equal halves, equal step times).  Step times are group 0's.
        if (!c->gskew) HIP_TRY(hipEventCreateWithFlags(&c->gskew, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(c->gfork, c->stream));
        for (int g = 0; g < ng; ++g) HIP_TRY(hipStreamWaitEvent(c->gst[g], c->gfork, 0));
        for (int n = 0; n < nsteps; ++n) {
            int rc = HOS_OK;
            if (n == 0)  // stream-ordered: the mark is an event between two launches
                rc = launch_group(c, 0, ng, c->gst[0], c->gskew, std::max(0, c->M - 2));
            else if (c->use_graph)
                HIP_TRY(hipGraphLaunch(c->gexec_by_active[graph_slot(c, ng, 0)], c->gst[0]));
            else
                rc = launch_group(c, 0, ng, c->gst[0]);
            if (rc) return rc;
            HIP_TRY(hipEventRecord(c->step_ev[n + 1], c->gst[0]));
        }
        for (int g = 1; g < ng; ++g) {
            HIP_TRY(hipStreamWaitEvent(c->gst[g], c->gskew, 0));
            for (int n = 0; n < nsteps; ++n) {
                if (c->use_graph) {
                    HIP_TRY(hipGraphLaunch(c->gexec_by_active[graph_slot(c, ng, g)], c->gst[g]));
                } else {
                    int rc = launch_group(c, g, ng, c->gst[g]);
                    if (rc) return rc;
                }
            }
        }
        for (int g = 0; g < ng; ++g) {
            HIP_TRY(hipEventRecord(c->gjoin[g], c->gst[g]));
            HIP_TRY(hipStreamWaitEvent(c->stream, c->gjoin[g], 0));
        }
    }
    const int serial = skew ? 0 : nsteps;  // steps still to enqueue
    for (int n = 0; n < serial && c->active > 0; ++n) {
        if (c->use_graph) {
            int rc = replay_step(c);
            if (rc) return rc;
        } else {
            int rc = step_once(c);
            if (rc) return rc;
        }
        HIP_TRY(hipEventRecord(c->step_ev[n + 1], c->stream));
    }
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipGetLastError());
    c->async_pending = true;
    c->async_steps = c->active > 0 ? nsteps : 0;
    return HOS_OK;
}

static int step_finish(hos_ctx *c)
{
    if (!c->async_pending) return HOS_OK;
    c->async_pending = false;
    const int nsteps = c->async_steps;
    HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = nsteps > 0 ? ms / nsteps : 0.0;
    c->step_ms.resize(nsteps);
    for (int n = 0; n < nsteps; ++n) {
        float sm = 0.f;
        HIP_TRY(hipEventElapsedTime(&sm, c->step_ev[n], c->step_ev[n + 1]));
        c->step_ms[n] = sm;
    }
    harvest_prof(c);
    for (int m = 0; m < c->active; ++m) c->host_t[m] += nsteps * c->host_dt[m];
    return HOS_OK;
}

// entry points that write the state, the time or the context's scratch: not
// while a hos_step_async is in flight (its kernels read them)
static int idle(const hos_ctx *c)
{
    return c->async_pending ? fail(HOS_E_STATE, "a hos_step_async is in flight (hos_wait first)") : HOS_OK;
}

static int upload_dt(hos_ctx *c, const double *dt)
{
    for (int m = 0; m < c->nm; ++m) c->host_dt[m] = dt[m];
    HIP_TRY(hipMemcpy(c->dt, c->host_dt.data(), sizeof(double) * c->nm, hipMemcpyHostToDevice));
    return HOS_OK;
}

extern "C" int hos_step_dt(hos_ctx *c, int nsteps, const double *dt)
{
    if (!c || !dt) return fail(HOS_E_INVALID, "null argument");
    // checked before the dt upload: a step in flight reads the device dt
    if (int e = idle(c)) return e;
    if (nsteps < 0) return fail(HOS_E_INVALID, "nsteps < 0");
    int rc = upload_dt(c, dt);
    if (!rc) rc = step_enqueue(c, nsteps);
    if (!rc) rc = step_finish(c);
    return rc;
}

extern "C" int hos_step(hos_ctx *c, int nsteps, double dt)
{
    if (!c) return fail(HOS_E_INVALID, "null context");
    std::vector<double> d(c->nm, dt);
    return hos_step_dt(c, nsteps, d.data());
}

extern "C" int hos_step_async(hos_ctx *c, int nsteps)
{
    if (!c) return fail(HOS_E_INVALID, "null context");
    return step_enqueue(c, nsteps);
}

extern "C" int hos_wait(hos_ctx *c)
{
    if (!c) return fail(HOS_E_INVALID, "null context");
    return step_finish(c);
}

extern "C" int hos_query(hos_ctx *c)
{
    if (!c) return fail(HOS_E_INVALID, "null context");
    if (!c->async_pending) return 0;
    const hipError_t e = hipEventQuery(c->ev1);
    if (e == hipErrorNotReady) return 1;
    if (e != hipSuccess) return fail(HOS_E_DEVICE, "hos_query: %s", hipGetErrorString(e));
    return step_finish(c);
}

extern "C" int hos_set_active(hos_ctx *c, int n)
{
    if (!c) return fail(HOS_E_INVALID, "null context");
    if (c->async_pending) return fail(HOS_E_STATE, "step in flight");
    if (n < 0 || n > c->nm) return fail(HOS_E_INVALID, "active count %d out of [0, %d]", n, c->nm);
    c->active = n;
    return HOS_OK;
}

extern "C" int hos_get_active(hos_ctx *c) { return c ? c->active : 0; }

// tanh(|k| h) of member m from its depth (device, FP64), synchronous
static int th_table(hos_ctx *c, int m)
{
    const long n = c->fs;
    hipLaunchKernelGGL(k_th_table, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, c->stream, ktab(c), c->p.Nx, m,
                       c->th);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    return HOS_OK;
}

extern "C" int hos_set_member(hos_ctx *c, int m, double depth, double dt, double t)
{
    if (!c) return fail(HOS_E_INVALID, "null context");
    if (m < 0 || m >= c->nm) return fail(HOS_E_INVALID, "member %d out of range [0, %d)", m, c->nm);
    if (c->async_pending) return fail(HOS_E_STATE, "step in flight");
    c->host_dt[m] = dt;
    c->host_t[m] = t;
    HIP_TRY(hipMemcpy(c->depth + m, &depth, sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->dt + m, &c->host_dt[m], sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->t + m, &c->host_t[m], sizeof(double), hipMemcpyHostToDevice));
    return th_table(c, m);
}

extern "C" int hos_copy_member(hos_ctx *c, int src, int dst)
{
    if (!c) return fail(HOS_E_INVALID, "null context");
    if (src < 0 || src >= c->nm || dst < 0 || dst >= c->nm) return fail(HOS_E_INVALID, "member out of range");
    if (c->async_pending) return fail(HOS_E_STATE, "step in flight");
    if (src == dst) return HOS_OK;
    const size_t st = 2 * c->fs * sizeof(cx<double>);
    HIP_TRY(hipMemcpyAsync(c->U + dst * 2 * c->fs, c->U + src * 2 * c->fs, st, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->Y + dst * 2 * c->fs, c->Y + src * 2 * c->fs, st, hipMemcpyDeviceToDevice, c->stream));
    double *scal[] = {c->t, c->dt, c->depth, c->mean0, c->H, c->vol};
    for (double *a : scal)
        HIP_TRY(hipMemcpyAsync(a + dst, a + src, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->th + dst * c->fs, c->th + src * c->fs, c->fs * sizeof(double), hipMemcpyDeviceToDevice,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->host_t[dst] = c->host_t[src];
    c->host_dt[dst] = c->host_dt[src];
    c->masked[dst] = c->masked[src];
    return HOS_OK;
}

extern "C" int hos_last_step_ms(hos_ctx *c, double *ms)
{
    if (!c || !ms) return fail(HOS_E_INVALID, "null argument");
    *ms = c->last_ms;
    return HOS_OK;
}

extern "C" int hos_step_times(hos_ctx *c, double *ms, int n)
{
    if (!c || (!ms && n > 0)) return fail(HOS_E_INVALID, "null argument");
    const int k = std::min<int>(n, (int)c->step_ms.size());
    for (int i = 0; i < k; ++i) ms[i] = c->step_ms[i];
    return (int)c->step_ms.size();
}

extern "C" int hos_profile(hos_ctx *c, int enable)
{
    if (!c) return fail(HOS_E_INVALID, "null context");
    c->prof_on = enable ? 1 : 0;
    for (int k = 0; k < HOS_K_NKINDS; ++k) {
        c->kms[k] = c->kbytes[k] = 0.0;
        c->kn[k] = 0;
    }
    return HOS_OK;
}

extern "C" int hos_kernel_stats(hos_ctx *c, int kind, double *ms_total, long *launches, double *bytes_total)
{
    if (!c) return fail(HOS_E_INVALID, "null context");
    if (kind < 0 || kind >= HOS_K_NKINDS) return fail(HOS_E_INVALID, "kernel kind %d out of range", kind);
    if (ms_total) *ms_total = c->kms[kind];
    if (launches) *launches = c->kn[kind];
    if (bytes_total) *bytes_total = c->kbytes[kind];
    return HOS_OK;
}

extern "C" int hos_sync(hos_ctx *c)
{
    if (!c) return fail(HOS_E_INVALID, "null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return HOS_OK;
}

extern "C" int hos_set_tile(hos_ctx *c, int kind, int tx)
{
    if (!c) return fail(HOS_E_INVALID, "null context");
    if (kind == HOS_TILE_PLAN) {
        if (tx < 0 || tx > NPLANS) return fail(HOS_E_INVALID, "FFT plan %d: 0 default, 1 .. %d", tx, NPLANS);
        if (tx > 1 && c->logny < PLAN_LOGN_MIN)
            return fail(HOS_E_INVALID, "FFT plans 2 .. %d need Ny >= %d", NPLANS, 1 << PLAN_LOGN_MIN);
        if (c->async_pending) return fail(HOS_E_STATE, "step in flight");
        c->plan = tx == 0 ? 0 : tx - 1;
        drop_graphs(c);
        return HOS_OK;
    }
    if (kind == HOS_TILE_BAND) {
        if (tx < 0 || tx > 2) return fail(HOS_E_INVALID, "band setting %d: 0 default, 1 off, 2 on", tx);
        if (c->async_pending) return fail(HOS_E_STATE, "step in flight");
        c->band_on = tx != 1;
        drop_graphs(c);
        return HOS_OK;
    }
    if (kind == HOS_TILE_FUSE) {
        if (tx < 0 || tx > 2) return fail(HOS_E_INVALID, "fused level passes %d: 0 default, 1 off, 2 on", tx);
        if (c->async_pending) return fail(HOS_E_STATE, "step in flight");
        c->fuse = tx != 1;
        drop_graphs(c);
        return HOS_OK;
    }
    if (kind == HOS_TILE_DIAG) {
        if (tx < 0 || tx > 2) return fail(HOS_E_INVALID, "diagonal passes %d: 0 default, 1 off, 2 on", tx);
        if (c->async_pending) return fail(HOS_E_STATE, "step in flight");
        c->diag = tx == 0 ? default_diag(c) : tx - 1;
        drop_graphs(c);
        return HOS_OK;
    }
    if (kind == HOS_TILE_XBLOCK) {
        if (tx < 0 || tx > 2) return fail(HOS_E_INVALID, "blocked x-intermediates %d: 0 default, 1 off, 2 on", tx);
        if (c->async_pending) return fail(HOS_E_STATE, "step in flight");
        c->xblock = tx != 1;
        drop_graphs(c);
        return HOS_OK;
    }
    if (kind == HOS_TILE_SKEW) {
        if (tx < 0 || tx > 2) return fail(HOS_E_INVALID, "offset member groups %d: 0 default, 1 off, 2 on", tx);
        if (c->async_pending) return fail(HOS_E_STATE, "step in flight");
        c->skew = tx != 1;
        return HOS_OK;
    }
    if (kind == HOS_TILE_GROUPS) {
        if (tx < 0 || tx > hos_ctx::MAXG) return fail(HOS_E_INVALID, "member groups %d: 0 default, 1 .. %d", tx, hos_ctx::MAXG);
        if (c->async_pending) return fail(HOS_E_STATE, "step in flight");
        c->ngroups = tx == 0 ? default_groups(c) : tx;
        drop_graphs(c);
        return HOS_OK;
    }
    if (kind == HOS_TILE_XPASS) {
        if (tx < 0 || tx > 2) return fail(HOS_E_INVALID, "x-pass mode %d: 0 default, 1 rows per workgroup, 2 row per wave", tx);
        if (tx == 2 && !x_wave_ok(c)) return fail(HOS_E_INVALID, "row-per-wave x passes need the FP32 form and Nx of 512 or 1024");
        if (c->async_pending) return fail(HOS_E_STATE, "step in flight");
        c->xwave = tx == 0 ? default_xwave(c) : tx - 1;
        drop_graphs(c);
        return HOS_OK;
    }
    if (kind < HOS_TILE_C2R || kind > HOS_TILE_ALL) return fail(HOS_E_INVALID, "unknown tile kind %d", kind);
    for (int k = HOS_TILE_C2R; k <= HOS_TILE_PAIR; ++k) {
        if (kind != HOS_TILE_ALL && kind != k) continue;
        const int v = tx ? tx : default_tx(c->f32, c->logny, k);
        if (!tx_fits(c->f32, c->logny, v))
            return fail(HOS_E_INVALID, "tile %d not available (4, 8 or 16 rows, <= 64 KiB LDS)", v);
        if (c->p.Nx % v) return fail(HOS_E_INVALID, "tile %d does not divide Nx", v);
    }
    if (c->async_pending) return fail(HOS_E_STATE, "step in flight");
    for (int k = HOS_TILE_C2R; k <= HOS_TILE_PAIR; ++k)
        if (kind == HOS_TILE_ALL || kind == k) c->tx[k] = tx ? tx : default_tx(c->f32, c->logny, k);
    drop_graphs(c);
    return HOS_OK;
}

extern "C" int hos_get_tile(hos_ctx *c, int kind)
{
    if (!c) return fail(HOS_E_INVALID, "null context");
    if (kind == HOS_TILE_XPASS) return c->xwave ? 2 : 1;
    if (kind == HOS_TILE_BAND) return c->band_on ? 2 : 1;
    if (kind == HOS_TILE_FUSE) return c->fuse ? 2 : 1;
    if (kind == HOS_TILE_SKEW) return c->skew ? 2 : 1;
    if (kind == HOS_TILE_DIAG) return c->diag ? 2 : 1;
    if (kind == HOS_TILE_XBLOCK) return c->xblock ? 2 : 1;
    if (kind == HOS_TILE_PLAN) return c->plan + 1;
    if (kind == HOS_TILE_GROUPS) return c->ngroups;
    if (kind < HOS_TILE_C2R || kind > HOS_TILE_PAIR) return fail(HOS_E_INVALID, "unknown tile kind %d", kind);
    return c->tx[kind];
}

// ---------------------------------------------------------------------------
// autotuner: y tiles per kernel kind, then the x-pass mode, timed on the
// context's own members
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int, int, int, int, int>, std::array<int, 7>> g_tune_cache;

extern "C" int hos_autotune(hos_ctx *c, int nsteps, int *tiles_out)
{
    if (!c) return fail(HOS_E_INVALID, "null context");
    if (c->async_pending) return fail(HOS_E_STATE, "step in flight");
    if (nsteps < 1) nsteps = 2;
    const auto key = std::make_tuple(c->p.Nx, c->p.Ny, c->M, (int)c->f32, c->active, (int)c->subdiag,
                                     (int)band_now(c));
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            for (int k = 0; k < 3; ++k) c->tx[k] = it->second[k];
            c->xwave = it->second[3];
            c->ngroups = it->second[4];
            c->plan = it->second[5];
            c->diag = it->second[6];
            drop_graphs(c);
            if (tiles_out)
                for (int k = 0; k < 3; ++k) tiles_out[k] = c->tx[k];
            return HOS_OK;
        }
    }
    // save what a trial step changes: U, Y, t, dt, H, volume
    const size_t st = (size_t)c->nm * 2 * c->fs * sizeof(cx<double>);
    std::vector<double> t_save = c->host_t, dt_save = c->host_dt, h_save(c->nm), vol_save(c->nm);
    HIP_TRY(hipMemcpy(h_save.data(), c->H, sizeof(double) * c->nm, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(vol_save.data(), c->vol, sizeof(double) * c->nm, hipMemcpyDeviceToHost));
    void *saveU = nullptr, *saveY = nullptr;
    HIP_TRY(hipMalloc(&saveU, st));
    if (hipMalloc(&saveY, st) != hipSuccess) {
        (void)hipFree(saveU);
        return fail(HOS_E_NOMEM, "autotune scratch");
    }
    (void)hipMemcpyAsync(saveU, c->U, st, hipMemcpyDeviceToDevice, c->stream);
    (void)hipMemcpyAsync(saveY, c->Y, st, hipMemcpyDeviceToDevice, c->stream);
    std::vector<double> dtev(c->nm);
    for (int m = 0; m < c->nm; ++m) dtev[m] = c->host_dt[m] > 0 ? c->host_dt[m] : 1e-3;
    int rc = upload_dt(c, dtev.data());
    const int cand[3] = {4, 8, 16};
    std::array<int, 7> best = {c->tx[0], c->tx[1], c->tx[2], c->xwave, c->ngroups, c->plan, c->diag};
    double best_ms = 1e30;
    const int saved_graph = c->use_graph;
    c->use_graph = 1;
    // one trial of the current launch shapes: capture and warm up, then the
    // faster of two timed runs of nsteps (one run is too noisy to rank shapes
    // a few percent apart)
    auto trial = [&](double &ms) {
        drop_graphs(c);
        int r = step_enqueue(c, 1);
        if (!r) r = step_finish(c);
        ms = 1e30;
        for (int rep = 0; rep < 2 && !r; ++rep) {
            r = step_enqueue(c, nsteps);
            if (!r) r = step_finish(c);
            if (!r) ms = std::min(ms, c->last_ms);
        }
        return r;
    };
    for (int a = 0; a < 3 && !rc; ++a)
        for (int b = 0; b < 3 && !rc; ++b)
            for (int d = 0; d < 3 && !rc; ++d) {
                const int tt[3] = {cand[a], cand[b], cand[d]};
                bool ok = true;
                for (int k = 0; k < 3; ++k) ok &= tx_fits(c->f32, c->logny, tt[k]) && c->p.Nx % tt[k] == 0;
                if (!ok) continue;
                for (int k = 0; k < 3; ++k) c->tx[k] = tt[k];
                double ms = 0.0;
                rc = trial(ms);
                if (!rc && ms < best_ms) {
                    best_ms = ms;
                    best = {tt[0], tt[1], tt[2], c->xwave, c->ngroups, c->plan, c->diag};
                }
            }
    if (!rc && x_wave_ok(c)) {
        // the other x-pass mode with the best tiles
        for (int k = 0; k < 3; ++k) c->tx[k] = best[k];
        c->xwave = !best[3];
        double ms = 0.0;
        rc = trial(ms);
        if (!rc && ms < best_ms) {
            best_ms = ms;
            best[3] = c->xwave;
        }
    }
    if (!rc) {
        // the other W-expansion form (diagonal passes or level passes) with the
        // best tiles and x-pass mode
        for (int k = 0; k < 3; ++k) c->tx[k] = best[k];
        c->xwave = best[3];
        c->diag = !best[6];
        double ms = 0.0;
        rc = trial(ms);
        if (!rc && ms < best_ms) {
            best_ms = ms;
            best[6] = c->diag;
        }
        c->diag = best[6];
    }
    if (!rc && c->logny >= PLAN_LOGN_MIN) {
        // the FFT radix plans of the RHS pass with the best tiles, x-pass mode
        // and W-expansion form
        for (int k = 0; k < 3; ++k) c->tx[k] = best[k];
        c->xwave = best[3];
        c->diag = best[6];
        const int tried = best[5];
        for (int pl = 0; pl < NPLANS && !rc; ++pl) {
            if (pl == tried) continue;
            c->plan = pl;
            double ms = 0.0;
            rc = trial(ms);
            if (!rc && ms < best_ms) {
                best_ms = ms;
                best[5] = pl;
            }
        }
        c->plan = best[5];
    }
    if (!rc && c->subdiag) {
        // every member-group count (each group at least two members) with
        // the best tiles, x-pass mode and FFT plan, in calls of at least 8
        // steps: two groups run offset over a call (HOS_TILE_SKEW), which
        // pays over several steps, not at the ends of a call
        for (int k = 0; k < 3; ++k) c->tx[k] = best[k];
        c->xwave = best[3];
        c->plan = best[5];
        c->diag = best[6];
        const int ns = nsteps;
        nsteps = std::max(nsteps, 8);
        double gbest = 1e30;
        for (int ng = 1; ng <= hos_ctx::MAXG && !rc; ++ng) {
            if (ng > 1 && c->active < 2 * ng) continue;
            c->ngroups = ng;
            double ms = 0.0;
            rc = trial(ms);
            if (!rc && ms < gbest) {
                gbest = ms;
                best[4] = ng;
            }
        }
        nsteps = ns;
    }
    c->use_graph = saved_graph;
    for (int k = 0; k < 3; ++k) c->tx[k] = best[k];
    c->xwave = best[3];
    c->ngroups = best[4];
    c->plan = best[5];
    c->diag = best[6];
    drop_graphs(c);
    (void)hipMemcpyAsync(c->U, saveU, st, hipMemcpyDeviceToDevice, c->stream);
    (void)hipMemcpyAsync(c->Y, saveY, st, hipMemcpyDeviceToDevice, c->stream);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(saveU);
    (void)hipFree(saveY);
    c->host_t = t_save;
    (void)hipMemcpy(c->t, c->host_t.data(), sizeof(double) * c->nm, hipMemcpyHostToDevice);
    (void)hipMemcpy(c->H, h_save.data(), sizeof(double) * c->nm, hipMemcpyHostToDevice);
    (void)hipMemcpy(c->vol, vol_save.data(), sizeof(double) * c->nm, hipMemcpyHostToDevice);
    {
        const int rd = upload_dt(c, dt_save.data());
        if (!rc) rc = rd;
    }
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tune_cache[key] = best;
    }
    if (tiles_out)
        for (int k = 0; k < 3; ++k) tiles_out[k] = best[k];
    return HOS_OK;
}

// ---------------------------------------------------------------------------
// state in / out
static int check_member(hos_ctx *c, int m)
{
    if (!c) return fail(HOS_E_INVALID, "null context");
    if (m < 0 || m >= c->nm) return fail(HOS_E_INVALID, "member %d out of range [0, %d)", m, c->nm);
    return HOS_OK;
}


extern "C" int hos_set_state(hos_ctx *c, int member, const double *u)
{
    int rc = check_member(c, member);
    if (rc) return rc;
    if (int e = idle(c)) return e;
    if (!u) return fail(HOS_E_INVALID, "null state");
    const size_t n = 2 * c->fs * sizeof(cx<double>);
    HIP_TRY(hipMemcpyAsync(c->U + member * 2 * c->fs, u, n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->Y + member * 2 * c->fs, u, n, hipMemcpyHostToDevice, c->stream));
    const double m0 = u[0];  // Re eta-hat(0,0)
    HIP_TRY(hipMemcpyAsync(c->mean0 + member, &m0, sizeof(double), hipMemcpyHostToDevice, c->stream));
    // is the state zero on every mode the mask removes (operators_par.c:309-322)?
    bool ok = true;
    for (int f = 0; f < 2 && ok; ++f)
        for (int j = 0; j < c->Nk && ok; ++j) {
            const double *row = u + 2 * ((long)f * c->fs + (long)j * c->p.Nx);
            for (int i = 0; i < c->p.Nx; ++i) {
                const bool cut = (i >= c->mx && i < c->p.Nx - c->mx + 1) || j >= c->my;
                if (cut && (row[2 * i] != 0.0 || row[2 * i + 1] != 0.0)) {
                    ok = false;
                    break;
                }
            }
        }
    c->masked[member] = ok ? 1 : 0;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return HOS_OK;
}

extern "C" int hos_get_state(hos_ctx *c, int member, double *u)
{
    int rc = check_member(c, member);
    if (rc) return rc;
    if (!u) return fail(HOS_E_INVALID, "null state");
    HIP_TRY(hipMemcpyAsync(u, c->U + member * 2 * c->fs, 2 * c->fs * sizeof(cx<double>), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return HOS_OK;
}

extern "C" int hos_set_time(hos_ctx *c, int member, double t)
{
    int rc = check_member(c, member);
    if (rc) return rc;
    if (int e = idle(c)) return e;
    c->host_t[member] = t;
    HIP_TRY(hipMemcpyAsync(c->t + member, &c->host_t[member], sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return HOS_OK;
}

extern "C" int hos_get_time(hos_ctx *c, int member, double *t)
{
    int rc = check_member(c, member);
    if (rc) return rc;
    if (!t) return fail(HOS_E_INVALID, "null argument");
    HIP_TRY(hipMemcpyAsync(t, c->t + member, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return HOS_OK;
}

extern "C" int hos_get_energy(hos_ctx *c, int member, double *H, double *volume)
{
    int rc = check_member(c, member);
    if (rc) return rc;
    if (H) HIP_TRY(hipMemcpyAsync(H, c->H + member, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (volume) HIP_TRY(hipMemcpyAsync(volume, c->vol + member, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return HOS_OK;
}

// ---------------------------------------------------------------------------
// single-evaluation entry points on member 0 (parity / reference shims).
// They use member 0's scratch and restore its stage input afterwards.
static int restore_member0(hos_ctx *c)
{
    HIP_TRY(hipMemcpyAsync(c->Y, c->U, 2 * c->fs * sizeof(cx<double>), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return HOS_OK;
}

template <class T>
static int rhs_impl(hos_ctx *c, double t)
{
    Eng<T> e{c, 1};
    RkArgs rk = rk_base(c);
    rk.mode = EPI_STORE_K;
    rk.stage = 1;  // no Hamiltonian partials
    rk.hpart = nullptr;
    rk.K = c->K;
    HIP_TRY(hipMemcpy(c->t, &t, sizeof(double), hipMemcpyHostToDevice));
    int rc = e.rhs(c->Y, rk);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(c->t, &c->host_t[0], sizeof(double), hipMemcpyHostToDevice));
    return rc;
}

extern "C" int hos_rhs(hos_ctx *c, const double *u, double *hrhs, double t)
{
    if (!c || !u || !hrhs) return fail(HOS_E_INVALID, "null argument");
    if (int e = idle(c)) return e;
    const size_t n = 2 * c->fs * sizeof(cx<double>);
    HIP_TRY(hipMemcpyAsync(c->Y, u, n, hipMemcpyHostToDevice, c->stream));
    int rc = c->f32 ? rhs_impl<float>(c, t) : rhs_impl<double>(c, t);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(hrhs, c->K, n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipGetLastError());
    return restore_member0(c);
}

// ---------------------------------------------------------------------------
// Allocation-free RHS (hos_rhs_noalloc): the buffers a context would allocate
// live in one file-scope device array per GPU.  In host code `scratch` names a
// host-side stand-in; its device address is taken once per device with
// hipGetSymbolAddress and only that address is ever passed to the GPU.
static __device__ __align__(256) unsigned char scratch[HOS_NOALLOC_SCRATCH_BYTES];
constexpr int NOALLOC_MAXDEV = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_dev[NOALLOC_MAXDEV];

// device address of this device's scratch array (looked up on the first call
// on the device, with it current)
static int scratch_on_device(unsigned char **out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAXDEV) return fail(HOS_E_DEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(HOS_E_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_dev[dev] = (unsigned char *)ptr;
    }
    *out = g_scratch_dev[dev];
    return HOS_OK;
}

// next 256-byte aligned region of `bytes` at offset off (advanced)
static size_t carve(size_t &off, size_t bytes)
{
    const size_t o = off;
    off = (off + bytes + 255) & ~(size_t)255;
    return o;
}

// the tables hos_create uploads, computed on the device: kx (the same
// expression as hos_create), twiddles exp(-2 pi i k/n) (sincospi in FP64,
// then rounded to T), and the scalars t, dt, depth.  grid ceil(max(Nx,Ny)/NT)
template <class T>
__global__ __launch_bounds__(NT) void k_noalloc_tables(double *kx, cx<T> *twx, cx<T> *twy, double *scal, int nx,
                                                       int ny, double kx0, int quirk, double t, double depth)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i < nx) {
        kx[i] = quirk ? ((i <= nx / 2 + 1) ? i * kx0 : -(nx - i) * kx0)
                      : ((i <= nx / 2) ? i * kx0 : -(nx - i) * kx0);
        double sn, cs;
        sincospi(-2.0 * i / nx, &sn, &cs);
        twx[i] = {T(cs), T(sn)};
    }
    if (i < ny) {
        double sn, cs;
        sincospi(-2.0 * i / ny, &sn, &cs);
        twy[i] = {T(cs), T(sn)};
    }
    if (i == 0) {
        scal[0] = t;
        scal[1] = 0.0;
        scal[2] = depth;
    }
}

template <class T>
static int rhs_noalloc_impl(hos_ctx *c, unsigned char *base, const double *d_u, double *d_rhs, double t)
{
    const size_t rb = sizeof(T);
    size_t off = 0;
    const size_t oI = carve(off, (size_t)c->nI * c->fsI * 2 * rb);
    const size_t oP = carve(off, (size_t)c->nP * c->fsP * rb);
    const size_t oth = carve(off, (size_t)c->fs * sizeof(double));
    const size_t okx = carve(off, (size_t)c->p.Nx * sizeof(double));
    const size_t otx = carve(off, (size_t)c->p.Nx * sizeof(cx<T>));
    const size_t oty = carve(off, (size_t)c->p.Ny * sizeof(cx<T>));
    const size_t osc = carve(off, 3 * sizeof(double));
    if (off > HOS_NOALLOC_SCRATCH_BYTES)
        return fail(HOS_E_INVALID,
                    "hos_rhs_noalloc: %dx%d at M=%d (%s) needs %zu bytes of scratch; the limit is %zu", c->p.Nx,
                    c->p.Ny, c->M, c->f32 ? "FP32" : "FP64", off, (size_t)HOS_NOALLOC_SCRATCH_BYTES);
    c->I = base + oI;
    c->P = base + oP;
    c->th = (double *)(base + oth);
    c->kx = (double *)(base + okx);
    c->twx = base + otx;
    c->twy = base + oty;
    double *scal = (double *)(base + osc);
    c->t = scal;
    c->dt = scal + 1;
    c->depth = scal + 2;
    const int n = std::max(c->p.Nx, c->p.Ny);
    hipLaunchKernelGGL(k_noalloc_tables<T>, dim3((n + NT - 1) / NT), dim3(NT), 0, c->stream, c->kx, (cx<T> *)c->twx,
                       (cx<T> *)c->twy, scal, c->p.Nx, c->p.Ny, 2 * M_PI / c->p.Lx, c->p.kx_quirk, t, c->p.depth);
    hipLaunchKernelGGL(k_th_table, dim3((unsigned)((c->fs + NT - 1) / NT)), dim3(NT), 0, c->stream, ktab(c),
                       c->p.Nx, 0, c->th);
    Eng<T> e{c, 1};
    RkArgs rk = rk_base(c);
    rk.mode = EPI_STORE_K;
    rk.stage = 1;  // no Hamiltonian partials
    rk.hpart = nullptr;
    rk.K = (cx<double> *)d_rhs;
    int rc = e.rhs((const cx<double> *)d_u, rk);
    if (!rc) HIP_TRY(hipGetLastError());
    return rc;
}

extern "C" int hos_rhs_noalloc(const hos_params *p, const double *d_u, double *d_rhs, double t, void *stream)
{
    if (!p || !d_u || !d_rhs) return fail(HOS_E_INVALID, "null argument");
    int rc = validate(p, 1);
    if (rc) return rc;
    hos_ctx c;  // host bookkeeping only; its device pointers go into the scratch array
    c.p = *p;
    c.nm = 1;
    c.active = 1;
    layout(&c);
    const size_t n = 2 * (size_t)c.fs * sizeof(cx<double>);
    const char *a = (const char *)d_u, *b = (const char *)d_rhs;
    if (a < b + n && b < a + n) return fail(HOS_E_INVALID, "hos_rhs_noalloc: d_rhs overlaps d_u");
    for (int k = HOS_TILE_C2R; k <= HOS_TILE_PAIR; ++k) c.tx[k] = default_tx(c.f32, c.logny, k);
    c.xwave = default_xwave(&c);
    c.stream = (hipStream_t)stream;
    unsigned char *base = nullptr;
    if ((rc = scratch_on_device(&base))) return rc;
    return c.f32 ? rhs_noalloc_impl<float>(&c, base, d_u, d_rhs, t) : rhs_noalloc_impl<double>(&c, base, d_u, d_rhs, t);
}

template <class T>
static int zvel_impl(hos_ctx *c)
{
    Eng<T> e{c, 1};
    int rc = e.w_levels(c->Y);
    RkArgs rk = rk_base(c);
    rk.mode = EPI_STORE_SPEC2;
    rk.hpart = nullptr;
    for (int which = 0; which < 2 && !rc; ++which) {
        ProZvel<T> pro{e.phys(), c->ix, which, e.comp()};
        rc = e.yr2c_pair(pro, 0, c->nP - 4);
        rk.K = c->K + (long)which * 2 * c->fs;
        if (!rc) rc = e.xfwd_rhs(0, -1, c->Y, rk);
    }
    return rc;
}

extern "C" int hos_zvel(hos_ctx *c, const double *u, double *wM, double *wM2, double *w2M, double *w2M2)
{
    if (!c || !u) return fail(HOS_E_INVALID, "null argument");
    if (int e = idle(c)) return e;
    HIP_TRY(hipMemcpyAsync(c->Y, u, 2 * c->fs * sizeof(cx<double>), hipMemcpyHostToDevice, c->stream));
    int rc = c->f32 ? zvel_impl<float>(c) : zvel_impl<double>(c);
    if (rc) return rc;
    const size_t n = c->fs * sizeof(cx<double>);
    double *outs[4] = {wM, wM2, w2M, w2M2};
    for (int q = 0; q < 4; ++q)
        if (outs[q]) HIP_TRY(hipMemcpyAsync(outs[q], c->K + (long)q * c->fs, n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipGetLastError());
    return restore_member0(c);
}

template <class T>
static int fft_impl(hos_ctx *c)
{
    Eng<T> e{c, 1};
    const long n = (long)c->p.Nx * c->p.Ny;
    hipLaunchKernelGGL(k_real_from_f64<T>, dim3(1024), dim3(NT), 0, c->stream, c->hostbuf, e.P(), n);
    ProRead<T> pro{e.phys(), 0};
    int rc = e.yr2c(pro, 0, 1);
    RkArgs rk = rk_base(c);
    rk.mode = EPI_STORE_SPEC2;
    rk.hpart = nullptr;
    rk.K = c->K;
    if (!rc) rc = e.xfwd_rhs(0, 0, c->Y, rk);
    return rc;
}

extern "C" int hos_fft_2d(hos_ctx *c, const double *u_padded, double *hu)
{
    if (!c || !u_padded || !hu) return fail(HOS_E_INVALID, "null argument");
    if (int e = idle(c)) return e;
    const int Nx = c->p.Nx, Ny = c->p.Ny;
    HIP_TRY(hipMemcpy2DAsync(c->hostbuf, sizeof(double) * Ny, u_padded, sizeof(double) * (Ny + 2),
                             sizeof(double) * Ny, Nx, hipMemcpyHostToDevice, c->stream));
    int rc = c->f32 ? fft_impl<float>(c) : fft_impl<double>(c);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(hu, c->K, c->fs * sizeof(cx<double>), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    return HOS_OK;
}

template <class T>
static int ifft_impl(hos_ctx *c)
{
    Eng<T> e{c, 1};
    const XinvList ops = {{0, OP_ID, 0, 0}};
    FieldMap fm{};
    fm.src[0] = 0;
    fm.dst[0] = 0;
    int rc = e.xinv(c->Y, ops);
    if (!rc) rc = e.yc2r(fm, 1);
    hipLaunchKernelGGL(k_real_to_f64<T>, dim3(1024), dim3(NT), 0, c->stream, e.P(), c->hostbuf,
                       (long)c->p.Nx * c->p.Ny);
    return rc;
}

extern "C" int hos_ifft_2d(hos_ctx *c, const double *hu, double *u_padded)
{
    if (!c || !u_padded || !hu) return fail(HOS_E_INVALID, "null argument");
    if (int e = idle(c)) return e;
    const int Nx = c->p.Nx, Ny = c->p.Ny;
    HIP_TRY(hipMemcpyAsync(c->Y, hu, c->fs * sizeof(cx<double>), hipMemcpyHostToDevice, c->stream));
    int rc = c->f32 ? ifft_impl<float>(c) : ifft_impl<double>(c);
    if (rc) return rc;
    HIP_TRY(hipMemcpy2DAsync(u_padded, sizeof(double) * (Ny + 2), c->hostbuf, sizeof(double) * Ny,
                             sizeof(double) * Ny, Nx, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int x = 0; x < Nx; ++x) u_padded[(size_t)x * (Ny + 2) + Ny] = u_padded[(size_t)x * (Ny + 2) + Ny + 1] = 0.0;
    return restore_member0(c);
}

// Mult (operators_par.c:210-235): c2r of a and b, product on the grid, r2c.
template <class T>
static int mult_impl(hos_ctx *c)
{
    Eng<T> e{c, 1};
    const XinvList ops = {{0, OP_ID, 0, 0}, {1, OP_ID, 0, 1}};
    FieldMap fm{};
    fm.src[0] = 0;
    fm.dst[0] = 0;
    fm.src[1] = 1;
    fm.dst[1] = 1;
    int rc = e.xinv(c->Y, ops);
    if (!rc) rc = e.yc2r(fm, 2);
    ProMul<T> pro{e.phys(), 0, 1};
    if (!rc) rc = e.yr2c(pro, 0, 2);
    RkArgs rk = rk_base(c);
    rk.mode = EPI_STORE_SPEC2;
    rk.hpart = nullptr;
    rk.K = c->K;
    if (!rc) rc = e.xfwd_rhs(0, 0, c->Y, rk);
    return rc;
}

extern "C" int hos_mult(hos_ctx *c, const double *a, const double *b, double *out)
{
    if (!c || !a || !b || !out) return fail(HOS_E_INVALID, "null argument");
    if (int e = idle(c)) return e;
    const size_t n = c->fs * sizeof(cx<double>);
    HIP_TRY(hipMemcpyAsync(c->Y, a, n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->Y + c->fs, b, n, hipMemcpyHostToDevice, c->stream));
    int rc = c->f32 ? mult_impl<float>(c) : mult_impl<double>(c);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(out, c->K, n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipGetLastError());
    return restore_member0(c);
}

extern "C" int hos_spectral_op(hos_ctx *c, int op, const double *a, const double *b, double *out)
{
    if (!c || !a || !out) return fail(HOS_E_INVALID, "null argument");
    if (int e = idle(c)) return e;
    if (op < HOS_OP_DX || op > HOS_OP_DEALIAS) return fail(HOS_E_INVALID, "unknown spectral op %d", op);
    if (op == HOS_OP_SUM && !b) return fail(HOS_E_INVALID, "sum needs two inputs");
    const size_t n = c->fs * sizeof(cx<double>);
    HIP_TRY(hipMemcpyAsync(c->K, a, n, hipMemcpyHostToDevice, c->stream));
    if (b) HIP_TRY(hipMemcpyAsync(c->K + c->fs, b, n, hipMemcpyHostToDevice, c->stream));
    const int grid = (int)std::min<long>((c->fs + NT - 1) / NT, 1024);
    hipLaunchKernelGGL(k_spec_op, dim3(grid), dim3(NT), 0, c->stream, c->K, c->K + c->fs, c->K + 2 * c->fs, op,
                       c->p.Nx, ktab(c));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, c->K + 2 * c->fs, n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return HOS_OK;
}

extern "C" int hos_hamiltonian(hos_ctx *c, const double *heta, const double *heta_t, const double *hphi,
                               double *H)
{
    if (!c || !heta || !heta_t || !hphi || !H) return fail(HOS_E_INVALID, "null argument");
    if (int e = idle(c)) return e;
    const size_t n = c->fs * sizeof(cx<double>);
    HIP_TRY(hipMemcpyAsync(c->K, heta, n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->K + c->fs, heta_t, n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->K + 2 * c->fs, hphi, n, hipMemcpyHostToDevice, c->stream));
    const int nparts = x_nparts(c) * c->nm;
    const int grid = std::min(nparts, 64);
    hipLaunchKernelGGL(k_hamiltonian, dim3(grid), dim3(NT), 0, c->stream, c->K, c->K + c->fs, c->K + 2 * c->fs,
                       c->p.g, c->p.Nx, c->Nk, c->hpart);
    HIP_TRY(hipGetLastError());
    std::vector<double> part(grid);
    HIP_TRY(hipMemcpyAsync(part.data(), c->hpart, sizeof(double) * grid, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (double v : part) s += v;
    *H = s;
    return HOS_OK;
}

// ---------------------------------------------------------------------------
// plain device copy of a known size (calibrates the HBM counters and gives
// the practical copy bandwidth next to the 8 TB/s peak)
extern "C" int hos_copy_bandwidth(size_t bytes, int reps, double *gbps)
{
    if (!gbps || reps < 1 || bytes < 4096) return fail(HOS_E_INVALID, "bad argument");
    const long n = (long)(bytes / sizeof(cx<double>));
    void *a = nullptr, *b = nullptr;
    HIP_TRY(hipMalloc(&a, n * sizeof(cx<double>)));
    if (hipMalloc(&b, n * sizeof(cx<double>)) != hipSuccess) {
        (void)hipFree(a);
        return fail(HOS_E_NOMEM, "copy buffers");
    }
    (void)hipMemset(a, 0, n * sizeof(cx<double>));
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int grid = 8192;
    hipLaunchKernelGGL(k_copy_c64, dim3(grid), dim3(NT), 0, nullptr, (const cx<double> *)a, (cx<double> *)b, n);
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy_c64, dim3(grid), dim3(NT), 0, nullptr, (const cx<double> *)a, (cx<double> *)b,
                           n);
    (void)hipEventRecord(e1, nullptr);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(HOS_E_DEVICE, "copy timing: %s", hipGetErrorString(e));
    *gbps = 2.0 * n * sizeof(cx<double>) * reps / (ms * 1e-3) / 1e9;
    return HOS_OK;
}

extern "C" int hos_device_name(char *buf, int len)
{
    if (!buf || len <= 0) return fail(HOS_E_INVALID, "bad buffer");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return HOS_OK;
}

extern "C" double hos_bytes_per_point_step(const hos_params *p)
{
    // F(M) = M(M+1)/2 + M + 6 two-dimensional real FFTs per RHS, each two
    // passes (row pass, transposed column pass) that read and write N*p bytes;
    // plus 64 B of FP64 RK state traffic per point per stage.
    const int M = p->nlevs + 1;
    const double F = M * (M + 1) / 2.0 + M + 6;
    const double pb = p->precision == HOS_FP32 ? 4.0 : 8.0;
    int S = 4;
    switch (p->scheme) {
    case 1: S = 7; break;
    case 3: S = 1; break;
    case 4:
    case 5: S = 2; break;
    default: S = 4;
    }
    return S * (F * 4 * pb + 64.0);
}
