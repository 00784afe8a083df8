// hos_ctx.hpp -- the solver context and the kernel launchers (Eng<T>),
// shared by hos_engine.hip (host orchestration, C ABI) and hos_levels.hip
// (the order-templated y r2c launches, built in parallel units).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <initializer_list>
#include <vector>

#include "../../include/hos_gpu.h"
#include "hos_kernels.hpp"

using namespace hos;

// records the message for hos_last_error() and returns code (hos_engine.hip)
__attribute__((visibility("hidden"))) int fail(int code, const char *fmt, ...);

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return fail(HOS_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));   \
    } while (0)


constexpr int LOGN_MIN = 4, LOGN_MAX = 11;
// the RHS pass's alternative FFT plans are built for Ny >= 512 only
constexpr int PLAN_LOGN_MIN = 9, NPLANS = 4;

struct hos_ctx {
    hos_params p{};
    int nm = 0, M = 0, L = 0, Nk = 0, lognx = 0, logny = 0, mx = 0, my = 0;
    int nI = 0, nP = 0;
    int tx[3] = {8, 4, 4};  // y tiles: c2r, r2c, pair
    int xwave = 0;          // x inverse passes one row per wave (where available)
    int plan = 0;           // FFT radix plan of the RHS pass (fft_stage_bits; plans 1-3 need Ny >= 512)
    int fuse = 1;           // a W level's y c2r carries the next level's y r2c (k_yc2r_phi)
    int diag = 0;           // W levels as diagonal passes (k_ydiag; HOS_TILE_DIAG)
    int xblock = 1;         // blocked x-intermediates in the diagonal form with wave x passes (HOS_TILE_XBLOCK)
    // member groups: with G > 1, the active members split into G nearly equal
    // ranges that run as independent launch chains on their own streams (gst),
    // forked from and joined to the context stream each step
    static constexpr int MAXG = 4;
    int ngroups = 1;
    hipStream_t gst[MAXG] = {};
    hipEvent_t gfork = nullptr, gjoin[MAXG] = {};
    hipEvent_t gskew = nullptr;  // group 0's mark that group 1 starts behind
    int skew = 1;                // two member groups run free within a call, offset (HOS_TILE_SKEW)
    bool f32 = false;
    long fs = 0;    // complex elements per spectral field  (Nk*Nx)
    long ip = 0;    // row pitch of the x-intermediates I    (Nx + 16)
    long fsI = 0;   // complex elements per x-intermediate   (xi_rows(Nk)*ip)
    long fsP = 0;   // real elements per physical field     (Nx*Ny + one spare row)
    hipStream_t stream = nullptr;
    // device memory
    cx<double> *U = nullptr, *A = nullptr, *Y = nullptr, *K = nullptr;
    void *I = nullptr, *P = nullptr;
    void *twx = nullptr, *twy = nullptr;
    double *kx = nullptr, *t = nullptr, *dt = nullptr, *H = nullptr, *vol = nullptr, *hpart = nullptr,
           *mean0 = nullptr, *depth = nullptr;
    double *th = nullptr;  // [nm][Nk][Nx] tanh(|k| h) per member
    int active = 0;  // members 0..active-1 are stepped
    // band of masked states: masked[m] = member m's state is zero on every
    // mode the dealiasing mask removes (checked by hos_set_state); with all
    // active members masked (and band_on), the RK steps skip those modes
    std::vector<char> masked;
    int band_on = 1;
    int graph_band = -1;  // the band setting the captured graphs were built with
    std::vector<hipGraphExec_t> gexec_by_active;  // [graph_slot(active, groups, group)]
    std::vector<hipGraph_t> graph_by_active;
    bool async_pending = false;
    int async_steps = 0;
    double *hostbuf = nullptr;  // device staging [Nx][Ny] double
    int nK = 0;                 // number of FP64 [2][Nk][Nx] slots in K (per member for RK)
    HosIdx ix{};
    // Butcher table, time_schemes_par.c:127-314
    int S = 0;
    double a[7] = {}, b[7][7] = {}, c[7] = {};
    bool subdiag = true;
    // timing / graph
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    int use_graph = 0;
    std::vector<double> host_t;   // host mirror of t
    std::vector<double> host_dt;  // host mirror of dt
    // per-step and per-kernel HIP-event timing
    std::vector<hipEvent_t> step_ev;
    std::vector<double> step_ms;
    int prof_on = 0;
    bool capturing = false;
    struct ProfRec {
        int kind;
        hipEvent_t a, b;
        double bytes;
    };
    std::vector<ProfRec> recs;
    std::vector<hipEvent_t> evpool;
    size_t evused = 0;
    double kms[HOS_K_NKINDS] = {}, kbytes[HOS_K_NKINDS] = {};
    long kn[HOS_K_NKINDS] = {};
};

inline hipEvent_t pool_event(hos_ctx *c)
{
    if (c->evused < c->evpool.size()) return c->evpool[c->evused++];
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    c->evpool.push_back(e);
    c->evused++;
    return e;
}

// Records a HIP event pair around one kernel launch on the launch's stream
// when kernel profiling is on (eager launches only).
struct Prof {
    hos_ctx *c;
    hipStream_t st;
    int kind;
    double bytes;
    hipEvent_t a = nullptr;
    Prof(hos_ctx *c_, hipStream_t s_, int k, double by) : c(c_), st(s_), kind(k), bytes(by)
    {
        if (c->prof_on && !c->capturing) {
            a = pool_event(c);
            (void)hipEventRecord(a, st);
        }
    }
    ~Prof()
    {
        if (a) {
            hipEvent_t b = pool_event(c);
            (void)hipEventRecord(b, st);
            c->recs.push_back({kind, a, b, bytes});
        }
    }
};

// kernel dispatch on (T, log2 N, tile)
inline KTab ktab(const hos_ctx *c)
{
    KTab k;
    k.kx = c->kx;
    k.ky0 = 2 * M_PI / c->p.Ly;
    k.depth = c->depth;
    k.th = c->th;
    k.ths = c->fs;
    k.nk = c->Nk;
    k.mx = c->mx;
    k.my = c->my;
    return k;
}

#define LOGN_SWITCH(v, ...)                                                                       \
    switch (v) {                                                                                   \
    case 4: { constexpr int LG = 4; __VA_ARGS__; } break;                                                 \
    case 5: { constexpr int LG = 5; __VA_ARGS__; } break;                                                 \
    case 6: { constexpr int LG = 6; __VA_ARGS__; } break;                                                 \
    case 7: { constexpr int LG = 7; __VA_ARGS__; } break;                                                 \
    case 8: { constexpr int LG = 8; __VA_ARGS__; } break;                                                 \
    case 9: { constexpr int LG = 9; __VA_ARGS__; } break;                                                 \
    case 10: { constexpr int LG = 10; __VA_ARGS__; } break;                                               \
    case 11: { constexpr int LG = 11; __VA_ARGS__; } break;                                               \
    default: return fail(HOS_E_INVALID, "unsupported size 2^%d", v);                                \
    }

// TXV = y tile; only tiles whose LDS fits 64 KiB are instantiated
// LDS a y tile may use: 64 KiB, or all 160 KiB of a gfx950 CU for the
// narrowest tile (FP64 rows of 2048)
constexpr int tile_lds_max(int tx) { return tx == 4 ? 163840 : 65536; }

#define TX_CASE(T, V, ...)                                                                         \
    case V: {                                                                                      \
        constexpr int TXV = V;                                                                     \
        if constexpr (Lds<T, (1 << LG)>::size(TXV / 2) * sizeof(cx<T>) <= tile_lds_max(TXV)) {     \
            __VA_ARGS__;                                                                           \
        } else {                                                                                   \
            return fail(HOS_E_INVALID, "tile %d exceeds the LDS limit at Ny = %d", V, 1 << LG);     \
        }                                                                                          \
    } break;
#define TX_SWITCH(T, tx, ...)                                                                      \
    switch (tx) {                                                                                  \
        TX_CASE(T, 4, __VA_ARGS__)                                                                 \
        TX_CASE(T, 8, __VA_ARGS__)                                                                 \
        TX_CASE(T, 16, __VA_ARGS__)                                                                \
    default: return fail(HOS_E_INVALID, "unsupported tile %d", tx);                                \
    }

// x inverse passes (k_xinv, k_xfwdinv): one row per wave (k_*_w) is
// available in the FP32 form for Nx = 512 and 1024 (at 2048 a wave's row
// needs more registers than a lane has); c->xwave selects it.
template <class T>
constexpr bool x_wave(int lognx)
{
    return sizeof(T) == 4 && (lognx == 9 || lognx == 10);
}
inline bool x_wave_ok(const hos_ctx *c) { return c->f32 ? x_wave<float>(c->lognx) : x_wave<double>(c->lognx); }
// stage-0 Hamiltonian partials per member written by the final x pass
inline int x_nparts(const hos_ctx *c)
{
    const int rows = c->lognx >= 10 ? 1 : (1 << (10 - c->lognx));
    return (c->Nk + rows - 1) / rows;
}

// the outputs of one x inverse pass, in order (fixed capacity: the launch
// path allocates nothing)
struct XinvList {
    XinvOp v[MAXOPS];
    int n = 0;
    XinvList() = default;
    XinvList(std::initializer_list<XinvOp> l)
    {
        for (const XinvOp &o : l) push_back(o);
    }
    void push_back(const XinvOp &o)
    {
        if (n < MAXOPS) v[n] = o;
        ++n;  // past MAXOPS: rejected by Eng::xinv
    }
    int size() const { return n; }
    const XinvOp *begin() const { return v; }
    const XinvOp *end() const { return v + (n < MAXOPS ? n : MAXOPS); }
};

template <class T>
struct Eng {
    hos_ctx *c;
    int nm;                  // members in this launch
    int m0 = 0;              // first member (the launch covers m0 .. m0+nm-1)
    hipStream_t st = nullptr;  // launch stream (null: the context's)
    int inflight = 0;        // members whose launches run concurrently with this one (0: nm)
    bool band = false;       // the stage input is a masked state: skip its masked modes
    hipEvent_t mark = nullptr;  // recorded on S() at point mark_at of the RHS (levels 0..M-1, M: after them)
    int mark_at = -1;
    // the x-intermediates of the RHS under way are blocked (xi_at): set by
    // rhs() for the diagonal form with wave x passes, which is the whole
    // chain of writers and readers of those slots
    int xb = 0;
    hipStream_t S() const { return st ? st : c->stream; }
    void mark_if(int at) const
    {
        if (mark && mark_at == at) (void)hipEventRecord(mark, S());
    }
    // spectral rows of the stage input that may be nonzero, and the column
    // band of the loads (0: every column)
    int nzr() const { return band ? c->my : c->Nk; }
    int cmx() const { return band ? c->mx : 0; }
    // fraction of the modes of a row that the column band keeps
    double colfrac() const { return band ? (2.0 * c->mx - 1.0) / c->p.Nx : 1.0; }
    // enough spectral rows in flight to fill the GPU twice over and to outgrow
    // its caches (8 members at 1024^2; one member at 1024^2 is not)
    bool many_rows() const { return (long)(inflight > 0 ? inflight : nm) * c->Nk >= 4096; }
    cx<T> *I() const { return (cx<T> *)c->I + (long)m0 * sI(); }
    T *P() const { return (T *)c->P + (long)m0 * sP(); }
    // wavenumber tables of members m0.. (the kernels index them by blockIdx.z)
    KTab kt() const
    {
        KTab k = ktab(c);
        k.depth += m0;
        k.th += (long)m0 * k.ths;
        return k;
    }
    long sI() const { return (long)c->nI * c->fsI; }
    long sP() const { return (long)c->nP * c->fsP; }
    long sY() const { return 2 * c->fs; }
    const cx<T> *twx() const { return (const cx<T> *)c->twx; }
    const cx<T> *twy() const { return (const cx<T> *)c->twy; }
    Phys<T> phys() const { return Phys<T>{P(), sP(), c->fsP, c->p.Ny}; }
    bool comp() const { return c->f32 && c->p.compensated; }

    // outputs list[i] (dst = intermediate index), grouped by source into
    // chunks of at most `per` outputs, one chunk per blockIdx.y.  Few rows in
    // flight (one member): small chunks for more workgroups; many members:
    // one chunk per source, so each source row is read once.  The rows in
    // flight count every member group stepping at the same time: two groups
    // of four 1024^2 members measured 3.01 ms per step with one chunk per
    // source against 3.08 with chunks of 3
    int xinv(const cx<double> *Y, const XinvList &list, int per = 0)
    {
        if (list.size() > MAXOPS) return fail(HOS_E_INVALID, "too many x-inverse outputs");
        // the rows this pass has in flight: a masked state's band only
        if (per <= 0) per = (long)(inflight > 0 ? inflight : nm) * nzr() >= 4096 ? MAXPERGRP : 3;
        XinvOps ops{};
        for (int s = 0; s < 2; ++s) {
            int k = 0;
            for (const XinvOp &o : list) {
                if (o.src != s) continue;
                if (k == 0) {
                    if (ops.ngroups == MAXGRP) return fail(HOS_E_INVALID, "too many x-inverse outputs");
                    ops.src[ops.ngroups] = s;
                    ops.n[ops.ngroups] = 0;
                    ops.ngroups++;
                }
                const int g = ops.ngroups - 1;
                ops.o[g][ops.n[g]++] = o;
                k = (k + 1) % std::min(per, MAXPERGRP);
            }
        }
        const int ngroups = ops.ngroups;
        const int nops = (int)list.size();
        const int nz = nzr();
        LOGN_SWITCH(c->lognx, {
            constexpr int ROWS = LG >= 10 ? 1 : (1 << (10 - LG));
            dim3 grid((nz + ROWS - 1) / ROWS, ngroups, nm);
            // FP64 form: the tanh(|k| h) table is read once per group
            const double thb = sizeof(T) == 8 ? 8.0 * ngroups : 0.0;
            const double rows = (double)nz * c->p.Nx;
            Prof pf(c, S(), HOS_K_XINV, (double)nm * rows * ((16.0 * ngroups + thb) * colfrac() + 2.0 * sizeof(T) * nops));
            bool waved = false;
            if constexpr (x_wave<T>(LG)) {
                if (c->xwave) {
                    dim3 gw((nz + XW - 1) / XW, ngroups, nm);
                    if (xb)
                        hipLaunchKernelGGL((k_xinv_w<T, LG, true>), gw, dim3(NT), 0, S(), Y, sY(), I(), sI(), c->fs, c->fsI,
                                           c->ip, ops, kt(), twx(), nz, cmx());
                    else
                        hipLaunchKernelGGL((k_xinv_w<T, LG, false>), gw, dim3(NT), 0, S(), Y, sY(), I(), sI(), c->fs, c->fsI,
                                           c->ip, ops, kt(), twx(), nz, cmx());
                    waved = true;
                }
            }
            if (!waved) {
                if (xb) return fail(HOS_E_INVALID, "blocked x-intermediates need the wave x passes");
                hipLaunchKernelGGL((k_xinv<T, LG, ROWS>), grid, dim3(NT), 0, S(), Y, sY(), I(), sI(), c->fs, c->fsI, c->ip,
                                   ops, kt(), twx(), nz, cmx());
            }
        });
        return HOS_OK;
    }
    // nz: rows of the x-intermediates that may be nonzero (0: all)
    int yc2r(const FieldMap &fm, int nf, int nz = 0)
    {
        if (nz <= 0) nz = c->Nk;
        const int tx = c->tx[HOS_TILE_C2R];
        LOGN_SWITCH(c->logny, {
            TX_SWITCH(T, tx, {
                dim3 grid(c->p.Nx / TXV, nf, nm);
                Prof pf(c, S(), HOS_K_YC2R, (double)nm * nf * sizeof(T) * (2.0 * nz * c->p.Nx + c->p.Nx * (double)c->p.Ny));
                hipLaunchKernelGGL((k_yc2r<T, LG, TXV>), grid, dim3(NT), 0, S(), I(), sI(), c->fsI, P(), sP(),
                                   c->fsP, fm, (int)c->ip, twy(), nz);
            })
        });
        return HOS_OK;
    }
    template <class PRO>
    int yr2c(const PRO &pro, int dst, int nin)
    {
        const int tx = c->tx[HOS_TILE_R2C];
        LOGN_SWITCH(c->logny, {
            TX_SWITCH(T, tx, {
                dim3 grid(c->p.Nx / TXV, 1, nm);
                Prof pf(c, S(), HOS_K_YR2C, (double)nm * sizeof(T) * (nin * (double)c->p.Nx * c->p.Ny + 2.0 * c->fs));
                hipLaunchKernelGGL((k_yr2c<T, LG, TXV, PRO>), grid, dim3(NT), 0, S(), pro, I(), sI(), c->fsI,
                                   dst, (int)c->ip, twy());
            })
        });
        return HOS_OK;
    }
    template <class PRO>
    int yr2c_pair(const PRO &pro, int da, int nin)
    {
        const int tx = c->tx[HOS_TILE_PAIR];
        LOGN_SWITCH(c->logny, {
            TX_SWITCH(T, tx, {
                constexpr int TXP = TXV / 2;
                dim3 grid(c->p.Nx / TXP, 1, nm);
                Prof pf(c, S(), HOS_K_YR2C_PAIR, (double)nm * sizeof(T) * (nin * (double)c->p.Nx * c->p.Ny + 4.0 * c->fs));
                hipLaunchKernelGGL((k_yr2c_pair<T, LG, TXP, PRO>), grid, dim3(NT), 0, S(), pro, I(), sI(),
                                   c->fsI, da, (int)c->ip, twy());
            })
        });
        return HOS_OK;
    }
    // outputs j = 1..ndst to fields dst0+j-1, the last one to dtop if dtop >= 0.
    // spec (ndst = 1, dtop >= 0): the output d phi/dz stays in x-spectral
    // space, rows ky < my only (what the mask keeps), for the RK epilogue
    int xfwdinv(int src, int dst0, int ndst, int dtop = -1, bool spec = false)
    {
        if (spec && (ndst != 1 || dtop < 0)) return fail(HOS_E_INVALID, "spectral x output: one output to dtop");
        const double scale = 1.0 / ((double)c->p.Nx * c->p.Ny);
        const int nrows = spec ? c->my : c->Nk;
        LOGN_SWITCH(c->lognx, {
            constexpr int ROWS = LG >= 10 ? 1 : (1 << (10 - LG));
            dim3 grid((nrows + ROWS - 1) / ROWS, 1, nm);
            Prof pf(c, S(), HOS_K_XFWDINV,
                    (double)nm * nrows * c->p.Nx * (2.0 * sizeof(T) * (1 + ndst) + (sizeof(T) == 8 ? 8.0 : 0.0)));
            bool waved = false;
            if constexpr (x_wave<T>(LG)) {
                if (c->xwave) {
                    dim3 gw((nrows + XW - 1) / XW, 1, nm);
                    if (xb)
                        hipLaunchKernelGGL((k_xfwdinv_w<T, LG, true>), gw, dim3(NT), 0, S(), I(), sI(), c->fsI, c->ip, src,
                                           dst0, ndst, dtop, scale, kt(), twx(), (int)spec, nrows);
                    else
                        hipLaunchKernelGGL((k_xfwdinv_w<T, LG, false>), gw, dim3(NT), 0, S(), I(), sI(), c->fsI, c->ip, src,
                                           dst0, ndst, dtop, scale, kt(), twx(), (int)spec, nrows);
                    waved = true;
                }
            }
            if (!waved) {
                if (xb) return fail(HOS_E_INVALID, "blocked x-intermediates need the wave x passes");
                hipLaunchKernelGGL((k_xfwdinv<T, LG, ROWS>), grid, dim3(NT), 0, S(), I(), sI(), c->fsI, c->ip, src, dst0,
                                   ndst, dtop, scale, kt(), twx(), (int)spec, nrows);
            }
        });
        return HOS_OK;
    }
    // rows of the final x pass: the mask's my for a masked state's RK stage
    int rhs_rows(const RkArgs &rk) const { return rk.band ? c->my : c->Nk; }
    // stage-0 Hamiltonian partials per member that pass writes
    int rhs_parts(const RkArgs &rk) const
    {
        const int rows = c->lognx >= 10 ? 1 : (1 << (10 - c->lognx));
        return (rhs_rows(rk) + rows - 1) / rows;
    }
    // ib < 0: the interleaved pair of yr2c_pair at field ia.  it >= 0: slot
    // of the x spectrum of d phi^(M)/dz (diagonal passes), added to N_eta
    int xfwd_rhs(int ia, int ib, const cx<double> *Y, const RkArgs &rk, int it = -1)
    {
        const double scale = 1.0 / ((double)c->p.Nx * c->p.Ny);
        const int nz = rhs_rows(rk);
        // FP64 [2][Nk][Nx] fields moved by the epilogue
        double f64 = 0.0;
        if (rk.mode == EPI_RK_SUBDIAG) {
            // stage 0 reads the state as its input; the last writes the state only
            const bool first = rk.stage == 0, last = rk.stage == rk.nstages - 1;
            f64 = 2 /*Y*/ + (first ? 0 : 4) /*u, acc*/ + 2 /*acc or u*/ + (last ? 0 : 2) /*Y next*/;
        }
        else if (rk.mode == EPI_STORE_K)
            f64 = 2 + 2 + (rk.hpart && rk.stage == 0 ? 2 : 0);
        else
            f64 = 2;
        // a masked state's epilogue moves the kept modes only
        const double kept = rk.band ? (double)c->my * (2.0 * c->mx - 1.0) : (double)c->fs;
        LOGN_SWITCH(c->lognx, {
            constexpr int ROWS = LG >= 10 ? 1 : (1 << (10 - LG));
            dim3 grid((nz + ROWS - 1) / ROWS, 1, nm);
            const double tbytes = it >= 0 ? 2.0 * sizeof(T) * std::min(nz, c->my) * c->p.Nx : 0.0;
            Prof pf(c, S(), HOS_K_XFWD_RHS,
                    (double)nm * (2.0 * 2.0 * sizeof(T) * nz * c->p.Nx + tbytes + 16.0 * f64 * kept));
            hipLaunchKernelGGL((k_xfwd_rhs<T, LG, ROWS>), grid, dim3(NT), 0, S(), I(), sI(), c->fsI, c->ip, ia,
                               ib < 0 ? ia : ib, ib < 0 ? 1 : 0, Y, sY(), scale, kt(), rk, twx(), nz, it);
        });
        return HOS_OK;
    }

    // x-intermediate slots of the fields the RHS pass transforms itself
    // (deferred; see k_yr2c_rhs): after the working slots 0 .. max(2, M)-1,
    // or with the diagonal passes after slots 0 (phi^(L+1)), 1 (eta) and the
    // M(M-1)/2 diagonal slots
    int dslot(int f) const { return (c->diag ? diag_base(c->M) : std::max(2, c->M)) + f; }
    static int diag_base(int M) { return 2 + M * (M - 1) / 2; }
    // diagonal slot of d^j phi^(a)/dz^j, a + j <= M: slots 2.. in order of a,
    // the M-a fields of each a consecutive (an x forward-inverse pass writes
    // its outputs j = 1.. to consecutive slots)
    int dg(int a, int j) const
    {
        int off = 2;
        for (int b = 1; b < a; ++b) off += c->M - b;
        return off + j - 1;
    }
    // slot of the W^(M) term k (d^{k+1} phi^(M-k)/dz^{k+1}) among them
    int dslot_w(int k) const { return dslot(4 + k); }

    // level 0: eta, eta_x, eta_y, phi_x, phi_y, d^j phi/dz^j (j = 1..M), the
    // outputs of the first x inverse pass; the ones bound for the physical
    // grid are listed in fm, their count returned.  defer: eta_x, eta_y,
    // phi_x, phi_y and d^M phi/dz^M stay x-intermediates for the RHS pass.
    int level0(XinvList &ops, FieldMap &fm, bool defer)
    {
        const HosIdx &ix = c->ix;
        const int M = c->M;
        int n = 0, slot = 0;
        auto add = [&](int src, int op, int pw, int pdst) {
            ops.push_back({src, op, pw, slot});
            fm.src[n] = slot++;
            fm.dst[n++] = pdst;
        };
        auto keep = [&](int src, int op, int pw, int dst) { ops.push_back({src, op, pw, dst}); };
        add(0, OP_ID, 0, ix.eta);
        if (defer) {
            keep(0, OP_DX, 0, dslot(0));
            keep(0, OP_DY, 0, dslot(1));
            keep(1, OP_DX, 0, dslot(2));
            keep(1, OP_DY, 0, dslot(3));
            for (int j = 1; j < M; ++j) add(1, OP_DZ, j, ix.d[1][j]);
            keep(1, OP_DZ, M, dslot_w(M - 1));
        } else {
            add(0, OP_DX, 0, ix.etax);
            add(0, OP_DY, 0, ix.etay);
            add(1, OP_DX, 0, ix.phix);
            add(1, OP_DY, 0, ix.phiy);
            for (int j = 1; j <= M; ++j) add(1, OP_DZ, j, ix.d[1][j]);
        }
        return n;
    }

    // y r2c of phi^(lev+1) (prologue instantiated per level) and the RHS
    // pass (instantiated per order; its plans 1-3 in their own unit); defined
    // in hos_levels.hip
    int phi_level(int lev);
    // y c2r of a level's outputs (fm, nf fields; field 0 d phi^(lev)/dz) with
    // the y r2c of phi^(lev+1) (k_yc2r_phi), lev = 2..7
    int yc2r_phi(const FieldMap &fm, int nf, int lev);
    int rhs_pair();
    int rhs_pair_plan();
    int rhs_pair_diag();
    int rhs_pair_diag_plan();
    // diagonal pass L (k_ydiag), L = 1..7; defined in hos_levels.hip
    int ydiag(int L);

    // the W expansion as diagonal passes (c->diag): level 0 as in w_levels;
    // then for L = 1..M-1 the diagonal pass L (phi^(L+1) to slot 0, W^(L) to
    // the physical grid, and in pass 1 the y c2r of eta) and the x forward-inverse pass of
    // phi^(L+1) (d^j phi^(L+1)/dz^j to the diagonal slots, the top one, a
    // W^(M) term, deferred).  The RHS pass then reads eta and W^(1) ..
    // W^(M-1).  Each d^j phi^(a) is y transformed once, by the pass that
    // consumes it, and never reaches the physical grid.
    int diag_levels(const cx<double> *Y)
    {
        const HosIdx &ix = c->ix;
        const int M = c->M;
        XinvList ops;
        ops.push_back({0, OP_ID, 0, 1});
        ops.push_back({0, OP_DX, 0, dslot(0)});
        ops.push_back({0, OP_DY, 0, dslot(1)});
        ops.push_back({1, OP_DX, 0, dslot(2)});
        ops.push_back({1, OP_DY, 0, dslot(3)});
        for (int j = 1; j < M; ++j) ops.push_back({1, OP_DZ, j, dg(1, j)});
        // the W^(M) terms k = 1..M-1 for the RHS pass (k = M-1: d^M phi^(1)/dz^M);
        // at M = 1 the only term, d phi/dz, cancels against W^(1) in N_eta
        if (M >= 2) ops.push_back({1, OP_DZ, M, dslot_w(M - 1)});
        int rc = xinv(Y, ops);
        if (M == 1 && !rc) {  // no diagonal pass: eta's y c2r on its own
            FieldMap fm{};
            fm.src[0] = 1;
            fm.dst[0] = ix.eta;
            rc = yc2r(fm, 1, nzr());
        }
        mark_if(0);
        for (int L = 1; L <= M - 1 && !rc; ++L) {
            rc = ydiag(L);
            // d^j phi^(L+1)/dz^j, j = 1..M-L: the W^(M) term j = M-L deferred.
            // At L = M-1 that term is d phi^(M)/dz, which enters N_eta alone
            // (no eta factor, no product): it stays in x-spectral space and the
            // RK epilogue adds it (xfwd_rhs)
            if (!rc) rc = xfwdinv(0, M - L > 1 ? dg(L + 1, 1) : 0, M - L, dslot_w(M - L - 1), L == M - 1);
            mark_if(L);
        }
        return rc;
    }

    // physical fields of the W expansion from the stage input Y; defer: the
    // fields only the RHS pass reads stay x-intermediates (dslot)
    int w_levels(const cx<double> *Y, bool defer = false)
    {
        const HosIdx &ix = c->ix;
        const int M = c->M;
        XinvList ops;
        FieldMap fm{};
        const int n = level0(ops, fm, defer);
        int rc = xinv(Y, ops);
        if (!rc) rc = yc2r(fm, n, nzr());
        mark_if(0);
        bool have_phi = false;  // phi^(lev+1) already transformed (by yc2r_phi)
        for (int lev = 1; lev <= M - 1 && !rc; ++lev) {
            if (!have_phi) rc = phi_level(lev);
            have_phi = false;
            const int nd = M - lev;
            // the top derivative d^{M-lev} phi^(lev+1) is the W^(M) term k = M-lev-1
            if (!rc) rc = xfwdinv(0, 1, nd, defer ? dslot_w(M - lev - 1) : -1);
            const int np = defer ? nd - 1 : nd;
            FieldMap f2{};
            for (int j = 1; j <= np; ++j) {
                f2.src[j - 1] = j;
                f2.dst[j - 1] = ix.d[lev + 1][j];
            }
            if (rc || np == 0) {
                mark_if(lev);
                continue;
            }
            // field 0 is d phi^(lev+1)/dz, which phi^(lev+2) reads
            if (c->fuse && lev + 1 <= M - 1) {
                rc = yc2r_phi(f2, np, lev + 1);
                have_phi = true;
            } else {
                rc = yc2r(f2, np);
            }
            mark_if(lev);
        }
        return rc;
    }

    // RHS: the W levels (each field its own workgroups; the fields only the
    // RHS pass reads are left as x-intermediates), the RHS pass (their y c2r,
    // N_eta and N_phi, y r2c), and the x forward pass with the RK epilogue.
    // (Fusing all of a level's y c2r fields with the next y r2c in one
    // workgroup serialises them and measured slower; k_yc2r_phi fuses the r2c
    // behind the one field it reads.)
    int rhs(const cx<double> *Y, const RkArgs &rk)
    {
        // blocked x-intermediates: the diagonal form's x passes (wave form)
        // write them, its y passes (k_ydiag, the RHS pass) read them and
        // k_ydiag's r2c output is read back by the wave x pass; the RHS
        // pass's output and the spectral top-order term stay row by row
        xb = c->xblock && c->diag && c->M >= 2 && c->xwave && x_wave<T>(c->lognx) ? 1 : 0;
        int rc = c->diag ? diag_levels(Y) : w_levels(Y, true);
        if (!rc) {
            mark_if(c->M);
            rc = rhs_pair();
        }
        if (!rc) rc = xfwd_rhs(0, -1, Y, rk, c->diag && c->M >= 2 ? dslot_w(0) : -1);
        xb = 0;
        return rc;
    }

    // the RHS pass of order MO (k_yr2c_rhs) with FFT plan PLAN; TXP = half the
    // pair tile, at most what keeps two 16-byte tasks per thread
    template <int MO, int PLAN, bool WF = false>
    int yr2c_rhs()
    {
        RhsDeferred<MO> dfr;
        // the gradients (0..3) and d^M phi^(1)/dz^M (term M-1) come from the
        // level-0 x inverse pass: a band of a masked state.  Over W fields
        // (WF) the term k = 0 is added in spectral space by the epilogue:
        // fields 4.. are the terms k = 1..M-1
        constexpr int K0 = WF ? 1 : 0;
        double drows = 0.0;
        for (int f = 0; f < MO + 4 - K0; ++f) {
            dfr.slot[f] = f < 4 ? dslot(f) : dslot(f + K0);
            dfr.nzr[f] = (f < 4 || f == MO + 3 - K0) ? nzr() : c->Nk;
            drows += dfr.nzr[f];
        }
        dfr.nzo = band ? c->my : c->Nk;
        dfr.blk = xb;
        // eta and d^j phi^(a), a + j <= M (or W^(1) .. W^(M-1)), from the physical grid
        const int nin = 1 + (WF ? MO - 1 : (MO - 1) * MO / 2);
        const int tx = c->tx[HOS_TILE_PAIR];
        LOGN_SWITCH(c->logny, {
            TX_SWITCH(T, tx, {
                constexpr int V = Vec<T>::N;
                constexpr int TXP0 = TXV / 2;
                constexpr int NTR = rhs_nt<T, WF>();
                constexpr int TXP = (TXP0 > 2 && TXP0 * (1 << LG) / V > 2 * NTR) ? 2 : TXP0;
                dim3 grid(c->p.Nx / TXP, 1, nm);
                Prof pf(c, S(), HOS_K_YR2C_PAIR,
                        (double)nm * sizeof(T) * c->p.Nx * (nin * (double)c->p.Ny + 2.0 * drows + 4.0 * dfr.nzo));
                if constexpr (PLAN != 0 && LG < PLAN_LOGN_MIN) {
                    return fail(HOS_E_INVALID, "FFT plan %d needs Ny >= %d", PLAN, 1 << PLAN_LOGN_MIN);
                } else {
                    hipLaunchKernelGGL((k_yr2c_rhs<T, LG, TXP, MO, PLAN, WF>), grid, dim3(NTR), 0, S(), phys(), c->ix,
                                       (int)comp(), I(), sI(), c->fsI, dfr, 0, (int)c->ip, twy());
                }
            })
        });
        return HOS_OK;
    }
};
