// hos_kernels.hpp -- HIP kernels of the HOS RHS / RK4 path for gfx950.
//
// Data layout in HBM (per member m, all row-major):
//   Y, U, A   complex double [2][Nk][Nx]   stage input, state, RK accumulator
//   I         complex T      [NI][Nk][Nx]  x-intermediates (y transformed,
//                                          x not yet), ky rows, x fastest
//   P         T              [NP][Nx][Ny]  physical fields, y fastest
// with Nk = Ny/2+1.  The spectral layout is the reference's transposed layout
// (index Nx*ky + kx, Csource/2dpar/operators_par.c:50-71).
//
// A 2-D c2r is k_xinv (x inverse on contiguous ky rows, with the spectral
// multiplier i kx, i ky or d^j/dz^j applied on load) followed by k_yc2r
// (y c2r on tiles of TX x-rows, reading the tile transposed).  A 2-D r2c is
// k_yr2c (the physical value is computed on the fly by a prologue functor,
// then y r2c, written transposed) followed by an x forward pass whose
// epilogue either re-inverts with d^j/dz^j (k_xfwdinv, one level of the W
// expansion) or forms the RHS and the RK stage update (k_xfwd_rhs).
#pragma once
#include "hos_fft.hpp"

namespace hos {

constexpr int NT = 256;   // threads per workgroup
constexpr int MAXM = 8;   // HOS order M = nlevs+1 up to 8
constexpr int MAXOPS = 16;

// 16-byte vector of T (4 floats or 2 doubles): the unit of the physical-field
// loads and stores, so one instruction moves 16 bytes per lane.
template <class T>
struct Vec;
template <>
struct Vec<float> {
    typedef float type __attribute__((ext_vector_type(4)));
    static constexpr int N = 4;
};
template <>
struct Vec<double> {
    typedef double type __attribute__((ext_vector_type(2)));
    static constexpr int N = 2;
};
template <class T>
using vec_t = typename Vec<T>::type;

// two adjacent complex values as 16-byte accesses (16-byte aligned address)
template <class T>
__device__ __forceinline__ void load2(const cx<T> *p, cx<T> &a, cx<T> &b)
{
    if constexpr (sizeof(T) == 4) {
        const vec_t<float> v = *reinterpret_cast<const vec_t<float> *>(p);
        a = {v[0], v[1]};
        b = {v[2], v[3]};
    } else {
        a = p[0];
        b = p[1];
    }
}
template <class T>
__device__ __forceinline__ void store2(cx<T> *p, cx<T> a, cx<T> b)
{
    if constexpr (sizeof(T) == 4) {
        *reinterpret_cast<vec_t<float> *>(p) = vec_t<float>{a.x, a.y, b.x, b.y};
    } else {
        p[0] = a;
        p[1] = b;
    }
}

// one complex value stored with the non-temporal hint (global_store ... nt):
// the x passes' whole-row outputs are read back only by a later pass, after
// other traffic, so their lines gain nothing from staying in L2.  Measured on
// the wave x passes of the 8 x 1024^2 M=5 FP32 batch: x inverse 107 -> 97 us,
// and the y r2c passes that follow run faster too (39 -> 33 us)
template <class T>
__device__ __forceinline__ void store_nt(cx<T> *p, cx<T> v)
{
    if constexpr (sizeof(T) == 4) {
        typedef float f2 __attribute__((ext_vector_type(2)));
        __builtin_nontemporal_store(f2{v.x, v.y}, reinterpret_cast<f2 *>(p));
    } else {
        typedef double d2 __attribute__((ext_vector_type(2)));
        __builtin_nontemporal_store(d2{v.x, v.y}, reinterpret_cast<d2 *>(p));
    }
}

// two adjacent complex values (16-byte aligned address) with the non-temporal hint
template <class T>
__device__ __forceinline__ void store2_nt(cx<T> *p, cx<T> a, cx<T> b)
{
    if constexpr (sizeof(T) == 4) {
        __builtin_nontemporal_store(vec_t<float>{a.x, a.y, b.x, b.y}, reinterpret_cast<vec_t<float> *>(p));
    } else {
        store_nt(p, a);
        store_nt(p + 1, b);
    }
}

// loads with the non-temporal hint, for data at its last use in a pass whose
// working set is far larger than the caches: the RK epilogue's reads of the
// RHS pair, the stage input and the accumulator when many rows are in flight
// (RkArgs::stream; 8 x 1024^2 FP32 members: 108 -> 100 us per launch).  One
// member's data partly stays in the last-level cache between passes, and
// there the hint costs hits
__device__ __forceinline__ cx<double> load_nt(const cx<double> *p)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 v = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(p));
    return {v[0], v[1]};
}
template <class T>
__device__ __forceinline__ void load2_nt(const cx<T> *p, cx<T> &a, cx<T> &b)
{
    if constexpr (sizeof(T) == 4) {
        const vec_t<float> v = __builtin_nontemporal_load(reinterpret_cast<const vec_t<float> *>(p));
        a = {v[0], v[1]};
        b = {v[2], v[3]};
    } else {
        a = load_nt(p);
        b = load_nt(p + 1);
    }
}

// ---------------------------------------------------------------------------
// wavenumbers: kx from a table (reference or signed), ky = j*Ky0,
// operators_par.c:50-71, 105-116, 168-192.
struct KTab {
    const double *kx;     // [Nx]
    double ky0;
    const double *depth;  // [members]; <= 0: deep water
    const double *th;     // [members][Nk][Nx] tanh(|k| h), 1 in deep water (k_th_table)
    long ths;             // member stride of th
    int nk;               // Ny/2+1
    int mx, my;           // dealias bounds, operators_par.c:290-291
    __device__ __forceinline__ double h(int m) const { return depth[m]; }
};

// |k| and tanh(|k| h) of one mode in the precision R of the multipliers: FP64
// from the table (thd, loaded by the caller with its other loads); FP32 computed
// on the fly, since the FP32 form rounds every multiplied value to FP32.
template <class R>
__device__ __forceinline__ void kz_th(double kxv, double kyv, double h, double thd, R &kz, R &th)
{
    if constexpr (sizeof(R) == 8) {
        kz = sqrt(kxv * kxv + kyv * kyv);
        th = thd;
    } else {
        kz = sqrtf(float(kxv * kxv + kyv * kyv));
        th = h > 0.0 ? tanhf(kz * float(h)) : 1.0f;
    }
}

// d^p/dz^p multiplier at |k| = kz: kz^p (deep), kz^p tanh(kz h) for odd p.
__device__ __forceinline__ double dz_mul(double kz, int p, double depth)
{
    double m = 1.0;
    for (int i = 0; i < p; ++i) m *= kz;
    if (depth > 0.0 && (p & 1)) m *= tanh(kz * depth);
    return m;
}

// zero (kx,ky) if (i>=mx && i<Nx-mx+1) || j>=my, operators_par.c:309-322
__device__ __forceinline__ bool masked(int i, int j, int nx, int mx, int my)
{
    return (i >= mx && i < nx - mx + 1) || (j >= my);
}

enum : int { OP_ID = 0, OP_DX = 1, OP_DY = 2, OP_DZ = 3 };
struct XinvOp {
    int src;  // 0: eta-hat, 1: phi-hat (of the FP64 stage input)
    int op;
    int pow;  // d^pow/dz^pow for OP_DZ
    int dst;  // intermediate index
};
// outputs grouped by source field: group g reads its source row once and
// forms up to MAXPERGRP outputs from it (one group per blockIdx.y)
constexpr int MAXGRP = 8, MAXPERGRP = 12;
struct XinvOps {
    XinvOp o[MAXGRP][MAXPERGRP];
    int n[MAXGRP];
    int src[MAXGRP];
    int ngroups;
};

// y passes with one row per wave (wave_fft): the FP32 form at Ny = 512, 1024
template <class T>
__host__ __device__ constexpr bool y_wave(int logny)
{
    return sizeof(T) == 4 && (logny == 9 || logny == 10);
}

// XCD-aware tile order: the 8 XCDs take workgroups round-robin, so give
// consecutive logical tiles to workgroups on the same XCD; tiles that share
// 128-byte lines of the transposed intermediate then meet in one L2.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    if ((nb & 7) != 0) return b;
    return (b & 7) * (nb >> 3) + (b >> 3);
}

// ---------------------------------------------------------------------------
// x passes: a workgroup owns ROWS ky-rows of NX (ROWS*NX = max(1024, NX)
// elements, PER = ROWS*NX/NT of them per thread) and keeps every load of its
// rows in flight at once (clamped rows, no branches between the loads).
// Outputs are transformed XNB at a time: XNB*ROWS rows share the stages and
// barriers of one LDS FFT.
template <class T, int LOGNX>
struct XGeom {
    static constexpr int NX = 1 << LOGNX;
    static constexpr int ROWS = LOGNX >= 10 ? 1 : (1 << (10 - LOGNX));
    static constexpr int NE = ROWS * NX;
    static constexpr int PER = NE / NT;
    static constexpr int BYTES = NE * (int)sizeof(cx<T>);
    static constexpr int XNB = BYTES <= 8192 ? 4 : (BYTES <= 16384 ? 2 : 1);
};

// Band of a masked state (operators_par.c:309-322 leaves modes kx < mx or
// kx > Nx-mx, ky < my only): a pass over spectral rows handles rows ky < nzr
// (my for a masked state, Nk otherwise) and, when cmx > 0, loads only the
// columns the mask keeps.  The modes it skips are exact zeros, which the
// transforms would only carry along, so results are unchanged.
__device__ __forceinline__ bool col_kept(int kx, int nx, int cmx) { return cmx == 0 || kx < cmx || kx > nx - cmx; }

// Blocked x-intermediates (blk): the rows ky in blocks of XKB, and inside a
// block the rows of each column pair x = 2c, 2c+1 side by side: element
// (ky, x) of the block starting at row b sits at b ip + c 2 XKB + (ky - b) 2
// + (x & 1).  A block occupies the XKB ip elements its rows occupy row by row
// (a field holds ceil(Nk / XKB) XKB rows, xi_rows).  A y pass reads or writes
// one column pair of every row: 16 bytes per row row by row, XKB*16
// contiguous bytes per block blocked (a quarter of the cache-line requests).
// The wave x passes, XKB rows per workgroup, store and load a block through
// LDS as one contiguous run.  Otherwise (ky ip + x) the rows lie one by one.
constexpr int XKB = 4;
constexpr int xi_rows(int nk) { return (nk + XKB - 1) / XKB * XKB; }
__device__ __forceinline__ long xi_at(int ky, int x, long ip, int blk)
{
    if (!blk) return (long)ky * ip + x;
    const int b = ky & ~(XKB - 1);
    return (long)b * ip + (long)(x >> 1) * (2 * XKB) + (ky - b) * 2 + (x & 1);
}

// x inverse pass: spectral (FP64 stage input, with multiplier) -> I.  The
// source row is read once into registers (FP64) and every output of the
// group is formed from it.  Rows ky < nzr only, columns by cmx (col_kept).
// grid (ceil(nzr/ROWS), ngroups, members)
template <class T, int LOGNX, int ROWS>
__global__ __launch_bounds__(NT) void k_xinv(const cx<double> *__restrict__ Y, long sY, cx<T> *__restrict__ I,
                                             long sI, long fs, long fsI, long ip, XinvOps ops, KTab kt,
                                             const cx<T> *__restrict__ twx, int nzr, int cmx)
{
    using G = XGeom<T, LOGNX>;
    static_assert(G::ROWS == ROWS, "x rows");
    constexpr int NX = G::NX, PER = G::PER, NB = G::XNB;
    using R = T;  // multiplier arithmetic: FP64 (FP64 form) or FP32 (FP32 form)
    using L = Lds<T, NX>;
    __shared__ cx<T> buf[L::size(NB * ROWS)];
    const int grp = blockIdx.y, m = blockIdx.z;
    const int ky0 = blockIdx.x * ROWS;
    const cx<double> *src = Y + m * sY + ops.src[grp] * fs;
    const double *tht = kt.th + m * kt.ths;
    const double h = kt.h(m);
    cx<double> s[PER];
    double kxv[PER], thd[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * NT;
        const int kx = e % NX, ky = min(ky0 + e / NX, nzr - 1);
        const bool keep = col_kept(kx, NX, cmx);
        s[q] = keep ? src[(long)ky * NX + kx] : cx<double>{0.0, 0.0};
        kxv[q] = kt.kx[kx];
        if constexpr (sizeof(R) == 8) thd[q] = keep ? tht[(long)ky * NX + kx] : 1.0;
        else thd[q] = 1.0;
    }
    cx<R> sr[PER];
    R kxr[PER], kz[PER], th[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * NT;
        const int ky = ky0 + e / NX;
        sr[q] = ky < nzr ? cx<R>{R(s[q].x), R(s[q].y)} : cx<R>{R(0), R(0)};
        kxr[q] = R(kxv[q]);
        kz_th<R>(kxv[q], ky * kt.ky0, h, thd[q], kz[q], th[q]);  // d^n/dz^n: k^n tanh(kh)^(n odd)
    }
    cx<T> *base = I + blockIdx.z * sI;
    const int nops = ops.n[grp];
    for (int o0 = 0; o0 < nops; o0 += NB) {
        if (o0) __syncthreads();  // the previous batch's stores have read buf
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if (o0 + b >= nops) break;
            const XinvOp op = ops.o[grp][o0 + b];
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int e = threadIdx.x + q * NT;
                const int ky = ky0 + e / NX;
                R re = sr[q].x, im = sr[q].y;
                if (op.op == OP_DX) {
                    re = -kxr[q] * sr[q].y;
                    im = kxr[q] * sr[q].x;
                } else if (op.op == OP_DY) {
                    const R k = R(ky * kt.ky0);
                    re = -k * sr[q].y;
                    im = k * sr[q].x;
                } else if (op.op == OP_DZ) {
                    R mm = R(1);
                    for (int i = 0; i < op.pow; ++i) mm *= kz[q];
                    if (op.pow & 1) mm *= th[q];
                    re = mm * sr[q].x;
                    im = mm * sr[q].y;
                }
                buf[L::at(b * ROWS + e / NX, e % NX)] = {T(re), T(im)};
            }
        }
        lds_fft<T, LOGNX, ROWS * NB, NT, true>(buf, twx);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if (o0 + b >= nops) break;
            cx<T> *dst = base + ops.o[grp][o0 + b].dst * fsI;
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int e = threadIdx.x + q * NT;
                const int kx = e % NX, ky = ky0 + e / NX;
                if (ky < nzr) dst[(long)ky * ip + kx] = buf[L::at(b * ROWS + e / NX, kx)];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// y c2r pass: I -> P for TX consecutive x-rows (two rows per complex FFT).
// The ky=0 and ky=Ny/2 values contribute their real part only.
// grid (Nx/TX, nfields, members)
struct FieldMap {
    int src[MAXOPS];
    int dst[MAXOPS];
};

// c2r along y of x-rows x0 .. x0+2*NPR-1 of one field: the tile of the
// x-intermediate src [Nk][nx] is read transposed, rows x0+2p and x0+2p+1 are
// the real and imaginary part of one complex inverse FFT, written to dst
// [nx][NY].  buf holds NPR*NY complex values.  Every load of the tile is
// issued before the first LDS write; one tile per workgroup.  Rows ky >= nzr
// are zero (not read).  KEEP: the transformed rows stay in buf (row pair p
// in row p, x0+2p the real parts), published by the caller's barrier.
template <class T, int LOGNY, int NPR, bool KEEP = false>
__device__ __forceinline__ void c2r_tile(const cx<T> *__restrict__ src, T *__restrict__ dst, int x0, int nx,
                                         cx<T> *buf, cx<T> *twl, const cx<T> *__restrict__ twy, int nzr)
{
    constexpr int NY = 1 << LOGNY, NK = NY / 2 + 1;
    using L = Lds<T, NY>;
    // task (p, ky): 16 consecutive ky of one row pair per 16 lanes, so the LDS
    // side is contiguous while each instruction still reads whole 16-byte
    // pieces of the same ky lines
    constexpr int NL = ((NK + 15) / 16) * 16 * NPR, PERL = (NL + NT - 1) / NT;
    TwLds<T, NY, NT> tws;  // twiddles read first, staged in LDS with the fill
    tws.load(twy);
    cx<T> A[PERL], B[PERL];
#pragma unroll
    for (int q = 0; q < PERL; ++q) {
        const int e = threadIdx.x + q * NT;
        const int p = (e / 16) % NPR;
        // past the end: reload the last row (in bounds, value unused)
        const int ky = min((e / (16 * NPR)) * 16 + e % 16, NK - 1);
        if (ky < nzr) {
            load2(src + (long)ky * nx + x0 + 2 * p, A[q], B[q]);
        } else {
            A[q] = {T(0), T(0)};
            B[q] = {T(0), T(0)};
        }
    }
#pragma unroll
    for (int q = 0; q < PERL; ++q) {
        const int e = threadIdx.x + q * NT;
        const int p = (e / 16) % NPR, ky = (e / (16 * NPR)) * 16 + e % 16;
        if (((NL % NT) == 0 || e < NL) && ky < NK) {
            cx<T> a = A[q], b = B[q];
            if (ky == 0 || ky == NY / 2) {
                a.y = T(0);
                b.y = T(0);
            }
            buf[L::at(p, ky)] = {a.x - b.y, a.y + b.x};
            if (ky > 0 && ky < NY / 2) buf[L::at(p, NY - ky)] = {a.x + b.y, b.x - a.y};
        }
    }
    tws.store(twl);
    if constexpr (y_wave<T>(LOGNY)) {
        // one complex row per wave: the wave FFT in registers (no workgroup
        // barrier between its stages), the two real rows stored from there
        __syncthreads();  // the rows are filled
        constexpr int E = NY / 64;
        const int lane = threadIdx.x & 63;
        for (int r = threadIdx.x >> 6; r < NPR; r += NT / 64) {
            cx<T> *row = buf + r * L::PITCH;
            cx<T> v[E];
#pragma unroll
            for (int m = 0; m < E; ++m) v[m] = row[L::at(0, lane + 64 * m)];
            wave_sync();  // read before the FFT's first exchange writes the row
            wave_fft<T, LOGNY, true>(v, row, twl, lane);
            T *d0 = dst + (long)(x0 + 2 * r) * NY, *d1 = d0 + NY;
#pragma unroll
            for (int m = 0; m < E; ++m) {
                d0[lane + 64 * m] = v[m].x;
                d1[lane + 64 * m] = v[m].y;
            }
            if constexpr (KEEP) {
                wave_sync();  // the FFT's last reads of the row precede these writes
#pragma unroll
                for (int m = 0; m < E; ++m) row[L::at(0, lane + 64 * m)] = v[m];
            }
        }
        return;
    }
    lds_fft<T, LOGNY, NPR, NT, true>(buf, twl);
    // V consecutive y of both rows per thread: 16-byte stores
    using F = vec_t<T>;
    constexpr int V = Vec<T>::N, NYV = NY / V, NS = NPR * NYV, PERS = (NS + NT - 1) / NT;
#pragma unroll
    for (int q = 0; q < PERS; ++q) {
        const int e = threadIdx.x + q * NT;
        if ((NS % NT) == 0 || e < NS) {
            const int p = e / NYV, y = (e % NYV) * V;
            F ra, rb;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const cx<T> z = buf[L::at(p, y) + v];
                ra[v] = z.x;
                rb[v] = z.y;
            }
            *reinterpret_cast<F *>(dst + (long)(x0 + 2 * p) * NY + y) = ra;
            *reinterpret_cast<F *>(dst + (long)(x0 + 2 * p + 1) * NY + y) = rb;
        }
    }
}

template <class T, int LOGNY, int TX>
__global__ __launch_bounds__(NT) void k_yc2r(const cx<T> *__restrict__ I, long sI, long fsI, T *__restrict__ P,
                                             long sP, long fsP, FieldMap fm, int nx,
                                             const cx<T> *__restrict__ twy, int nzr)
{
    constexpr int NY = 1 << LOGNY, NPR = TX / 2;
    __shared__ cx<T> buf[Lds<T, NY>::size(NPR)];
    __shared__ cx<T> twl[NY / 2];
    const int x0 = xcd_tile(blockIdx.x, gridDim.x) * TX;
    c2r_tile<T, LOGNY, NPR>(I + blockIdx.z * sI + fm.src[blockIdx.y] * fsI,
                            P + blockIdx.z * sP + fm.dst[blockIdx.y] * fsP, x0, nx, buf, twl, twy, nzr);
}

// ---------------------------------------------------------------------------
// physical-space prologues for the y r2c pass
template <class T>
struct Phys {
    const T *P;
    long sP, fsP;
    int ny;
    __device__ __forceinline__ T at(int m, int f, int x, int y) const
    {
        return P[m * sP + f * fsP + (long)x * ny + y];
    }
    // y .. y+V-1 (y a multiple of V)
    __device__ __forceinline__ vec_t<T> vat(int m, int f, int x, int y) const
    {
        return *reinterpret_cast<const vec_t<T> *>(P + m * sP + f * fsP + (long)x * ny + y);
    }
};

// Kahan-compensated (COMP) or plain sum, lane-wise over a vector type
template <class F, bool COMP>
struct KSum {
    F s = F(0), c = F(0);
    __device__ __forceinline__ void add(F v)
    {
        if constexpr (COMP) {
            const F y = v - c;
            const F t = s + y;
            c = (t - s) - y;
            s = t;
        } else {
            s += v;
        }
    }
    __device__ __forceinline__ F value() const { return s - c; }
};

// field indices of the physical fields of one member
struct HosIdx {
    int M;
    int eta, etax, etay, phix, phiy;
    int d[MAXM + 1][MAXM + 1];  // d[a][j] = d^j phi^(a) / dz^j, a=1..M, j=1..M-a+1
    int w[MAXM + 1];            // W^(a), a = 1..M-1 (the diagonal passes' physical outputs)
    double invfact[MAXM + 1];   // 1/k!
};

// order sums; comp = Kahan-compensated (FP32 form)
template <class T>
struct Sum {
    T s = T(0), c = T(0);
    bool comp;
    __device__ __forceinline__ explicit Sum(bool k) : comp(k) {}
    __device__ __forceinline__ void add(T v)
    {
        if (comp) {
            const T y = v - c;
            const T t = s + y;
            c = (t - s) - y;
            s = t;
        } else {
            s += v;
        }
    }
    __device__ __forceinline__ T value() const { return s - c; }
};

// scalar order sums (the run-time-order Zvel prologue)
// W^(a) = sum_{k=0}^{a-1} eta^k/k! d^{k+1} phi^(a-k)/dz^{k+1}, a = 1..M
// (the reference's hwn, hosm_2d_par.c:322-357)
template <class T>
__device__ __forceinline__ void w_orders(const Phys<T> &ph, const HosIdx &ix, int m, int x, int y, T e, T *W,
                                         bool comp)
{
    for (int a = 1; a <= ix.M; ++a) {
        T ek = T(1);
        Sum<T> s(comp);
        for (int k = 0; k < a; ++k) {
            s.add(ek * T(ix.invfact[k]) * ph.at(m, ix.d[a - k][k + 1], x, y));
            ek *= e;
        }
        W[a] = s.value();
    }
}

// (W^2)^(K) = sum_{a,b>=1, a+b<=K} W^(a) W^(b)  (hosm_2d_par.c:397-442)
template <class T>
__device__ __forceinline__ T w2_upto(const T *W, int K, bool comp)
{
    Sum<T> s(comp);
    for (int a = 1; a < K; ++a)
        for (int b = 1; a + b <= K; ++b) s.add(W[a] * W[b]);
    return s.value();
}

// sum_{a=1}^{K} W^(a), from a = a0
template <class T>
__device__ __forceinline__ T w_upto(const T *W, int a0, int K, bool comp)
{
    Sum<T> s(comp);
    for (int a = K; a >= a0; --a) s.add(W[a]);  // small terms first
    return s.value();
}

// Prologues of the y r2c pass.  A prologue computes V = Vec<T>::N
// consecutive y values of one x-row: single-output ones return them, pair
// ones write two vectors.  The order-templated prologues issue every field
// load of the V points before the first use.

// plain read of one field
template <class T>
struct ProRead {
    Phys<T> ph;
    int f;
    __device__ __forceinline__ vec_t<T> operator()(int m, int x, int y) const { return ph.vat(m, f, x, y); }
};

// phi^(LEV+1) = - sum_{k=1}^{LEV} eta^k/k! d^k phi^(LEV+1-k)/dz^k
// (West et al. 1987; the reference's hphin update, hosm_2d_par.c:302-317)
template <class T, int LEV>
struct ProPhi {
    Phys<T> ph;
    HosIdx ix;
    bool comp;
    using F = vec_t<T>;
    template <bool C>
    __device__ __forceinline__ F calc(const F *f) const
    {
        const F e = f[0];
        F ek = F(1);
        KSum<F, C> s;
#pragma unroll
        for (int k = 1; k <= LEV; ++k) {
            ek *= e;
            s.add(ek * T(ix.invfact[k]) * f[k]);
        }
        return -s.value();
    }
    __device__ __forceinline__ F operator()(int m, int x, int y) const
    {
        F f[LEV + 1];
        f[0] = ph.vat(m, ix.eta, x, y);
#pragma unroll
        for (int k = 1; k <= LEV; ++k) f[k] = ph.vat(m, ix.d[LEV + 1 - k][k], x, y);
        return comp ? calc<true>(f) : calc<false>(f);
    }
    // the same with the k = 1 field d phi^(LEV)/dz given rather than read
    __device__ __forceinline__ F operator()(int m, int x, int y, F d1) const
    {
        F f[LEV + 1];
        f[0] = ph.vat(m, ix.eta, x, y);
        f[1] = d1;
#pragma unroll
        for (int k = 2; k <= LEV; ++k) f[k] = ph.vat(m, ix.d[LEV + 1 - k][k], x, y);
        return comp ? calc<true>(f) : calc<false>(f);
    }
};

// position of d^j phi^(a)/dz^j among the M(M+1)/2 fields a = 1..M, j = 1..M-a+1
template <int M>
__host__ __device__ constexpr int d_pos(int a, int j)
{
    int p = 0;
    for (int b = 1; b < a; ++b) p += M - b + 1;
    return p + j - 1;
}

// W^(a) = sum_{k=0}^{a-1} eta^k/k! d^{k+1} phi^(a-k)/dz^{k+1}, a = 1..M
// (the reference's hwn, hosm_2d_par.c:322-357); d holds the fields in d_pos order
template <class F, class T, int M, bool C>
__device__ __forceinline__ void w_orders_v(const HosIdx &ix, F e, const F *d, F *W)
{
#pragma unroll
    for (int a = 1; a <= M; ++a) {
        F ek = F(1);
        KSum<F, C> s;
#pragma unroll
        for (int k = 0; k < a; ++k) {
            s.add(ek * T(ix.invfact[k]) * d[d_pos<M>(a - k, k + 1)]);
            ek *= e;
        }
        W[a] = s.value();
    }
}

// (W^2)^(K) = sum_{a,b>=1, a+b<=K} W^(a) W^(b)  (hosm_2d_par.c:397-442).
// The reference sums both orders of every pair; W^(a) W^(b) + W^(b) W^(a) is
// exactly 2 W^(a) W^(b) in floating point, so each pair a < b enters once,
// doubled (6 products instead of 10 at K = 5), the orders a + b descending
// (the smaller terms first)
template <class F, int K, bool C>
__device__ __forceinline__ F w2_upto_v(const F *W)
{
    KSum<F, C> s;
#pragma unroll
    for (int n = K; n >= 2; --n)
#pragma unroll
        for (int a = 1; 2 * a <= n; ++a) {
            const int b = n - a;
            s.add(a == b ? W[a] * W[a] : (W[a] + W[a]) * W[b]);
        }
    return s.value();
}

// sum_{a=A0}^{K} W^(a), small terms first
template <class F, int A0, int K, bool C>
__device__ __forceinline__ F w_upto_v(const F *W)
{
    KSum<F, C> s;
#pragma unroll
    for (int a = K; a >= A0; --a) s.add(W[a]);
    return s.value();
}

// Zvel outputs: which=0 -> (W^(M), W^(M-2)), which=1 -> ((W^2)^(M), (W^2)^(M-2)),
// order M at run time, one point at a time (the Zvel entry point, not the step)
template <class T>
struct ProZvel {
    Phys<T> ph;
    HosIdx ix;
    int which;
    bool comp;
    __device__ __forceinline__ void point(int m, int x, int y, T &a, T &b) const
    {
        const int M = ix.M;
        const T e = ph.at(m, ix.eta, x, y);
        T W[MAXM + 1];
        w_orders<T>(ph, ix, m, x, y, e, W, comp);
        if (which == 0) {
            a = w_upto<T>(W, 1, M, comp);
            b = (M - 2 >= 1) ? w_upto<T>(W, 1, M - 2, comp) : T(0);
        } else {
            a = w2_upto<T>(W, M, comp);
            b = w2_upto<T>(W, M - 2, comp);
        }
    }
    __device__ __forceinline__ void operator()(int m, int x, int y, vec_t<T> &a, vec_t<T> &b) const
    {
#pragma unroll
        for (int v = 0; v < Vec<T>::N; ++v) {
            T sa, sb;
            point(m, x, y + v, sa, sb);
            a[v] = sa;
            b[v] = sb;
        }
    }
};
// the half spectra of the two real rows of each pair p < NPR from the
// complex FFT in buf (Z(ky) and Z(NY-ky)), written transposed to out [Nk][nx]
// (x0+2p, x0+2p+1 side by side); 16 consecutive ky of one pair per 16 lanes
// (contiguous in LDS)
template <class T, int LOGNY, int NPR, int NTT = NT>
__device__ __forceinline__ void r2c_store(const cx<T> *buf, cx<T> *__restrict__ out, int x0, int nx, int nzo = -1,
                                          int blk = 0)
{
    constexpr int NY = 1 << LOGNY, NK = NY / 2 + 1;
    constexpr int NO = ((NK + 15) / 16) * 16 * NPR;
    if (nzo < 0 || nzo > NK) nzo = NK;  // default: every row
    for (int e = threadIdx.x; e < NO; e += NTT) {
        const int p = (e / 16) % NPR, ky = (e / (16 * NPR)) * 16 + e % 16;
        if (ky >= nzo) continue;  // nzo <= NK
        const cx<T> zk = buf[Lds<T, NY>::at(p, ky)];
        const cx<T> zm = buf[Lds<T, NY>::at(p, (NY - ky) & (NY - 1))];
        store2(out + xi_at(ky, x0 + 2 * p, nx, blk), cx<T>{T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)},
               cx<T>{T(0.5) * (zk.y + zm.y), T(-0.5) * (zk.x - zm.x)});
    }
}

// r2c along y of x-rows x0 .. x0+2*NPR-1 of the field the prologue computes
// on the fly: rows x0+2p and x0+2p+1 are the real and imaginary part of one
// complex FFT; the half spectra are written transposed into out [Nk][nx].
// One tile per workgroup.
template <class T, int LOGNY, int NPR, class PRO>
__device__ __forceinline__ void r2c_tile(const PRO &pro, int m, cx<T> *__restrict__ out, int x0, int nx,
                                         cx<T> *buf, cx<T> *twl, const cx<T> *__restrict__ twy)
{
    using F = vec_t<T>;
    constexpr int V = Vec<T>::N;
    constexpr int NY = 1 << LOGNY, NYV = NY / V, NTASK = NPR * NYV;
    constexpr int PER = (NTASK + NT - 1) / NT;
    TwLds<T, NY, NT> tws;  // twiddles read first, staged in LDS with the fill
    tws.load(twy);
    // one task = V consecutive y of the row pair; the loop is unrolled so the
    // loads of several tasks are in flight together
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * NT;
        if ((NTASK % NT) == 0 || e < NTASK) {
            const int p = e / NYV, y = (e % NYV) * V;
            const F ra = pro(m, x0 + 2 * p, y);
            const F rb = pro(m, x0 + 2 * p + 1, y);
            cx<T> *d = buf + Lds<T, NY>::at(p, y);
#pragma unroll
            for (int v = 0; v < V; ++v) d[v] = {ra[v], rb[v]};
        }
    }
    tws.store(twl);
    lds_fft<T, LOGNY, NPR, NT, false>(buf, twl);
    r2c_store<T, LOGNY, NPR>(buf, out, x0, nx);
}

// r2c along y of x-rows x0 .. x0+NROW-1 of two fields the prologue computes
// together (the real and imaginary part of one complex FFT per x-row).  The
// two half spectra are written interleaved, o[(ky*nx + x)*2 + {0,1}], so each
// (ky, x) is one 16-byte piece and a tile's piece of a ky line is contiguous
// (separate fields would leave 8-byte pieces per tile, which cost whole
// write sectors).
template <class T, int LOGNY, int NROW, class PRO>
__device__ __forceinline__ void r2c_pair_tile(const PRO &pro, int m, cx<T> *__restrict__ o, int x0, int nx,
                                              cx<T> *buf, cx<T> *twl, const cx<T> *__restrict__ twy)
{
    using F = vec_t<T>;
    constexpr int V = Vec<T>::N;
    constexpr int NY = 1 << LOGNY, NK = NY / 2 + 1, NYV = NY / V, NTASK = NROW * NYV;
    constexpr int PER = (NTASK + NT - 1) / NT;
    TwLds<T, NY, NT> tws;  // twiddles read first, staged in LDS with the fill
    tws.load(twy);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * NT;
        if ((NTASK % NT) == 0 || e < NTASK) {
            const int p = e / NYV, y = (e % NYV) * V;
            F a, b;
            pro(m, x0 + p, y, a, b);
            cx<T> *d = buf + Lds<T, NY>::at(p, y);
#pragma unroll
            for (int v = 0; v < V; ++v) d[v] = {a[v], b[v]};
        }
    }
    tws.store(twl);
    lds_fft<T, LOGNY, NROW, NT, false>(buf, twl);
    constexpr int NO = ((NK + 15) / 16) * 16 * NROW;
    for (int e = threadIdx.x; e < NO; e += NT) {
        const int p = (e / 16) % NROW, ky = (e / (16 * NROW)) * 16 + e % 16;
        if (ky >= NK) continue;
        const cx<T> zk = buf[Lds<T, NY>::at(p, ky)];
        const cx<T> zm = buf[Lds<T, NY>::at(p, (NY - ky) & (NY - 1))];
        store2(o + 2 * ((long)ky * nx + x0 + p), cx<T>{T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)},
               cx<T>{T(0.5) * (zk.y + zm.y), T(-0.5) * (zk.x - zm.x)});
    }
}

// y r2c pass, one output field.  grid (Nx/TX, 1, members)
template <class T, int LOGNY, int TX, class PRO>
__global__ __launch_bounds__(NT) void k_yr2c(PRO pro, cx<T> *__restrict__ I, long sI, long fsI, int dst, int nx,
                                             const cx<T> *__restrict__ twy)
{
    constexpr int NY = 1 << LOGNY, NPR = TX / 2;
    __shared__ cx<T> buf[Lds<T, NY>::size(NPR)];
    __shared__ cx<T> twl[NY / 2];
    const int x0 = xcd_tile(blockIdx.x, gridDim.x) * TX;
    const int m = blockIdx.z;
    r2c_tile<T, LOGNY, NPR, PRO>(pro, m, I + m * sI + dst * fsI, x0, nx, buf, twl, twy);
}

// y r2c pass, two output fields at the same x, written interleaved into the
// space of fields dst, dst+1.  grid (Nx/TXP, 1, members)
template <class T, int LOGNY, int TXP, class PRO>
__global__ __launch_bounds__(NT) void k_yr2c_pair(PRO pro, cx<T> *__restrict__ I, long sI, long fsI, int dst, int nx,
                                                  const cx<T> *__restrict__ twy)
{
    constexpr int NY = 1 << LOGNY;
    __shared__ cx<T> buf[Lds<T, NY>::size(TXP)];
    __shared__ cx<T> twl[NY / 2];
    const int x0 = xcd_tile(blockIdx.x, gridDim.x) * TXP;
    const int m = blockIdx.z;
    r2c_pair_tile<T, LOGNY, TXP, PRO>(pro, m, I + m * sI + dst * fsI, x0, nx, buf, twl, twy);
}

// y c2r pass of a W level's outputs with the next level's y r2c fused behind
// field 0.  Field 0 is d phi^(LEV)/dz, the one output of the level that
// phi^(LEV+1) (ProPhi<LEV>) reads: its workgroups keep the transformed tile in
// LDS, form phi^(LEV+1) on the same x-rows from it, eta and the earlier
// levels' fields, and do that field's y r2c into slot dst; the other fields'
// workgroups do their c2r only.  Same arithmetic as k_yc2r then k_yr2c with
// equal tiles, so the same bits; one launch and one read of field 0 fewer.
// grid (Nx/TX, nfields, members)
template <class T, int LOGNY, int TX, int LEV>
__global__ __launch_bounds__(NT) void k_yc2r_phi(cx<T> *__restrict__ I, long sI, long fsI, T *__restrict__ P,
                                                 long sP, long fsP, FieldMap fm, int nx,
                                                 const cx<T> *__restrict__ twy, int nzr, ProPhi<T, LEV> pro, int dst)
{
    using F = vec_t<T>;
    constexpr int V = Vec<T>::N;
    constexpr int NY = 1 << LOGNY, NPR = TX / 2, NYV = NY / V, NTASK = NPR * NYV;
    constexpr int PER = (NTASK + NT - 1) / NT;
    using L = Lds<T, NY>;
    __shared__ cx<T> buf[L::size(NPR)];
    __shared__ cx<T> twl[NY / 2];
    const int x0 = xcd_tile(blockIdx.x, gridDim.x) * TX;
    const int m = blockIdx.z;
    const cx<T> *src = I + m * sI + fm.src[blockIdx.y] * fsI;
    T *pd = P + m * sP + fm.dst[blockIdx.y] * fsP;
    if (blockIdx.y != 0) {
        c2r_tile<T, LOGNY, NPR>(src, pd, x0, nx, buf, twl, twy, nzr);
        return;
    }
    c2r_tile<T, LOGNY, NPR, true>(src, pd, x0, nx, buf, twl, twy, nzr);
    __syncthreads();  // the tile of d phi^(LEV)/dz in buf
    // as r2c_tile's prologue; each task reads and rewrites its own elements
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * NT;
        if ((NTASK % NT) == 0 || e < NTASK) {
            const int p = e / NYV, y = (e % NYV) * V;
            cx<T> *d = buf + L::at(p, y);
            F da, db;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                da[v] = d[v].x;
                db[v] = d[v].y;
            }
            const F ra = pro(m, x0 + 2 * p, y, da);
            const F rb = pro(m, x0 + 2 * p + 1, y, db);
#pragma unroll
            for (int v = 0; v < V; ++v) d[v] = {ra[v], rb[v]};
        }
    }
    lds_fft<T, LOGNY, NPR, NT, false>(buf, twl);
    r2c_store<T, LOGNY, NPR>(buf, I + m * sI + dst * fsI, x0, nx);
}

// ---------------------------------------------------------------------------
// The RHS pass with the deferred fields.  Nine of the physical fields (at M=5)
// are read by nothing but the RHS prologue: the gradients eta_x, eta_y, phi_x,
// phi_y and the M terms of W^(M), d^{k+1} phi^(M-k)/dz^{k+1}, k = 0..M-1 (the
// fields with order + derivative = M+1).  Instead of a y c2r to the physical
// grid and a read back, this kernel does their y c2r itself, three fields at a
// time in LDS (the next three tiles load during the current FFT), and keeps at
// each of its points only the four gradients and the compensated sum W^(M).
// The other fields (eta and d^j phi^(a) with a+j <= M) come from the physical
// grid (see rhs_early for when they are read).  Then the nonlinear RHS in physical space
// (rhs_hos, hosm_2d_par.c:149-206) without the two linear terms W^(1) and
// -g eta, which the x epilogue adds in FP64:
//   N_eta = -(eta_x phi_x + eta_y phi_y) + (W^(M) - W^(1)) + |grad eta|^2 W^(M-2)
//   N_phi = 1/2[(W^2)^(M) + |grad eta|^2 (W^2)^(M-2)] - 1/2 |grad phi|^2
// with W^(a) (hosm_2d_par.c:322-357) and the W sums (:362-442) as in
// w_orders_v / w_upto_v / w2_upto_v, and their y r2c, written interleaved as
// in k_yr2c_pair.  grid (Nx/TXP, 1, members)
template <int M>
struct RhsDeferred {
    int slot[M + 4];  // x-intermediate slot of eta_x, eta_y, phi_x, phi_y, then W^(M) term k = 0..M-1
    int nzr[M + 4];   // rows that may be nonzero (a band of a masked state, or Nk)
    int nzo;          // output rows written (my: the final x pass masks the rest)
    int blk;          // the x-intermediates are blocked (xi_at)
};

// y c2r tile loads of NF fields at once (x-rows x0 .. x0+2 NPR-1 of each,
// read transposed as in c2r_tile: 16 consecutive ky of one row pair per 16
// lanes).  The NF fields' index ranges lie end to end over the workgroup, so a
// thread holds ceil(NF NL / NTT) pieces rather than NF ceil(NL / NTT) (with
// 512 threads and NL = 528, 4 rather than 6 for three fields)
template <class T, int LOGNY, int NPR, int NTT, int NF>
struct C2rBatch {
    static constexpr int NY = 1 << LOGNY, NK = NY / 2 + 1;
    static constexpr int NL = ((NK + 15) / 16) * 16 * NPR, NLB = NF * NL, PERL = (NLB + NTT - 1) / NTT;
    cx<T> A[PERL], B[PERL];
    __device__ __forceinline__ void load(const cx<T> *__restrict__ base, long fs, const int *slot, const int *nzr,
                                         int x0, int nx, int blk)
    {
        const int tid = opaque_tid<NTT>();
#pragma unroll
        for (int q = 0; q < PERL; ++q) {
            const int e = min(tid + q * NTT, NLB - 1);  // past the end: reload, unused
            const int i = e / NL, el = e % NL;
            const int p = (el / 16) % NPR;
            const int ky = min((el / (16 * NPR)) * 16 + el % 16, NK - 1);
            int sl = slot[0], nz = nzr[0];  // slot[i] by selects (no indexed private array)
#pragma unroll
            for (int k = 1; k < NF; ++k)
                if (i == k) {
                    sl = slot[k];
                    nz = nzr[k];
                }
            if (ky < nz) {
                load2(base + sl * fs + xi_at(ky, x0 + 2 * p, nx, blk), A[q], B[q]);
            } else {  // a row of a masked state's band that is zero
                A[q] = {T(0), T(0)};
                B[q] = {T(0), T(0)};
            }
        }
    }
    // Hermitian fill: field i of the batch to rows i NPR .. i NPR + NPR-1 of buf
    __device__ __forceinline__ void fill(cx<T> *buf) const
    {
        using L = Lds<T, NY>;
        const int tid = opaque_tid<NTT>();
#pragma unroll
        for (int q = 0; q < PERL; ++q) {
            const int e = tid + q * NTT;
            const int i = e / NL, el = e % NL;
            const int p = (el / 16) % NPR, ky = (el / (16 * NPR)) * 16 + el % 16;
            if (((NLB % NTT) == 0 || e < NLB) && ky < NK) {
                cx<T> a = A[q], b = B[q];
                if (ky == 0 || ky == NY / 2) {
                    a.y = T(0);
                    b.y = T(0);
                }
                cx<T> *row = buf + (i * NPR + p) * L::PITCH;
                row[ky + (ky >> 4)] = {a.x - b.y, a.y + b.x};
                if (ky > 0 && ky < NY / 2) row[L::at(0, NY - ky)] = {a.x + b.y, b.x - a.y};
            }
        }
    }
};

// the parts of the RHS sums that need W^(1) .. W^(M-1) only, from eta and the
// physical fields d^j phi^(a), a + j <= M (no deferred field): the sum of
// W^(M-1) .. W^(2) (W^(M) is added last, after phase A), W^(M-2) and
// (W^2)^(K) for K = M, M-2
template <class F, bool C>
struct RhsPre {
    KSum<F, C> nl;
    F m2, w2, w22;
};

// the physical fields the RHS pass reads besides eta, at one task's points:
// d^j phi^(a), a + j <= M (d_pos order), or with WF (the diagonal passes ran)
// W^(1) .. W^(M-1) themselves
template <class T, int M, bool WF>
struct RhsPhys {
    static constexpr int NF = WF ? M - 1 : (M - 1) * M / 2;
    vec_t<T> d[NF > 0 ? NF : 1];
    __device__ __forceinline__ void load(const Phys<T> &ph, const HosIdx &ix, int m, int x, int y)
    {
        if constexpr (WF) {
#pragma unroll
            for (int a = 1; a < M; ++a) d[a - 1] = ph.vat(m, ix.w[a], x, y);
        } else {
#pragma unroll
            for (int a = 1; a < M; ++a)
#pragma unroll
                for (int j = 1; a + j <= M; ++j) d[d_pos<M - 1>(a, j)] = ph.vat(m, ix.d[a][j], x, y);
        }
    }
};

template <class T, int M, bool C, bool WF>
__device__ __forceinline__ RhsPre<vec_t<T>, C> rhs_pre(const HosIdx &ix, const RhsPhys<T, M, WF> &ps, vec_t<T> e)
{
    using F = vec_t<T>;
    const F *d = ps.d;
    F W[M + 1];
#pragma unroll
    for (int a = 1; a < M; ++a) {
        if constexpr (WF) {
            W[a] = d[a - 1];
        } else {
            F w = F(1);
            KSum<F, C> sum;
#pragma unroll
            for (int k = 0; k < a; ++k) {
                sum.add(w * T(ix.invfact[k]) * d[d_pos<M - 1>(a - k, k + 1)]);
                w *= e;
            }
            W[a] = sum.value();
        }
    }
    RhsPre<F, C> r;
#pragma unroll
    for (int a = M - 1; a >= 2; --a) r.nl.add(W[a]);
    r.m2 = F(0);
    r.w22 = F(0);
    if constexpr (M - 2 >= 1) {
        r.m2 = w_upto_v<F, 1, M - 2, C>(W);
        r.w22 = w2_upto_v<F, M - 2, C>(W);
    }
    r.w2 = w2_upto_v<F, M, C>(W);
    return r;
}

// phase B of the RHS pass: N_eta, N_phi at the tasks' points from the sums of
// rhs_pre and the deferred values, then the y r2c and the interleaved store
template <class T, int LOGNY, int TXP, int M, bool C, int NTT, int PER, int PLAN>
__device__ __forceinline__ void rhs_phase_b(cx<T> *__restrict__ out, int x0, int nx, int nzo, cx<T> *buf,
                                            const cx<T> *__restrict__ twy, const int *tp, const int *ty,
                                            const RhsPre<vec_t<T>, C> *pre, const vec_t<T> *ex, const vec_t<T> *ey,
                                            const vec_t<T> *px, const vec_t<T> *py, const KSum<vec_t<T>, C> *wm)
{
    using F = vec_t<T>;
    constexpr int V = Vec<T>::N;
    constexpr int NY = 1 << LOGNY, NK = NY / 2 + 1, NYV = NY / V;
    constexpr int NTASK = TXP * NYV;
    using L = Lds<T, NY>;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * NTT;
        if ((NTASK % NTT) != 0 && e >= NTASK) continue;
        // sum_{a=2}^{M} W^(a): W^(M) joins the compensated sum of W^(M-1) .. W^(2)
        KSum<F, C> nl = pre[q].nl;
        if constexpr (M >= 2) nl.add(wm[q].value());
        const F G = ex[q] * ex[q] + ey[q] * ey[q];
        const F na = -(ex[q] * px[q] + ey[q] * py[q]) + nl.value() + G * pre[q].m2;
        const F nb = T(0.5) * (pre[q].w2 + G * pre[q].w22) - T(0.5) * (px[q] * px[q] + py[q] * py[q]);
        cx<T> *o = buf + L::at(tp[q], ty[q]);
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] = {na[v], nb[v]};
    }
    lds_fft<T, LOGNY, TXP, NTT, false, PLAN>(buf, twy);
    constexpr int NO = ((NK + 15) / 16) * 16 * TXP;
    for (int e = threadIdx.x; e < NO; e += NTT) {
        const int p = (e / 16) % TXP, ky = (e / (16 * TXP)) * 16 + e % 16;
        if (ky >= nzo) continue;  // nzo <= NK
        const cx<T> zk = buf[L::at(p, ky)];
        const cx<T> zm = buf[L::at(p, (NY - ky) & (NY - 1))];
        store2(out + 2 * ((long)ky * nx + x0 + p), cx<T>{T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)},
               cx<T>{T(0.5) * (zk.y + zm.y), T(-0.5) * (zk.x - zm.x)});
    }
}


// Deferred fields per LDS batch of the RHS pass: four in the FP32 form over
// W fields at Ny >= 1024 (8 x 1024^2, M = 5, blocked x-intermediates: two
// batches of four 101-103 us per launch against 108-110 with batches of 3, 3
// and 2; the same VGPRs, 39 KB of LDS), three otherwise.  There a workgroup
// holds one column pair whatever the y tile (TXP = 2), so the batch and the
// radix plan of its row FFTs (fft_stage_bits: the same for 2 to 4 rows of 1024
// or 2048 points) do not depend on the tile.  At Ny = 256, 512 a four-row
// batch would take another plan than a three-row one, and tiles of 4 and 8
// column pairs other plans again.
template <class T, bool WF, int LOGNY>
constexpr int rhs_bf()
{
    return WF && sizeof(T) == 4 && LOGNY >= 10 ? 4 : 3;
}

// the physical fields are read before phase A (with the first deferred tiles,
// one memory latency less per tile) where they are few: FP32 form, at most
// ten fields (M <= 5, or W^(a) fields at any M); their sums then wait in LDS
// (4 vectors per task, 5 compensated)
template <class T, int M, bool WF = false>
constexpr bool rhs_early()
{
    // over W^(a) fields the sums are a few products: read them after phase A
    // and keep the LDS free (141 against 145 us per launch at 8 x 1024^2, M = 5)
    return sizeof(T) == 4 && !WF && RhsPhys<T, M, WF>::NF <= 10;
}

template <class T, int LOGNY, int TXP, int M, bool C, int NTT, int PLAN, bool WF>
__device__ __forceinline__ void rhs_tile(const Phys<T> &ph, const HosIdx &ix, const cx<T> *__restrict__ Im,
                                         long fsI, const RhsDeferred<M> &dfr, cx<T> *__restrict__ out, int m,
                                         int x0, int nx, cx<T> *buf, vec_t<T> *pbuf, cx<T> *twl,
                                         const cx<T> *__restrict__ twy)
{
    using F = vec_t<T>;
    constexpr int V = Vec<T>::N;
    constexpr int NY = 1 << LOGNY, NYV = NY / V, NPR = TXP / 2;
    constexpr int NTASK = TXP * NYV, PER = (NTASK + NTT - 1) / NTT;
    // deferred fields: the gradients and the W^(M) terms k = K0..M-1 (over W
    // fields the term k = 0, d phi^(M)/dz, is added in spectral space by the
    // RK epilogue)
    constexpr int K0 = WF ? 1 : 0, ND = M + 4 - K0;
    constexpr bool EARLY = rhs_early<T, M, WF>();
    using L = Lds<T, NY>;
    // this thread's tasks: V consecutive y of x-row p (clamped past the end)
    int tp[PER], ty[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = min((int)threadIdx.x + q * NTT, NTASK - 1);
        tp[q] = e / NYV;
        ty[q] = (e % NYV) * V;
    }
    F eta[PER], ex[PER], ey[PER], px[PER], py[PER], wek[PER];
    KSum<F, C> wm[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) wek[q] = F(1);
    RhsPre<F, C> pre[PER];
    // deferred fields in batches of BF, the last one of NFL
    constexpr int BF = rhs_bf<T, WF, LOGNY>(), NB = (ND + BF - 1) / BF, NFL = ND - (NB - 1) * BF;
    C2rBatch<T, LOGNY, NPR, NTT, BF> tb;   // batches 0 .. NB-2
    C2rBatch<T, LOGNY, NPR, NTT, NFL> tl;  // batch NB-1
    // the twiddles to LDS (TwLds): the FFTs run while the next batch loads
    TwLds<T, NY, NTT> tws;
    tws.load(twy);
#pragma unroll
    for (int q = 0; q < PER; ++q) eta[q] = ph.vat(m, ix.eta, x0 + tp[q], ty[q]);
    if constexpr (EARLY) {
        // the physical fields first, the first batch of deferred tiles right
        // after them (loads complete in order: the sums wait for the physical
        // fields only, the tiles are still in flight)
        RhsPhys<T, M, WF> ps[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) ps[q].load(ph, ix, m, x0 + tp[q], ty[q]);
        if (NB > 1)
            tb.load(Im, fsI, dfr.slot, dfr.nzr, x0, nx, dfr.blk);
        else
            tl.load(Im, fsI, dfr.slot, dfr.nzr, x0, nx, dfr.blk);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = threadIdx.x + q * NTT;
            if ((NTASK % NTT) != 0 && e >= NTASK) continue;
            const RhsPre<F, C> r = rhs_pre<T, M, C, WF>(ix, ps[q], eta[q]);
            F *o = pbuf + e;  // own slots only: no barrier
            o[0] = r.nl.s;
            o[NTASK] = r.m2;
            o[2 * NTASK] = r.w2;
            o[3 * NTASK] = r.w22;
            if constexpr (C) o[4 * NTASK] = r.nl.c;
        }
    } else {
        if (NB > 1)
            tb.load(Im, fsI, dfr.slot, dfr.nzr, x0, nx, dfr.blk);
        else
            tl.load(Im, fsI, dfr.slot, dfr.nzr, x0, nx, dfr.blk);
    }
    tws.store(twl);
    // phase A: y c2r of the deferred fields, BF at a time (BF*NPR rows share
    // the stages and barriers of one LDS FFT); the next batch's tiles are
    // loaded into the same registers during the current FFT.  Unrolled, so
    // each batch's row count is a constant and no tile register is copied
    // (a copy would wait for the loads in flight)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int nf = b + 1 < NB ? BF : NFL;
        if (b + 1 < NB)
            tb.fill(buf);
        else
            tl.fill(buf);
        if (b + 2 < NB)
            tb.load(Im, fsI, dfr.slot + (b + 1) * BF, dfr.nzr + (b + 1) * BF, x0, nx, dfr.blk);
        else if (b + 1 < NB)
            tl.load(Im, fsI, dfr.slot + (b + 1) * BF, dfr.nzr + (b + 1) * BF, x0, nx, dfr.blk);
        if (nf == 4)
            lds_fft<T, LOGNY, 4 * NPR, NTT, true, PLAN>(buf, twl);
        else if (nf == 3)
            lds_fft<T, LOGNY, 3 * NPR, NTT, true, PLAN>(buf, twl);
        else if (nf == 2)
            lds_fft<T, LOGNY, 2 * NPR, NTT, true, PLAN>(buf, twl);
        else
            lds_fft<T, LOGNY, NPR, NTT, true, PLAN>(buf, twl);
#pragma unroll
        for (int i = 0; i < BF; ++i) {
            if (i >= nf) break;
            const int f = b * BF + i;
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                F val;
                const cx<T> *z = buf + L::at(i * NPR + (tp[q] >> 1), ty[q]);
#pragma unroll
                for (int v = 0; v < V; ++v) val[v] = (tp[q] & 1) ? z[v].y : z[v].x;
                if (f == 0) ex[q] = val;
                else if (f == 1) ey[q] = val;
                else if (f == 2) px[q] = val;
                else if (f == 3) py[q] = val;
                else {
                    // W^(M) term k = f-4+K0: eta^k/k! d^{k+1} phi^(M-k)/dz^{k+1}, k
                    // ascending (wek: eta^k, the running product, the same bits
                    // as a product of k factors from 1)
                    if (f > 4 || K0) wek[q] *= eta[q];
                    wm[q].add(wek[q] * T(ix.invfact[f - 4 + K0]) * val);
                }
            }
        }
        __syncthreads();  // buf is refilled next
    }
    if constexpr (EARLY) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = threadIdx.x + q * NTT;
            if ((NTASK % NTT) != 0 && e >= NTASK) continue;
            const F *o = pbuf + e;
            pre[q].nl.s = o[0];
            pre[q].m2 = o[NTASK];
            pre[q].w2 = o[2 * NTASK];
            pre[q].w22 = o[3 * NTASK];
            if constexpr (C) pre[q].nl.c = o[4 * NTASK];
        }
    } else {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            RhsPhys<T, M, WF> ps;
            ps.load(ph, ix, m, x0 + tp[q], ty[q]);
            pre[q] = rhs_pre<T, M, C, WF>(ix, ps, eta[q]);
        }
    }
    rhs_phase_b<T, LOGNY, TXP, M, C, NTT, PER, PLAN>(out, x0, nx, dfr.nzo, buf, twl, tp, ty, pre, ex, ey, px, py, wm);
}

// 512 threads: half the per-thread points (and accumulators) of a 256-thread
// workgroup on the same tile
constexpr int NT_RHS = 512;
// The RHS pass over W^(a) fields (after the diagonal passes) holds far fewer
// values per point, so in the FP32 form it runs in 256-thread workgroups (two
// tasks per thread): more, smaller workgroups per CU hide its latency better
// (8 x 1024^2, M = 5: 145 -> 133 us per launch); FP64 keeps 512 threads
template <class T, bool WF>
constexpr int rhs_nt()
{
    return WF && sizeof(T) == 4 ? 256 : NT_RHS;
}

// PLAN: the radix family of its row FFTs (fft_stage_bits; the autotuner's choice).
// WF: the physical inputs are W^(1) .. W^(M-1) (written by the diagonal
// passes, k_ydiag) rather than the d^j phi^(a), a + j <= M
template <class T, int LOGNY, int TXP, int M, int PLAN, bool WF = false>
__global__ __launch_bounds__((rhs_nt<T, WF>())) void k_yr2c_rhs(Phys<T> ph, HosIdx ix, int comp, cx<T> *__restrict__ I, long sI,
                                                     long fsI, RhsDeferred<M> dfr, int dst, int nx,
                                                     const cx<T> *__restrict__ twy)
{
    static_assert(TXP % 2 == 0, "x-rows in pairs");
    constexpr int NY = 1 << LOGNY;
    // the r2c of TXP rows, or a batch of rhs_bf deferred fields of TXP/2 rows each
    constexpr int BROWS = rhs_bf<T, WF, LOGNY>() * (TXP / 2);  // rows of the largest deferred batch
    __shared__ cx<T> buf[Lds<T, NY>::size(TXP > BROWS ? TXP : BROWS)];
    // the early sums (rhs_early): up to 5 vectors per task
    constexpr int NTASK = TXP * (NY / Vec<T>::N);
    __shared__ vec_t<T> pbuf[rhs_early<T, M, WF>() ? 5 * NTASK : 1];
    __shared__ cx<T> twl[NY / 2];
    const int x0 = xcd_tile(blockIdx.x, gridDim.x) * TXP;
    const int m = blockIdx.z;
    const cx<T> *Im = I + m * sI;
    cx<T> *out = I + m * sI + dst * fsI;
    if (comp)
        rhs_tile<T, LOGNY, TXP, M, true, rhs_nt<T, WF>(), PLAN, WF>(ph, ix, Im, fsI, dfr, out, m, x0, nx, buf, pbuf, twl,
                                                                     twy);
    else
        rhs_tile<T, LOGNY, TXP, M, false, rhs_nt<T, WF>(), PLAN, WF>(ph, ix, Im, fsI, dfr, out, m, x0, nx, buf, pbuf,
                                                                      twl, twy);
}

// ---------------------------------------------------------------------------
// Diagonal pass L of the W expansion (L = 1 .. M-1).  The fields with order +
// derivative = L+1, F_k = d^k phi^(L+1-k)/dz^k (k = 1..L), are exactly the
// ones both phi^(L+1) and W^(L) are made of:
//   phi^(L+1) = - sum_{k=1}^{L} eta^k/k! F_k          (hosm_2d_par.c:302-317)
//   W^(L)     =   sum_{k=1}^{L} eta^(k-1)/(k-1)! F_k  (hosm_2d_par.c:322-357)
// so no F_k needs to reach the physical grid: this pass does their y c2r
// itself straight from the x-intermediates (three at a time in LDS, the next
// tiles loading during the current FFT, as the RHS pass does its deferred
// fields), sums both at each point with terms in the reference's order (k
// ascending), writes W^(L) to the physical grid for the RHS pass, and does
// the y r2c of phi^(L+1) into slot dst for the next x forward-inverse pass.
// Each F_k is then written once (as an x-intermediate) and read once, against
// a y c2r write and two reads of the physical field (the level pass and the
// RHS pass) otherwise.  grid (Nx/TXP, 1, members)
// Pass 1 also does the y c2r of eta (field 0 of its batch), which every later
// pass and the RHS pass read from the physical grid.
template <int L>
struct DiagArgs {
    static constexpr int ETA = L == 1 ? 1 : 0;
    static constexpr int NF = L + ETA;  // fields transformed: [eta,] F_1 .. F_L
    // at most three fields per LDS batch, the batches as even as they come
    static constexpr int NB = (NF + 2) / 3, BF = (NF + NB - 1) / NB, NFL = NF - (NB - 1) * BF;
    int slot[NF];  // x-intermediate slots ([eta,] F_k)
    int nzr[NF];   // rows that may be nonzero (a band of a masked state, or Nk)
    int dst;       // slot of the y r2c of phi^(L+1)
    int nzo;       // rows of that r2c written (the last pass: the rows the mask keeps)
    int wf;        // physical field of W^(L)
    int etaf;      // physical field of eta (pass 1)
    int blk;       // the x-intermediates are blocked (xi_at)
};

// 256-thread workgroups: 53.7 against 56.6 us per pass at 8 x 1024^2, M = 5
// with 512 (more independent workgroups per CU; their stages are barriers)
constexpr int NT_DIAG = 256;

template <class T, int LOGNY, int TXP, int LV, bool C, int NTT>
__device__ __forceinline__ void diag_tile(const Phys<T> &ph, T *__restrict__ P, const HosIdx &ix,
                                          const cx<T> *__restrict__ Im, long fsI, const DiagArgs<LV> &da,
                                          cx<T> *__restrict__ out, int m, int x0, int nx, cx<T> *buf, cx<T> *twl,
                                          const cx<T> *__restrict__ twy)
{
    using F = vec_t<T>;
    constexpr int V = Vec<T>::N;
    constexpr int NY = 1 << LOGNY, NYV = NY / V, NPR = TXP / 2;
    constexpr int NTASK = TXP * NYV, PER = (NTASK + NTT - 1) / NTT;
    using Lay = Lds<T, NY>;
    // this thread's tasks: V consecutive y of x-row p (clamped past the end)
    int tp[PER], ty[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = min((int)threadIdx.x + q * NTT, NTASK - 1);
        tp[q] = e / NYV;
        ty[q] = (e % NYV) * V;
    }
    using DA = DiagArgs<LV>;
    constexpr int ETA = DA::ETA, NB = DA::NB, BF = DA::BF, NFL = DA::NFL;
    C2rBatch<T, LOGNY, NPR, NTT, BF> tb;   // batches 0 .. NB-2
    C2rBatch<T, LOGNY, NPR, NTT, NFL> tl;  // batch NB-1
    TwLds<T, NY, NTT> tws;
    tws.load(twy);
    if (NB > 1)
        tb.load(Im, fsI, da.slot, da.nzr, x0, nx, da.blk);
    else
        tl.load(Im, fsI, da.slot, da.nzr, x0, nx, da.blk);
    F eta[PER], ek[PER];
    KSum<F, C> ps[PER], ws[PER];  // the sums of phi^(L+1) (negated last) and W^(L)
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        if constexpr (!ETA) eta[q] = ph.vat(m, ix.eta, x0 + tp[q], ty[q]);
        ek[q] = F(1);  // eta^(k-1) before field k
    }
    tws.store(twl);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int nf = b + 1 < NB ? BF : NFL;
        if (b + 1 < NB)
            tb.fill(buf);
        else
            tl.fill(buf);
        if (b + 2 < NB)
            tb.load(Im, fsI, da.slot + (b + 1) * BF, da.nzr + (b + 1) * BF, x0, nx, da.blk);
        else if (b + 1 < NB)
            tl.load(Im, fsI, da.slot + (b + 1) * BF, da.nzr + (b + 1) * BF, x0, nx, da.blk);
        if (nf == 3)
            lds_fft<T, LOGNY, 3 * NPR, NTT, true>(buf, twl);
        else if (nf == 2)
            lds_fft<T, LOGNY, 2 * NPR, NTT, true>(buf, twl);
        else
            lds_fft<T, LOGNY, NPR, NTT, true>(buf, twl);
#pragma unroll
        for (int i = 0; i < BF; ++i) {
            if (i >= nf) break;
            const int k = b * BF + i + 1 - ETA;  // field F_k (k = 0: eta)
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                F val;
                const cx<T> *z = buf + Lay::at(i * NPR + (tp[q] >> 1), ty[q]);
#pragma unroll
                for (int v = 0; v < V; ++v) val[v] = (tp[q] & 1) ? z[v].y : z[v].x;
                if (ETA && k == 0) {
                    eta[q] = val;
                    continue;
                }
                ws[q].add(ek[q] * T(ix.invfact[k - 1]) * val);
                ek[q] *= eta[q];
                ps[q].add(ek[q] * T(ix.invfact[k]) * val);
            }
        }
        __syncthreads();  // buf is refilled next
    }
    // W^(L) (and eta) to the physical grid; phi^(L+1) of x-rows x0+2p (real
    // part) and x0+2p+1 (imaginary part) into row p of the r2c tile
    T *wp = P + m * ph.sP + da.wf * ph.fsP;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * NTT;
        if ((NTASK % NTT) != 0 && e >= NTASK) continue;
        const long at = (long)(x0 + tp[q]) * ph.ny + ty[q];
        *reinterpret_cast<F *>(wp + at) = ws[q].value();
        if constexpr (ETA) *reinterpret_cast<F *>(P + m * ph.sP + da.etaf * ph.fsP + at) = eta[q];
        const F phv = -ps[q].value();
        cx<T> *o = buf + Lay::at(tp[q] >> 1, ty[q]);
#pragma unroll
        for (int v = 0; v < V; ++v) {
            if (tp[q] & 1)
                o[v].y = phv[v];
            else
                o[v].x = phv[v];
        }
    }
    lds_fft<T, LOGNY, NPR, NTT, false>(buf, twl);
    r2c_store<T, LOGNY, NPR, NTT>(buf, out, x0, nx, da.nzo, da.blk);
}

template <class T, int LOGNY, int TXP, int LV>
__global__ __launch_bounds__(NT_DIAG) void k_ydiag(Phys<T> ph, T *__restrict__ P, HosIdx ix, int comp,
                                                  cx<T> *__restrict__ I, long sI, long fsI, DiagArgs<LV> da, int nx,
                                                  const cx<T> *__restrict__ twy)
{
    static_assert(TXP % 2 == 0, "x-rows in pairs");
    constexpr int NY = 1 << LOGNY, NPR = TXP / 2, NBUF = DiagArgs<LV>::BF;  // rows of the largest batch
    __shared__ cx<T> buf[Lds<T, NY>::size(NBUF * NPR)];
    __shared__ cx<T> twl[NY / 2];
    const int x0 = xcd_tile(blockIdx.x, gridDim.x) * TXP;
    const int m = blockIdx.z;
    const cx<T> *Im = I + m * sI;
    cx<T> *out = I + m * sI + da.dst * fsI;
    if (comp)
        diag_tile<T, LOGNY, TXP, LV, true, NT_DIAG>(ph, P, ix, Im, fsI, da, out, m, x0, nx, buf, twl, twy);
    else
        diag_tile<T, LOGNY, TXP, LV, false, NT_DIAG>(ph, P, ix, Im, fsI, da, out, m, x0, nx, buf, twl, twy);
}

// ---------------------------------------------------------------------------
// x forward pass of phi^(lev+1), then d^j/dz^j for j = 1..ndst and the x
// inverse of each, written to fields dst0 .. dst0+ndst-1 (the last one to
// field dtop instead when dtop >= 0): the spectral phi^(lev+1) never leaves the chip (each
// thread keeps its modes in registers; the inverses run XNB rows at a time).
// grid (ceil(Nk/ROWS), 1, members)
template <class T, int LOGNX, int ROWS>
__global__ __launch_bounds__(NT) void k_xfwdinv(cx<T> *__restrict__ I, long sI, long fs, long ip, int src, int dst0,
                                                int ndst, int dtop, double scale, KTab kt,
                                                const cx<T> *__restrict__ twx, int spec, int nrows)
{
    using G = XGeom<T, LOGNX>;
    static_assert(G::ROWS == ROWS, "x rows");
    constexpr int NX = G::NX, PER = G::PER, NB = G::XNB;
    using L = Lds<T, NX>;
    __shared__ cx<T> W[L::size(NB * ROWS)];
    const int ky0 = blockIdx.x * ROWS;
    using R = T;  // multiplier arithmetic: FP64 (FP64 form) or FP32 (FP32 form)
    const int m = blockIdx.z;
    cx<T> *base = I + m * sI;
    const cx<T> *in = base + src * fs;
    const double *tht = kt.th + m * kt.ths;
    cx<T> s[PER];
    double kxv[PER], thd[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * NT;
        const int kx = e % NX, ky = min(ky0 + e / NX, kt.nk - 1);
        s[q] = in[(long)ky * ip + kx];
        kxv[q] = kt.kx[kx];
        if constexpr (sizeof(R) == 8) thd[q] = tht[(long)ky * NX + kx];
        else thd[q] = 1.0;
    }
    const double h = kt.h(m);
    R kz[PER], th[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * NT;
        const int ky = ky0 + e / NX;
        W[L::at(e / NX, e % NX)] = (ky < kt.nk) ? s[q] : cx<T>{T(0), T(0)};
        kz_th<R>(kxv[q], ky * kt.ky0, h, thd[q], kz[q], th[q]);
    }
    lds_fft<T, LOGNX, ROWS, NT, false>(W, twx);
#pragma unroll
    for (int q = 0; q < PER; ++q) s[q] = W[L::at((threadIdx.x + q * NT) / NX, (threadIdx.x + q * NT) % NX)];
    if (spec) {
        // the one output, d phi/dz (ndst = 1), left in x-spectral space for
        // the RK epilogue: rows ky < nrows only
        cx<T> *out = base + dtop * fs;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = threadIdx.x + q * NT;
            const int kx = e % NX, ky = ky0 + e / NX;
            const T mul = T(R(scale) * kz[q] * th[q]);
            if (ky < nrows) out[(long)ky * ip + kx] = {mul * s[q].x, mul * s[q].y};
        }
        return;
    }
    R kj[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) kj[q] = R(scale);
    for (int j0 = 1; j0 <= ndst; j0 += NB) {
        __syncthreads();  // W has been read (forward result or the previous batch)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = j0 + b;
            if (j > ndst) break;
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                kj[q] *= kz[q];  // scale * |k|^j
                const T mul = T((j & 1) ? kj[q] * th[q] : kj[q]);  // odd j: tanh(|k| h)
                const int e = threadIdx.x + q * NT;
                W[L::at(b * ROWS + e / NX, e % NX)] = {mul * s[q].x, mul * s[q].y};
            }
        }
        lds_fft<T, LOGNX, ROWS * NB, NT, true>(W, twx);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if (j0 + b > ndst) break;
            const int j = j0 + b;
            cx<T> *out = base + ((j == ndst && dtop >= 0) ? dtop : dst0 + j - 1) * fs;
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int e = threadIdx.x + q * NT;
                const int kx = e % NX, ky = ky0 + e / NX;
                if (ky < kt.nk) out[(long)ky * ip + kx] = W[L::at(b * ROWS + e / NX, kx)];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// final x forward pass: N_eta, N_phi -> K = RHS (linear terms added in FP64,
// ramp, dealias) -> RK stage update.  The two inputs are fields iA and iB, or
// (ilv) the interleaved pair that k_yr2c_pair wrote at field iA.
enum : int { EPI_RK_SUBDIAG = 0, EPI_STORE_K = 1, EPI_STORE_SPEC2 = 2 };
struct RkArgs {
    int mode;
    int stage, nstages;
    double c_s;       // c[s]
    double b_next;    // b[s+1][s] (sub-diagonal schemes)
    double g;
    int rampflg;
    double tramp;
    const double *t;  // per-member time t_n
    const double *dt; // per-member dt
    cx<double> *U;    // state
    cx<double> *A;    // accumulator
    cx<double> *Yn;   // next stage input (may alias the current stage input)
    cx<double> *K;    // EPI_STORE_K / EPI_STORE_SPEC2 output
    long sK;          // member stride of K
    double *hpart;    // stage-0 Hamiltonian partials [members][gridDim.x]
    double *vol;      // [members]
    int mean_level;
    const double *mean0;  // [members]
    int stream;           // 1: last-use loads with the non-temporal hint
    int band;             // 1: masked states (EPI_RK_SUBDIAG): the masked modes of U, A, Y are
                          // neither read nor written (they stay zero)
};

// RHS at spectral point idx = (ky, kx) of member m from the forward-transformed
// nonlinear terms na (eta_t) and nb (phi_t) and the RK stage update
// (time_schemes_par.c:38-67 / sol_update_RK), FP64.  hsum accumulates the
// stage-0 Hamiltonian terms of the point.
// ta: the x spectrum of d phi^(M)/dz, scaled (the diagonal passes leave this
// term of W^(M), the one that enters N_eta linearly with no eta factor, in
// spectral space; zero otherwise)
template <class T>
__device__ __forceinline__ void rk_epilogue(const RkArgs &rk, const KTab &kt, int m, long idx, int kx, int ky, int NX,
                                            cx<T> na, cx<T> nb, double scale, double R, double dt,
                                            const cx<double> *y, long fsd, double &hsum, cx<T> ta)
{
    if (rk.mode == EPI_STORE_SPEC2) {
        cx<double> *k = rk.K + m * rk.sK;
        k[idx] = {scale * double(na.x), scale * double(na.y)};
        k[fsd + idx] = {scale * double(nb.x), scale * double(nb.y)};
        return;
    }
    // a masked state's masked modes stay zero: K is zero there, so u, the
    // accumulator and the next stage input keep their zeros without a store
    if (rk.band && masked(kx, ky, NX, kt.mx, kt.my)) return;
    const cx<double> ye = rk.stream ? load_nt(y + idx) : y[idx];
    const cx<double> yp = rk.stream ? load_nt(y + fsd + idx) : y[fsd + idx];
    const double kxv = kt.kx[kx], kyv = ky * kt.ky0, h = kt.h(m);
    // |k| tanh(|k| h), the expression of k_th_table (same value): computed,
    // not read, since this pass is bound by its FP64 traffic
    const double kz = sqrt(kxv * kxv + kyv * kyv);
    const double w1 = kz * (h > 0.0 ? tanh(kz * h) : 1.0);
    // eta_t = W^(1) + R*(rest), phi_t = -g eta + R*(rest)  (hosm_2d_par.c:211-229)
    cx<double> ke = {w1 * yp.x + R * (scale * double(na.x) + double(ta.x)),
                     w1 * yp.y + R * (scale * double(na.y) + double(ta.y))};
    cx<double> kp = {-rk.g * ye.x + R * scale * double(nb.x), -rk.g * ye.y + R * scale * double(nb.y)};
    if (masked(kx, ky, NX, kt.mx, kt.my)) {
        ke = {0.0, 0.0};
        kp = {0.0, 0.0};
    }
    if (rk.mode == EPI_STORE_K) {
        cx<double> *k = rk.K + m * rk.sK;
        k[idx] = ke;
        k[fsd + idx] = kp;
        if (rk.stage == 0 && rk.hpart) {
            const cx<double> *u = rk.U + m * 2 * fsd;
            const cx<double> ue = u[idx], up = u[fsd + idx];
            const double w = (ky == 0) ? 1.0 : 2.0;
            hsum += w * (rk.g * (ue.x * ue.x + ue.y * ue.y) + up.x * ke.x + up.y * ke.y);
            if (idx == 0) rk.vol[m] = ue.x;
        }
        return;
    }
    // EPI_RK_SUBDIAG: Y_{s+1} = u + dt b[s+1][s] K_s ; acc += dt c[s] K_s.
    // Stage 0's input is the state itself (y == u, not read twice); the last
    // stage writes the state only, which the next step's stage 0 reads.
    cx<double> *u = rk.U + m * 2 * fsd;
    cx<double> *acc = rk.A + m * 2 * fsd;
    cx<double> *yn = rk.Yn + m * 2 * fsd;
    const cx<double> ue = rk.stage == 0 ? ye : u[idx], up = rk.stage == 0 ? yp : u[fsd + idx];
    if (rk.stage == 0 && rk.hpart) {
        // Hamiltonian, hosm_2d_par.c:462-523: weight 1 on row j=0, 2 elsewhere
        const double w = (ky == 0) ? 1.0 : 2.0;
        hsum += w * (rk.g * (ue.x * ue.x + ue.y * ue.y) + up.x * ke.x + up.y * ke.y);
        if (idx == 0) rk.vol[m] = ue.x;
    }
    const double fc = dt * rk.c_s;
    cx<double> ae, ap;
    if (rk.stage == 0) {
        ae = ue;
        ap = up;
    } else {
        ae = rk.stream ? load_nt(acc + idx) : acc[idx];
        ap = rk.stream ? load_nt(acc + fsd + idx) : acc[fsd + idx];
    }
    ae = {ae.x + fc * ke.x, ae.y + fc * ke.y};
    ap = {ap.x + fc * kp.x, ap.y + fc * kp.y};
    if (rk.stage == rk.nstages - 1) {
        if (rk.mean_level && idx == 0) ae = {rk.mean0[m], 0.0};
        u[idx] = ae;
        u[fsd + idx] = ap;
    } else {
        acc[idx] = ae;
        acc[fsd + idx] = ap;
        const double fb = dt * rk.b_next;
        yn[idx] = {ue.x + fb * ke.x, ue.y + fb * ke.y};
        yn[fsd + idx] = {up.x + fb * kp.x, up.y + fb * kp.y};
    }
}

// Rows ky < nzr only (my with RkArgs::band: the RHS is masked beyond).
// grid (ceil(nzr/ROWS), 1, members); the stage-0 Hamiltonian partials are
// [members][gridDim.x].
template <class T, int LOGNX, int ROWS>
__global__ __launch_bounds__(NT) void k_xfwd_rhs(const cx<T> *__restrict__ I, long sI, long fs, long ip, int iA,
                                                 int iB, int ilv, const cx<double> *Y, long sY, double scale,
                                                 KTab kt, RkArgs rk, const cx<T> *__restrict__ twx, int nzr, int iT)
{
    using G = XGeom<T, LOGNX>;
    static_assert(G::ROWS == ROWS, "x rows");
    constexpr int NX = G::NX, PER = G::PER;
    using L = Lds<T, NX>;
    __shared__ cx<T> S[L::size(2 * ROWS)];
    __shared__ double red[NT / 64];
    const int ky0 = blockIdx.x * ROWS;
    const int m = blockIdx.z;
    const cx<T> *ia = I + m * sI + iA * fs;
    const cx<T> *ib = I + m * sI + iB * fs;
    {
        cx<T> va[PER], vb[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = threadIdx.x + q * NT;
            const long idx = (long)min(ky0 + e / NX, nzr - 1) * ip + e % NX;  // x-intermediate row
            if (ilv) {
                // the interleaved pair of k_yr2c_pair
                if (rk.stream)
                    load2_nt(ia + 2 * idx, va[q], vb[q]);
                else
                    load2(ia + 2 * idx, va[q], vb[q]);
            } else {
                va[q] = ia[idx];
                vb[q] = ib[idx];
            }
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = threadIdx.x + q * NT;
            const bool ok = ky0 + e / NX < nzr;
            S[L::at(e / NX, e % NX)] = ok ? va[q] : cx<T>{T(0), T(0)};
            S[L::at(ROWS + e / NX, e % NX)] = ok ? vb[q] : cx<T>{T(0), T(0)};
        }
    }
    lds_fft<T, LOGNX, 2 * ROWS, NT, false>(S, twx);

    const long fsd = (long)kt.nk * NX;  // field stride of the FP64 [2][Nk][Nx] arrays
    const cx<double> *y = Y + m * sY;
    double hsum = 0.0;
    double R = 1.0;
    if (rk.rampflg == 1) {
        const double tt = rk.t[m];
        R = 1.0 - exp(-tt * tt / rk.tramp / rk.tramp);  // RampFun, hosm_2d_par.c:526-530
    }
    const double dt = (rk.mode == EPI_RK_SUBDIAG) ? rk.dt[m] : 0.0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * NT;
        const int r = e / NX, kx = e % NX, ky = ky0 + r;
        if (ky >= nzr) continue;
        const long idx = (long)ky * NX + kx;
        // the spectral term (iT >= 0) exists on the rows the mask keeps
        cx<T> ta = {T(0), T(0)};
        if (iT >= 0 && ky < kt.my) ta = I[m * sI + iT * fs + (long)ky * ip + kx];
        rk_epilogue(rk, kt, m, idx, kx, ky, NX, S[L::at(r, kx)], S[L::at(ROWS + r, kx)], scale, R, dt, y, fsd, hsum,
                    ta);
    }
    if (rk.hpart && rk.stage == 0 && rk.mode != EPI_STORE_SPEC2) {
        // deterministic block reduction: wave shuffle, then fixed-order sum of waves
        for (int off = 32; off > 0; off >>= 1) hsum += __shfl_down(hsum, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = hsum;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (int w = 0; w < NT / 64; ++w) s += red[w];
            rk.hpart[m * gridDim.x + blockIdx.x] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// Wave-per-row x inverse passes (the FP32 form at Nx = 512, 1024): each 64-lane
// wave owns one ky row, lane l the modes kx = l + 64 i (i = 0..E-1), which
// wave_fft transforms in registers, so the loads and stores of a row are
// coalesced 64-lane accesses and no workgroup barrier is needed.  A workgroup
// is XW waves on consecutive rows; grid (ceil(Nk/XW), groups, members); a wave
// past row Nk-1 returns at once, unless the x-intermediates are blocked
// (xi_at): then a workgroup's XW = XKB rows are one block, which the
// workgroup stores and loads together (xblock_store, xblock_load), and every
// wave takes part (a wave past the last row transforms data it never stores).
constexpr int XW = NT / 64;
static_assert(XW == XKB, "a wave x-pass workgroup holds one block of rows");

// units of 16 bytes (one column pair of one row) per thread in a block
template <int LOGNX>
constexpr int xblock_units() { return XKB * (1 << LOGNX) / 2 / NT; }

// the workgroup's rows ky0 .. ky0+XKB-1 from the waves' registers (v, this
// wave's row) to the blocked field dst: each wave puts its row in its LDS
// row, then consecutive threads store consecutive 16-byte units of the
// block (unit u = c XKB + r: column pair c of row ky0+r); rows >= nrows are
// left out.  Called by every wave of the workgroup (two barriers).
template <class T, int LOGNX>
__device__ __forceinline__ void xblock_store(const cx<T> *v, cx<T> *row, const cx<T> *lds, cx<T> *__restrict__ dst,
                                             int ky0, int nrows, long ip, int lane)
{
    constexpr int NX = 1 << LOGNX, E = NX / 64, NU = xblock_units<LOGNX>();
    using L = Lds<T, NX>;
    wave_sync();  // the FFT's last reads of the row precede these writes
#pragma unroll
    for (int i = 0; i < E; ++i) row[L::at(0, lane + 64 * i)] = v[i];
    __syncthreads();
    const int nv = min(XKB, nrows - ky0);
    cx<T> *b = dst + (long)ky0 * ip;
#pragma unroll 2
    for (int q = 0; q < NU; ++q) {
        const int u = threadIdx.x + q * NT, c = u / XKB, r = u % XKB;
        if (r < nv) {
            const cx<T> *rr = lds + r * L::PITCH;
            store2_nt(b + 2 * u, rr[L::at(0, 2 * c)], rr[L::at(0, 2 * c + 1)]);
        }
    }
    __syncthreads();  // the rows are free for the next transform
}

// the reverse: the block of rows ky0 .. ky0+XKB-1 of the blocked field src
// (rows < nrows) through the LDS rows into each wave's registers s (natural
// order, lane l holding x = l + 64 i).  Called by every wave of the workgroup.
template <class T, int LOGNX>
__device__ __forceinline__ void xblock_load(cx<T> *s, cx<T> *row, cx<T> *lds, const cx<T> *__restrict__ src, int ky0,
                                            int nrows, long ip, int lane)
{
    constexpr int NX = 1 << LOGNX, E = NX / 64, NU = xblock_units<LOGNX>();
    using L = Lds<T, NX>;
    const int nv = min(XKB, nrows - ky0);
    const cx<T> *b = src + (long)ky0 * ip;
    cx<T> a[NU], bb[NU];
#pragma unroll
    for (int q = 0; q < NU; ++q) {
        const int u = threadIdx.x + q * NT;
        if (u % XKB < nv) load2(b + 2 * u, a[q], bb[q]);
    }
#pragma unroll
    for (int q = 0; q < NU; ++q) {
        const int u = threadIdx.x + q * NT, c = u / XKB, r = u % XKB;
        if (r < nv) {
            cx<T> *rr = lds + r * L::PITCH;
            rr[L::at(0, 2 * c)] = a[q];
            rr[L::at(0, 2 * c + 1)] = bb[q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < E; ++i) s[i] = row[L::at(0, lane + 64 * i)];
    wave_sync();  // read before the FFT's exchanges write the row
}

template <class T, int LOGNX, bool BLK>
__global__ __launch_bounds__(NT) void k_xinv_w(const cx<double> *__restrict__ Y, long sY, cx<T> *__restrict__ I,
                                               long sI, long fs, long fsI, long ip, XinvOps ops, KTab kt,
                                               const cx<T> *__restrict__ twx, int nzr, int cmx)
{
    constexpr int NX = 1 << LOGNX, E = NX / 64;
    using L = Lds<T, NX>;
    using R = T;  // multiplier arithmetic in the precision of the form
    __shared__ cx<T> lds[XW * L::PITCH];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ky = blockIdx.x * XW + w;
    const bool live = ky < nzr;  // rows of the band only (col_kept)
    if (!BLK && !live) return;
    cx<T> *row = lds + w * L::PITCH;
    const int grp = blockIdx.y, m = blockIdx.z;
    const long r0 = (long)ky * NX;
    const cx<double> *src = Y + m * sY + ops.src[grp] * fs + r0;
    const double *tht = kt.th + m * kt.ths + r0;
    const double h = kt.h(m), kyv = ky * kt.ky0;
    cx<R> sr[E];
    R kxr[E], kz[E], th[E];
    {
        cx<double> sd[E];
        double kxv[E], thd[E];
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int kx = lane + 64 * i;
            const bool keep = live && col_kept(kx, NX, cmx);
            sd[i] = keep ? src[kx] : cx<double>{0.0, 0.0};
            kxv[i] = kt.kx[kx];
            if constexpr (sizeof(R) == 8) thd[i] = keep ? tht[kx] : 1.0;
            else thd[i] = 1.0;
        }
#pragma unroll
        for (int i = 0; i < E; ++i) {
            sr[i] = {R(sd[i].x), R(sd[i].y)};
            kxr[i] = R(kxv[i]);
            kz_th<R>(kxv[i], kyv, h, thd[i], kz[i], th[i]);  // d^n/dz^n: k^n tanh(kh)^(n odd)
        }
    }
    const R kyr = R(kyv);
    const int nops = ops.n[grp];
    for (int o = 0; o < nops; ++o) {
        const XinvOp op = ops.o[grp][o];
        cx<T> v[E];
        // the operator is uniform: one branch per output, not a select per mode
        if (op.op == OP_DX) {
#pragma unroll
            for (int i = 0; i < E; ++i) v[i] = {T(-kxr[i] * sr[i].y), T(kxr[i] * sr[i].x)};
        } else if (op.op == OP_DY) {
#pragma unroll
            for (int i = 0; i < E; ++i) v[i] = {T(-kyr * sr[i].y), T(kyr * sr[i].x)};
        } else if (op.op == OP_DZ) {
            R mm[E];
#pragma unroll
            for (int i = 0; i < E; ++i) mm[i] = (op.pow & 1) ? th[i] : R(1);
            for (int p = 0; p < op.pow; ++p)
#pragma unroll
                for (int i = 0; i < E; ++i) mm[i] *= kz[i];
#pragma unroll
            for (int i = 0; i < E; ++i) v[i] = {T(mm[i] * sr[i].x), T(mm[i] * sr[i].y)};
        } else {
#pragma unroll
            for (int i = 0; i < E; ++i) v[i] = {T(sr[i].x), T(sr[i].y)};
        }
        wave_fft<T, LOGNX, true>(v, row, twx, lane);
        cx<T> *dst = I + m * sI + op.dst * fsI;
        if constexpr (BLK) {
            xblock_store<T, LOGNX>(v, row, lds, dst, ky - w, nzr, ip, lane);
        } else {
#pragma unroll
            for (int i = 0; i < E; ++i) store_nt(dst + (long)ky * ip + lane + 64 * i, v[i]);
        }
    }
}

// x forward pass of phi^(lev+1), d^j/dz^j for j = 1..ndst and the x inverse of
// each, one row per wave: the spectral row stays in registers throughout
template <class T, int LOGNX, bool BLK>
__global__ __launch_bounds__(NT) void k_xfwdinv_w(cx<T> *__restrict__ I, long sI, long fs, long ip, int src,
                                                  int dst0, int ndst, int dtop, double scale, KTab kt,
                                                  const cx<T> *__restrict__ twx, int spec, int nrows)
{
    constexpr int NX = 1 << LOGNX, E = NX / 64;
    using L = Lds<T, NX>;
    using R = T;
    __shared__ cx<T> lds[XW * L::PITCH];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ky = blockIdx.x * XW + w;
    const bool live = ky < nrows;
    if (!BLK && !live) return;
    cx<T> *row = lds + w * L::PITCH;
    const int m = blockIdx.z;
    const long r0 = (long)ky * NX, rI = (long)ky * ip;  // tanh table row, x-intermediate row
    cx<T> *base = I + m * sI;
    const cx<T> *in = base + src * fs + rI;
    const double *tht = kt.th + m * kt.ths + r0;
    const double h = kt.h(m), kyv = ky * kt.ky0;
    cx<T> s[E];
    R kz[E], th[E];
    if constexpr (BLK) xblock_load<T, LOGNX>(s, row, lds, base + src * fs, ky - w, nrows, ip, lane);
    {
        double kxv[E], thd[E];
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int kx = lane + 64 * i;
            if constexpr (!BLK) s[i] = in[kx];
            kxv[i] = kt.kx[kx];
            if constexpr (sizeof(R) == 8) thd[i] = live ? tht[kx] : 1.0;
            else thd[i] = 1.0;
        }
#pragma unroll
        for (int i = 0; i < E; ++i) kz_th<R>(kxv[i], kyv, h, thd[i], kz[i], th[i]);
    }
    wave_fft<T, LOGNX, false, true>(s, row, twx, lane);
    if (spec) {
        // the one output, d phi/dz (ndst = 1), left in x-spectral space for
        // the RK epilogue (row by row, as that pass reads it)
        if (!live) return;
        cx<T> *out = base + dtop * fs + rI;
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const T mul = T(R(scale) * kz[i] * th[i]);
            out[lane + 64 * i] = {mul * s[i].x, mul * s[i].y};
        }
        return;
    }
    R kj[E];
#pragma unroll
    for (int i = 0; i < E; ++i) kj[i] = R(scale);
    for (int j = 1; j <= ndst; ++j) {
        cx<T> v[E];
#pragma unroll
        for (int i = 0; i < E; ++i) {
            kj[i] *= kz[i];  // scale * |k|^j
            const T mul = T((j & 1) ? kj[i] * th[i] : kj[i]);  // odd j: tanh(|k| h)
            v[i] = {mul * s[i].x, mul * s[i].y};
        }
        wave_fft<T, LOGNX, true, true>(v, row, twx, lane);
        cx<T> *out = base + ((j == ndst && dtop >= 0) ? dtop : dst0 + j - 1) * fs;
        if constexpr (BLK) {
            xblock_store<T, LOGNX>(v, row, lds, out, ky - w, nrows, ip, lane);
        } else {
#pragma unroll
            for (int i = 0; i < E; ++i) store_nt(out + rI + lane + 64 * i, v[i]);
        }
    }
}

}  // namespace hos
