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

// ---------------------------------------------------------------------------
// wavenumbers: kx from a table (reference or signed), ky = j*Ky0,
// operators_par.c:50-71, 105-116, 168-192.
struct KTab {
    const double *kx;     // [Nx]
    double ky0;
    const double *depth;  // [members]; <= 0: deep water
    int nk;               // Ny/2+1
    int mx, my;           // dealias bounds, operators_par.c:290-291
    __device__ __forceinline__ double h(int m) const { return depth[m]; }
};

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

// XCD-aware tile order: the 8 XCDs take workgroups round-robin, so give
// consecutive logical tiles to workgroups on the same XCD; tiles that share
// 128-byte lines of the transposed intermediate then meet in one L2.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    if ((nb & 7) != 0) return b;
    return (b & 7) * (nb >> 3) + (b >> 3);
}

// ---------------------------------------------------------------------------
// x inverse pass: spectral (FP64 stage input, with multiplier) -> I.  The
// source row is read once into registers (FP64) and every output of the
// group is formed from it.  grid (ceil(Nk/ROWS), ngroups, members)
template <class T, int LOGNX, int ROWS>
__global__ __launch_bounds__(NT) void k_xinv(const cx<double> *__restrict__ Y, long sY, cx<T> *__restrict__ I,
                                             long sI, long fs, XinvOps ops, KTab kt,
                                             const cx<T> *__restrict__ twx)
{
    constexpr int NX = 1 << LOGNX;
    constexpr int PER = (ROWS * NX + NT - 1) / NT;
    __shared__ cx<T> buf[ROWS * NX];
    const int grp = blockIdx.y;
    const int ky0 = blockIdx.x * ROWS;
    const cx<double> *src = Y + blockIdx.z * sY + ops.src[grp] * fs;
    cx<double> s[PER];
    double kxv[PER], kz[PER], th[PER];
    const double h = kt.h(blockIdx.z);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * NT;
        const int r = e / NX, kx = e % NX, ky = ky0 + r;
        s[q] = {0.0, 0.0};
        kxv[q] = 0.0;
        kz[q] = 0.0;
        th[q] = 1.0;
        if (e < ROWS * NX && ky < kt.nk) {
            s[q] = src[(long)ky * NX + kx];
            kxv[q] = kt.kx[kx];
            const double kyv = ky * kt.ky0;
            kz[q] = sqrt(kxv[q] * kxv[q] + kyv * kyv);
            if (h > 0.0) th[q] = tanh(kz[q] * h);  // d^n/dz^n at finite depth: k^n tanh(kh)^(n odd)
        }
    }
    cx<T> *base = I + blockIdx.z * sI;
    for (int o = 0; o < ops.n[grp]; ++o) {
        const XinvOp op = ops.o[grp][o];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = threadIdx.x + q * NT;
            if (e >= ROWS * NX) break;
            const int ky = ky0 + e / NX;
            double re = s[q].x, im = s[q].y;
            if (op.op == OP_DX) {
                re = -kxv[q] * s[q].y;
                im = kxv[q] * s[q].x;
            } else if (op.op == OP_DY) {
                const double k = ky * kt.ky0;
                re = -k * s[q].y;
                im = k * s[q].x;
            } else if (op.op == OP_DZ) {
                double m = 1.0;
                for (int i = 0; i < op.pow; ++i) m *= kz[q];
                if (op.pow & 1) m *= th[q];
                re = m * s[q].x;
                im = m * s[q].y;
            }
            buf[e] = {T(re), T(im)};
        }
        lds_fft<T, LOGNX, ROWS, NT, true>(buf, twx);
        cx<T> *dst = base + op.dst * fs;
        for (int e = threadIdx.x; e < ROWS * NX; e += NT) {
            const int r = e / NX, kx = e % NX, ky = ky0 + r;
            if (ky < kt.nk) dst[(long)ky * NX + kx] = buf[e];
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
// [nx][NY].  buf holds NPR*NY complex values.
template <class T, int LOGNY, int NPR>
__device__ __forceinline__ void c2r_tile(const cx<T> *__restrict__ src, T *__restrict__ dst, int x0, int nx,
                                         cx<T> *buf, const cx<T> *__restrict__ twy)
{
    constexpr int NY = 1 << LOGNY, NK = NY / 2 + 1;
    __syncthreads();  // buf may still be read by a previous tile's store loop
    for (int e = threadIdx.x; e < NK * NPR; e += NT) {
        const int p = e % NPR, ky = e / NPR;
        cx<T> A = src[(long)ky * nx + x0 + 2 * p];
        cx<T> B = src[(long)ky * nx + x0 + 2 * p + 1];
        if (ky == 0 || ky == NY / 2) {
            A.y = T(0);
            B.y = T(0);
        }
        buf[p * NY + ky] = {A.x - B.y, A.y + B.x};
        if (ky > 0 && ky < NY / 2) buf[p * NY + NY - ky] = {A.x + B.y, B.x - A.y};
    }
    lds_fft<T, LOGNY, NPR, NT, true>(buf, twy);
    for (int e = threadIdx.x; e < NPR * NY; e += NT) {
        const int p = e / NY, y = e % NY;
        const cx<T> z = buf[e];
        dst[(long)(x0 + 2 * p) * NY + y] = z.x;
        dst[(long)(x0 + 2 * p + 1) * NY + y] = z.y;
    }
}

template <class T, int LOGNY, int TX>
__global__ __launch_bounds__(NT) void k_yc2r(const cx<T> *__restrict__ I, long sI, long fsI, T *__restrict__ P,
                                             long sP, long fsP, FieldMap fm, int nx,
                                             const cx<T> *__restrict__ twy)
{
    constexpr int NY = 1 << LOGNY, NPR = TX / 2;
    __shared__ cx<T> buf[NPR * NY];
    const int x0 = xcd_tile(blockIdx.x, gridDim.x) * TX;
    c2r_tile<T, LOGNY, NPR>(I + blockIdx.z * sI + fm.src[blockIdx.y] * fsI,
                            P + blockIdx.z * sP + fm.dst[blockIdx.y] * fsP, x0, nx, buf, twy);
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
};

// field indices of the physical fields of one member
struct HosIdx {
    int M;
    int eta, etax, etay, phix, phiy;
    int d[MAXM + 1][MAXM + 1];  // d[a][j] = d^j phi^(a) / dz^j, a=1..M, j=1..M-a+1
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

// plain read of one field
template <class T>
struct ProRead {
    Phys<T> ph;
    int f;
    __device__ __forceinline__ T operator()(int m, int x, int y) const { return ph.at(m, f, x, y); }
};

// phi^(lev+1) = - sum_{k=1}^{lev} eta^k/k! d^k phi^(lev+1-k)/dz^k
// (West et al. 1987; the reference's hphin update, hosm_2d_par.c:302-317)
template <class T>
struct ProPhi {
    Phys<T> ph;
    HosIdx ix;
    int lev;
    bool comp;
    __device__ __forceinline__ T operator()(int m, int x, int y) const
    {
        const T e = ph.at(m, ix.eta, x, y);
        T ek = T(1);
        Sum<T> s(comp);
        for (int k = 1; k <= lev; ++k) {
            ek *= e;
            s.add(ek * T(ix.invfact[k]) * ph.at(m, ix.d[lev + 1 - k][k], x, y));
        }
        return -s.value();
    }
};

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

// nonlinear RHS in physical space (rhs_hos, hosm_2d_par.c:149-206) without
// the two linear terms W^(1) and -g eta, which the x epilogue adds in FP64:
//   N_eta = -(eta_x phi_x + eta_y phi_y) + (W^(M) - W^(1)) + |grad eta|^2 W^(M-2)
//   N_phi = 1/2[(W^2)^(M) + |grad eta|^2 (W^2)^(M-2)] - 1/2 |grad phi|^2
template <class T>
struct ProRhs {
    Phys<T> ph;
    HosIdx ix;
    bool comp;
    __device__ __forceinline__ void operator()(int m, int x, int y, T &a, T &b) const
    {
        const int M = ix.M;
        const T e = ph.at(m, ix.eta, x, y);
        const T ex = ph.at(m, ix.etax, x, y), ey = ph.at(m, ix.etay, x, y);
        const T px = ph.at(m, ix.phix, x, y), py = ph.at(m, ix.phiy, x, y);
        T W[MAXM + 1];
        w_orders<T>(ph, ix, m, x, y, e, W, comp);
        const T G = ex * ex + ey * ey;
        const T wM2 = (M - 2 >= 1) ? w_upto<T>(W, 1, M - 2, comp) : T(0);
        const T wMnl = w_upto<T>(W, 2, M, comp);
        a = -(ex * px + ey * py) + wMnl + G * wM2;
        const T w2M = w2_upto<T>(W, M, comp);
        const T w2M2 = w2_upto<T>(W, M - 2, comp);
        b = T(0.5) * (w2M + G * w2M2) - T(0.5) * (px * px + py * py);
    }
};

// Zvel outputs: which=0 -> (W^(M), W^(M-2)), which=1 -> ((W^2)^(M), (W^2)^(M-2))
template <class T>
struct ProZvel {
    Phys<T> ph;
    HosIdx ix;
    int which;
    bool comp;
    __device__ __forceinline__ void operator()(int m, int x, int y, T &a, T &b) const
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
};

// ---------------------------------------------------------------------------
// r2c along y of x-rows x0 .. x0+2*NPR-1 of the field the prologue computes
// on the fly: rows x0+2p and x0+2p+1 are the real and imaginary part of one
// complex FFT; the half spectra are written transposed into out [Nk][nx].
template <class T, int LOGNY, int NPR, class PRO>
__device__ __forceinline__ void r2c_tile(const PRO &pro, int m, cx<T> *__restrict__ out, int x0, int nx,
                                         cx<T> *buf, const cx<T> *__restrict__ twy)
{
    constexpr int NY = 1 << LOGNY, NK = NY / 2 + 1, TX = 2 * NPR;
    __syncthreads();
    // compile-time trip count, unrolled: the loads of several points are in
    // flight together (the prologue reads up to M+1 fields per point)
    constexpr int PER = (NPR * NY + NT - 1) / NT;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * NT;
        if (e < NPR * NY) {
            const int p = e / NY, y = e % NY;
            buf[e] = {pro(m, x0 + 2 * p, y), pro(m, x0 + 2 * p + 1, y)};
        }
    }
    lds_fft<T, LOGNY, NPR, NT, false>(buf, twy);
    for (int e = threadIdx.x; e < NK * TX; e += NT) {
        const int t = e % TX, ky = e / TX, p = t >> 1;
        const cx<T> zk = buf[p * NY + ky];
        const cx<T> zm = buf[p * NY + ((NY - ky) & (NY - 1))];
        cx<T> v;
        if ((t & 1) == 0)
            v = {T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
        else
            v = {T(0.5) * (zk.y + zm.y), T(-0.5) * (zk.x - zm.x)};
        out[(long)ky * nx + x0 + t] = v;
    }
}

// r2c along y of x-rows x0 .. x0+NROW-1 of two fields the prologue computes
// together (the real and imaginary part of one complex FFT per x-row).
template <class T, int LOGNY, int NROW, class PRO>
__device__ __forceinline__ void r2c_pair_tile(const PRO &pro, int m, cx<T> *__restrict__ oa,
                                              cx<T> *__restrict__ ob, int x0, int nx, cx<T> *buf,
                                              const cx<T> *__restrict__ twy)
{
    constexpr int NY = 1 << LOGNY, NK = NY / 2 + 1;
    __syncthreads();
    constexpr int PER = (NROW * NY + NT - 1) / NT;
#pragma unroll 2
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * NT;
        if (e < NROW * NY) {
            const int p = e / NY, y = e % NY;
            T a, b;
            pro(m, x0 + p, y, a, b);
            buf[e] = {a, b};
        }
    }
    lds_fft<T, LOGNY, NROW, NT, false>(buf, twy);
    for (int e = threadIdx.x; e < NK * NROW * 2; e += NT) {
        const int t = e % (2 * NROW), ky = e / (2 * NROW);
        const int p = t % NROW, which = t / NROW;
        const cx<T> zk = buf[p * NY + ky];
        const cx<T> zm = buf[p * NY + ((NY - ky) & (NY - 1))];
        if (which == 0)
            oa[(long)ky * nx + x0 + p] = {T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
        else
            ob[(long)ky * nx + x0 + p] = {T(0.5) * (zk.y + zm.y), T(-0.5) * (zk.x - zm.x)};
    }
}

// y r2c pass, one output field.  grid (Nx/TX, 1, members)
template <class T, int LOGNY, int TX, class PRO>
__global__ __launch_bounds__(NT) void k_yr2c(PRO pro, cx<T> *__restrict__ I, long sI, long fsI, int dst, int nx,
                                             const cx<T> *__restrict__ twy)
{
    constexpr int NY = 1 << LOGNY, NPR = TX / 2;
    __shared__ cx<T> buf[NPR * NY];
    const int x0 = xcd_tile(blockIdx.x, gridDim.x) * TX;
    const int m = blockIdx.z;
    r2c_tile<T, LOGNY, NPR, PRO>(pro, m, I + m * sI + dst * fsI, x0, nx, buf, twy);
}

// y r2c pass, two output fields at the same x.  grid (Nx/TXP, 1, members)
template <class T, int LOGNY, int TXP, class PRO>
__global__ __launch_bounds__(NT) void k_yr2c_pair(PRO pro, cx<T> *__restrict__ I, long sI, long fsI, int dstA,
                                                  int dstB, int nx, const cx<T> *__restrict__ twy)
{
    constexpr int NY = 1 << LOGNY;
    __shared__ cx<T> buf[TXP * NY];
    const int x0 = xcd_tile(blockIdx.x, gridDim.x) * TXP;
    const int m = blockIdx.z;
    r2c_pair_tile<T, LOGNY, TXP, PRO>(pro, m, I + m * sI + dstA * fsI, I + m * sI + dstB * fsI, x0, nx, buf, twy);
}

// ---------------------------------------------------------------------------
// x forward pass of phi^(lev+1), then d^j/dz^j for j = 1..ndst and the x
// inverse of each: the spectral phi^(lev+1) never leaves LDS.
// grid (ceil(Nk/ROWS), 1, members)
template <class T, int LOGNX, int ROWS>
__global__ __launch_bounds__(NT) void k_xfwdinv(cx<T> *__restrict__ I, long sI, long fs, int src, int dst0, int ndst,
                                                double scale, KTab kt, const cx<T> *__restrict__ twx)
{
    constexpr int NX = 1 << LOGNX;
    constexpr int PER = (ROWS * NX + NT - 1) / NT;
    __shared__ cx<T> S[ROWS * NX];
    __shared__ cx<T> W[ROWS * NX];
    const int ky0 = blockIdx.x * ROWS;
    cx<T> *base = I + blockIdx.z * sI;
    const cx<T> *in = base + src * fs;
    for (int e = threadIdx.x; e < ROWS * NX; e += NT) {
        const int r = e / NX, kx = e % NX, ky = ky0 + r;
        S[e] = (ky < kt.nk) ? in[(long)ky * NX + kx] : cx<T>{T(0), T(0)};
    }
    // |k| and tanh(|k| h) of this thread's modes, FP64
    double kz[PER], th[PER];
    const double h = kt.h(blockIdx.z);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * NT;
        const int kx = e % NX, ky = ky0 + e / NX;
        const double kxv = kt.kx[kx], kyv = ky * kt.ky0;
        kz[q] = sqrt(kxv * kxv + kyv * kyv);
        th[q] = h > 0.0 ? tanh(kz[q] * h) : 1.0;
    }
    lds_fft<T, LOGNX, ROWS, NT, false>(S, twx);
    double kj[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) kj[q] = scale;
    for (int j = 1; j <= ndst; ++j) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = threadIdx.x + q * NT;
            if (e >= ROWS * NX) break;
            kj[q] *= kz[q];  // scale * |k|^j
            const T mul = T((j & 1) ? kj[q] * th[q] : kj[q]);
            const cx<T> s = S[e];
            W[e] = {mul * s.x, mul * s.y};
        }
        lds_fft<T, LOGNX, ROWS, NT, true>(W, twx);
        cx<T> *out = base + (dst0 + j - 1) * fs;
        for (int e = threadIdx.x; e < ROWS * NX; e += NT) {
            const int r = e / NX, kx = e % NX, ky = ky0 + r;
            if (ky < kt.nk) out[(long)ky * NX + kx] = W[e];
        }
    }
}

// ---------------------------------------------------------------------------
// final x forward pass: N_eta, N_phi -> K = RHS (linear terms added in FP64,
// ramp, dealias) -> RK stage update.
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
};

template <class T, int LOGNX, int ROWS>
__global__ __launch_bounds__(NT) void k_xfwd_rhs(const cx<T> *__restrict__ I, long sI, long fs, int iA, int iB,
                                                 const cx<double> *Y, long sY, double scale, KTab kt,
                                                 RkArgs rk, const cx<T> *__restrict__ twx)
{
    constexpr int NX = 1 << LOGNX;
    __shared__ cx<T> S[2 * ROWS * NX];
    __shared__ double red[NT / 64];
    const int ky0 = blockIdx.x * ROWS;
    const int m = blockIdx.z;
    const cx<T> *ia = I + m * sI + iA * fs;
    const cx<T> *ib = I + m * sI + iB * fs;
    for (int e = threadIdx.x; e < ROWS * NX; e += NT) {
        const int r = e / NX, kx = e % NX, ky = ky0 + r;
        const bool ok = ky < kt.nk;
        S[e] = ok ? ia[(long)ky * NX + kx] : cx<T>{T(0), T(0)};
        S[ROWS * NX + e] = ok ? ib[(long)ky * NX + kx] : cx<T>{T(0), T(0)};
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
    for (int e = threadIdx.x; e < ROWS * NX; e += NT) {
        const int r = e / NX, kx = e % NX, ky = ky0 + r;
        if (ky >= kt.nk) continue;
        const long idx = (long)ky * NX + kx;
        const cx<T> na = S[e], nb = S[ROWS * NX + e];
        if (rk.mode == EPI_STORE_SPEC2) {
            cx<double> *k = rk.K + m * rk.sK;
            k[idx] = {scale * double(na.x), scale * double(na.y)};
            k[fsd + idx] = {scale * double(nb.x), scale * double(nb.y)};
            continue;
        }
        const cx<double> ye = y[idx], yp = y[fsd + idx];
        const double kxv = kt.kx[kx], kyv = ky * kt.ky0;
        const double w1 = dz_mul(sqrt(kxv * kxv + kyv * kyv), 1, kt.h(m));
        // eta_t = W^(1) + R*(rest), phi_t = -g eta + R*(rest)  (hosm_2d_par.c:211-229)
        cx<double> ke = {w1 * yp.x + R * scale * double(na.x), w1 * yp.y + R * scale * double(na.y)};
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
            continue;
        }
        // EPI_RK_SUBDIAG: Y_{s+1} = u + dt b[s+1][s] K_s ; acc += dt c[s] K_s
        cx<double> *u = rk.U + m * 2 * fsd;
        cx<double> *acc = rk.A + m * 2 * fsd;
        cx<double> *yn = rk.Yn + m * 2 * fsd;
        const cx<double> ue = u[idx], up = u[fsd + idx];
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
            ae = acc[idx];
            ap = acc[fsd + idx];
        }
        ae = {ae.x + fc * ke.x, ae.y + fc * ke.y};
        ap = {ap.x + fc * kp.x, ap.y + fc * kp.y};
        if (rk.stage == rk.nstages - 1) {
            if (rk.mean_level && idx == 0) ae = {rk.mean0[m], 0.0};
            u[idx] = ae;
            u[fsd + idx] = ap;
            yn[idx] = ae;
            yn[fsd + idx] = ap;
        } else {
            acc[idx] = ae;
            acc[fsd + idx] = ap;
            const double fb = dt * rk.b_next;
            yn[idx] = {ue.x + fb * ke.x, ue.y + fb * ke.y};
            yn[fsd + idx] = {up.x + fb * kp.x, up.y + fb * kp.y};
        }
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
// end of step: fixed-order sum of the Hamiltonian partials; t += dt.
// grid (members), block NT
__global__ __launch_bounds__(NT) void k_end_step(const double *__restrict__ hpart, int nparts, double *H,
                                                 double *t, const double *dt, int advance)
{
    __shared__ double red[NT / 64];
    const int m = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += NT) s += hpart[m * nparts + i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int w = 0; w < NT / 64; ++w) tot += red[w];
        H[m] = tot;
        if (advance) t[m] += dt[m];
    }
}

// generic Butcher combination for schemes with a full lower triangle:
// out = u + dt * sum_p coef[p] K_p   (time_schemes_par.c:38-67, 98-116)
struct CombArgs {
    double coef[8];
    int np;
};
__global__ __launch_bounds__(NT) void k_combine(const cx<double> *U, const cx<double> *__restrict__ K,
                                                long sK_stage, long sM, cx<double> *out, long n, const double *dt,
                                                CombArgs ca)
{
    const int m = blockIdx.y;
    const double d = dt[m];
    for (long i = blockIdx.x * (long)NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
        cx<double> v = U[m * sM + i];
        for (int p = 0; p < ca.np; ++p) {
            const cx<double> k = K[p * sK_stage + m * sM + i];
            v.x += d * ca.coef[p] * k.x;
            v.y += d * ca.coef[p] * k.y;
        }
        out[m * sM + i] = v;
    }
}

// pointwise spectral operators on FP64 fields [Nk][Nx]
// (operators_par.c: Dx 50-71, Dy 105-116, Dz 168-192, Sum 266-276, Dealias 309-322)
enum : int { SOP_DX = 1, SOP_DY = 2, SOP_DZ = 3, SOP_SUM = 4, SOP_DEALIAS = 5 };
__global__ __launch_bounds__(NT) void k_spec_op(const cx<double> *__restrict__ a, const cx<double> *__restrict__ b,
                                                cx<double> *__restrict__ out, int op, int nx, KTab kt)
{
    const long n = (long)kt.nk * nx;
    for (long i = blockIdx.x * (long)NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
        const int kx = (int)(i % nx), ky = (int)(i / nx);
        const cx<double> s = a[i];
        cx<double> r = s;
        if (op == SOP_DX) {
            const double k = kt.kx[kx];
            r = {-k * s.y, k * s.x};
        } else if (op == SOP_DY) {
            const double k = ky * kt.ky0;
            r = {-k * s.y, k * s.x};
        } else if (op == SOP_DZ) {
            const double kxv = kt.kx[kx], kyv = ky * kt.ky0;
            const double m = dz_mul(sqrt(kxv * kxv + kyv * kyv), 1, kt.h(0));
            r = {m * s.x, m * s.y};
        } else if (op == SOP_SUM) {
            r = {s.x + b[i].x, s.y + b[i].y};
        } else if (op == SOP_DEALIAS) {
            if (masked(kx, ky, nx, kt.mx, kt.my)) r = {0.0, 0.0};
        }
        out[i] = r;
    }
}

// Hamiltonian of three FP64 spectral fields (hosm_2d_par.c:462-523):
// partial sums per block, row j=0 weight 1, others 2.
__global__ __launch_bounds__(NT) void k_hamiltonian(const cx<double> *__restrict__ eta,
                                                    const cx<double> *__restrict__ etat,
                                                    const cx<double> *__restrict__ phi, double g, int nx, int nk,
                                                    double *part)
{
    __shared__ double red[NT / 64];
    const long n = (long)nk * nx;
    double h = 0.0;
    for (long i = blockIdx.x * (long)NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
        const double w = (i < nx) ? 1.0 : 2.0;
        const cx<double> e = eta[i], t = etat[i], p = phi[i];
        h += w * (g * (e.x * e.x + e.y * e.y) + p.x * t.x + p.y * t.y);
    }
    for (int off = 32; off > 0; off >>= 1) h += __shfl_down(h, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = h;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < NT / 64; ++w) s += red[w];
        part[blockIdx.x] = s;
    }
}

// product of two physical fields (the pointwise step of Mult, operators_par.c:222-231)
template <class T>
struct ProMul {
    Phys<T> ph;
    int fa, fb;
    __device__ __forceinline__ T operator()(int m, int x, int y) const
    {
        return ph.at(m, fa, x, y) * ph.at(m, fb, x, y);
    }
};

// FP64 spectral field [Nk][Nx] (row stride nx) copy, optionally from a
// different source; used for state set-up on the device.
__global__ __launch_bounds__(NT) void k_copy_c64(const cx<double> *__restrict__ src, cx<double> *__restrict__ dst,
                                                 long n)
{
    for (long i = blockIdx.x * (long)NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) dst[i] = src[i];
}

// real [Nx][Ny] (T) <-> padded double host layout helpers
template <class T>
__global__ __launch_bounds__(NT) void k_real_from_f64(const double *__restrict__ src, T *__restrict__ dst, long n)
{
    for (long i = blockIdx.x * (long)NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) dst[i] = T(src[i]);
}
template <class T>
__global__ __launch_bounds__(NT) void k_real_to_f64(const T *__restrict__ src, double *__restrict__ dst, long n)
{
    for (long i = blockIdx.x * (long)NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) dst[i] = double(src[i]);
}

}  // namespace hos
