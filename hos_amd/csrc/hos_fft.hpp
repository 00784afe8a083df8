// hos_fft.hpp -- device building blocks of the batched 2-D real FFTs (gfx950).
//
// A row FFT of length N = 2^LOGN runs in LDS as a Stockham autosort sequence
// of radix-4 stages (one radix-2 stage for odd LOGN; radix-2 only for N = 8),
// all threads of the workgroup cooperating on NROWS rows at once.  Each stage
// reads its R inputs into registers, barriers, and writes R outputs in place,
// so one buffer of NROWS*N complex values is all the LDS a transform needs.
// Twiddles come from a per-length table tw[k] = exp(-2 pi i k/N) computed on
// the host in long double; the inverse uses the conjugate.  Transforms are
// unnormalised (the 1/(Nx*Ny) of fft_2d is applied by the r2c x-pass).
#pragma once
#include <hip/hip_runtime.h>

namespace hos {

template <class T>
struct alignas(2 * sizeof(T)) cx {
    T x, y;
};

template <class T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }

// stage plan: radix of stage s for length 2^logn.  The first and last stage
// always share a radix.
__host__ __device__ constexpr int fft_nstages(int logn)
{
    return logn == 0 ? 0 : logn == 1 ? 1 : logn == 3 ? 3 : (logn % 2 == 0 ? logn / 2 : (logn - 1) / 2 + 1);
}
__host__ __device__ constexpr int fft_radix(int logn, int s)
{
    return logn == 1 ? 2 : logn == 3 ? 2 : (logn % 2 == 0 ? 4 : (s == 1 ? 2 : 4));
}

template <class T, int R, bool INV>
__device__ __forceinline__ void dft_small(cx<T> *v)
{
    if constexpr (R == 2) {
        cx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else {
        static_assert(R == 4, "radix");
        cx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        cx<T> t2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        // forward: t3 = -i*d ; inverse: t3 = +i*d
        cx<T> t3 = INV ? cx<T>{-d.y, d.x} : cx<T>{d.y, -d.x};
        v[0] = cadd(t0, t2);
        v[1] = cadd(t1, t3);
        v[2] = csub(t0, t2);
        v[3] = csub(t1, t3);
    }
}

template <class T, int R, int LOGN, int NROWS, int NT, bool INV, int NS>
__device__ __forceinline__ void fft_stage(cx<T> *buf, const cx<T> *__restrict__ tw)
{
    constexpr int N = 1 << LOGN;
    constexpr int NR = N / R;
    constexpr int NBF = NROWS * NR;
    constexpr int Q = (NBF + NT - 1) / NT;
    cx<T> v[Q][R];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int bf = threadIdx.x + q * NT;
        if ((NBF % NT) == 0 || bf < NBF) {
            const int row = bf / NR, j = bf % NR;
            const cx<T> *src = buf + row * N;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = src[j + r * NR];
            if constexpr (NS > 1) {
                // w^r for r = 1..R-1 from one table entry w = exp(-+2 pi i k/(NS R))
                const int k = j % NS;
                cx<T> w = tw[k * (N / (NS * R))];
                if (INV) w.y = -w.y;
                cx<T> wr = w;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    v[q][r] = cmul(v[q][r], wr);
                    if (r + 1 < R) wr = cmul(wr, w);
                }
            }
            dft_small<T, R, INV>(v[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int bf = threadIdx.x + q * NT;
        if ((NBF % NT) == 0 || bf < NBF) {
            const int row = bf / NR, j = bf % NR;
            const int idxD = (j / NS) * NS * R + (j % NS);
            cx<T> *dst = buf + row * N;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[idxD + r * NS] = v[q][r];
        }
    }
    __syncthreads();
}

template <class T, int LOGN, int NROWS, int NT, bool INV, int S = 0, int NS = 1>
__device__ __forceinline__ void fft_stages(cx<T> *buf, const cx<T> *__restrict__ tw)
{
    if constexpr (S < fft_nstages(LOGN)) {
        constexpr int R = fft_radix(LOGN, S);
        fft_stage<T, R, LOGN, NROWS, NT, INV, NS>(buf, tw);
        fft_stages<T, LOGN, NROWS, NT, INV, S + 1, NS * R>(buf, tw);
    }
}

// In-place FFT of NROWS contiguous rows of length 2^LOGN held in LDS.  Starts
// and ends with a workgroup barrier, so callers may write buf just before and
// read it just after.
template <class T, int LOGN, int NROWS, int NT, bool INV>
__device__ __forceinline__ void lds_fft(cx<T> *buf, const cx<T> *__restrict__ tw)
{
    __syncthreads();
    fft_stages<T, LOGN, NROWS, NT, INV>(buf, tw);
}

}  // namespace hos
