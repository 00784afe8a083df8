// hos_fft.hpp -- device building blocks of the batched 2-D real FFTs (gfx950).
//
// A row FFT of length N = 2^LOGN runs in LDS as a Stockham autosort sequence
// of radix-16/8/4/2 stages (fft_stage_bits: 1024 = 16*16*4 when enough rows
// share the workgroup, 8*8*4*4 for one row), all threads of the workgroup
// cooperating on NROWS rows at once.  Each stage reads its R inputs into
// registers, does the R-point DFT there (radix 16 = 4 x 4), barriers, and
// writes R outputs in place, so one buffer of NROWS rows (Lds<T, N> layout,
// padded against bank conflicts) is all the LDS a transform needs and each
// stage is one LDS round trip.
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

// LDS layout of rows of N complex values: one pad element after every 16, so
// the strided Stockham writes (stride R = 16 in the first stage) land in
// distinct banks, and a row pitch of N + N/16 plus 128 bytes.
template <class T, int N>
struct Lds {
    static_assert(N >= 16, "rows of at least 16");
    static constexpr int SH = 128 / (int)sizeof(cx<T>);
    static constexpr int PITCH = N + N / 16 + SH;
    static constexpr int size(int rows) { return rows * PITCH; }
    __device__ static __forceinline__ int at(int row, int i) { return row * PITCH + i + (i >> 4); }
};

// Stage plan for NROWS rows of length 2^logn on nt threads: radix 16, 8, 4 or
// 2 per stage, the largest radix that keeps at least min(nt/2, NROWS*N/4)
// butterflies per stage (so most of the workgroup stays busy) and does not
// leave a single radix-2 bit behind.  Fewer stages = fewer LDS round trips.
// Above 256 threads the bound is min(nt/2, NROWS*N/8), and the first stage,
// which has no twiddles, needs half of it (512 threads, one to three rows of
// 1024: 16 8 8).  That is plan 0; the autotuner may pick another family for
// the RHS pass:
//   plan 1: radices of at most 8 under the same rule (1024: 8 8 4 4);
//   plan 2: radix 4 only (1024: 4 4 4 4 4);
//   plan 3: the largest radix at every stage, no occupancy rule (16 16 4).
__host__ __device__ constexpr int fft_stage_bits(int logn, int nrows, int s, int nt = 256, int plan = 0)
{
    const int cap = plan == 1 ? 3 : (plan == 2 ? 2 : 4);
    const long total = (long)nrows << logn;
    const long part = nt > 256 ? total / 8 : total / 4;
    const long thr1 = plan == 3 ? 0 : (part < nt / 2 ? part : nt / 2);
    int rem = logn, bits = 0;
    for (int i = 0; i <= s; ++i) {
        const long thr = (i == 0 && nt > 256) ? thr1 / 2 : thr1;
        int b = rem == 1 ? 1 : 2;
        if (cap >= 4 && 4 <= rem && rem - 4 != 1 && (total >> 4) >= thr)
            b = 4;
        else if (cap >= 3 && 3 <= rem && rem - 3 != 1 && (total >> 3) >= thr)
            b = 3;
        if ((rem == 2 || rem == 3) && b < rem && rem <= cap && (total >> rem) >= thr) b = rem;
        bits = b;
        rem -= b;
        if (rem <= 0 && i < s) return 0;
    }
    return bits;
}
__host__ __device__ constexpr int fft_nstages(int logn, int nrows, int nt = 256, int plan = 0)
{
    int n = 0;
    while (fft_stage_bits(logn, nrows, n, nt, plan) > 0) ++n;
    return n;
}
// the stage radices of a plan as a number (16 8 8 -> 1688), for tests and notes
__host__ __device__ constexpr long fft_plan_code(int logn, int nrows, int nt, int plan)
{
    long c = 0;
    for (int s = 0; s < fft_nstages(logn, nrows, nt, plan); ++s) {
        const int r = 1 << fft_stage_bits(logn, nrows, s, nt, plan);
        c = c * (r >= 10 ? 100 : 10) + r;
    }
    return c;
}
static_assert(fft_plan_code(10, 3, 512, 0) == 1688, "plan 0 at 1024, three rows, 512 threads");
static_assert(fft_plan_code(10, 3, 512, 1) == 8844, "plan 1");
static_assert(fft_plan_code(10, 3, 512, 2) == 44444, "plan 2");
static_assert(fft_plan_code(10, 3, 512, 3) == 16164, "plan 3");
static_assert(fft_plan_code(10, 2, 512, 0) == 1688, "plan 0, two rows");

// e^{-2 pi i m/16}, m = 0..15 (cos and sin of multiples of pi/8)
template <class T>
__device__ __forceinline__ cx<T> w16(int m)
{
    constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, h = 0.70710678118654752440;
    constexpr double re[16] = {1, c1, h, s1, 0, -s1, -h, -c1, -1, -c1, -h, -s1, 0, s1, h, c1};
    constexpr double im[16] = {0, -s1, -h, -c1, -1, -c1, -h, -s1, 0, s1, h, c1, 1, c1, h, s1};
    return {T(re[m & 15]), T(im[m & 15])};
}

template <class T, bool INV>
__device__ __forceinline__ void dft4(cx<T> &a0, cx<T> &a1, cx<T> &a2, cx<T> &a3)
{
    const cx<T> t0 = cadd(a0, a2), t1 = csub(a0, a2);
    const cx<T> t2 = cadd(a1, a3), d = csub(a1, a3);
    // forward: t3 = -i*d ; inverse: t3 = +i*d
    const cx<T> t3 = INV ? cx<T>{-d.y, d.x} : cx<T>{d.y, -d.x};
    a0 = cadd(t0, t2);
    a1 = cadd(t1, t3);
    a2 = csub(t0, t2);
    a3 = csub(t1, t3);
}

template <class T, bool INV>
__device__ __forceinline__ cx<T> tw16(cx<T> v, int m)
{
    cx<T> w = w16<T>(m);
    if (INV) w.y = -w.y;
    return cmul(v, w);
}

// in-register DFT of size R, natural order in and out
template <class T, int R, bool INV>
__device__ __forceinline__ void dft_small(cx<T> *v)
{
    if constexpr (R == 2) {
        const cx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        dft4<T, INV>(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 8) {
        // even / odd halves, then X[k] = E[k] + w8^k O[k], X[k+4] = E[k] - w8^k O[k]
        cx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
        cx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
        dft4<T, INV>(e0, e1, e2, e3);
        dft4<T, INV>(o0, o1, o2, o3);
        o1 = tw16<T, INV>(o1, 2);
        o2 = tw16<T, INV>(o2, 4);
        o3 = tw16<T, INV>(o3, 6);
        v[0] = cadd(e0, o0);
        v[4] = csub(e0, o0);
        v[1] = cadd(e1, o1);
        v[5] = csub(e1, o1);
        v[2] = cadd(e2, o2);
        v[6] = csub(e2, o2);
        v[3] = cadd(e3, o3);
        v[7] = csub(e3, o3);
    } else {
        static_assert(R == 16, "radix");
        // 16 = 4 x 4: DFT4 over n1 for each n2, twiddle w16^(n2 k1), DFT4 over n2
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) dft4<T, INV>(v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]);
        // v[n2 + 4 k1] now holds A[n2][k1]
#pragma unroll
        for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
            for (int k1 = 1; k1 < 4; ++k1) v[n2 + 4 * k1] = tw16<T, INV>(v[n2 + 4 * k1], n2 * k1);
        cx<T> o[16];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            cx<T> b0 = v[4 * k1], b1 = v[1 + 4 * k1], b2 = v[2 + 4 * k1], b3 = v[3 + 4 * k1];
            dft4<T, INV>(b0, b1, b2, b3);
            o[k1] = b0;
            o[k1 + 4] = b1;
            o[k1 + 8] = b2;
            o[k1 + 12] = b3;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = o[i];
    }
}

// Stockham twiddles of butterfly j of a stage (NS = product of the earlier
// radices): input r is multiplied by w^r, w = exp(-+2 pi i (j mod NS)/(NS R)).
// Powers below 4 by repeated products of w, higher ones as (w^4)^q * w^b with
// w^4 a second table entry, so no power is more than three products away from
// a table value.
template <class T, int R, int N, int NS, bool INV>
__device__ __forceinline__ void stage_twiddles(cx<T> *v, int j, const cx<T> *__restrict__ tw)
{
    if constexpr (NS > 1) {
        const int k = j % NS;
        constexpr int STRIDE = N / (NS * R);
        cx<T> w1 = tw[k * STRIDE];
        if (INV) w1.y = -w1.y;
        cx<T> wb[4];
        wb[0] = {T(1), T(0)};
        wb[1] = w1;
        if constexpr (R > 2) {
            wb[2] = cmul(w1, w1);
            wb[3] = cmul(wb[2], w1);
        }
        cx<T> wq[4];
        wq[0] = {T(1), T(0)};
        if constexpr (R > 4) {
            cx<T> w4 = tw[4 * k * STRIDE];
            if (INV) w4.y = -w4.y;
            wq[1] = w4;
            if constexpr (R > 8) {
                wq[2] = cmul(w4, w4);
                wq[3] = cmul(wq[2], w4);
            }
        }
#pragma unroll
        for (int r = 1; r < R; ++r) {
            const cx<T> w = (r >> 2) == 0 ? wb[r & 3] : ((r & 3) == 0 ? wq[r >> 2] : cmul(wq[r >> 2], wb[r & 3]));
            v[r] = cmul(v[r], w);
        }
    }
}

template <class T, int R, int LOGN, int NROWS, int NT, bool INV, int NS>
__device__ __forceinline__ void fft_stage(cx<T> *buf, const cx<T> *__restrict__ tw, int tid)
{
    constexpr int N = 1 << LOGN;
    constexpr int NR = N / R;
    constexpr int NBF = NROWS * NR;
    constexpr int Q = (NBF + NT - 1) / NT;
    using L = Lds<T, N>;
    cx<T> v[Q][R];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int bf = tid + q * NT;
        if ((NBF % NT) == 0 || bf < NBF) {
            const int row = bf / NR, j = bf % NR;
            // L::at(row, j + r NR) as a per-thread base plus a constant per r:
            // NR is a power of two and j < NR, so (j + r NR) >> 4 equals
            // (j >> 4) + ((r NR) >> 4)
            const int rb = row * L::PITCH + j + (j >> 4);
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = buf[rb + r * NR + ((r * NR) >> 4)];
            stage_twiddles<T, R, N, NS, INV>(v[q], j, tw);
            dft_small<T, R, INV>(v[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int bf = tid + q * NT;
        if ((NBF % NT) == 0 || bf < NBF) {
            const int row = bf / NR, j = bf % NR;
            const int idxD = (j / NS) * NS * R + (j % NS);
            // likewise (idxD + r NS) >> 4 = (idxD >> 4) + ((r NS) >> 4): idxD is a
            // multiple of NS R plus j % NS < NS
            const int wb = row * L::PITCH + idxD + (idxD >> 4);
#pragma unroll
            for (int r = 0; r < R; ++r) buf[wb + r * NS + ((r * NS) >> 4)] = v[q][r];
        }
    }
    __syncthreads();
}

template <class T, int LOGN, int NROWS, int NT, bool INV, int PLAN = 0, int S = 0, int NS = 1>
__device__ __forceinline__ void fft_stages(cx<T> *buf, const cx<T> *__restrict__ tw, int tid)
{
    if constexpr (S < fft_nstages(LOGN, NROWS, NT, PLAN)) {
        constexpr int R = 1 << fft_stage_bits(LOGN, NROWS, S, NT, PLAN);
        fft_stage<T, R, LOGN, NROWS, NT, INV, NS>(buf, tw, tid);
        fft_stages<T, LOGN, NROWS, NT, INV, PLAN, S + 1, NS * R>(buf, tw, tid);
    }
}

// threadIdx.x through an opaque copy: what a caller derives from it (LDS and
// global addresses) is recomputed at each use, a few integer operations,
// instead of being shared across the kernel and held in registers throughout.
// The copy keeps its range, 0 <= tid < NT (the workgroup size every caller's
// plan is built for), so divisions and remainders by tile lengths fold: 9%
// fewer instructions in the RHS pass, whose LDS FFT phases are issue-bound
template <int NT = 1024>
__device__ __forceinline__ int opaque_tid()
{
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    __builtin_assume(tid >= 0);
    __builtin_assume(tid < NT);
    return tid;
}

// In-place FFT of NROWS rows of length 2^LOGN held in LDS in the Lds<T, N>
// layout.  Starts and ends with a workgroup barrier, so callers may write buf
// just before and read it just after.  tw: e^{-2 pi i m/N}; every plan reads
// m < N/2 only (stage twiddle indices stay below N/R_last, or 4 N/R_last for a
// last radix of 8 or 16)
template <class T, int LOGN, int NROWS, int NT, bool INV, int PLAN = 0>
__device__ __forceinline__ void lds_fft(cx<T> *buf, const cx<T> *__restrict__ tw)
{
    const int tid = opaque_tid<NT>();  // stage addresses recomputed per call
    __syncthreads();
    fft_stages<T, LOGN, NROWS, NT, INV, PLAN>(buf, tw, tid);
}

// The twiddles an lds_fft reads (the first N/2 of the table) staged in LDS.
// From global memory, the twiddle loads of the stages wait behind whatever
// the workgroup still has in flight (tile prefetches, stores): loads complete
// in order.  load() before the tile loads, store() once they are waited for
// (with the tile's LDS fill); the FFT's opening barrier publishes the copy.
template <class T, int N, int NTT>
struct TwLds {
    static constexpr int NTW = N / 2, P = (NTW + NTT - 1) / NTT;
    cx<T> v[P];
    __device__ __forceinline__ void load(const cx<T> *__restrict__ tw)
    {
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = tw[min((int)threadIdx.x + q * NTT, NTW - 1)];
    }
    __device__ __forceinline__ void store(cx<T> *l) const
    {
#pragma unroll
        for (int q = 0; q < P; ++q)
            if ((NTW % NTT) == 0 || (int)threadIdx.x + q * NTT < NTW) l[threadIdx.x + q * NTT] = v[q];
    }
};

// ---------------------------------------------------------------------------
// Wave FFT: one 64-lane wave transforms one row of N = 64 E complex values
// held in registers, lane l owning elements l + 64 m (m = 0..E-1), in natural
// order on input and on output.  Stockham stages of radix 16/8/4 in which
// every stage has N/R >= 64 butterflies (one or more per lane): the first
// stage reads its inputs straight from the registers (butterfly j = l + 64 q
// needs elements j + r N/R, all owned by lane l), the last one writes its
// outputs back to them (Stockham's last stage stores element j + r N/R), and
// the stages in between exchange through the wave's own padded LDS row.  Only
// the wave's own lanes touch that row and a wave's LDS operations complete in
// issue order, so no workgroup barrier is needed, only compiler ordering.
__host__ __device__ constexpr int wave_stage_bits(int logn, int s)
{
    // 512 = 8 8 8, 1024 = 16 16 4, 2048 = 16 16 8
    return logn == 9 ? (s < 3 ? 3 : 0)
         : logn == 10 ? (s < 2 ? 4 : (s == 2 ? 2 : 0))
         : logn == 11 ? (s < 2 ? 4 : (s == 2 ? 3 : 0)) : 0;
}
__host__ __device__ constexpr int wave_nstages(int logn)
{
    int n = 0;
    while (wave_stage_bits(logn, n) > 0) ++n;
    return n;
}

// orders the LDS accesses of the wave's lanes on either side (no instruction
// is needed: the wave's LDS operations complete in order)
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T, int LOGN, bool INV, int S, int NS>
__device__ __forceinline__ void wave_stage(cx<T> *v, cx<T> *row, const cx<T> *__restrict__ tw, int lane)
{
    constexpr int N = 1 << LOGN, R = 1 << wave_stage_bits(LOGN, S), NR = N / R, B = NR / 64;
    constexpr bool FIRST = S == 0, LAST = S + 1 == wave_nstages(LOGN);
    static_assert(NR % 64 == 0, "a butterfly per lane at least");
    cx<T> u[B][R];
#pragma unroll
    for (int q = 0; q < B; ++q) {
        const int j = lane + 64 * q;
        const int rb = j + (j >> 4);  // Lds<T, N>::at(0, j + r NR), as in fft_stage
#pragma unroll
        for (int r = 0; r < R; ++r) u[q][r] = FIRST ? v[q + r * B] : row[rb + r * NR + ((r * NR) >> 4)];
    }
    if constexpr (!FIRST) wave_sync();  // every lane has read before any lane writes
#pragma unroll
    for (int q = 0; q < B; ++q) {
        const int j = lane + 64 * q;
        stage_twiddles<T, R, N, NS, INV>(u[q], j, tw);
        dft_small<T, R, INV>(u[q]);
        if constexpr (LAST) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[q + r * B] = u[q][r];
        } else {
            const int idxD = (j / NS) * NS * R + (j % NS);
            const int wb = idxD + (idxD >> 4);
#pragma unroll
            for (int r = 0; r < R; ++r) row[wb + r * NS + ((r * NS) >> 4)] = u[q][r];
        }
    }
    if constexpr (!LAST) wave_sync();  // the writes precede the next stage's reads
}

template <class T, int LOGN, bool INV, int S = 0, int NS = 1>
__device__ __forceinline__ void wave_stages(cx<T> *v, cx<T> *row, const cx<T> *__restrict__ tw, int lane)
{
    if constexpr (S < wave_nstages(LOGN)) {
        wave_stage<T, LOGN, INV, S, NS>(v, row, tw, lane);
        wave_stages<T, LOGN, INV, S + 1, NS * (1 << wave_stage_bits(LOGN, S))>(v, row, tw, lane);
    }
}

// unnormalised FFT of the wave's row v[E] (see above); row is the wave's LDS
// row of Lds<T, N>::PITCH elements
template <class T, int LOGN, bool INV, bool OPQ = false>
__device__ __forceinline__ void wave_fft(cx<T> *v, cx<T> *row, const cx<T> *__restrict__ tw, int lane)
{
    static_assert(wave_nstages(LOGN) > 0, "wave FFT lengths: 512, 1024, 2048");
    // OPQ: the lane index through an opaque copy, so the addresses are
    // recomputed per call rather than held in registers between the calls of
    // a loop (fewer registers, one more wave per SIMD for k_xfwdinv_w)
    if constexpr (OPQ) asm volatile("" : "+v"(lane));
    wave_stages<T, LOGN, INV>(v, row, tw, lane);
}

}  // namespace hos
