// hos_misc_kernels.hpp -- the non-template kernels (end of step, Butcher
// combination, spectral operators of the C ABI, Hamiltonian, copies, the
// finite-depth table).  Included by hos_engine.hip only, so each kernel is
// defined in one translation unit.
#pragma once
#include "hos_kernels.hpp"

namespace hos {

// tanh(|k| h) of every mode of member m (1 where h <= 0), FP64; the table the
// finite-depth operators read instead of evaluating tanh per launch.
// grid (ceil(Nk*Nx/NT)), block NT
__global__ __launch_bounds__(NT) void k_th_table(KTab kt, int nx, int m, double *th)
{
    const long n = (long)kt.nk * nx;
    const long i = blockIdx.x * (long)NT + threadIdx.x;
    if (i >= n) return;
    const int kx = (int)(i % nx), ky = (int)(i / nx);
    const double kxv = kt.kx[kx], kyv = ky * kt.ky0, h = kt.h(m);
    const double kz = sqrt(kxv * kxv + kyv * kyv);
    th[m * kt.ths + i] = h > 0.0 ? tanh(kz * h) : 1.0;
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
// out = u + dt * sum_p coef[p] K_p   (time_schemes_par.c:38-67, 98-116).
// mean0 (non-null on the last stage with the mean-level correction on): the
// new state's eta-hat(0,0) is held at mean0[m], as the sub-diagonal epilogue
// does.
struct CombArgs {
    double coef[8];
    int np;
};
__global__ __launch_bounds__(NT) void k_combine(const cx<double> *U, const cx<double> *__restrict__ K,
                                                long sK_stage, long sM, cx<double> *out, long n, const double *dt,
                                                CombArgs ca, const double *mean0)
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
        if (mean0 && i == 0) v = {mean0[m], 0.0};
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
    __device__ __forceinline__ vec_t<T> operator()(int m, int x, int y) const
    {
        return ph.vat(m, fa, x, y) * ph.vat(m, fb, x, y);
    }
};

// FP64 spectral field [Nk][Nx] (row stride nx) copy, optionally from a
// different source; used for state set-up on the device.
__global__ __launch_bounds__(NT) void k_copy_c64(const cx<double> *__restrict__ src, cx<double> *__restrict__ dst,
                                                 long n)
{
    // four 16-byte loads in flight per thread per iteration
    constexpr int U = 4;
    const long step = (long)gridDim.x * NT * U;
    for (long i0 = blockIdx.x * (long)NT * U + threadIdx.x; i0 < n; i0 += step) {
        cx<double> v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long i = i0 + (long)u * NT;
            if (i < n) v[u] = src[i];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long i = i0 + (long)u * NT;
            if (i < n) dst[i] = v[u];
        }
    }
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
