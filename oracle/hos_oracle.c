/*
 * hos_oracle.c -- FP64 CPU restatement of the HOS time-step path of
 * EsterDoesProgramming/hos (Csource/2dpar).  TEST INFRASTRUCTURE ONLY.
 *
 * This file is the parity checker for the HIP library in hos_amd/.  Only the
 * tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and only as the checker / the CPU baseline; the product path never calls it.
 *
 * It follows the reference's own structure: every pseudo-spectral product is a
 * Mult (c2r, c2r, pointwise product, r2c), Zvel runs the same triple loop with
 * in-place Dz, and sol_update_RK uses the same Butcher tables and the same
 * stage-time semantics.  The reference's FFTW-MPI plans are replaced by a
 * self-contained FP64 FFT (FFTW is not available here); its conventions are
 * kept: forward e^{-i}, 1/(Nx*Ny) on r2c, unnormalised c2r, transposed
 * spectral layout [Ny/2+1][Nx] (index Nx*j + i), padded real layout
 * [Nx][Ny+2].
 *
 * Policy for c2r inputs that are not Hermitian on the ky=0 / ky=Ny/2 columns
 * (the reference's kx table can produce those, operators_par.c:55,173): the
 * x-inverse is done per ky column, then the real part of the ky=0 and ky=Ny/2
 * values is used in the y-inverse.  This is c2r of the Hermitian projection.
 *
 * Pins (see tests/): brute-force DFT and numpy.fft.rfft2 for the FFT,
 * closed forms per index for Dx/Dy/Dz/Dealias, single-mode products for Mult,
 * the exact harmonic potential for Zvel (series convergence in M), the
 * single-mode closed form for rhs_hos, the RK4 stability polynomial for the
 * linear problem, Parseval for the Hamiltonian, and Stokes-wave phase speed.
 */
#include <complex.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex cplx;

/* ------------------------------------------------------------------------ */
/* context: the reference keeps all of this in globals (euler_2d_par.h:40-98,
 * fft_routines_2d_par.h:17-38, hosm_2d_par.c:12-26, time_schemes_par.c:13-25) */
typedef struct or_ctx {
    int Nx, Ny, nlevs;           /* NLevs, euler_2d_par.h:24 ; M = NLevs+1 */
    double Lx, Ly, Kx0, Ky0, g;  /* Kx0 = 2pi/Lx, hdf5_routines_2d_par.c:98 */
    int rampflg;
    double Tramp;
    double depth;                /* <= 0: deep water (the reference's live Dz) */
    int kx_quirk;                /* 1: reference kx table (i <= Nx/2+1 -> +i) */
    ptrdiff_t local_N, alloc_local;
    /* FFT scratch */
    cplx *rowbuf, *colbuf, *tmp;  /* tmp: [Ny/2+1][Nx] */
    cplx *twx, *twy;
    /* Mult scratch (operators_par.c:218-233 uses global temp1/temp2) */
    double *temp1, *temp2;
    /* rhs_hos_setup scratch (hosm_2d_par.c:57-95) */
    cplx *htemp1, *htemp2, *hetan, *hphin, *hwn, *hwM, *hwM2, *hw2M, *hw2M2;
    double *Coeff;
    int *dzorder;                /* derivative order held by hphin[m] (finite depth) */
    /* time scheme (time_schemes_par.c:13-25) */
    int Nstages;
    double a[7], b[7][7], c[7];
    cplx *fun, *htemp;
} or_ctx;

/* ------------------------------------------------------------------------ */
/* 1-D complex FFT, FP64.  Power-of-two lengths: iterative radix-2 with a
 * twiddle table whose entries are computed directly (no recurrences).  Other
 * lengths: direct DFT with exact index reduction.  sign = -1 forward, +1
 * inverse, unnormalised (FFTW convention). */
static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static cplx *make_twiddles(int n)
{
    cplx *tw = malloc(sizeof(cplx) * (size_t)n);
    for (int k = 0; k < n; k++) {
        /* exp(-2 pi i k / n), argument reduced to the first octant for accuracy */
        long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
        tw[k] = (double)cosl(ang) + I * (double)sinl(ang);
    }
    return tw;
}

static void fft1d(cplx *x, int n, int sign, const cplx *tw, cplx *scratch)
{
    if (n == 1) return;
    if (!is_pow2(n)) {
        for (int k = 0; k < n; k++) {
            cplx s = 0;
            for (int m = 0; m < n; m++) {
                long idx = ((long)k * m) % n;
                cplx w = tw[idx];
                if (sign > 0) w = conj(w);
                s += x[m] * w;
            }
            scratch[k] = s;
        }
        memcpy(x, scratch, sizeof(cplx) * (size_t)n);
        return;
    }
    /* bit reversal */
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cplx t = x[i]; x[i] = x[j]; x[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; k++) {
                cplx w = tw[k * step];
                if (sign > 0) w = conj(w);
                cplx u = x[i + k], v = x[i + k + half] * w;
                x[i + k] = u + v;
                x[i + k + half] = u - v;
            }
        }
    }
}

/* fft_2d, fft_routines_2d_par.c:21-58: padded real u [Nx][Ny+2] -> hu
 * [Ny/2+1][Nx] (FFTW_MPI_TRANSPOSED_OUT), forward, then hu = hf/Nx/Ny. */
/* The row and column loops below run on OpenMP threads (or_set_threads):
 * every 1-D transform is independent and done by one thread with its own
 * scratch, so the results do not depend on the thread count. */
void or_fft_2d(or_ctx *C, const double *u, cplx *hu)
{
    const int Nx = C->Nx, Ny = C->Ny, Nk = Ny / 2 + 1;
    const int nmax = Nx > Ny ? Nx : Ny;
#pragma omp parallel
    {
        cplx *rowbuf = malloc(sizeof(cplx) * (size_t)nmax), *colbuf = malloc(sizeof(cplx) * (size_t)nmax);
#pragma omp for schedule(static)
        for (int x = 0; x < Nx; x++) {
            for (int y = 0; y < Ny; y++) rowbuf[y] = u[(size_t)(Ny + 2) * x + y];
            fft1d(rowbuf, Ny, -1, C->twy, colbuf);
            for (int k = 0; k < Nk; k++) C->tmp[(size_t)Nx * k + x] = rowbuf[k];
        }
        const double s = 1.0 / Nx / Ny;
#pragma omp for schedule(static)
        for (int k = 0; k < Nk; k++) {
            cplx *row = &C->tmp[(size_t)Nx * k];
            fft1d(row, Nx, -1, C->twx, colbuf);
            for (int x = 0; x < Nx; x++) hu[(size_t)Nx * k + x] = row[x] * s;
        }
        free(rowbuf);
        free(colbuf);
    }
}

/* ifft_2d, fft_routines_2d_par.c:68-103: hu [Ny/2+1][Nx] -> padded real u,
 * unnormalised backward transform (FFTW_MPI_TRANSPOSED_IN). */
void or_ifft_2d(or_ctx *C, const cplx *hu, double *u)
{
    const int Nx = C->Nx, Ny = C->Ny, Nk = Ny / 2 + 1;
    const int nmax = Nx > Ny ? Nx : Ny;
#pragma omp parallel
    {
    cplx *rowbuf = malloc(sizeof(cplx) * (size_t)nmax), *colbuf = malloc(sizeof(cplx) * (size_t)nmax);
#pragma omp for schedule(static)
    for (int k = 0; k < Nk; k++) {
        cplx *row = &C->tmp[(size_t)Nx * k];
        memcpy(row, &hu[(size_t)Nx * k], sizeof(cplx) * (size_t)Nx);
        fft1d(row, Nx, +1, C->twx, colbuf);
    }
#pragma omp for schedule(static)
    for (int x = 0; x < Nx; x++) {
        cplx *r = rowbuf;
        r[0] = creal(C->tmp[x]);
        for (int k = 1; k < Nk; k++) {
            cplx v = C->tmp[(size_t)Nx * k + x];
            if (2 * k == Ny) { r[k] = creal(v); continue; }
            r[k] = v;
            r[Ny - k] = conj(v);
        }
        fft1d(r, Ny, +1, C->twy, colbuf);
        for (int y = 0; y < Ny; y++) u[(size_t)(Ny + 2) * x + y] = creal(r[y]);
        u[(size_t)(Ny + 2) * x + Ny] = 0.0;
        u[(size_t)(Ny + 2) * x + Ny + 1] = 0.0;
    }
    free(rowbuf);
    free(colbuf);
    }
}

/* ------------------------------------------------------------------------ */
/* wavenumber tables, operators_par.c:50-71 (Dx), 105-116 (Dy), 168-192 (Dz) */
static double kx_of(const or_ctx *C, int i)
{
    const int Nx = C->Nx;
    if (C->kx_quirk)  /* operators_par.c:55: i <= Nx/2+1 -> +i*Kx0 */
        return (i <= Nx / 2 + 1) ? i * C->Kx0 : -(Nx - i) * C->Kx0;
    /* signed table: 0..Nx/2 positive (Nyquist +), Nx/2+1.. negative */
    return (i <= Nx / 2) ? i * C->Kx0 : -(Nx - i) * C->Kx0;
}

void or_Dx(or_ctx *C, const cplx *hu, cplx *hu_x)
{
    for (int j = 0; j < C->Ny / 2 + 1; j++)
        for (int i = 0; i < C->Nx; i++) {
            size_t idx = (size_t)C->Nx * j + i;
            hu_x[idx] = I * kx_of(C, i) * hu[idx];
        }
}

void or_Dy(or_ctx *C, const cplx *hu, cplx *hu_y)
{
    for (int j = 0; j < C->Ny / 2 + 1; j++)
        for (int i = 0; i < C->Nx; i++) {
            size_t idx = (size_t)C->Nx * j + i;
            hu_y[idx] = I * (j * C->Ky0) * hu[idx];
        }
}

/* Dz step taken from derivative order n-1 to n.  Deep water (the reference's
 * live branch, operators_par.c:168-192): multiply by |k| every time.  Finite
 * depth (build-defined; the reference's tanh branch :138-166 is dead code):
 * d^n/dz^n of cosh(k(z+h))/cosh(kh) at z=0 is k^n tanh(kh) for odd n and k^n
 * for even n, so an odd step multiplies by k tanh(kh), an even step by
 * k / tanh(kh). */
static void dz_step(const or_ctx *C, const cplx *hu, cplx *out, int n)
{
    for (int j = 0; j < C->Ny / 2 + 1; j++) {
        double ky = j * C->Ky0;
        for (int i = 0; i < C->Nx; i++) {
            size_t idx = (size_t)C->Nx * j + i;
            double kx = kx_of(C, i);
            double kz = sqrt(kx * kx + ky * ky);
            double m = kz;
            if (C->depth > 0) {
                double t = tanh(kz * C->depth);
                if (n & 1) m = kz * t;
                else m = (kz > 0) ? kz / t : 0.0;
            }
            out[idx] = m * hu[idx];
        }
    }
}

void or_Dz(or_ctx *C, const cplx *hu, cplx *hu_z) { dz_step(C, hu, hu_z, 1); }

/* Mult, operators_par.c:210-235: c2r both, product on the real grid, r2c. */
void or_Mult(or_ctx *C, const cplx *hu1, const cplx *hu2, cplx *hprod)
{
    or_ifft_2d(C, hu1, C->temp1);
    or_ifft_2d(C, hu2, C->temp2);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < C->Nx; i++)
        for (int j = 0; j < C->Ny; j++) {
            size_t idx = (size_t)(C->Ny + 2) * i + j;
            C->temp1[idx] = C->temp1[idx] * C->temp2[idx];
        }
    or_fft_2d(C, C->temp1, hprod);
}

/* Sum, operators_par.c:247-279 */
void or_Sum(or_ctx *C, const cplx *a, const cplx *b, cplx *s)
{
    for (ptrdiff_t i = 0; i < C->local_N; i++) s[i] = a[i] + b[i];
}

/* Dealias, operators_par.c:283-325: zero (i,j) if (i>=mx && i<Nx-mx+1) || j>=my */
int or_dealias_mx(int N, int nlevs) { return (int)floor(0.5 * N / (1 + 0.5 * nlevs)); }

void or_Dealias(or_ctx *C, cplx *hu)
{
    const int mx = or_dealias_mx(C->Nx, C->nlevs), my = or_dealias_mx(C->Ny, C->nlevs);
    for (int j = 0; j < C->Ny / 2 + 1; j++)
        for (int i = 0; i < C->Nx; i++)
            if ((i >= mx && i < C->Nx - mx + 1) || (j >= my)) hu[(size_t)C->Nx * j + i] = 0.0;
}

/* ------------------------------------------------------------------------ */
/* rhs_hos_setup, hosm_2d_par.c:51-97 */
static void rhs_hos_setup(or_ctx *C)
{
    const size_t al = (size_t)C->alloc_local, L1 = (size_t)C->nlevs + 1;
    C->temp1 = calloc(2 * al + 2 * (size_t)C->Nx, sizeof(double));
    C->temp2 = calloc(2 * al + 2 * (size_t)C->Nx, sizeof(double));
    C->htemp1 = calloc(2 * al, sizeof(cplx));
    C->htemp2 = &C->htemp1[al];
    C->hetan = calloc(L1 * al, sizeof(cplx));
    C->hphin = calloc(L1 * al, sizeof(cplx));
    C->hwn = calloc(L1 * al, sizeof(cplx));
    C->hwM = calloc(al, sizeof(cplx));
    C->hwM2 = calloc(al, sizeof(cplx));
    C->hw2M = calloc(al, sizeof(cplx));
    C->hw2M2 = calloc(al, sizeof(cplx));
    C->Coeff = malloc(sizeof(double) * (L1 + 1));
    C->dzorder = calloc(L1, sizeof(int));
    C->Coeff[0] = 1.0;
    C->Coeff[1] = 1.0;
    for (int n = 2; n <= C->nlevs; n++) C->Coeff[n] = C->Coeff[n - 1] * n;
}

/* Zvel, hosm_2d_par.c:240-445 (West et al. 1987 W expansion). */
void or_Zvel(or_ctx *C, const cplx *hu, cplx *hZvelM, cplx *hZvelM2, cplx *hZvel2M, cplx *hZvel2M2,
             double t)
{
    (void)t;
    const ptrdiff_t N = C->local_N, al = C->alloc_local;
    const int L = C->nlevs;
    for (int n = 0; n <= L; n++)
        for (ptrdiff_t i = 0; i < N; i++) {
            C->hwn[n * al + i] = 0.0;
            C->hphin[n * al + i] = 0.0;
        }
    for (ptrdiff_t i = 0; i < N; i++) {
        hZvelM[i] = hZvelM2[i] = hZvel2M[i] = hZvel2M2[i] = 0.0;
        C->hetan[i] = 0.0;
    }
    C->hetan[0] = 1.0;                       /* :276-278, unit mean */
    dz_step(C, &hu[al], C->hphin, 1);        /* :280 */
    C->dzorder[0] = 1;
    for (ptrdiff_t i = 0; i < N; i++) C->hwn[i] = C->hphin[i]; /* :283-288, W^(1) */

    for (int n = 1; n <= L; n++) {
        const ptrdiff_t Nindex = n * al;
        or_Mult(C, hu, &C->hetan[(n - 1) * al], &C->hetan[Nindex]); /* :297 eta^n */
        /* phi^(n+1), :302-317 */
        for (int m = 0; m <= n - 1; m++) {
            const ptrdiff_t Mindex = m * al;
            or_Mult(C, &C->hphin[Mindex], &C->hetan[Nindex - Mindex], C->htemp1);
            for (ptrdiff_t i = 0; i < N; i++) C->hphin[Nindex + i] -= C->htemp1[i] / C->Coeff[n - m];
        }
        C->dzorder[n] = 0;
        /* W^(n+1), :322-357, Dz in place on hphin[m] */
        for (int m = 0; m <= n; m++) {
            const ptrdiff_t Mindex = m * al;
            C->dzorder[m] += 1;
            dz_step(C, &C->hphin[Mindex], &C->hphin[Mindex], C->dzorder[m]);
            or_Mult(C, &C->hphin[Mindex], &C->hetan[Nindex - Mindex], C->htemp1);
            for (ptrdiff_t i = 0; i < N; i++) C->hwn[Nindex + i] += C->htemp1[i] / C->Coeff[n - m];
        }
    }
    /* W^(M-2), :362-370 */
    for (int n = 0; n <= L - 2; n++)
        for (ptrdiff_t i = 0; i < N; i++) hZvelM2[i] += C->hwn[n * al + i];
    /* W^(M), :373-394 (n >= 0 guard :382) */
    for (ptrdiff_t i = 0; i < N; i++) hZvelM[i] = hZvelM2[i];
    for (int n = L - 1; n <= L; n++)
        if (n >= 0)
            for (ptrdiff_t i = 0; i < N; i++) hZvelM[i] += C->hwn[n * al + i];
    /* (W^2)^(M-2), :397-416 */
    for (int n = 0; n <= L - 3; n++)
        for (int m = 0; m <= n; m++) {
            or_Mult(C, &C->hwn[m * al], &C->hwn[(n - m) * al], C->htemp1);
            or_Sum(C, hZvel2M2, C->htemp1, hZvel2M2);
        }
    /* (W^2)^(M), :420-442 */
    for (ptrdiff_t i = 0; i < N; i++) hZvel2M[i] = hZvel2M2[i];
    for (int n = L - 2; n <= L - 1; n++)
        for (int m = 0; m <= n; m++) {
            or_Mult(C, &C->hwn[m * al], &C->hwn[(n - m) * al], C->htemp1);
            or_Sum(C, hZvel2M, C->htemp1, hZvel2M);
        }
}

/* ZvelLinear, hosm_2d_par.c:448-455 */
void or_ZvelLinear(or_ctx *C, const cplx *hu, cplx *out) { dz_step(C, &hu[C->alloc_local], out, 1); }

/* RampFun, hosm_2d_par.c:526-530 */
double or_RampFun(or_ctx *C, double t) { return 1 - exp(-t * t / C->Tramp / C->Tramp); }

/* rhs_hos, hosm_2d_par.c:138-235 */
void or_rhs_hos(or_ctx *C, cplx *hrhs, const cplx *hu, double t)
{
    const ptrdiff_t N = C->local_N, al = C->alloc_local;
    cplx *h1 = C->htemp1, *h2 = C->htemp2;
    /* eta_x phi_x + eta_y phi_y, :150-161 */
    or_Dx(C, hu, h1);
    or_Dx(C, &hu[al], h2);
    or_Mult(C, h1, h2, hrhs);
    or_Dy(C, hu, h1);
    or_Dy(C, &hu[al], h2);
    or_Mult(C, h1, h2, h1);
    for (ptrdiff_t i = 0; i < N; i++) hrhs[i] = -hrhs[i] - h1[i];
    /* :164 */
    or_Zvel(C, hu, C->hwM, C->hwM2, C->hw2M, C->hw2M2, t);
    /* :166-176 */
    or_Dx(C, hu, h1);
    or_Mult(C, h1, h1, h1);
    or_Dy(C, hu, h2);
    or_Mult(C, h2, h2, h2);
    or_Sum(C, h1, h2, h1);
    or_Mult(C, h1, C->hwM2, h2);
    or_Sum(C, C->hwM, h2, h2);
    or_Sum(C, hrhs, h2, hrhs);
    /* phi part, :181-189 */
    or_Mult(C, h1, C->hw2M2, h2);
    or_Sum(C, C->hw2M, h2, h2);
    for (ptrdiff_t i = 0; i < N; i++) hrhs[al + i] = -C->g * hu[i] + 0.5 * h2[i];
    /* -0.5 |grad phi|^2, :192-206 */
    or_Dx(C, &hu[al], h1);
    or_Mult(C, h1, h1, h1);
    or_Dy(C, &hu[al], h2);
    or_Mult(C, h2, h2, h2);
    or_Sum(C, h1, h2, h1);
    for (ptrdiff_t i = 0; i < N; i++) hrhs[al + i] -= 0.5 * h1[i];
    /* Dommermuth ramp, :211-229 */
    if (C->rampflg == 1) {
        const double R = or_RampFun(C, t);
        or_ZvelLinear(C, hu, h1);
        for (ptrdiff_t i = 0; i < N; i++) hrhs[i] = (hrhs[i] - h1[i]) * R + h1[i];
        for (ptrdiff_t i = 0; i < N; i++) hrhs[al + i] = (hrhs[al + i] + C->g * hu[i]) * R - C->g * hu[i];
    }
    /* :232-233 */
    or_Dealias(C, hrhs);
    or_Dealias(C, &hrhs[al]);
}

/* Hamiltonian, hosm_2d_par.c:462-523 (1 rank: row j=0 weight 1, j>=1 weight 2) */
double or_Hamiltonian(or_ctx *C, const cplx *heta, const cplx *heta_t, const cplx *hphi)
{
    double H = 0.0;
    const double g = C->g;
    for (int j = 0; j < C->Ny / 2 + 1; j++) {
        const double w = (j == 0) ? 1.0 : 2.0;
        for (int i = 0; i < C->Nx; i++) {
            size_t idx = (size_t)C->Nx * j + i;
            H += w * g * creal(heta[idx]) * creal(heta[idx]);
            H += w * g * cimag(heta[idx]) * cimag(heta[idx]);
            H += w * creal(hphi[idx]) * creal(heta_t[idx]);
            H += w * cimag(hphi[idx]) * cimag(heta_t[idx]);
        }
    }
    return H;
}

/* ------------------------------------------------------------------------ */
/* Setup_TimeScheme, time_schemes_par.c:127-314 */
int or_Setup_TimeScheme(or_ctx *C, int scheme_flg)
{
    int ns;
    switch (scheme_flg) {
    case 1: ns = 7; break;
    case 2: ns = 4; break;
    case 3: ns = 1; break;
    case 4: ns = 2; break;
    case 5: ns = 2; break;
    default: return -1;
    }
    memset(C->a, 0, sizeof C->a);
    memset(C->b, 0, sizeof C->b);
    memset(C->c, 0, sizeof C->c);
    C->Nstages = ns;
    switch (scheme_flg) {
    case 1: /* Dormand-Prince, 4th-order weights, :181-238 */
        C->a[1] = 1.0 / 5; C->a[2] = 3.0 / 10; C->a[3] = 4.0 / 5; C->a[4] = 8.0 / 9; C->a[5] = 1; C->a[6] = 1;
        C->b[1][0] = 1.0 / 5;
        C->b[2][0] = 3.0 / 40; C->b[2][1] = 9.0 / 40;
        C->b[3][0] = 44.0 / 45; C->b[3][1] = -56.0 / 15; C->b[3][2] = 32.0 / 9;
        C->b[4][0] = 19372.0 / 6561; C->b[4][1] = -25360.0 / 2187; C->b[4][2] = 64448.0 / 6561;
        C->b[4][3] = -212.0 / 729;
        C->b[5][0] = 9017.0 / 3168; C->b[5][1] = -355.0 / 33; C->b[5][2] = 46732.0 / 5247;
        C->b[5][3] = 49.0 / 176; C->b[5][4] = -5103.0 / 18656;
        C->b[6][0] = 35.0 / 384; C->b[6][1] = 0; C->b[6][2] = 500.0 / 1113; C->b[6][3] = 125.0 / 192;
        C->b[6][4] = -2187.0 / 6784; C->b[6][5] = 11.0 / 84;
        C->c[0] = 1951.0 / 21600; C->c[1] = 0; C->c[2] = 22642.0 / 50085; C->c[3] = 451.0 / 720;
        C->c[4] = -12231.0 / 42400; C->c[5] = 649.0 / 6300; C->c[6] = 1.0 / 60;
        break;
    case 2: /* classical RK4, :241-264 */
        C->a[1] = 0.5; C->a[2] = 0.5; C->a[3] = 1.0;
        C->b[1][0] = 0.5; C->b[2][1] = 0.5; C->b[3][2] = 1.0;
        C->c[0] = 1.0 / 6; C->c[1] = 1.0 / 3; C->c[2] = 1.0 / 3; C->c[3] = 1.0 / 6;
        break;
    case 3: C->c[0] = 1.0; break;                                      /* Euler :268-276 */
    case 4: C->a[1] = 0.5; C->b[1][0] = 0.5; C->c[0] = 0; C->c[1] = 1; break; /* midpoint :278-290 */
    case 5: C->a[1] = 1; C->b[1][0] = 1; C->c[0] = 0.5; C->c[1] = 0.5; break; /* Heun :292-304 */
    }
    free(C->fun);
    free(C->htemp);
    C->fun = calloc((size_t)2 * ns * C->alloc_local, sizeof(cplx));
    C->htemp = calloc((size_t)2 * C->alloc_local, sizeof(cplx));
    return 0;
}

/* sol_update_RK, time_schemes_par.c:28-121.  Every stage is evaluated at t_n
 * (:70).  Returns the stage-0 Hamiltonian that the reference logs (:73-93). */
double or_sol_update_RK(or_ctx *C, cplx *u, double *t, double dt)
{
    const ptrdiff_t N = C->local_N, al = C->alloc_local;
    double Ham = 0.0;
    for (int s = 0; s < C->Nstages; s++) {
        for (ptrdiff_t i = 0; i < N; i++) {
            C->htemp[i] = u[i];
            C->htemp[al + i] = u[al + i];
        }
        for (int p = 0; p < s; p++)
            for (ptrdiff_t i = 0; i < N; i++) {
                C->htemp[i] += dt * C->b[s][p] * C->fun[2 * al * p + i];
                C->htemp[al + i] += dt * C->b[s][p] * C->fun[2 * al * p + al + i];
            }
        or_rhs_hos(C, &C->fun[2 * al * s], C->htemp, t[0]);
        if (s == 0) Ham = or_Hamiltonian(C, u, C->fun, &u[al]);
    }
    for (int s = 0; s < C->Nstages; s++)
        for (ptrdiff_t i = 0; i < N; i++) {
            u[i] += dt * C->c[s] * C->fun[2 * al * s + i];
            u[al + i] += dt * C->c[s] * C->fun[2 * al * s + al + i];
        }
    t[0] += dt;
    return Ham;
}

/* nsteps calls of sol_update_RK; H per step into Hlog (may be NULL) */
void or_run(or_ctx *C, cplx *u, double *t, double dt, int nsteps, double *Hlog)
{
    for (int n = 0; n < nsteps; n++) {
        double H = or_sol_update_RK(C, u, t, dt);
        if (Hlog) Hlog[n] = H;
    }
}

/* ------------------------------------------------------------------------ */
or_ctx *or_create(int Nx, int Ny, int nlevs, double Lx, double Ly, double g, int rampflg, double Tramp,
                  double depth, int kx_quirk)
{
    if (Nx < 2 || Ny < 2 || (Ny & 1) || nlevs < 0) return NULL;
    or_ctx *C = calloc(1, sizeof(or_ctx));
    C->Nx = Nx; C->Ny = Ny; C->nlevs = nlevs;
    C->Lx = Lx; C->Ly = Ly; C->g = g;
    C->Kx0 = 2 * M_PI / Lx; C->Ky0 = 2 * M_PI / Ly;
    C->rampflg = rampflg; C->Tramp = Tramp; C->depth = depth; C->kx_quirk = kx_quirk;
    C->local_N = (ptrdiff_t)Nx * (Ny / 2 + 1);
    C->alloc_local = C->local_N;
    const int nmax = Nx > Ny ? Nx : Ny;
    C->rowbuf = malloc(sizeof(cplx) * (size_t)nmax);
    C->colbuf = malloc(sizeof(cplx) * (size_t)nmax);
    C->tmp = malloc(sizeof(cplx) * (size_t)C->local_N);
    C->twx = make_twiddles(Nx);
    C->twy = make_twiddles(Ny);
    rhs_hos_setup(C);
    or_Setup_TimeScheme(C, 2);
    return C;
}

void or_destroy(or_ctx *C)
{
    if (!C) return;
    free(C->rowbuf); free(C->colbuf); free(C->tmp); free(C->twx); free(C->twy);
    free(C->temp1); free(C->temp2); free(C->htemp1); free(C->hetan); free(C->hphin); free(C->hwn);
    free(C->hwM); free(C->hwM2); free(C->hw2M); free(C->hw2M2); free(C->Coeff); free(C->dzorder);
    free(C->fun); free(C->htemp);
    free(C);
}

ptrdiff_t or_local_N(or_ctx *C) { return C->local_N; }

/* OpenMP threads of the transforms (n <= 0: the runtime's default) */
#ifdef _OPENMP
#include <omp.h>
int or_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}
#else
int or_set_threads(int n) { (void)n; return 1; }
#endif

/* Butcher tableau of the current scheme (for order-condition tests). */
int or_get_tableau(or_ctx *C, double *a, double *b, double *c)
{
    for (int s = 0; s < 7; s++) {
        a[s] = C->a[s];
        c[s] = C->c[s];
        for (int p = 0; p < 7; p++) b[7 * s + p] = C->b[s][p];
    }
    return C->Nstages;
}
