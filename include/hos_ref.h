/*
 * hos_ref.h -- the reference solver's own call boundary (Csource/2dpar, the
 * USE_DOUBLES build), implemented by libhos_ref.so on top of the native API in
 * hos_gpu.h.  A driver written against Csource/2dpar (main_2d_par.c) links
 * against libhos_ref.so (and libhos_gpu.so, with -fcommon) instead of
 * time_schemes_par.o, hosm_2d_par.o, operators_par.o and fft_routines_2d_par.o.
 *
 * Types: hos_cplx is hos_real[2]: double[2], the same type as fftw_complex,
 * or float[2] = fftwf_complex in the float build, so pointers declared
 * fftw_complex* (fftwf_complex*) pass unchanged.  hos_plan stands in for
 * fftw_plan; the FFTW-MPI plans are replaced outright and the argument is
 * ignored.  The float build converts its arrays to FP64 at the boundary,
 * keeps the GPU state in FP64 between steps, and computes in the FP32 form
 * (hos_precision defaults to HOS_FP32 there).
 *
 * Globals: the parameters the reference keeps in header-defined globals
 * (euler_2d_par.h:40-71, fft_routines_2d_par.h:28-33) are defined here; a
 * driver built from the reference's headers with -fcommon merges its
 * tentative definitions with these.  The compile-time choices of the
 * reference (NLevs, euler_2d_par.h:24; precision, :33-37) become the runtime
 * globals hos_nlevs and hos_precision, read when rhs_hos_setup runs.
 *
 * State residency: sol_update_RK keeps the state on the GPU between calls.
 * The host copy of u is refreshed lazily, before any call here reads host
 * memory inside u (ifft_2d of heta / hphi at output time, main_2d_par.c:310),
 * or by hos_ref_sync().  A caller that writes u itself between steps calls
 * hos_ref_host_modified().
 *
 * Errors: every entry point prints the library message and exits with
 * EXIT_FAILURE, the reference's fail-hard convention (time_schemes_par.c:81-86).
 */
#ifndef HOS_REF_H
#define HOS_REF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* TYPE_REAL of the reference (euler_2d_par.h:33-37): double in its
 * USE_DOUBLES build (libhos_ref.so), float in the other one (define
 * HOS_REF_FLOAT and link libhos_ref_f32.so: the same names over
 * fftwf_complex / float, hosm_2d_par.h:31-38, operators_par.h:24-30,
 * fft_routines_2d_par.h:45-47, time_schemes_par.h:22-24). */
#ifdef HOS_REF_FLOAT
typedef float hos_real;
#else
typedef double hos_real;
#endif
typedef hos_real hos_cplx[2];
typedef void *hos_plan;

/* euler_2d_par.h:40-71 */
extern int Nx, Ny, runsubid, saveflg, rampflg;
extern hos_real Lx, Ly, Kx0, Ky0, g, T, dtsave, Tramp;
/* fft_routines_2d_par.h:28-33 (one GPU: the whole grid is local) */
extern ptrdiff_t alloc_local, local_Nx, local_Nyhpo, local_N, local_0_start, local_1_start;
extern int mpi_rank, mpi_size;

/* runtime forms of the reference's compile-time switches */
extern int hos_nlevs;        /* NLevs (default 2, euler_2d_par.h:24)              */
extern int hos_precision;    /* HOS_FP64 (USE_DOUBLES) or HOS_FP32                */
extern double hos_depth;     /* <= 0: deep water (the reference's live Dz)        */
extern int hos_kx_quirk;     /* 1: reference kx table (operators_par.c:55)        */
extern int hos_runtime_log;  /* 1: append "t H" to runtime.dat every step, as
                                time_schemes_par.c:73-93                          */

/* replaces fftw_mpi_local_size_2d_transposed (main_2d_par.c:143): n0 = Nx,
 * n1 = Ny/2+1; sets the layout globals and returns alloc_local. */
ptrdiff_t hos_local_size_2d_transposed(ptrdiff_t n0, ptrdiff_t n1, ptrdiff_t *local_n0, ptrdiff_t *local_0_start_,
                                       ptrdiff_t *local_n1, ptrdiff_t *local_1_start_);

/* time_schemes_par.h:20-26 */
void Setup_TimeScheme(int scheme_flg);
void sol_update_RK(hos_cplx *u, hos_real *t, hos_real dt, char *dtflag);

/* hosm_2d_par.h:19-40 */
void rhs_hos_setup(void);
void rhs_hos_clean(void);
void rhs_hos(hos_cplx *hrhs, hos_cplx *hu, hos_real t);
void Zvel(hos_cplx *hu, hos_cplx *hZvelM, hos_cplx *hZvelM2, hos_cplx *hZvel2M, hos_cplx *hZvel2M2, hos_real t);
void ZvelLinear(const hos_cplx *hu, hos_cplx *hZvelLinear);
hos_real Hamiltonian(const hos_cplx *heta, const hos_cplx *heta_t, const hos_cplx *hphi);
hos_real RampFun(const hos_real t);

/* operators_par.h:17-23 */
void Dx(hos_cplx *hu, hos_cplx *hu_x);
void Dy(hos_cplx *hu, hos_cplx *hu_y);
void Dz(const hos_cplx *hu, hos_cplx *hu_z);
void Mult(hos_cplx *hu1, hos_cplx *hu2, hos_cplx *hprod);
void Sum(hos_cplx *hu1, hos_cplx *hu2, hos_cplx *hsum);
void Dealias(hos_cplx *hu);

/* fft_routines_2d_par.h:40-50 */
void fft_2d(hos_real *u, hos_cplx *hu, hos_plan plan);
void ifft_2d(hos_cplx *hu, hos_real *u, hos_plan plan);

/* additions of this boundary */
void hos_ref_sync(void);               /* refresh the host copy of the state */
void hos_ref_host_modified(void);      /* the caller changed u on the host   */
struct hos_ctx *hos_ref_context(void); /* the native context (NULL before setup) */

#ifdef __cplusplus
}
#endif
#endif
