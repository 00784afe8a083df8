/*
 * hos_ic.h -- seeded JONSWAP initial conditions (host C), the recipe of the
 * reference's python/CreateSpectra.py:11-44 (create_jonswap) with
 * python/OmegaTheta.py:10-41, python/JONSWAP.py:28-63 and python/DtoF.py:6-63:
 *
 *   omega = sqrt(g k tanh(k h))                        (OmegaTheta.py:33-34)
 *   S(omega) = alpha g^2 omega^-5 exp(-5/4 (omega_p/omega)^4)
 *              gamma^exp(-(omega-omega_p)^2 / (2 sigma^2 omega_p^2)),
 *   sigma = 0.07 (omega <= omega_p) or 0.09              (JONSWAP.py:1-9,40-49)
 *   D(theta) = 2/Theta cos^2(pi theta/Theta), |theta| < Theta/2, Theta = pi
 *                                                        (JONSWAP.py:21-24)
 *   A(k) = sqrt(S D) exp(i 2 pi U(k))                    (CreateSpectra.py:25)
 *   eta-hat(k) = 1/2 norm(k) (A(k) + conj A(-k)),
 *   phi-hat(k) = -i/2 norm(k) (g/omega) (A(k) - conj A(-k))   (DtoF.py:36-56)
 *   k = 0 and the Nyquist lines zeroed                   (DtoF.py:38-43)
 *   rescaled so that std(eta) = Hs/4                     (CreateSpectra.py:40-42)
 *
 * Additions that the reference leaves open: U(k) comes from a counter-based
 * generator (SplitMix64 of seed and mode index), so the same field is produced
 * in C and in numpy; modes with |k| > kcut are zeroed; Hs = 2 eps / kp.  The
 * output is the spectral state in the reference layout, [2][Ny/2+1][Nx]
 * complex double, normalised as fft_2d (fft_routines_2d_par.c:52-56), before
 * any dealiasing.
 */
#ifndef HOS_IC_H
#define HOS_IC_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hos_jonswap {
    int Nx, Ny;
    double Lx, Ly;
    double g;
    double depth;    /* <= 0: deep water */
    double kp;       /* peak wavenumber */
    double gamma;    /* peak enhancement */
    double eps;      /* steepness kp*Hs/2 */
    double theta_p;  /* mean direction (rad) */
    double kcut;     /* zero modes with |k| > kcut (<= 0: no cut) */
    unsigned long long seed;
} hos_jonswap;

/* u: [2][Ny/2+1][Nx] complex double (interleaved re, im).  Returns 0, or -1 on
 * invalid input. */
int hos_ic_jonswap(const hos_jonswap *p, double *u);

/* The uniform variate U in [0,1) that hos_ic_jonswap uses for full-grid mode
 * (i, j), i in [0,Nx), j in [0,Ny) (FFT index order). */
double hos_ic_uniform(unsigned long long seed, int i, int j);

/* Peak period Tp = 2 pi / omega_p with omega_p = sqrt(g kp tanh(kp h)). */
double hos_ic_peak_period(double g, double kp, double depth);

/* Apply the reference's dealiasing mask (operators_par.c:309-322) in place to
 * both fields of a state (main_2d_par.c:261-262). */
void hos_dealias_state(int Nx, int Ny, int nlevs, double *u);

#ifdef __cplusplus
}
#endif
#endif
