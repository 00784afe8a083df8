/*
 * hos_ic.c -- seeded JONSWAP initial condition (include/hos_ic.h).
 *
 * Follows python/CreateSpectra.py:11-44, python/JONSWAP.py:28-63,
 * python/OmegaTheta.py:10-41 and python/DtoF.py:6-63 of the reference; the
 * additions (counter-based phases, |k| cut, Hs = 2 eps / kp) are listed in
 * the header.
 */
#include "../../include/hos_ic.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

double hos_ic_uniform(unsigned long long seed, int i, int j)
{
    uint64_t z = (uint64_t)seed * 0x100000000000ULL + (uint64_t)j * 0x200000ULL + (uint64_t)i;
    /* SplitMix64 finaliser */
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

double hos_ic_peak_period(double g, double kp, double depth)
{
    const double t = depth > 0 ? tanh(kp * depth) : 1.0;
    return 2 * M_PI / sqrt(g * kp * t);
}

/* omega(k), OmegaTheta.py:33-34 */
static double omega_of(double g, double k, double h) { return sqrt(g * k * (h > 0 ? tanh(k * h) : 1.0)); }

/* JONSWAP S(omega) without alpha g^2 (the final rescale removes it), JONSWAP.py:40-49 */
static double jonswap(double w, double wp, double gamma)
{
    if (w <= 0) return 0.0;
    const double sigma = (w <= wp) ? 0.07 : 0.09;
    const double delta = exp(-(w - wp) * (w - wp) / (2 * sigma * sigma * wp * wp));
    return pow(w, -5.0) * exp(-1.25 * pow(wp / w, 4.0)) * pow(gamma, delta);
}

/* cos^2 spreading with Theta = pi, JONSWAP.py:21-24 */
static double spread(double th)
{
    while (th > M_PI) th -= 2 * M_PI;
    while (th <= -M_PI) th += 2 * M_PI;
    if (fabs(th) < 0.5 * M_PI) return 2.0 / M_PI * cos(th) * cos(th);
    return 0.0;
}

/* Jacobian factor "norm" of DtoF.py:33-36 (deep) / :38-41 (finite depth) */
static double norm_of(double g, double k, double h)
{
    if (h <= 0) return sqrt(sqrt(g) / 2 / pow(k, 1.5));
    const double kh = k * h, th = tanh(kh), sech = 1.0 / cosh(kh);
    return sqrt(0.5 * sqrt(g) * (sqrt(th / (k * k * k)) + h * (1.0 / sqrt(k * th)) * sech * sech));
}

/* complex amplitude A(k) = sqrt(S D) e^{i 2 pi U} of full-grid mode (i, j) */
static void amp(const hos_jonswap *p, int i, int j, double wp, double *re, double *im)
{
    const double kx = ((i <= p->Nx / 2) ? i : i - p->Nx) * 2 * M_PI / p->Lx;
    const double ky = ((j <= p->Ny / 2) ? j : j - p->Ny) * 2 * M_PI / p->Ly;
    const double k = sqrt(kx * kx + ky * ky);
    *re = *im = 0.0;
    if (k == 0.0) return;
    if (p->kcut > 0 && k > p->kcut) return;
    const double w = omega_of(p->g, k, p->depth);
    const double a = sqrt(jonswap(w, wp, p->gamma) * spread(atan2(ky, kx) - p->theta_p));
    const double ph = 2 * M_PI * hos_ic_uniform(p->seed, i, j);
    *re = a * cos(ph);
    *im = a * sin(ph);
}

int hos_ic_jonswap(const hos_jonswap *p, double *u)
{
    if (!p || !u || p->Nx < 2 || p->Ny < 2 || (p->Nx & 1) || (p->Ny & 1) || !(p->Lx > 0) || !(p->Ly > 0) ||
        !(p->g > 0) || !(p->kp > 0) || !(p->gamma >= 1.0) || !(p->eps > 0))
        return -1;
    const int Nx = p->Nx, Ny = p->Ny, Nk = Ny / 2 + 1;
    const size_t fs = (size_t)Nx * Nk;
    double *eta = u, *phi = u + 2 * fs;
    memset(u, 0, sizeof(double) * 4 * fs);
    const double wp = omega_of(p->g, p->kp, p->depth);
    double var = 0.0;
    for (int j = 0; j < Nk; j++) {
        for (int i = 0; i < Nx; i++) {
            if (i == Nx / 2 || j == Ny / 2) continue; /* Nyquist lines, DtoF.py:38-43 */
            const double kx = ((i <= Nx / 2) ? i : i - Nx) * 2 * M_PI / p->Lx;
            const double ky = j * 2 * M_PI / p->Ly;
            const double k = sqrt(kx * kx + ky * ky);
            if (k == 0.0) continue;
            double ar, ai, br, bi;
            amp(p, i, j, wp, &ar, &ai);                             /* A(k)  */
            amp(p, (Nx - i) % Nx, (Ny - j) % Ny, wp, &br, &bi);     /* A(-k) */
            const double nrm = norm_of(p->g, k, p->depth);
            const double alpha = p->g / omega_of(p->g, k, p->depth);
            /* eta = 1/2 norm (A + conj B); phi = -i/2 norm alpha (A - conj B) */
            const double er = 0.5 * nrm * (ar + br), ei = 0.5 * nrm * (ai - bi);
            const double dr = ar - br, di = ai + bi;
            const double pr = 0.5 * nrm * alpha * di, pi = -0.5 * nrm * alpha * dr;
            const size_t idx = (size_t)Nx * j + i;
            eta[2 * idx] = er;
            eta[2 * idx + 1] = ei;
            phi[2 * idx] = pr;
            phi[2 * idx + 1] = pi;
            var += (j == 0 ? 1.0 : 2.0) * (er * er + ei * ei);
        }
    }
    if (!(var > 0)) return -1;
    const double target = p->eps / (2 * p->kp); /* std(eta) = Hs/4, Hs = 2 eps/kp */
    const double s = target / sqrt(var);
    for (size_t q = 0; q < 4 * fs; q++) u[q] *= s;
    return 0;
}

void hos_dealias_state(int Nx, int Ny, int nlevs, double *u)
{
    const int mx = (int)floor(0.5 * Nx / (1 + 0.5 * nlevs)), my = (int)floor(0.5 * Ny / (1 + 0.5 * nlevs));
    const size_t fs = (size_t)Nx * (Ny / 2 + 1);
    for (int c = 0; c < 2; c++)
        for (int j = 0; j < Ny / 2 + 1; j++)
            for (int i = 0; i < Nx; i++)
                if ((i >= mx && i < Nx - mx + 1) || j >= my) {
                    const size_t idx = c * fs + (size_t)Nx * j + i;
                    u[2 * idx] = u[2 * idx + 1] = 0.0;
                }
}
