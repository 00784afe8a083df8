// hos_levels.hip -- the y r2c launches whose prologue is instantiated per
// HOS level (phi^(lev+1), lev = 1..7: k_yr2c, and k_yc2r_phi, which fuses it
// behind the previous level's y c2r) or per order (the RHS pass k_yr2c_rhs,
// M = 1..8).
// Compiled once per precision (HOS_T) and part (HOS_PART 0: levels, 1: the
// RHS pass with the default FFT plan, 2: its other plans) by the Makefile, so
// the several hundred kernel instantiations build in parallel translation
// units.
#include "hos_ctx.hpp"

#ifndef HOS_T
#error "the Makefile sets HOS_T (float or double) and HOS_PART (0, 1 or 2)"
#endif

#if HOS_PART == 0
template <class T, int LEV>
static int phi_level_t(Eng<T> &e)
{
    ProPhi<T, LEV> pro{e.phys(), e.c->ix, e.comp()};
    return e.yr2c(pro, 0, LEV + 1);
}

template <class T>
int Eng<T>::phi_level(int lev)
{
    static_assert(MAXM == 8, "levels 1..MAXM-1");
    switch (lev) {
    case 1: return phi_level_t<T, 1>(*this);
    case 2: return phi_level_t<T, 2>(*this);
    case 3: return phi_level_t<T, 3>(*this);
    case 4: return phi_level_t<T, 4>(*this);
    case 5: return phi_level_t<T, 5>(*this);
    case 6: return phi_level_t<T, 6>(*this);
    case 7: return phi_level_t<T, 7>(*this);
    }
    return fail(HOS_E_INVALID, "level %d out of range", lev);
}
template int Eng<HOS_T>::phi_level(int);

template <class T, int LEV>
static int yc2r_phi_t(Eng<T> &e, const FieldMap &fm, int nf)
{
    hos_ctx *c = e.c;
    ProPhi<T, LEV> pro{e.phys(), c->ix, e.comp()};
    const int tx = c->tx[HOS_TILE_C2R];
    LOGN_SWITCH(c->logny, {
        TX_SWITCH(T, tx, {
            dim3 grid(c->p.Nx / TXV, nf, e.nm);
            // the c2r bytes of nf fields, and the r2c's: LEV inputs (field 0
            // comes from LDS) and its output
            Prof pf(c, e.S(), HOS_K_YC2R_PHI,
                    (double)e.nm * sizeof(T) *
                        (nf * (2.0 * c->Nk * c->p.Nx + c->p.Nx * (double)c->p.Ny) +
                         LEV * (double)c->p.Nx * c->p.Ny + 2.0 * c->fs));
            hipLaunchKernelGGL((k_yc2r_phi<T, LG, TXV, LEV>), grid, dim3(NT), 0, e.S(), e.I(), e.sI(), c->fsI, e.P(),
                               e.sP(), c->fsP, fm, (int)c->ip, e.twy(), c->Nk, pro, 0);
        })
    });
    return HOS_OK;
}

template <class T>
int Eng<T>::yc2r_phi(const FieldMap &fm, int nf, int lev)
{
    switch (lev) {
    case 2: return yc2r_phi_t<T, 2>(*this, fm, nf);
    case 3: return yc2r_phi_t<T, 3>(*this, fm, nf);
    case 4: return yc2r_phi_t<T, 4>(*this, fm, nf);
    case 5: return yc2r_phi_t<T, 5>(*this, fm, nf);
    case 6: return yc2r_phi_t<T, 6>(*this, fm, nf);
    case 7: return yc2r_phi_t<T, 7>(*this, fm, nf);
    }
    return fail(HOS_E_INVALID, "fused level %d out of range", lev);
}
template int Eng<HOS_T>::yc2r_phi(const FieldMap &, int, int);
#else
template <class T, int MO, int PLAN>
static int rhs_pair_t(Eng<T> &e)
{
    return e.template yr2c_rhs<MO, PLAN>();
}

template <class T, int PLAN>
static int rhs_pair_m(Eng<T> &e)
{
    switch (e.c->M) {
    case 1: return rhs_pair_t<T, 1, PLAN>(e);
    case 2: return rhs_pair_t<T, 2, PLAN>(e);
    case 3: return rhs_pair_t<T, 3, PLAN>(e);
    case 4: return rhs_pair_t<T, 4, PLAN>(e);
    case 5: return rhs_pair_t<T, 5, PLAN>(e);
    case 6: return rhs_pair_t<T, 6, PLAN>(e);
    case 7: return rhs_pair_t<T, 7, PLAN>(e);
    case 8: return rhs_pair_t<T, 8, PLAN>(e);
    }
    return fail(HOS_E_INVALID, "order %d out of range", e.c->M);
}

#if HOS_PART == 1
// plan 0 here; the other plans (RHS pass at Ny >= 512) in part 2
template <class T>
int Eng<T>::rhs_pair()
{
    if (c->plan != 0) return rhs_pair_plan();
    return rhs_pair_m<T, 0>(*this);
}
template int Eng<HOS_T>::rhs_pair();
#else
template <class T>
int Eng<T>::rhs_pair_plan()
{
    static_assert(NPLANS == 4, "plans 1..3");
    switch (c->plan) {
    case 1: return rhs_pair_m<T, 1>(*this);
    case 2: return rhs_pair_m<T, 2>(*this);
    case 3: return rhs_pair_m<T, 3>(*this);
    }
    return fail(HOS_E_INVALID, "FFT plan %d out of range", c->plan);
}
template int Eng<HOS_T>::rhs_pair_plan();
#endif
#endif
