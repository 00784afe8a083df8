// hos_levels.hip -- the y r2c launches whose prologue is instantiated per
// HOS level (phi^(lev+1), lev = 1..7) or per order (the RHS pass
// k_yr2c_rhs, M = 1..8).
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
