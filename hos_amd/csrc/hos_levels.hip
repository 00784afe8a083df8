// hos_levels.hip -- the y r2c launches whose prologue is instantiated per
// HOS level (phi^(lev+1), lev = 1..7: k_yr2c, and k_yc2r_phi, which fuses it
// behind the previous level's y c2r) or per order (the RHS pass k_yr2c_rhs,
// M = 1..8).
// Compiled once per precision (HOS_T) and part (HOS_PART 0: levels, 1: the
// RHS pass with the default FFT plan, 2: its other plans, 3 and 4: the same
// over W^(a) fields after the diagonal passes, 5: the diagonal passes) by the
// Makefile, so the several hundred kernel instantiations build in parallel
// translation units.
#include "hos_ctx.hpp"

#ifndef HOS_T
#error "the Makefile sets HOS_T (float or double) and HOS_PART (0 .. 5)"
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
#elif HOS_PART <= 4
template <class T, int MO, int PLAN, bool WF>
static int rhs_pair_t(Eng<T> &e)
{
    return e.template yr2c_rhs<MO, PLAN, WF>();
}

template <class T, int PLAN, bool WF>
static int rhs_pair_m(Eng<T> &e)
{
    switch (e.c->M) {
    case 1: return rhs_pair_t<T, 1, PLAN, WF>(e);
    case 2: return rhs_pair_t<T, 2, PLAN, WF>(e);
    case 3: return rhs_pair_t<T, 3, PLAN, WF>(e);
    case 4: return rhs_pair_t<T, 4, PLAN, WF>(e);
    case 5: return rhs_pair_t<T, 5, PLAN, WF>(e);
    case 6: return rhs_pair_t<T, 6, PLAN, WF>(e);
    case 7: return rhs_pair_t<T, 7, PLAN, WF>(e);
    case 8: return rhs_pair_t<T, 8, PLAN, WF>(e);
    }
    return fail(HOS_E_INVALID, "order %d out of range", e.c->M);
}

// plans 1..3 of the RHS pass (Ny >= 512)
template <class T, bool WF>
static int rhs_pair_plans(Eng<T> &e)
{
    static_assert(NPLANS == 4, "plans 1..3");
    switch (e.c->plan) {
    case 1: return rhs_pair_m<T, 1, WF>(e);
    case 2: return rhs_pair_m<T, 2, WF>(e);
    case 3: return rhs_pair_m<T, 3, WF>(e);
    }
    return fail(HOS_E_INVALID, "FFT plan %d out of range", e.c->plan);
}

#if HOS_PART == 1
// plan 0 here; the other plans (RHS pass at Ny >= 512) in part 2, and the
// RHS pass over W^(a) fields (after the diagonal passes) in parts 3 and 4
template <class T>
int Eng<T>::rhs_pair()
{
    if (c->diag) return rhs_pair_diag();
    if (c->plan != 0) return rhs_pair_plan();
    return rhs_pair_m<T, 0, false>(*this);
}
template int Eng<HOS_T>::rhs_pair();
#elif HOS_PART == 2
template <class T>
int Eng<T>::rhs_pair_plan()
{
    return rhs_pair_plans<T, false>(*this);
}
template int Eng<HOS_T>::rhs_pair_plan();
#elif HOS_PART == 3
template <class T>
int Eng<T>::rhs_pair_diag()
{
    if (c->plan != 0) return rhs_pair_diag_plan();
    return rhs_pair_m<T, 0, true>(*this);
}
template int Eng<HOS_T>::rhs_pair_diag();
#else
template <class T>
int Eng<T>::rhs_pair_diag_plan()
{
    return rhs_pair_plans<T, true>(*this);
}
template int Eng<HOS_T>::rhs_pair_diag_plan();
#endif
#else
// HOS_PART 5: the diagonal passes k_ydiag, L = 1..7
template <class T, int LV>
static int ydiag_t(Eng<T> &e)
{
    hos_ctx *c = e.c;
    using DA = DiagArgs<LV>;
    DA da;
    double rows = 0.0;
    if (DA::ETA) {  // eta's x-intermediate (slot 1), a band of a masked state
        da.slot[0] = 1;
        da.nzr[0] = e.nzr();
        rows += da.nzr[0];
    }
    for (int k = 1; k <= LV; ++k) {
        // F_k = d^k phi^(a)/dz^k with a = LV+1-k; the level-0 ones (a = 1) are
        // a band of a masked state
        const int a = LV + 1 - k;
        da.slot[k - 1 + DA::ETA] = e.dg(a, k);
        da.nzr[k - 1 + DA::ETA] = a == 1 ? e.nzr() : c->Nk;
        rows += da.nzr[k - 1 + DA::ETA];
    }
    da.dst = 0;
    // the last pass's phi^(M) feeds only d phi^(M)/dz, which stays in x-spectral
    // space on the rows the mask keeps (Eng::diag_levels)
    da.nzo = LV == c->M - 1 ? c->my : c->Nk;
    da.wf = c->ix.w[LV];
    da.etaf = c->ix.eta;
    da.blk = e.xb;
    const int tx = c->tx[HOS_TILE_PAIR];
    LOGN_SWITCH(c->logny, {
        TX_SWITCH(T, tx, {
            constexpr int V = Vec<T>::N;
            constexpr int TXP0 = TXV / 2;
            constexpr int TXP = (TXP0 > 2 && TXP0 * (1 << LG) / V > 2 * NT_DIAG) ? 2 : TXP0;
            dim3 grid(c->p.Nx / TXP, 1, e.nm);
            // eta read (pass 1: written) and W^(L) written (physical), the
            // tiles read, phi^(L+1) written (spectral rows)
            Prof pf(c, e.S(), HOS_K_YDIAG,
                    (double)e.nm * sizeof(T) * c->p.Nx * (2.0 * c->p.Ny + 2.0 * rows + 2.0 * da.nzo));
            hipLaunchKernelGGL((k_ydiag<T, LG, TXP, LV>), grid, dim3(NT_DIAG), 0, e.S(), e.phys(), e.P(), c->ix,
                               (int)e.comp(), e.I(), e.sI(), c->fsI, da, (int)c->ip, e.twy());
        })
    });
    return HOS_OK;
}

template <class T>
int Eng<T>::ydiag(int L)
{
    static_assert(MAXM == 8, "diagonals 1..MAXM-1");
    switch (L) {
    case 1: return ydiag_t<T, 1>(*this);
    case 2: return ydiag_t<T, 2>(*this);
    case 3: return ydiag_t<T, 3>(*this);
    case 4: return ydiag_t<T, 4>(*this);
    case 5: return ydiag_t<T, 5>(*this);
    case 6: return ydiag_t<T, 6>(*this);
    case 7: return ydiag_t<T, 7>(*this);
    }
    return fail(HOS_E_INVALID, "diagonal %d out of range", L);
}
template int Eng<HOS_T>::ydiag(int);
#endif
