// FP64 HIP kernels for gfx950 of the dual-time backward-Euler iteration with
// the point-implicit LU-SGS (backwardEuler_LU_SGS, Time_BackwardEuler.f90:
// 58-324; implicit_time_method = 3).
//
// Per inner iteration, after TimeStep, reconstruction and flux_sourceterm:
//   k_be_prepare  the diagonal dd = sum 0.5*alpaF + vol*(1/step + 1/dt), the
//                 dual-time right-hand side hWA - vol/dt*WA + fixed (fixed =
//                 vol/dt*WA saved at the first inner iteration)
//   k_be_factor   every cell's Nvar x Nvar block dU/dPA*dd - vol*d(omega)/dPA,
//                 equilibrated and LU-factored as LAPACK's dgesvx with
//                 FACT='E' does (advanced_linear_solve_lapack,
//                 la_math.f90:120-213: dgeequ, dlaqge, dgetf2), all cells at
//                 once: the block depends only on the cell's own state, so
//                 both sweeps solve with the same factors
//   k_be_sweep    forward levels, then backward levels: eight lanes per
//                 cell (one per row of its system) add the linearised Rusanov
//                 terms of its lower (forward) or upper (backward)
//                 neighbours, solve with the stored factors (dgetrs) and
//                 refine (dgerfs: at most 5 steps while the componentwise
//                 backward error exceeds eps and halves)
//   k_be_update   PA += dPA, WA = U(PA), the cell-average cache
//
// Each value is formed with the operations, in the order, of the restated
// routines, so each cell's solution is that of the serial dgesvx call.
#include "rc_items.h"

namespace rc {

#define FIX_(k, i) S.fixed[(k) * M.sc + (i)]

constexpr int kBeLanes = 8;                       // lanes per cell in the sweeps (one per row)
constexpr int kPiCells = 32;                      // cells per thread block of a sweep level
constexpr int kPiThreads = kPiCells * kBeLanes;

// stored factors of cell c (arrays of sc): the equilibrated block AE, its LU
// factors (rows in pivoted order), row and column scales, pivots and flags
// (bit 0 rows scaled, bit 1 columns scaled, bit 2 singular)
constexpr int kBeAE = 0, kBeLU = kNvar * kNvar, kBeRS = 2 * kNvar * kNvar, kBeCS = kBeRS + kNvar;
constexpr int kBeFac = kBeCS + kNvar;          // doubles per cell
constexpr int kBePiv = kNvar + 1;              // ints per cell: pivots, flags
constexpr int kBeNb = 8;                        // cache record per cell: rr, Rcpcv, Cp, hh, h_k (4 species)
static_assert(kBeNb >= 4 + kNsp, "backward-Euler cache record too small");
static_assert(kBeLU == kBeAE + kNvar * kNvar, "AE and LU are contiguous");
#define BEF_(k, c) fac[(size_t)(k) * sc + (c)]
#define BEP_(k, c) piv[(size_t)(k) * sc + (c)]

// LAPACK machine constants (dlamch): eps = 2^-53, safe minimum = 2^-1022
constexpr double kEps = 1.1102230246251565e-16;
constexpr double kSfmin = 2.2250738585072014e-308;
constexpr double kBignum = 1.0 / kSfmin;
constexpr double kSmallEq = kSfmin / (2 * kEps);  // dlaqge SMALL = safe minimum / precision
constexpr double kLargeEq = 1.0 / kSmallEq;
constexpr double kThresh = 0.1;

// the neighbour term of backwardEuler_LU_SGS (Time_BackwardEuler.f90:
// 131-176 forward, 222-267 backward): the neighbour's primitive increment
// dpa -> conserved increment at its state pa, then the linearised Rusanov
// flux change 0.5*(-alpaF*dU + dF(dU).S).  The neighbour's gas parameters
// and enthalpies come from its cache record nc (rr, Rcpcv, Cp, hh, h_k),
// formed from the same state by k_be_factor with the same arithmetic.
template <int NSP, class TT>
__device__ __forceinline__ void be_term(TT &th, const double *nc, const double *pa, const double *dpa, double dx,
                                        double dy, double af, double *c) {
    const double pp = pa[0], uu = pa[1], vv = pa[2], tt = pa[3];
    double yy[NSP];
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) yy[k] = pa[kNFlow + k];
    const double rr = nc[0], rcp = nc[1], cp = nc[2], hh = nc[3];
    const double rou = rr * uu, rov = rr * vv;
    const double roe = rr * (hh + 0.5 * (uu * uu + vv * vv)) - pp;
    const double dpp = dpa[0], duu = dpa[1], dvv = dpa[2], dtt = dpa[3];
    double dyy[NSP], s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) { dyy[k] = dpa[kNFlow + k]; s = s + dyy[k]; }
    dyy[NSP - 1] = -s;
    double s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) s1 = s1 + th.rsp[k] * dyy[k];
#pragma unroll
    for (int k = 0; k < NSP; ++k) s2 = s2 + nc[4 + k] * dyy[k];
    const double drr = (dpp - rr * rcp * dtt - rr * tt * s1) / (rcp * tt);
    const double dru = rr * duu + uu * drr;
    const double drv = rr * dvv + vv * drr;
    const double dre = rr * cp * dtt + rr * s2 + hh * drr + rr * (uu * duu + vv * dvv) +
                       0.5 * (uu * uu + vv * vv) * drr - dpp;
    const double un = uu * dx + vv * dy, dun = duu * dx + dvv * dy;
    c[0] = 0.5 * (-af * drr + dru * dx + drv * dy);
    c[1] = 0.5 * (-af * dru + dru * un + rou * dun + dpp * dx);
    c[2] = 0.5 * (-af * drv + drv * un + rov * dun + dpp * dy);
    c[3] = 0.5 * (-af * dre + dun * (roe + pp) + un * (dre + dpp));
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) {
        const double dry = rr * dyy[k] + yy[k] * drr;
        c[kNFlow + k] = 0.5 * (-af * dry + dry * un + rr * yy[k] * dun);
    }
}

// dkf/dPA of compute_derivative_production_rate (chemical_kinetics.f90:
// 167-265): model 0 the derivative of A exp(-E/RT) [H2]^0.2 [O2]^1.3, model
// 1 that of 1e4 [H2] above 710 K (the reference's model-1 branch has dkf/dT
// = 0 and leaves dkf/dY(3) unassigned; both are 0 here); zero where Y(1) <
// 1e-12.  rho is p/(R T) of the unclipped state, y the clipped fractions.
template <int NSP, class TT>
__device__ __forceinline__ void dkf_dpa(TT &th, double p, double T, double rho, double rcp, const double *y,
                                        double *dkf) {
#pragma unroll
    for (int k = 0; k < kNvar; ++k) dkf[k] = 0.0;
    if (y[0] < 1.0e-12) return;
    double af = th.a_f * exp(-th.e_act / (th.rconstant * T));
    const double r1 = y[0] * rho / (th.mw[0] * 1000), dr1dr = y[0] / (th.mw[0] * 1000),
                 dr1dy1 = rho / (th.mw[0] * 1000);
    const double r2 = y[1] * rho / (th.mw[1] * 1000), dr2dr = y[1] / (th.mw[1] * 1000),
                 dr2dy2 = rho / (th.mw[1] * 1000);
    const double drdp = rho / p, drdT = -rho / T;
    double drdy[NSP - 1];
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) drdy[k] = -(th.rsp[k] - th.rsp[NSP - 1]) * rho / rcp;
    double dkfdr, dkfdp, dkfdT, dkfdy[NSP - 1];
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) dkfdy[k] = 0.0;
    if (th.model == 0) {
        const double dter1_dr = 0.2 * pow(rho, -0.8) * pow(dr1dr, 0.2);
        dkfdr = af * (pow(r1, 0.2) * 1.3 * pow(r2, 0.3) * dr2dr + dter1_dr * pow(r2, 1.3));
        dkfdp = dkfdr * drdp;
        dkfdT = af * pow(r1, 0.2) * pow(r2, 1.3) * th.e_act / (th.rconstant * T * T) + dkfdr * drdT;
        dkfdy[0] = af * 0.2 * pow(r1, -0.8) * dr1dy1 * pow(r2, 1.3) + dkfdr * drdy[0];
        dkfdy[1] = af * pow(r1, 0.2) * 1.3 * pow(r2, 0.3) * dr2dy2 + dkfdr * drdy[1];
#pragma unroll
        for (int k = 2; k < NSP - 1; ++k) dkfdy[k] = dkfdr * drdy[k];
    } else {
        // integer powers as the reference's x**0 = 1, x**1 = x
        af = 0.0;
        if (T > 710.0) af = 1e4;
        const double dter1_dr = 1 * 1.0 * dr1dr;
        dkfdr = af * (r1 * 0 * 1.0 * dr2dr + dter1_dr * 1.0);
        dkfdp = dkfdr * drdp;
        dkfdT = 0.;
        dkfdy[0] = af * 1 * 1.0 * dr1dy1 * 1.0 + dkfdr * drdy[0];
        dkfdy[1] = af * r1 * 0 * 1.0 * dr2dy2 + dkfdr * drdy[1];
    }
    dkf[0] = dkfdp;
    dkf[1] = 0.;
    dkf[2] = 0.;
    dkf[3] = dkfdT;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) dkf[kNFlow + k] = dkfdy[k];
}

// dd, fixed and the dual-time right-hand side (Time_BackwardEuler.f90:73-100)
// With wp (BDF2_LU_SGS, Time_BDF2.f90:88-110): fixed from the states at the
// start of the last two steps, wp and wp2 (Nvar arrays of sc), weighted
// with phi; dd and the right-hand side carry (1+phi)/dt.
template <int NSP, bool B>
__global__ void __launch_bounds__(256) k_be_prepare(ITEM_ARGS, double dt, int first, const double *__restrict__ wp,
                                                    const double *__restrict__ wp2, double phi) {
    ITEM_PROLOGUE(i);
    double d = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int cf = M.cface[k * M.sc + i];
        const double al = ALPHA_(cf >= 0 ? cf : ~cf);
        S.lu_af[k * M.sc + i] = al;
        d = d + 0.5 * al;
    }
    const double v = M.vol[i];
    if (wp) {
        d = d + v * (1. / S.step[i] + (1 + phi) / dt);
        S.dd[i] = d;
        const RC_G double *p1 = gp(wp), *p2 = gp(wp2);
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            if (first) FIX_(k, i) = v / dt * ((1 + 2 * phi) * p1[k * M.sc + i] - phi * p2[k * M.sc + i]);
            HWA_(k, i) = HWA_(k, i) - (1 + phi) * v / dt * WA_(k, i) + FIX_(k, i);
        }
        return;
    }
    d = d + v * (1. / S.step[i] + 1. / dt);
    S.dd[i] = d;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double w = WA_(k, i);
        if (first) FIX_(k, i) = v / dt * w;
        HWA_(k, i) = HWA_(k, i) - v / dt * w + FIX_(k, i);
    }
}

// the block of cell i: dU/dPA (Compute_Con_to_Pri_Jaco,
// Con_to_Pri_Jacobian.f90:1-87) * dd, minus vol * d(omega)/dPA in the
// species rows (Time_BackwardEuler.f90:106-116); false if the temperature is
// out of range
template <int NSP, class TT>
__device__ __forceinline__ bool cell_block(TT &th, const double *pa, double d, double vol, double (&J)[kNvar][kNvar],
                                           double *nc) {
    constexpr int NV = kNvar;
    const double p = pa[0], u = pa[1], v = pa[2], T = pa[3];
    double y[NSP], yc[NSP], s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) { y[k] = pa[kNFlow + k]; s = s + y[k]; }
    y[NSP - 1] = 1.0 - s;
    double cp, gam, rcp, entha[NSP];
    bool ok = gas_param<NSP>(th, y, T, cp, gam, rcp);
    ok = enthalpy<NSP>(th, T, entha) && ok;
    const double rho = p / (rcp * T);
    // the cache record the neighbour terms read (Time_BackwardEuler.f90:
    // 139-152): rr, Rcpcv, Cp, hh = sum Y_k h_k, h_k
    {
        double hh = 0.0;
#pragma unroll
        for (int k = 0; k < NSP; ++k) hh = hh + y[k] * entha[k];
        nc[0] = rho;
        nc[1] = rcp;
        nc[2] = cp;
        nc[3] = hh;
#pragma unroll
        for (int k = 0; k < NSP; ++k) nc[4 + k] = entha[k];
    }
    double h = entha[NSP - 1], dhdy[NSP - 1];
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) {
        h = h + y[k] * (entha[k] - entha[NSP - 1]);
        dhdy[k] = entha[k] - entha[NSP - 1];
    }
    const double drdp = 1. / (rcp * T), drdt = -p / (rcp * T * T);
    double drdy[NSP - 1];
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) drdy[k] = -(th.rsp[k] - th.rsp[NSP - 1]) * p / (T * rcp * rcp);
    const double e2 = h + 0.5 * (u * u + v * v);
#pragma unroll
    for (int r = 0; r < NV; ++r)
#pragma unroll
        for (int k = 0; k < NV; ++k) J[r][k] = 0.0;
    J[0][0] = drdp; J[0][3] = drdt;
    J[1][0] = u * drdp; J[1][1] = rho; J[1][3] = u * drdt;
    J[2][0] = v * drdp; J[2][2] = rho; J[2][3] = v * drdt;
    J[3][0] = e2 * drdp - 1.; J[3][1] = rho * u; J[3][2] = rho * v; J[3][3] = rho * cp + e2 * drdt;
#pragma unroll
    for (int j = 0; j < NSP - 1; ++j) {
        J[0][kNFlow + j] = drdy[j];
        J[1][kNFlow + j] = u * drdy[j];
        J[2][kNFlow + j] = v * drdy[j];
        J[3][kNFlow + j] = e2 * drdy[j] + rho * dhdy[j];
    }
#pragma unroll
    for (int i = 0; i < NSP - 1; ++i) {
        J[kNFlow + i][0] = drdp * y[i];
        J[kNFlow + i][3] = drdt * y[i];
#pragma unroll
        for (int j = 0; j < NSP - 1; ++j) J[kNFlow + i][kNFlow + j] = i == j ? rho + y[i] * drdy[j] : y[i] * drdy[j];
    }
#pragma unroll
    for (int r = 0; r < NV; ++r)
#pragma unroll
        for (int k = 0; k < NV; ++k) J[r][k] = J[r][k] * d;
    // compute_derivative_production_rate's state: rho of the unclipped
    // fractions, then the fractions clipped to [0, 1]
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) yc[k] = fmin(1., fmax(0., y[k]));
    double dkf[NV];
    dkf_dpa<NSP>(th, p, T, rho, rcp, yc, dkf);
#pragma unroll
    for (int sp = 0; sp < NSP - 1; ++sp) {
        const double wfac = 1000 * th.mw[sp] * th.dnu[sp];
#pragma unroll
        for (int k = 0; k < NV; ++k) J[kNFlow + sp][k] = J[kNFlow + sp][k] - vol * (wfac * dkf[k]);
    }
    return ok;
}

// k_be_factor: dgeequ + dlaqge + dgetf2 of every cell's block
template <int NSP, bool B>
__global__ void __launch_bounds__(256) k_be_factor(ITEM_ARGS, double *__restrict__ fac_, int *__restrict__ piv_,
                                                   double *__restrict__ nbc_) {
    ITEM_PROLOGUE_LDS(i);
    constexpr int N = NV;
    const DevThermo &th = TH;
    const int sc = M.sc;
    RC_G double *fac = gp(fac_);
    RC_G int *piv = gp(piv_);
    double pa[NV], a[N][N], nc[kBeNb];
#pragma unroll
    for (int k = 0; k < NV; ++k) pa[k] = PA_(k, i);
    if (!cell_block<NSP>(th, pa, S.dd[i], M.vol[i], a, nc)) flag_error(S.err, RC_ERR_TEMPERATURE, i, K_BE_FWD);
    RC_G double *nbc = gp(nbc_) + (size_t)i * kBeNb;
#pragma unroll
    for (int k = 0; k < kBeNb; ++k) nbc[k] = nc[k];
    int flags = 0;
    // dgeequ
    double r[N], c[N];
    double rcmax = 0.0, rcmin = kBignum;
#pragma unroll
    for (int ii = 0; ii < N; ++ii) {
        r[ii] = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) r[ii] = fmax(r[ii], fabs(a[ii][j]));
        rcmax = fmax(rcmax, r[ii]);
        rcmin = fmin(rcmin, r[ii]);
    }
    const double amax = rcmax;
    bool singular = rcmin == 0.0;
#pragma unroll
    for (int ii = 0; ii < N; ++ii) r[ii] = 1.0 / fmin(fmax(r[ii], kSfmin), kBignum);
    const double rowcnd = fmax(rcmin, kSfmin) / fmin(rcmax, kBignum);
    double ccmin = kBignum, ccmax = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        c[j] = 0.0;
#pragma unroll
        for (int ii = 0; ii < N; ++ii) c[j] = fmax(c[j], fabs(a[ii][j]) * r[ii]);
        ccmin = fmin(ccmin, c[j]);
        ccmax = fmax(ccmax, c[j]);
    }
    singular = singular || ccmin == 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) c[j] = 1.0 / fmin(fmax(c[j], kSfmin), kBignum);
    const double colcnd = fmax(ccmin, kSfmin) / fmin(ccmax, kBignum);
    // dlaqge
    bool rows = false, cols;
    if (rowcnd >= kThresh && amax >= kSmallEq && amax <= kLargeEq) {
        cols = colcnd < kThresh;
    } else {
        rows = true;
        cols = colcnd < kThresh;
    }
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
        for (int ii = 0; ii < N; ++ii) {
            if (rows && cols) a[ii][j] = c[j] * r[ii] * a[ii][j];
            else if (rows) a[ii][j] = r[ii] * a[ii][j];
            else if (cols) a[ii][j] = c[j] * a[ii][j];
        }
#pragma unroll
    for (int ii = 0; ii < N; ++ii)
#pragma unroll
        for (int j = 0; j < N; ++j) BEF_(kBeAE + ii * N + j, i) = a[ii][j];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        BEF_(kBeRS + k, i) = r[k];
        BEF_(kBeCS + k, i) = c[k];
    }
    flags = (rows ? 1 : 0) | (cols ? 2 : 0);
    // dgetf2: partial pivoting (the first row of largest |a(i,j)|), rows
    // exchanged in registers by selects
#pragma unroll
    for (int j = 0; j < N; ++j) {
        int p = j;
        double pmax = fabs(a[j][j]);
#pragma unroll
        for (int ii = j + 1; ii < N; ++ii) {
            const double t = fabs(a[ii][j]);
            if (t > pmax) { pmax = t; p = ii; }
        }
        BEP_(j, i) = p;
        double pval = a[j][j];
#pragma unroll
        for (int ii = j + 1; ii < N; ++ii) pval = p == ii ? a[ii][j] : pval;
        singular = singular || pval == 0.0;
#pragma unroll
        for (int ii = j + 1; ii < N; ++ii) {
            const bool sw = p == ii;
#pragma unroll
            for (int k = 0; k < N; ++k) {
                const double t = a[j][k];
                a[j][k] = sw ? a[ii][k] : t;
                a[ii][k] = sw ? t : a[ii][k];
            }
        }
        if (fabs(a[j][j]) >= kSfmin) {
            const double rp = 1.0 / a[j][j];
#pragma unroll
            for (int ii = j + 1; ii < N; ++ii) a[ii][j] = rp * a[ii][j];
        } else {
#pragma unroll
            for (int ii = j + 1; ii < N; ++ii) a[ii][j] = a[ii][j] / a[j][j];
        }
#pragma unroll
        for (int k = j + 1; k < N; ++k) {
            const double t = -a[j][k];
#pragma unroll
            for (int ii = j + 1; ii < N; ++ii) a[ii][k] = a[ii][k] + t * a[ii][j];
        }
    }
#pragma unroll
    for (int ii = 0; ii < N; ++ii)
#pragma unroll
        for (int j = 0; j < N; ++j) BEF_(kBeLU + ii * N + j, i) = a[ii][j];
    if (singular) {
        flags |= 4;
        flag_error(S.err, RC_ERR_SINGULAR, i, K_BE_FWD);
    }
    BEP_(N, i) = flags;
}

// dgetrs on a vector held one element per lane (lane r of the cell's group
// holds element r; src = the lane whose element the pivot interchanges
// bring to row r; lu = row r of the LU factors).  Every element is formed
// with the operations, in the order, of the serial routine: the
// interchanges, then column-oriented unit-lower and upper substitution,
// a column skipped where its solution element is zero.
template <int N>
__device__ __forceinline__ double be_getrs_rows(double v, int r, int src, const double (&lu)[N]) {
    v = __shfl(v, src, kBeLanes);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const double vk = __shfl(v, k, kBeLanes);
        if (vk != 0.0 && r > k) v = v - vk * lu[k];
    }
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
        const int nz = v != 0.0;
        if (r == k && nz) v = v / lu[k];
        const double vk = __shfl(v, k, kBeLanes);
        if (__shfl(nz, k, kBeLanes) && r < k) v = v - vk * lu[k];
    }
    return v;
}

// one level of the forward (FWD) or backward point-implicit sweep
// (Time_BackwardEuler.f90:102-193 / 195-285) over cells pi_cells[t] of the
// work item: kBeLanes lanes per cell, lane r owning row r of the cell's
// Nvar x Nvar system (lane Nvar has no row and writes nothing).  Lanes 0..2
// form the terms of the cell's three faces; the rows gather them in face
// order, then solve with the stored factors (dgetrs) and refine (dgerfs)
// row-parallel, the broadcasts by shuffles within the group.
template <int NSP, bool FWD, bool B>
__global__ void __launch_bounds__(kPiThreads) k_be_sweep(ITEM_ARGS, const double *__restrict__ fac_,
                                                        const int *__restrict__ piv_,
                                                        const double *__restrict__ nbc_) {
    __shared__ DevThermo th_lds_;
    const RC_G CaseDev *Cp_;
    int t, tend;
    if constexpr (B) {
        const int4 w_ = work[blockIdx.x];
        Cp_ = gp(cases) + w_.x;
        t = w_.y + (int)threadIdx.x / kBeLanes;
        tend = w_.z;
    } else {
        Cp_ = gp(cases);
        t = lo + (int)(blockIdx.x * blockDim.x + threadIdx.x) / kBeLanes;
        tend = hi;
    }
    const RC_G CaseDev &C = *Cp_;
    {
        const RC_G double *src_ = (const RC_G double *)gp(C.th);
        double *dst_ = reinterpret_cast<double *>(&th_lds_);
        for (int k_ = threadIdx.x; k_ < (int)(sizeof(DevThermo) / 8); k_ += blockDim.x) dst_[k_] = src_[k_];
    }
    __syncthreads();
    if (t >= tend) return;               // the group's lanes share t: groups leave whole
    const GMesh M = gview(C.M);
    const GState S = gview(C.S);
    const DevThermo &th = th_lds_;
    constexpr int N = NV;
    static_assert(N < kBeLanes, "one lane per row");
    const int r = (int)threadIdx.x % kBeLanes;
    const bool row = r < N;
    const int rr = row ? r : N - 1;      // the row-less lane mirrors row N-1's loads
    const int sc = M.sc;
    const RC_G double *fac = gp(fac_);
    const RC_G int *piv = gp(piv_);
    const RC_G double *nbc = gp(nbc_);
    const int i = M.pi_cells[t];
    const int flags = BEP_(N, i);
    if (flags & 4) return;               // singular block: reported by k_be_factor
    double lu[N], ae[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        lu[j] = BEF_(kBeLU + rr * N + j, i);
        ae[j] = BEF_(kBeAE + rr * N + j, i);
    }
    // the pivot interchanges as a lane permutation: row r of the pivoted
    // vector is element src of the unpivoted one
    int src = r;
    {
        int idx[N];
#pragma unroll
        for (int k = 0; k < N; ++k) idx[k] = k;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const int p = BEP_(j, i);
#pragma unroll
            for (int ii = j + 1; ii < N; ++ii) {
                const bool sw = p == ii;
                const int tj = idx[j];
                idx[j] = sw ? idx[ii] : tj;
                idx[ii] = sw ? tj : idx[ii];
            }
        }
#pragma unroll
        for (int k = 0; k < N; ++k) src = r == k ? idx[k] : src;
    }
    // lane ks < 3: the term of face ks when that neighbour takes part
    double c[N];
#pragma unroll
    for (int k = 0; k < N; ++k) c[k] = 0.0;
    int use = 0;
    if (r < 3) {
        const int nr = M.lu_nb[r * sc + i];
        use = nr >= 0 && (FWD ? nr < i : nr > i);
        if (use) {
            double pj[N], dj[N], nc[kBeNb];
#pragma unroll
            for (int k = 0; k < N; ++k) { pj[k] = PA_(k, nr); dj[k] = HWA_(k, nr); }
#pragma unroll
            for (int k = 0; k < kBeNb; ++k) nc[k] = nbc[(size_t)nr * kBeNb + k];
            be_term<NSP>(th, nc, pj, dj, M.lu_n[(r * 2) * sc + i], M.lu_n[(r * 2 + 1) * sc + i],
                         S.lu_af[r * sc + i], c);
        }
    }
    double dflux = 0.0;
#pragma unroll
    for (int ks = 0; ks < 3; ++ks) {
        const int u = __shfl(use, ks, kBeLanes);
        double mine = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const double v = __shfl(c[k], ks, kBeLanes);
            mine = r == k ? v : mine;
        }
        dflux = u ? dflux + mine : dflux;
    }
    const double hw = HWA_(rr, i);
    const double rhs = FWD ? hw - dflux : dflux;
    const double b = (flags & 1) ? BEF_(kBeRS + rr, i) * rhs : rhs;
    double x = be_getrs_rows<N>(b, r, src, lu);
    // dgerfs
    constexpr double safe1 = (N + 1) * kSfmin, safe2 = safe1 / kEps;
    double lstres = 3.0;
    for (int count = 1; count <= 5; ++count) {
        double xs[N], w = b, den = fabs(b);
#pragma unroll
        for (int j = 0; j < N; ++j) xs[j] = __shfl(x, j, kBeLanes);
#pragma unroll
        for (int j = 0; j < N; ++j) w = w + (-xs[j]) * ae[j];
#pragma unroll
        for (int j = 0; j < N; ++j) den = den + fabs(ae[j]) * fabs(xs[j]);
        double berr = row ? (den > safe2 ? fabs(w) / den : (fabs(w) + safe1) / (den + safe1)) : 0.0;
#pragma unroll
        for (int m = 1; m < kBeLanes; m <<= 1) berr = fmax(berr, __shfl_xor(berr, m, kBeLanes));
        if (!(berr > kEps && 2.0 * berr <= lstres)) break;
        w = be_getrs_rows<N>(w, r, src, lu);
        x = x + w;
        lstres = berr;
    }
    if (flags & 2) x = BEF_(kBeCS + rr, i) * x;
    if (row) HWA_(r, i) = FWD ? x : hw - x;
}

// PA += dPA (Time_BackwardEuler.f90:288-292), primitve_to_conserved_variables
// (update_flow_variables.f90:50-84) and the cell-average cache of the new state
template <int NSP, bool B>
__global__ void __launch_bounds__(256) k_be_update(ITEM_ARGS) {
    ITEM_PROLOGUE_LDS(i);
    const DevThermo &th = TH;
    double pa[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        pa[k] = PA_(k, i) + HWA_(k, i);
        PA_(k, i) = pa[k];
    }
    if (!cell_prim_to_cons<NSP>(M, S, th, i, pa)) flag_error(S.err, RC_ERR_TEMPERATURE, i, K_BE_UPDATE);
}

void launch_be_prepare(const LaunchSet &L, double dt, bool first, const double *wp, const double *wp2, double phi,
                       hipStream_t st) {
    const int f = first ? 1 : 0;
    if (L.single) {
        const int n = L.cell.hi - L.cell.lo;
        if (n > 0)
            k_be_prepare<kNsp, false><<<(n + kChunk - 1) / kChunk, kChunk, 0, st>>>(L.cases, nullptr, L.cell.lo,
                                                                                    L.cell.hi, dt, f, wp, wp2, phi);
    } else if (L.cell.n) {
        k_be_prepare<kNsp, true><<<L.cell.n, kChunk, 0, st>>>(L.cases, L.cell.w, 0, 0, dt, f, wp, wp2, phi);
    }
}

void launch_be_factor(const LaunchSet &L, double *fac, int *piv, double *nbc, hipStream_t st) {
    if (L.single) {
        const int n = L.cell.hi - L.cell.lo;
        if (n > 0)
            k_be_factor<kNsp, false><<<(n + kChunk - 1) / kChunk, kChunk, 0, st>>>(L.cases, nullptr, L.cell.lo,
                                                                                   L.cell.hi, fac, piv, nbc);
    } else if (L.cell.n) {
        k_be_factor<kNsp, true><<<L.cell.n, kChunk, 0, st>>>(L.cases, L.cell.w, 0, 0, fac, piv, nbc);
    }
}

void launch_be_level(const LaunchSet &L, bool fwd, int l, const double *fac, const int *piv, const double *nbc,
                     hipStream_t st) {
    const WorkList &w = L.pi_level[l];
    if (L.single) {
        const int n = w.hi - w.lo;
        if (n <= 0) return;
        const int nb = (n + kPiCells - 1) / kPiCells;
        if (fwd) k_be_sweep<kNsp, true, false><<<nb, kPiThreads, 0, st>>>(L.cases, nullptr, w.lo, w.hi, fac, piv, nbc);
        else k_be_sweep<kNsp, false, false><<<nb, kPiThreads, 0, st>>>(L.cases, nullptr, w.lo, w.hi, fac, piv, nbc);
    } else if (w.n) {
        if (fwd) k_be_sweep<kNsp, true, true><<<w.n, kPiThreads, 0, st>>>(L.cases, w.w, 0, 0, fac, piv, nbc);
        else k_be_sweep<kNsp, false, true><<<w.n, kPiThreads, 0, st>>>(L.cases, w.w, 0, 0, fac, piv, nbc);
    }
}

void launch_be_update(const LaunchSet &L, hipStream_t st) { LAUNCH(L.cell, kChunk, st, k_be_update, kNsp); }

int pi_chunk() { return kPiCells; }
size_t be_factor_doubles() { return kBeFac; }
size_t be_factor_ints() { return kBePiv; }
size_t be_cache_doubles() { return kBeNb; }

}  // namespace rc
