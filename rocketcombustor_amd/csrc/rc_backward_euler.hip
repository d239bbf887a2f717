// FP64 HIP kernels for gfx950 of the dual-time backward-Euler iteration with
// the point-implicit LU-SGS (backwardEuler_LU_SGS, Time_BackwardEuler.f90:
// 58-324; implicit_time_method = 3).
//
// Per inner iteration, after TimeStep, reconstruction and flux_sourceterm:
//   k_be_prepare  the diagonal dd = sum 0.5*alpaF + vol*(1/step + 1/dt), the
//                 dual-time right-hand side hWA - vol/dt*WA + fixed (fixed =
//                 vol/dt*WA saved at the first inner iteration)
//   k_be_sweep    forward levels, then backward levels: each cell solves its
//                 Nvar x Nvar block  dU/dPA*dd - vol*d(omega)/dPA  for the
//                 primitive increment, with the linearised Rusanov terms of
//                 its lower (forward) or upper (backward) neighbours
//   k_be_update   PA += dPA, WA = U(PA), the cell-average cache
//
// A cell of the sweep is solved by a group of 8 lanes, lane r owning row r
// of the block: lanes 0-2 form the terms of the cell's three face slots at
// once, every lane forms its row of the block, and the group solves the
// system as LAPACK's dgesvx with FACT='E' does (advanced_linear_solve_lapack,
// la_math.f90:120-213): equilibration (dgeequ, dlaqge), LU with partial
// pivoting, triangular solves and iterative refinement (dgerfs), unscaling.
// Row operations are lane-local; pivot search, row exchanges, pivot-row
// broadcasts and the solution vector go by shuffles within the group.  Each
// value is formed with the operations, in the order, of the restated
// routines, so the result is that of the serial solve.
#include "rc_items.h"

namespace rc {

#define FIX_(k, i) S.fixed[(k) * M.sc + (i)]

constexpr int kPiLanes = 8;                       // lanes per cell
constexpr int kPiChunk = kChunk / kPiLanes;       // cells per thread block
static_assert(kNvar <= kPiLanes, "one lane per row of the block");

// LAPACK machine constants (dlamch): eps = 2^-53, safe minimum = 2^-1022
constexpr double kEps = 1.1102230246251565e-16;
constexpr double kSfmin = 2.2250738585072014e-308;
constexpr double kBignum = 1.0 / kSfmin;
constexpr double kSmallEq = kSfmin / (2 * kEps);  // dlaqge SMALL = safe minimum / precision
constexpr double kLargeEq = 1.0 / kSmallEq;
constexpr double kThresh = 0.1;

__device__ __forceinline__ double gsh(double v, int src) { return __shfl(v, src, kPiLanes); }
__device__ __forceinline__ int gsh(int v, int src) { return __shfl(v, src, kPiLanes); }
__device__ __forceinline__ double gmax(double v) {
#pragma unroll
    for (int o = 1; o < kPiLanes; o <<= 1) v = fmax(v, __shfl_xor(v, o, kPiLanes));
    return v;
}
__device__ __forceinline__ double gmin(double v) {
#pragma unroll
    for (int o = 1; o < kPiLanes; o <<= 1) v = fmin(v, __shfl_xor(v, o, kPiLanes));
    return v;
}
template <int N>
__device__ __forceinline__ double pick(const double (&a)[N], int k) {
    double v = a[0];
#pragma unroll
    for (int j = 1; j < N; ++j) v = k == j ? a[j] : v;
    return v;
}

// dgetrs of the group's LU factors on one right-hand side held by row
// (lane r: entry r of the original row order): the row interchanges as one
// permutation (src = original row now in lane r), unit-lower then upper
// triangular solves.  Returns entry r of the solution.
template <int N>
__device__ __forceinline__ double group_getrs(int r, const double (&lu)[N], const double (&ukk)[N], int src,
                                              double v) {
    v = gsh(v, src);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const double vk = gsh(v, k);
        if (vk != 0.0 && r > k && r < N) v = v - vk * lu[k];
    }
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
        double vk = gsh(v, k);
        if (vk != 0.0) {
            vk = vk / ukk[k];
            if (r == k) v = vk;
            else if (r < k) v = v - vk * lu[k];
        }
    }
    return v;
}

// dgesvx('E', 'N', N, 1, A, ...) by a group of 8 lanes: lane r < N holds row
// r of A (arow) and entry r of b; returns entry r of x in x.  False when A
// has a zero row or column or a zero pivot (dgesvx INFO in 1..N); the
// result is then not formed.
template <int N>
__device__ bool group_dgesvx(int r, const double (&arow)[N], double b, double &x) {
    const bool row = r < N;
    double a[N];
#pragma unroll
    for (int k = 0; k < N; ++k) a[k] = row ? arow[k] : 0.0;
    b = row ? b : 0.0;
    // dgeequ: row scale factors, then column scale factors of the row-scaled A
    double rs = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) rs = fmax(rs, fabs(a[k]));
    const double rcmax = gmax(row ? rs : 0.0);
    const double rcmin = gmin(row ? rs : kBignum);
    const double amax = rcmax;
    if (rcmin == 0.0) return false;
    rs = 1.0 / fmin(fmax(rs, kSfmin), kBignum);
    const double rowcnd = fmax(rcmin, kSfmin) / fmin(rcmax, kBignum);
    double cs[N];
    double ccmin = kBignum, ccmax = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        cs[j] = gmax(row ? fabs(a[j]) * rs : 0.0);
        ccmin = fmin(ccmin, cs[j]);
        ccmax = fmax(ccmax, cs[j]);
    }
    if (ccmin == 0.0) return false;
#pragma unroll
    for (int j = 0; j < N; ++j) cs[j] = 1.0 / fmin(fmax(cs[j], kSfmin), kBignum);
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
    for (int j = 0; j < N; ++j) {
        if (rows && cols) a[j] = cs[j] * rs * a[j];
        else if (rows) a[j] = rs * a[j];
        else if (cols) a[j] = cs[j] * a[j];
    }
    if (rows) b = rs * b;
    // dgetf2: the group's rows move with the interchanges; src tracks the
    // original row held by each lane
    double lu[N];
#pragma unroll
    for (int k = 0; k < N; ++k) lu[k] = a[k];
    int src = r;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        // pivot: the first row (from j on) of largest |a(i,j)|
        double pv = (r >= j && row) ? fabs(lu[j]) : -1.0;
        int pi = r;
#pragma unroll
        for (int o = 1; o < kPiLanes; o <<= 1) {
            const double ov = __shfl_xor(pv, o, kPiLanes);
            const int oi = __shfl_xor(pi, o, kPiLanes);
            if (ov > pv || (ov == pv && oi < pi)) { pv = ov; pi = oi; }
        }
        const int p = pi;
        if (gsh(lu[j], p) == 0.0) return false;
        if (p != j) {
            const int partner = r == j ? p : (r == p ? j : r);
#pragma unroll
            for (int k = 0; k < N; ++k) lu[k] = gsh(lu[k], partner);
            src = gsh(src, partner);
        }
        double prow[N];
#pragma unroll
        for (int k = j; k < N; ++k) prow[k] = gsh(lu[k], j);
        if (r > j && row) {
            double l;
            if (fabs(prow[j]) >= kSfmin) {
                const double rp = 1.0 / prow[j];
                l = rp * lu[j];
            } else {
                l = lu[j] / prow[j];
            }
            lu[j] = l;
#pragma unroll
            for (int k = j + 1; k < N; ++k) lu[k] = lu[k] + (-prow[k]) * l;
        }
    }
    double ukk[N];
#pragma unroll
    for (int k = 0; k < N; ++k) ukk[k] = gsh(lu[k], k);
    // dgetrs, then dgerfs: at most 5 refinement steps while the componentwise
    // backward error exceeds eps and halves
    double xv = group_getrs<N>(r, lu, ukk, src, b);
    const double safe1 = (N + 1) * kSfmin, safe2 = safe1 / kEps;
    double lstres = 3.0;
    for (int count = 1; count <= 5; ++count) {
        double w = b, den = fabs(b);
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const double xj = gsh(xv, j);
            w = w + (-xj) * a[j];
        }
#pragma unroll
        for (int j = 0; j < N; ++j) den = den + fabs(a[j]) * fabs(gsh(xv, j));
        const double e = !row ? 0.0 : (den > safe2 ? fabs(w) / den : (fabs(w) + safe1) / (den + safe1));
        const double berr = gmax(e);
        if (!(berr > kEps && 2.0 * berr <= lstres)) break;
        const double dx = group_getrs<N>(r, lu, ukk, src, w);
        xv = xv + dx;
        lstres = berr;
    }
    if (cols) xv = pick(cs, r) * xv;
    x = xv;
    return true;
}

// the neighbour term of backwardEuler_LU_SGS (Time_BackwardEuler.f90:
// 131-176 forward, 222-267 backward): the neighbour's primitive increment
// dpa -> conserved increment at its state pa, then the linearised Rusanov
// flux change 0.5*(-alpaF*dU + dF(dU).S); false if pa's temperature is out
// of range
template <int NSP, class TT>
__device__ __forceinline__ bool be_term(TT &th, const double *pa, const double *dpa, double dx, double dy,
                                        double af, double *c) {
    const double pp = pa[0], uu = pa[1], vv = pa[2], tt = pa[3];
    double yy[NSP], s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) { yy[k] = pa[kNFlow + k]; s = s + yy[k]; }
    yy[NSP - 1] = 1.0 - s;
    double cp, gam, rcp, entha[NSP];
    bool ok = gas_param<NSP>(th, yy, tt, cp, gam, rcp);
    ok = enthalpy<NSP>(th, tt, entha) && ok;
    double hh = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) hh = hh + yy[k] * entha[k];
    const double rr = pp / (rcp * tt), rou = rr * uu, rov = rr * vv;
    const double roe = rr * (hh + 0.5 * (uu * uu + vv * vv)) - pp;
    const double dpp = dpa[0], duu = dpa[1], dvv = dpa[2], dtt = dpa[3];
    double dyy[NSP];
    s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) { dyy[k] = dpa[kNFlow + k]; s = s + dyy[k]; }
    dyy[NSP - 1] = -s;
    double s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) s1 = s1 + th.rsp[k] * dyy[k];
#pragma unroll
    for (int k = 0; k < NSP; ++k) s2 = s2 + entha[k] * dyy[k];
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
    return ok;
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
template <int NSP, bool B>
__global__ void __launch_bounds__(256) k_be_prepare(ITEM_ARGS, double dt, int first) {
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
    d = d + v * (1. / S.step[i] + 1. / dt);
    S.dd[i] = d;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double w = WA_(k, i);
        if (first) FIX_(k, i) = v / dt * w;
        HWA_(k, i) = HWA_(k, i) - v / dt * w + FIX_(k, i);
    }
}

// one level of the forward (FWD) or backward point-implicit sweep
// (Time_BackwardEuler.f90:102-193 / 195-285) over cells pi_cells[t] of the
// work item, one 8-lane group per cell
template <int NSP, bool FWD, bool B>
__global__ void __launch_bounds__(256) k_be_sweep(ITEM_ARGS) {
    constexpr int G = kPiLanes;
    __shared__ DevThermo th_lds_;
    const RC_G CaseDev *Cp_;
    int t, tend;
    if constexpr (B) {
        const int4 w_ = work[blockIdx.x];
        Cp_ = gp(cases) + w_.x;
        t = w_.y + (int)threadIdx.x / G;
        tend = w_.z;
    } else {
        Cp_ = gp(cases);
        t = lo + (int)(blockIdx.x * blockDim.x + threadIdx.x) / G;
        tend = hi;
    }
    const RC_G CaseDev &C = *Cp_;
    {
        const RC_G double *src_ = (const RC_G double *)gp(C.th);
        double *dst_ = reinterpret_cast<double *>(&th_lds_);
        for (int k_ = threadIdx.x; k_ < (int)(sizeof(DevThermo) / 8); k_ += blockDim.x) dst_[k_] = src_[k_];
    }
    __syncthreads();
    if (t >= tend) return;             // the whole group
    const GMesh M = gview(C.M);
    const GState S = gview(C.S);
    const DevThermo &th = th_lds_;
    const int r = threadIdx.x % G;
    const int sc = M.sc;
    const int i = M.pi_cells[t];
    const int kerr = FWD ? K_BE_FWD : K_BE_BWD;

    // lanes 0-2: the term of face slot r, if its neighbour is on the swept side
    double c[NV];
    int use = 0;
    {
        const int ks = r < 3 ? r : 0;
        const int nr = M.lu_nb[ks * sc + i];
        use = (r < 3 && nr >= 0 && (FWD ? nr < i : nr > i)) ? 1 : 0;
        const int j = use ? nr : i;
        double pj[NV], dj[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) { pj[k] = PA_(k, j); dj[k] = HWA_(k, j); }
        const bool ok = be_term<NSP>(th, pj, dj, M.lu_n[(ks * 2) * sc + i], M.lu_n[(ks * 2 + 1) * sc + i],
                                     S.lu_af[ks * sc + i], c);
        if (use && !ok) flag_error(S.err, RC_ERR_TEMPERATURE, i, kerr);
    }
    // dflux(r): the terms in face order
    double dflux = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double mine = 0.0;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double x = gsh(c[v], k);
            mine = r == v ? x : mine;
        }
        dflux = dflux + (gsh(use, k) ? mine : 0.0);
    }

    // row r of the block: dU/dPA (Con_to_Pri_Jacobian.f90:1-87) * dd, minus
    // vol * d(omega)/dPA in the species rows
    double pa[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) pa[k] = PA_(k, i);
    const double p = pa[0], u = pa[1], v = pa[2], T = pa[3];
    double y[NSP], yc[NSP], s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) { y[k] = pa[kNFlow + k]; s = s + y[k]; }
    y[NSP - 1] = 1.0 - s;
    double cp, gam, rcp, entha[NSP];
    bool ok = gas_param<NSP>(th, y, T, cp, gam, rcp);
    ok = enthalpy<NSP>(th, T, entha) && ok;
    if (!ok) flag_error(S.err, RC_ERR_TEMPERATURE, i, kerr);
    const double rho = p / (rcp * T);
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
    double jr[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) jr[k] = 0.0;
    if (r == 0) {
        jr[0] = drdp;
        jr[3] = drdt;
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k) jr[kNFlow + k] = drdy[k];
    } else if (r == 1) {
        jr[0] = u * drdp;
        jr[1] = rho;
        jr[3] = u * drdt;
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k) jr[kNFlow + k] = u * drdy[k];
    } else if (r == 2) {
        jr[0] = v * drdp;
        jr[2] = rho;
        jr[3] = v * drdt;
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k) jr[kNFlow + k] = v * drdy[k];
    } else if (r == 3) {
        jr[0] = e2 * drdp - 1.;
        jr[1] = rho * u;
        jr[2] = rho * v;
        jr[3] = rho * cp + e2 * drdt;
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k) jr[kNFlow + k] = e2 * drdy[k] + rho * dhdy[k];
    } else if (r < NV) {
        const int sp = r - kNFlow;
        double ys = y[0];
#pragma unroll
        for (int k = 1; k < NSP - 1; ++k) ys = sp == k ? y[k] : ys;
        jr[0] = drdp * ys;
        jr[3] = drdt * ys;
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k) jr[kNFlow + k] = sp == k ? rho + ys * drdy[k] : ys * drdy[k];
    }
    const double d = S.dd[i];
#pragma unroll
    for (int k = 0; k < NV; ++k) jr[k] = jr[k] * d;
    // compute_derivative_production_rate's state: rho of the unclipped
    // fractions, then the fractions clipped to [0, 1]
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) yc[k] = fmin(1., fmax(0., y[k]));
    double dkf[NV];
    dkf_dpa<NSP>(th, p, T, rho, rcp, yc, dkf);
    if (r >= kNFlow && r < NV) {
        const int sp = r - kNFlow;
        const double wfac = 1000 * th.mw[sp] * th.dnu[sp];
        const double vol = M.vol[i];
#pragma unroll
        for (int k = 0; k < NV; ++k) jr[k] = jr[k] - vol * (wfac * dkf[k]);
    }

    // right-hand side and solve
    const double hw = r < NV ? HWA_(r, i) : 0.0;
    const double rhs = FWD ? hw - dflux : dflux;
    double sol = 0.0;
    if (!group_dgesvx<NV>(r, jr, rhs, sol)) {
        if (r == 0) flag_error(S.err, RC_ERR_SINGULAR, i, kerr);
        return;
    }
    if (r < NV) HWA_(r, i) = FWD ? sol : hw - sol;
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

void launch_be_prepare(const LaunchSet &L, double dt, bool first, hipStream_t st) {
    const int f = first ? 1 : 0;
    if (L.single) {
        const int n = L.cell.hi - L.cell.lo;
        if (n > 0)
            k_be_prepare<kNsp, false><<<(n + kChunk - 1) / kChunk, kChunk, 0, st>>>(L.cases, nullptr, L.cell.lo,
                                                                                    L.cell.hi, dt, f);
    } else if (L.cell.n) {
        k_be_prepare<kNsp, true><<<L.cell.n, kChunk, 0, st>>>(L.cases, L.cell.w, 0, 0, dt, f);
    }
}

void launch_be_level(const LaunchSet &L, bool fwd, int l, hipStream_t st) {
    if (fwd) LAUNCH(L.pi_level[l], kPiChunk, st, k_be_sweep, kNsp, true);
    else LAUNCH(L.pi_level[l], kPiChunk, st, k_be_sweep, kNsp, false);
}

void launch_be_update(const LaunchSet &L, hipStream_t st) { LAUNCH(L.cell, kChunk, st, k_be_update, kNsp); }

int pi_chunk() { return kPiChunk; }

}  // namespace rc
