// FP64 HIP kernels for gfx950 of the steady residual + LU-SGS step.
//
// Every kernel is a gather: a thread owns one face or one cell and writes
// only its own outputs, so no atomics are needed and results are
// deterministic.  Per-cell sums over faces run over the cell's faces in
// ascending face number (NeighF order), which is the order in which the
// reference's face loops accumulate into the same cell; the LU-SGS sweeps
// run level by level over wavefronts of the given cell order.
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>

#include "rc_items.h"

namespace rc {

// TimeStep face loop (TimeStep.f90:16-92, inviscid): spectral radius of face
// f from the cell averages of its two cells (one cell on a boundary)
template <int NSP, class TT>
__device__ __forceinline__ double face_spectral_radius(const GMesh &M, const GState &S, TT &th,
                                                       int f, int nl, int nr, double dx, double dy) {
    double pp, uu, vv, tt, y[NSP], s = 0.0;
    if (nr >= 0) {
        pp = 0.5 * (PA_(0, nl) + PA_(0, nr));
        uu = 0.5 * (PA_(1, nl) + PA_(1, nr));
        vv = 0.5 * (PA_(2, nl) + PA_(2, nr));
        tt = 0.5 * (PA_(3, nl) + PA_(3, nr));
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k) { y[k] = 0.5 * (PA_(kNFlow + k, nl) + PA_(kNFlow + k, nr)); s = s + y[k]; }
    } else {
        pp = PA_(0, nl); uu = PA_(1, nl); vv = PA_(2, nl); tt = PA_(3, nl);
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k) { y[k] = PA_(kNFlow + k, nl); s = s + y[k]; }
    }
    y[NSP - 1] = 1.0 - s;
    double cp, gam, r;
    if (!gas_param<NSP>(th, y, tt, cp, gam, r)) flag_error(S.err, RC_ERR_TEMPERATURE, f, K_TS_FACE);
    const double rr = pp / (r * tt);
    if (rr < 0. || pp < 0. || tt < 0.) flag_error(S.err, RC_ERR_NEGATIVE, f, K_TS_FACE);
    const double un = uu * dx + vv * dy;
    return fabs(un) + sqrt(gam * pp / rr) * sqrt(dx * dx + dy * dy);
}

template <int NSP, bool B>
__global__ void __launch_bounds__(256) k_timestep_face(ITEM_ARGS) {
    ITEM_PROLOGUE(f);
    ALPHA_(f) = face_spectral_radius<NSP>(M, S, TH, f, M.fnl[f], M.fnr[f], M.vecx[f], M.vecx[M.sf + f]);
}

// TimeStep cell loop (TimeStep.f90:94-96; viscous: :77-81) and the LU-SGS diagonal
// dd(i) = sum 0.5*alpaF + vol/step (Time_LUSGS_steady.f90:46-53)
// (stored when `store`; the face alpaF values and dd are also returned)
// (cf: the cell's three face codes, cface)
__device__ __forceinline__ double cell_step_dd(const GMesh &M, const GState &S, GParams &P, int i,
                                               const int *cfs, double *al, bool store = true) {
    const double cfl = P.cfl;
    const double v = M.vol[i];
    double a = 0.0, d = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int cf = cfs[k];
        const int f = cf >= 0 ? cf : ~cf;
        al[k] = ALPHA_(f);
        // viscous: alpaC(i) += alpa + alpaVis/vol(i) for either side of the
        // face (TimeStep.f90:79-81); alpaF stays alpa + alpaVis/vol(NL)
        a = P.viscous ? a + FV_(0, f) + FV_(kRecF - 1, f) / v : a + al[k];
        if (store) S.lu_af[k * M.sc + i] = al[k];
    }
    if (P.periodic) {
        // a periodic face adds its alpha to both cells (TimeStep.f90:84-91),
        // so a cell that is the neighbour of periodic faces sums more than
        // its own three, all in face order
        a = 0.0;
#pragma unroll
        for (int k = 0; k < kTsFaces; ++k) {
            const int f = M.tsx[k * M.sc + i];
            if (f >= 0) a = P.viscous ? a + FV_(0, f) + FV_(kRecF - 1, f) / v : a + ALPHA_(f);
        }
    }
    const double st = cfl * v / a;
#pragma unroll
    for (int k = 0; k < 3; ++k) d = d + 0.5 * al[k];
    d = d + v / st;
    if (store) {
        S.step[i] = st;
        S.dd[i] = d;
    }
    return d;
}
__device__ __forceinline__ double cell_step_dd(const GMesh &M, const GState &S, GParams &P, int i, double *al,
                                               bool store = true) {
    int cfs[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) cfs[k] = M.cface[k * M.sc + i];
    return cell_step_dd(M, S, P, i, cfs, al, store);
}

template <int NSP, bool B>
__global__ void __launch_bounds__(256) k_timestep_cell(ITEM_ARGS) {
    ITEM_PROLOGUE(i);
    double al[3];
    cell_step_dd(M, S, PAR, i, al);
}

// advances each case's device step counter used in error reports
// and clears the item counters of the dataflow sweep that follows.  (The
// fused steady step without the dataflow sweep advances it in its first and
// last kernels instead: the gradient kernel's thread of cell 0 sets err[4]
// = err[5] + 1 and its flags record err[5] + 1, the update's thread of cell
// 0 sets err[5] = err[4]; the kernels in between record err[4].  No kernel
// reads a word that another thread of the same launch writes.)
__global__ void k_step_begin(const CaseDev *__restrict__ cases, int n, int *__restrict__ ctr) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) {
        RC_G int *err = gp(gp(cases)[c].S.err);
        err[4] = err[4] + 1;
        err[5] = err[4];
    }
    if (ctr && c < kXcdLists) gp(ctr)[c] = 0;
}

// update_ghostcell_variables (update_flow_variables.f90:90-269): the state
// of ghost g mirrored from its host cell h's primitive state pl[]; the
// host's gamma and density (ComputeGasParameter of its own state) come from
// the cell-average cache
template <int NSP, bool SF = false, class TT>
__device__ __forceinline__ void ghost_state(const GMesh &M, const GState &S, TT &th,
                                            GParams &P, int g, int h, const double *pl_, double *out) {
    const double pl = pl_[0], ul = pl_[1], vl = pl_[2], tl = pl_[3];
    double yl[NSP], yr[NSP], s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) { yl[k] = pl_[kNFlow + k]; s = s + yl[k]; }
    yl[NSP - 1] = 1.0 - s;
    const double gam = LC_(LC_GAM, h);
    const double rl = LC_(LC_RR, h);
    const double s1n = M.gnrm[g], s2n = M.gnrm[M.sg + g];
    const double ac = sqrt(gam * pl / rl);
    double pr, ur, vr, tr;
    const int type = M.ghost_type[g];
    if (type == 1) {
        pr = pl; tr = tl; ur = -ul; vr = -vl;
#pragma unroll
        for (int k = 0; k < NSP; ++k) yr[k] = yl[k];
    } else if (type == 2 || type == 6) {
        pr = pl; tr = tl;
        const double vn = ul * s1n + vl * s2n;
        ur = ul - 2 * vn * s1n;
        vr = vl - 2 * vn * s2n;
#pragma unroll
        for (int k = 0; k < NSP; ++k) yr[k] = yl[k];
    } else if (type == 3 || type == 4) {
        // oxidizer_inlet_bc / fuel_inlet_bc (oxidizer.f90:53-68, fuel.f90:54-69)
        const RC_G double *yin = type == 3 ? P.ox_y : P.fu_y;
#pragma unroll
        for (int k = 0; k < NSP; ++k) yr[k] = yin[k];
        tr = type == 3 ? P.ox_t : P.fu_t;
        double c2, g2, r2;
        if (!gas_param<NSP>(th, yr, tr, c2, g2, r2)) {
            if (SF) flag_error_at(S.err, RC_ERR_TEMPERATURE, g, K_GHOST, S.err[5] + 1);
            else flag_error(S.err, RC_ERR_TEMPERATURE, g, K_GHOST);
        }
        pr = pl;
        const double rr = pr / (r2 * tr);
        ur = (type == 3 ? P.ox_g : P.fu_g) / rr;
        vr = 0.0;
    } else {  // 5: outlet, supersonic copy or fixed p_out
        const double vn = ul * s1n + vl * s2n;
        pr = (fabs(vn) > ac) ? pl : P.p_out;
        ur = ul; vr = vl; tr = tl;
#pragma unroll
        for (int k = 0; k < NSP; ++k) yr[k] = yl[k];
    }
    out[0] = pr; out[1] = ur; out[2] = vr; out[3] = tr;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) out[kNFlow + k] = yr[k];
}

template <int NSP, bool B>
__global__ void __launch_bounds__(256) k_ghost(ITEM_ARGS) {
    ITEM_PROLOGUE(g);
    const int i = M.ghost_host[g];
    double pl[NV], out[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) pl[k] = PA_(k, i);
    ghost_state<NSP>(M, S, TH, PAR, g, i, pl, out);
#pragma unroll
    for (int k = 0; k < NV; ++k) GPA_(k, g) = out[k];
}

// gradient (gradient.f90:4-45) and/or Limiter_Barth (limiter.f90:5-107),
// one thread per cell.  GRAD: compute GradC = AMLS * (wdis * dPA) over the
// 3 stencil slots (interior neighbours, then ghosts); the ghost states of a
// boundary cell depend only on the cell itself, so they are formed here (and
// stored for the limiter-only entry point).  LIM: scale each variable's
// gradient by the Barth factor over the 3 face midpoints (LIM_ALWAYS), or
// when the case's limit_ID > 0 (LIM_CASE), as reconstruction does.
enum { LIM_NONE = 0, LIM_ALWAYS = 1, LIM_CASE = 2 };
// (the work of one cell i of case C; k_gradlim and the persistent step call
// it).  SF: the first kernel of the fused step, which also begins the step
// on the device counter (see k_step_begin).
template <int NSP, bool GRAD, int LIM, bool SF = false>
__device__ __forceinline__ void gradlim_cell(const RC_G CaseDev &C, int i) {
    const GMesh M = gview(C.M);
    const GState S = gview(C.S);
    GThermo &TH = *gp(C.th);
    GParams &PAR = *gp(C.par);
    if (SF && i == 0) S.err[4] = S.err[5] + 1;
    const bool LIMIT = LIM == LIM_ALWAYS || (LIM == LIM_CASE && PAR.limit_id > 0);
    double pi[NV], pmin[NV], pmax[NV], g0[NV], g1[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) { pi[k] = PA_(k, i); pmin[k] = pi[k]; pmax[k] = pi[k]; g0[k] = 0.0; g1[k] = 0.0; }
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const int j = M.stencil[s * M.sc + i];
        const double w = M.wdis[s * M.sc + i];
        const double a0 = M.amls[(s * 2 + 0) * M.sc + i];
        const double a1 = M.amls[(s * 2 + 1) * M.sc + i];
        double q[NV];
        if (j >= 0) {
#pragma unroll
            for (int k = 0; k < NV; ++k) q[k] = PA_(k, j);
        } else if (GRAD) {
            ghost_state<NSP, SF>(M, S, TH, PAR, -j - 1, i, pi, q);
#pragma unroll
            for (int k = 0; k < NV; ++k) GPA_(k, -j - 1) = q[k];
        } else {
#pragma unroll
            for (int k = 0; k < NV; ++k) q[k] = GPA_(k, -j - 1);
        }
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const double pj = q[k];
            if (GRAD) {
                const double cc = w * (pj - pi[k]);
                g0[k] = g0[k] + a0 * cc;
                g1[k] = g1[k] + a1 * cc;
            }
            if (LIMIT) { pmin[k] = fmin(pmin[k], pj); pmax[k] = fmax(pmax[k], pj); }
        }
    }
    if (!GRAD) {
#pragma unroll
        for (int k = 0; k < NV; ++k) { g0[k] = GR_(k, 0, i); g1[k] = GR_(k, 1, i); }
    }
    if (LIMIT) {
        double v0[3], v1[3];
#pragma unroll
        for (int kf = 0; kf < 3; ++kf) { v0[kf] = M.lvv[(kf * 2) * M.sc + i]; v1[kf] = M.lvv[(kf * 2 + 1) * M.sc + i]; }
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double phi = 1.0;
#pragma unroll
            for (int kf = 0; kf < 3; ++kf) {
                const double fv = pi[k] + (g0[k] * v0[kf] + g1[k] * v1[kf]);
                double pf;
                if (fv > pi[k]) pf = fmin(1.0, (pmax[k] - pi[k]) / (fv - pi[k]));
                else if (fv < pi[k]) pf = fmin(1.0, (pmin[k] - pi[k]) / (fv - pi[k]));
                else pf = 1.0;
                phi = fmin(phi, pf);
            }
            g0[k] = phi * g0[k];
            g1[k] = phi * g1[k];
        }
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) { GR_(k, 0, i) = g0[k]; GR_(k, 1, i) = g1[k]; }
    // the record's unused slots too: whole 128-byte lines go to memory
#pragma unroll
    for (int k = 2 * NV; k < kRecGR; ++k) S.gradc[(size_t)i * kRecGR + k] = 0.0;
}

template <int NSP, bool GRAD, int LIM, bool SF, bool B>
__global__ void __launch_bounds__(256) k_gradlim(ITEM_ARGS) {
    ITEM_CASE(i);
    gradlim_cell<NSP, GRAD, LIM, SF>(*Cp_, i);
}

// flux_lax (flux_lax.f90:17-296), one thread per face: reconstructed L/R
// states at the face point, boundary branch, Rusanov flux -> fflux(7,f).
// TS: also the face spectral radius of TimeStep, which depends only on the
// cell averages this thread loads anyway (the values equal a separate
// TimeStep pass: PA does not change between TimeStep and flux_lax).
template <int NSP, bool TS, bool B>
__global__ void __launch_bounds__(256) k_flux(ITEM_ARGS) {
    ITEM_PROLOGUE_LDS(f);
    const DevThermo &th = TH;
    const int nl = M.fnl[f], nr = M.fnr[f];
    const double dx = M.vecx[f], dy = M.vecx[M.sf + f];
    if (TS) ALPHA_(f) = face_spectral_radius<NSP>(M, S, th, f, nl, nr, dx, dy);
    const double sav1 = dx, sav2 = dy;
    const double sav = fmax(sqrt(dx * dx + dy * dy), kSmall);
    const double s1n = sav1 / sav, s2n = sav2 / sav;
    const double v0 = M.fvl[f], v1 = M.fvl[M.sf + f];
    double ql[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) ql[k] = PA_(k, nl) + (GR_(k, 0, nl) * v0 + GR_(k, 1, nl) * v1);
    double pl = ql[0], ul = ql[1], vl = ql[2], tl = ql[3];
    double yl[NSP], yr[NSP], s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) { yl[k] = ql[kNFlow + k]; s = s + yl[k]; }
    yl[NSP - 1] = 1.0 - s;
    double cpl, gaml, rcl;
    if (!gas_param<NSP>(th, yl, tl, cpl, gaml, rcl)) flag_error(S.err, RC_ERR_TEMPERATURE, f, K_FLUX);
    double rl = pl / (rcl * tl);
    double pr, ur, vr, tr, rr, cpr, gamr, rcr;
    if (nr >= 0) {
        const double w0 = M.fvr[f], w1 = M.fvr[M.sf + f];
        double qr[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) qr[k] = PA_(k, nr) + (GR_(k, 0, nr) * w0 + GR_(k, 1, nr) * w1);
        pr = qr[0]; ur = qr[1]; vr = qr[2]; tr = qr[3];
        s = 0.0;
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k) { yr[k] = qr[kNFlow + k]; s = s + yr[k]; }
        yr[NSP - 1] = 1.0 - s;
        if (!gas_param<NSP>(th, yr, tr, cpr, gamr, rcr)) flag_error(S.err, RC_ERR_TEMPERATURE, f, K_FLUX);
        rr = pr / (rcr * tr);
    } else {
        GParams &P = PAR;
        const int type = -nr;
        cpr = cpl; gamr = gaml; rcr = rcl;
#pragma unroll
        for (int k = 0; k < NSP; ++k) yr[k] = yl[k];
        if (type == 1) {
            rr = rl; pr = pl; tr = tl; ur = -ul; vr = -vl;
        } else if (type == 2 || type == 6) {
            rr = rl; pr = pl; tr = tl;
            const double vn = ul * s1n + vl * s2n;
            ur = ul - 2 * vn * s1n;
            vr = vl - 2 * vn * s2n;
        } else if (type == 3 || type == 4) {
            const RC_G double *yin = type == 3 ? P.ox_y : P.fu_y;
#pragma unroll
            for (int k = 0; k < NSP; ++k) yr[k] = yin[k];
            tr = type == 3 ? P.ox_t : P.fu_t;
            if (!gas_param<NSP>(th, yr, tr, cpr, gamr, rcr)) flag_error(S.err, RC_ERR_TEMPERATURE, f, K_FLUX);
            pr = pl;
            rr = pr / (rcr * tr);
            ur = (type == 3 ? P.ox_g : P.fu_g) / rr;
            vr = 0.0;
            pl = pr; ul = ur; vl = vr; tl = tr; rl = rr;
#pragma unroll
            for (int k = 0; k < NSP; ++k) yl[k] = yr[k];
            cpl = cpr; gaml = gamr; rcl = rcr;
        } else {  // 5: nozzle outlet
            const double ac = sqrt(gaml * pl / rl);
            const double vn = ul * s1n + vl * s2n;
            if (fabs(vn) > ac) {
                rr = rl; ur = ul; vr = vl; pr = pl; tr = tl;
            } else {
                pr = P.p_out;
                ur = ul + s1n * (pl - pr) / (rl * ac);
                vr = vl + s2n * (pl - pr) / (rl * ac);
                rr = rl + (pr - pl) / (ac * ac);
                tr = pr / (rcl * rr);
                pl = pr; ul = ur; vl = vr; tl = tr; rl = rr;
            }
        }
    }
    if (rl < 0. || pl < 0. || (nr >= 0 && (rr < 0. || pr < 0.))) flag_error(S.err, RC_ERR_NEGATIVE, f, K_FLUX);
    double h[NSP], hl = 0.0, hr = 0.0;
    if (!enthalpy<NSP>(th, tl, h)) flag_error(S.err, RC_ERR_TEMPERATURE, f, K_FLUX);
#pragma unroll
    for (int k = 0; k < NSP; ++k) hl = hl + yl[k] * h[k];
    if (!enthalpy<NSP>(th, tr, h)) flag_error(S.err, RC_ERR_TEMPERATURE, f, K_FLUX);
#pragma unroll
    for (int k = 0; k < NSP; ++k) hr = hr + yr[k] * h[k];
    const double vaml = ul * ul + vl * vl, vamr = ur * ur + vr * vr;
    hl = hl + 0.5 * vaml;
    hr = hr + 0.5 * vamr;
    const double el = hl - pl / rl, er = hr - pr / rr;
    const double acl = sqrt(gaml * pl / rl), acr = sqrt(gamr * pr / rr);
    const double uln = sav1 * ul + sav2 * vl, urn = sav1 * ur + sav2 * vr;
    const double coex = fmax(fabs(uln) + acl * sav, fabs(urn) + acr * sav);
    const double ruln = rl * uln, rurn = rr * urn;
    F_(0, f) = 0.5 * (ruln + rurn - coex * (rr - rl));
    F_(1, f) = 0.5 * (ruln * ul + rurn * ur + sav1 * (pl + pr) - coex * (rr * ur - rl * ul));
    F_(2, f) = 0.5 * (ruln * vl + rurn * vr + sav2 * (pl + pr) - coex * (rr * vr - rl * vl));
    F_(3, f) = 0.5 * (ruln * hl + rurn * hr - coex * (rr * er - rl * el));
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k)
        F_(kNFlow + k, f) = 0.5 * (ruln * yl[k] + rurn * yr[k] - coex * (rr * yr[k] - rl * yl[k]));
}

// flux_lax with two lanes per face (lanes 2k, 2k+1 of a wave take face k of
// the block): lane 0 forms the left state, lane 1 the right one (on a
// boundary face both lanes form the left state and the boundary branch, and
// each keeps its side), so a lane holds one cell's loads and one thermo
// evaluation at a time; the two sides are exchanged by shuffles and each
// lane writes half of the face record.  TS: lane 1 also forms the face
// spectral radius from the cell averages.  Every value is formed by the
// expressions of k_flux; a face's failed checks are reported once, by lane
// 0, in k_flux's order (so the first error recorded is the same).
// (the work of lane `side` of face f of case C, the thermo table staged in
// LDS; k_flux2 and the persistent step call it)
template <int NSP, bool TS>
__device__ __forceinline__ void flux2_face(const RC_G CaseDev &C, int f, int side, const DevThermo &th) {
    const GMesh M = gview(C.M);
    const GState S = gview(C.S);
    GParams &P = *gp(C.par);
    const int nl = M.fnl[f], nr = M.fnr[f];
    const bool inner = nr >= 0;
    const double dx = M.vecx[f], dy = M.vecx[M.sf + f];
    const double sav1 = dx, sav2 = dy;
    const double sav = fmax(sqrt(dx * dx + dy * dy), kSmall);
    const double s1n = sav1 / sav, s2n = sav2 / sav;
    // failed checks of this lane: 1 TS temperature, 2 TS negative, 4 own
    // thermo, 8 inlet thermo, 16 negative state, 32 own enthalpy
    int bad = 0;
    // the lane's cell: the right one for lane 1 of an inner face, else the left
    const int c = (side && inner) ? nr : nl;
    const RC_G double *fv = (side && inner) ? M.fvr : M.fvl;
    const double w0 = fv[f], w1 = fv[M.sf + f];
    double pa[NV], q[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        pa[k] = PA_(k, c);
        q[k] = pa[k] + (GR_(k, 0, c) * w0 + GR_(k, 1, c) * w1);
    }
    double pl = q[0], ul = q[1], vl = q[2], tl = q[3];
    double yl[NSP], s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) { yl[k] = q[kNFlow + k]; s = s + yl[k]; }
    yl[NSP - 1] = 1.0 - s;
    double cpl, gaml, rcl;
    if (!gas_param<NSP>(th, yl, tl, cpl, gaml, rcl)) bad |= 4;
    double rl = pl / (rcl * tl);
    // (pl .. rcl: the lane's own state so far; on a boundary face the left one)
    if (!inner) {
        double pr, ur, vr, tr, rr, cpr, gamr, rcr, yr[NSP];
        const int type = -nr;
        cpr = cpl; gamr = gaml; rcr = rcl;
#pragma unroll
        for (int k = 0; k < NSP; ++k) yr[k] = yl[k];
        if (type == 1) {
            rr = rl; pr = pl; tr = tl; ur = -ul; vr = -vl;
        } else if (type == 2 || type == 6) {
            rr = rl; pr = pl; tr = tl;
            const double vn = ul * s1n + vl * s2n;
            ur = ul - 2 * vn * s1n;
            vr = vl - 2 * vn * s2n;
        } else if (type == 3 || type == 4) {
            const RC_G double *yin = type == 3 ? P.ox_y : P.fu_y;
#pragma unroll
            for (int k = 0; k < NSP; ++k) yr[k] = yin[k];
            tr = type == 3 ? P.ox_t : P.fu_t;
            if (!gas_param<NSP>(th, yr, tr, cpr, gamr, rcr)) bad |= 8;
            pr = pl;
            rr = pr / (rcr * tr);
            ur = (type == 3 ? P.ox_g : P.fu_g) / rr;
            vr = 0.0;
            pl = pr; ul = ur; vl = vr; tl = tr; rl = rr;
#pragma unroll
            for (int k = 0; k < NSP; ++k) yl[k] = yr[k];
            cpl = cpr; gaml = gamr; rcl = rcr;
        } else {  // 5: nozzle outlet
            const double ac = sqrt(gaml * pl / rl);
            const double vn = ul * s1n + vl * s2n;
            if (fabs(vn) > ac) {
                rr = rl; ur = ul; vr = vl; pr = pl; tr = tl;
            } else {
                pr = P.p_out;
                ur = ul + s1n * (pl - pr) / (rl * ac);
                vr = vl + s2n * (pl - pr) / (rl * ac);
                rr = rl + (pr - pl) / (ac * ac);
                tr = pr / (rcl * rr);
                pl = pr; ul = ur; vl = vr; tl = tr; rl = rr;
            }
        }
        if (rl < 0. || pl < 0.) bad |= 16;
        if (side) {
            // lane 1 keeps the right side
            pl = pr; ul = ur; vl = vr; tl = tr; rl = rr;
#pragma unroll
            for (int k = 0; k < NSP; ++k) yl[k] = yr[k];
            cpl = cpr; gaml = gamr; rcl = rcr;
        }
    } else {
        if (rl < 0. || pl < 0.) bad |= 16;
    }
    // the lane's side: total enthalpy, energy, sound speed, normal velocity
    double h[NSP], hs = 0.0;
    if (!enthalpy<NSP>(th, tl, h)) bad |= 32;
#pragma unroll
    for (int k = 0; k < NSP; ++k) hs = hs + yl[k] * h[k];
    const double vam = ul * ul + vl * vl;
    hs = hs + 0.5 * vam;
    const double es = hs - pl / rl;
    const double acs = sqrt(gaml * pl / rl);
    const double uns = sav1 * ul + sav2 * vl;
    // the other side
    const double po = __shfl_xor(pl, 1), uo = __shfl_xor(ul, 1), vo = __shfl_xor(vl, 1), ro = __shfl_xor(rl, 1);
    const double ho = __shfl_xor(hs, 1), eo = __shfl_xor(es, 1), aco = __shfl_xor(acs, 1);
    const double uno = __shfl_xor(uns, 1);
    double yo[NSP - 1];
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) yo[k] = __shfl_xor(yl[k], 1);
    // left (L) and right (R) by side
    const double rL = side ? ro : rl, rR = side ? rl : ro;
    const double uL = side ? uo : ul, uR = side ? ul : uo, vL = side ? vo : vl, vR = side ? vl : vo;
    const double pL = side ? po : pl, pR = side ? pl : po;
    const double hL = side ? ho : hs, hR = side ? hs : ho, eL = side ? eo : es, eR = side ? es : eo;
    const double uln = side ? uno : uns, urn = side ? uns : uno;
    const double acl = side ? aco : acs, acr = side ? acs : aco;
    const double coex = fmax(fabs(uln) + acl * sav, fabs(urn) + acr * sav);
    const double ruln = rL * uln, rurn = rR * urn;
    if (side == 0) {
        F_(0, f) = 0.5 * (ruln + rurn - coex * (rR - rL));
        F_(1, f) = 0.5 * (ruln * uL + rurn * uR + sav1 * (pL + pR) - coex * (rR * uR - rL * uL));
        F_(2, f) = 0.5 * (ruln * vL + rurn * vR + sav2 * (pL + pR) - coex * (rR * vR - rL * vL));
        F_(3, f) = 0.5 * (ruln * hL + rurn * hR - coex * (rR * eR - rL * eL));
    } else {
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k)
            F_(kNFlow + k, f) = 0.5 * (ruln * yo[k] + rurn * yl[k] - coex * (rR * yl[k] - rL * yo[k]));
    }
    // TimeStep's face part, by lane 1, from the cell averages
    if constexpr (TS) {
        double pa_o[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) pa_o[k] = __shfl_xor(pa[k], 1);
        if (side) {
        // own pa is the right cell's on an inner face, the left cell's on a
        // boundary face; pa_o is the left cell's
            double pp, uu, vv, tt, y[NSP], sy = 0.0;
            if (inner) {
                pp = 0.5 * (pa_o[0] + pa[0]);
                uu = 0.5 * (pa_o[1] + pa[1]);
                vv = 0.5 * (pa_o[2] + pa[2]);
                tt = 0.5 * (pa_o[3] + pa[3]);
#pragma unroll
                for (int k = 0; k < NSP - 1; ++k) { y[k] = 0.5 * (pa_o[kNFlow + k] + pa[kNFlow + k]); sy = sy + y[k]; }
            } else {
                pp = pa[0]; uu = pa[1]; vv = pa[2]; tt = pa[3];
#pragma unroll
                for (int k = 0; k < NSP - 1; ++k) { y[k] = pa[kNFlow + k]; sy = sy + y[k]; }
            }
            y[NSP - 1] = 1.0 - sy;
            double cp, gam, r;
            if (!gas_param<NSP>(th, y, tt, cp, gam, r)) bad |= 1;
            const double rr = pp / (r * tt);
            if (rr < 0. || pp < 0. || tt < 0.) bad |= 2;
            const double un = uu * dx + vv * dy;
            ALPHA_(f) = fabs(un) + sqrt(gam * pp / rr) * sqrt(dx * dx + dy * dy);
        }
    }
    // report the face's first failed check in k_flux's order
    const int bad_o = __shfl_xor(bad, 1);
    if (side == 0 && (bad | bad_o)) {
        const int b0 = bad, b1 = bad_o;
        if (b1 & 1) flag_error(S.err, RC_ERR_TEMPERATURE, f, K_TS_FACE);
        else if (b1 & 2) flag_error(S.err, RC_ERR_NEGATIVE, f, K_TS_FACE);
        else if (b0 & 4) flag_error(S.err, RC_ERR_TEMPERATURE, f, K_FLUX);
        else if (inner && (b1 & 4)) flag_error(S.err, RC_ERR_TEMPERATURE, f, K_FLUX);
        else if (!inner && (b0 & 8)) flag_error(S.err, RC_ERR_TEMPERATURE, f, K_FLUX);
        else if ((b0 & 16) || (inner && (b1 & 16))) flag_error(S.err, RC_ERR_NEGATIVE, f, K_FLUX);
        else if ((b0 & 32) || (b1 & 32)) flag_error(S.err, RC_ERR_TEMPERATURE, f, K_FLUX);
    }
}

template <int NSP, bool TS, bool B>
__global__ void __launch_bounds__(512) k_flux2(ITEM_ARGS) {
    __shared__ DevThermo th_lds_;
    const RC_G CaseDev *Cp_;
    int f, end_;
    const int side = (int)threadIdx.x & 1;
    if constexpr (B) {
        const int4 w_ = work[blockIdx.x];
        Cp_ = gp(cases) + w_.x;
        f = w_.y + (int)(threadIdx.x >> 1);
        end_ = w_.z;
    } else {
        Cp_ = gp(cases);
        f = lo + (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 1);
        end_ = hi;
    }
    stage_thermo(*Cp_, th_lds_);
    __syncthreads();
    if (f >= end_) return;      // both lanes of a face leave together
    flux2_face<NSP, TS>(*Cp_, f, side, th_lds_);
}

// residual assembly of flux_sourceterm (fvm2d.f90:201-219) of cell i: h =
// sum of -F (owner) / +F (neighbour) over the cell's faces in face order,
// then the chemical source vol*omega of sourceterm (sourceterm.f90:2-70)
template <int NSP, class TT>
__device__ __forceinline__ void cell_residual(const GMesh &M, const GState &S, TT &th, GParams &P, int i,
                                              const int *cfs, double *h) {
#pragma unroll
    for (int k = 0; k < NV; ++k) h[k] = 0.0;
#pragma unroll
    for (int kf = 0; kf < 3; ++kf) {
        const int cf = cfs[kf];
        const int f = cf >= 0 ? cf : ~cf;
        if (cf >= 0) {
#pragma unroll
            for (int k = 0; k < NV; ++k) h[k] = h[k] - F_(k, f);
        } else {
#pragma unroll
            for (int k = 0; k < NV; ++k) h[k] = h[k] + F_(k, f);
        }
    }
    if (P.viscous) {
        // flux_viscous after flux_lax (fvm2d.f90:212), face order again;
        // its mass flux is zero
#pragma unroll
        for (int kf = 0; kf < 3; ++kf) {
            const int cf = cfs[kf];
            const int f = cf >= 0 ? cf : ~cf;
            if (cf >= 0) {
                h[0] = h[0] - 0.0;
#pragma unroll
                for (int k = 1; k < NV; ++k) h[k] = h[k] - FV_(k, f);
            } else {
                h[0] = h[0] + 0.0;
#pragma unroll
                for (int k = 1; k < NV; ++k) h[k] = h[k] + FV_(k, f);
            }
        }
    }
    if (P.combustion && P.surrogate) {
        // the learned-chemistry source of the cell (accepted surrogate or
        // its exact fallback), formed before the sweep
        const double v = M.vol[i];
#pragma unroll
        for (int k = 0; k < kNFlow; ++k) h[k] = h[k] + v * 0.0;
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k) h[kNFlow + k] = h[kNFlow + k] + v * S.src[k * M.sc + i];
    } else if (P.combustion) {
        const double t = PA_(3, i);
        double y[NSP], s = 0.0;
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k) { y[k] = PA_(kNFlow + k, i); s = s + y[k]; }
        y[NSP - 1] = 1.0 - s;
        // rho = p/(t*R) of sourceterm.f90:35, formed from the cell's own
        // primitive record as the cell-average cache forms it (p/(R*t), R
        // from the same mass fractions): the cached value bit for bit,
        // without reading the cache record
        const double rho = PA_(0, i) / (gas_rconst<NSP>(th, y) * t);
        double w[NSP];
        production_rate<NSP>(th, t, rho, y, w);
        const double v = M.vol[i];
#pragma unroll
        for (int k = 0; k < kNFlow; ++k) h[k] = h[k] + v * 0.0;
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k) h[kNFlow + k] = h[kNFlow + k] + v * w[k];
    }
}
template <int NSP, class TT>
__device__ __forceinline__ void cell_residual(const GMesh &M, const GState &S, TT &th, GParams &P, int i,
                                              double *h) {
    int cfs[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) cfs[k] = M.cface[k * M.sc + i];
    cell_residual<NSP>(M, S, th, P, i, cfs, h);
}

// the residual of flux_sourceterm into hWA; TS: also the cell part of
// TimeStep (step, dd) from the same faces
template <int NSP, bool TS, bool B>
__global__ void __launch_bounds__(256) k_residual(ITEM_ARGS) {
    ITEM_PROLOGUE(i);
    if (TS) {
        double al[3];
        cell_step_dd(M, S, PAR, i, al);
    }
    double h[NV];
    cell_residual<NSP>(M, S, TH, PAR, i, h);
#pragma unroll
    for (int k = 0; k < NV; ++k) HWA_(k, i) = h[k];
}

// cell-average quantities of one neighbour read by the LU-SGS term
struct LuNb {
    double rr, rou, rov, roep, uu, vv, tt, gama, h[kNsp], ry[kNsp - 1];
};

__device__ __forceinline__ LuNb load_lunb(const GState &S, int nr) {
    LuNb n;
    n.rr = LC_(LC_RR, nr); n.rou = LC_(LC_ROU, nr); n.rov = LC_(LC_ROV, nr); n.roep = LC_(LC_ROEP, nr);
    n.uu = LC_(LC_U, nr); n.vv = LC_(LC_V, nr); n.tt = LC_(LC_T, nr); n.gama = LC_(LC_GAM, nr);
#pragma unroll
    for (int k = 0; k < kNsp; ++k) n.h[k] = LC_(LC_H0 + k, nr);
#pragma unroll
    for (int k = 0; k < kNsp - 1; ++k) n.ry[k] = LC_(LC_RY0 + k, nr);
    return n;
}

// value k of three (k = 0, 1, 2)
__device__ __forceinline__ double sel3(int k, double a, double b, double c) { return k == 0 ? a : (k == 1 ? b : c); }

// the neighbour term of both LU-SGS sweeps (Time_LUSGS_steady.f90:88-111):
// the linearised Rusanov flux change 0.5*(-alpaF*dW + dF(dW).S) of a
// neighbour's increment dw; the caller adds it to dflux in face order
template <int NSP, class TT>
__device__ __forceinline__ void lusgs_term(TT &th, const LuNb &n, const double *dw, double dx,
                                           double dy, double af, double *c) {
    const double drr = dw[0], dru = dw[1], drv = dw[2], dre = dw[3];
    double dry[NSP], s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) { dry[k] = dw[kNFlow + k]; s = s + dry[k]; }
    dry[NSP - 1] = drr - s;
    const double duu = (dru - n.uu * drr) / n.rr;
    const double dvv = (drv - n.vv * drr) / n.rr;
    double s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) s1 = s1 + n.h[k] * dry[k];
#pragma unroll
    for (int k = 0; k < NSP; ++k) s2 = s2 + th.rsp[k] * dry[k];
    const double dpp = (n.gama - 1) * (dre - s1 + n.tt * s2 * n.gama / (n.gama - 1)
                                       - 0.5 * (dru * n.uu + duu * n.rou + drv * n.vv + dvv * n.rov));
    const double un = n.uu * dx + n.vv * dy;
    const double dun = duu * dx + dvv * dy;
    c[0] = 0.5 * (-af * drr + dru * dx + drv * dy);
    c[1] = 0.5 * (-af * dru + dru * un + n.rou * dun + dpp * dx);
    c[2] = 0.5 * (-af * drv + drv * un + n.rov * dun + dpp * dy);
    c[3] = 0.5 * (-af * dre + dun * n.roep + un * (dre + dpp));
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) c[kNFlow + k] = 0.5 * (-af * dry[k] + dry[k] * un + n.ry[k] * dun);
}

// Loads and stores of the dataflow sweep: increments and progress flags
// pass between thread blocks of one launch, which may run on different XCDs
// (whose L2 caches are not coherent with each other), so they are
// agent-scope atomic accesses (the L2 of the reading XCD does not answer
// them from a stale line); plain arrays written by earlier launches are
// read as usual.
__device__ __forceinline__ double ld_agent(const RC_G double *p) {
    return __hip_atomic_load((RC_G double *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(RC_G double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_agent(const RC_G int *p) {
    return __hip_atomic_load((RC_G int *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(RC_G int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// polls of one dependency wait before the sweep gives up with an error
// (a wait normally ends within the sweep's own duration, tens of us)
constexpr int kSpinMax = 1 << 18;

// The forward (FWD) or backward sweep of LU_SGS_steady
// (Time_LUSGS_steady.f90:56-120 / 127-190) over runs t .. tend-1 of case C.
// A run is up to kRunMax consecutive cells swept in order; it is handled by
// a group of kRunMax lanes, lane p owning cell first+p.  Each lane loads the
// records of its cell's neighbours at once and forms the terms of
// neighbours outside the run (swept before: lower levels in the forward
// sweep, higher in the backward one).  Then the run's cells are finished one
// step at a time in the sweep direction: the lane of the step adds the terms
// of neighbours inside the run, whose new increments the group has received
// by shuffles in the steps before, and forms its increment.  Terms enter
// dflux in face order; a face that does not take part adds +0.0, which
// leaves the reference's sums unchanged.
// RES (forward sweep of the fused step): each lane first forms its cell's
// TimeStep cell part and residual (what k_residual stores) and sweeps with
// them from registers.
// FLOW (the dataflow sweep, all levels in one launch): before it reads a
// neighbour's increment (and, in the backward sweep, its own forward one) a
// lane waits until that cell's progress flag shows the sweep that wrote it
// done in step `epoch`; increments are written before the flag (a wait for
// the stores in between).  The backward sweep forms the TimeStep cell part
// again from the face records instead of reading what the forward sweep of
// another block stored (same arithmetic, same values).
// TOP (forward sweep of the highest level): the lanes then also run that
// level's backward sweep and store its increments.  A cell of the highest
// level has no higher neighbour outside its run (that neighbour's run would
// lie in a higher level), so its backward terms come from its run alone:
// the forward increments stay in registers, the run's backward increments
// pass by shuffles, and every value is the one the backward launch forms.
template <int NSP, bool FWD, bool RES, bool FLOW, bool TOP = false>
__device__ __forceinline__ void lusgs_runs(const RC_G CaseDev &C, int t, int tend, int epoch, double *tl) {
    constexpr int R = kRunMax;
    const GMesh M = gview(C.M);
    const GState S = gview(C.S);
    GThermo &th = *gp(C.th);
    GParams &PAR = *gp(C.par);
    const int lane = threadIdx.x % R;
    const int sc = M.sc;
    // every lane stays to the end: the group's shuffles need all four; the
    // lane's indices come from its slot (one load phase, no run table)
    const int q = 4 * (t < tend ? t : 0) + lane, ss = M.ss;
    const int first = M.slot[SL_FIRST * ss + q];
    const int cnt = t < tend ? M.slot[SL_CNT * ss + q] : 0;
    const bool live = lane < cnt;
    const int i = M.slot[SL_CELL * ss + q];
    int nb[3], cfs[3];
    double nx[3], ny[3], af[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        cfs[k] = M.slot[(SL_CF + k) * ss + q];
        nb[k] = M.slot[(SL_NB + k) * ss + q];
        nx[k] = M.slot_n[(k * 2) * ss + q];
        ny[k] = M.slot_n[(k * 2 + 1) * ss + q];
    }
    double d, rhs[NV];
    if (RES) {
        d = cell_step_dd(M, S, PAR, i, cfs, af, live);
        cell_residual<NSP>(M, S, th, PAR, i, cfs, rhs);
    } else if (FLOW) {
        d = cell_step_dd(M, S, PAR, i, cfs, af, false);
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) af[k] = S.lu_af[k * sc + i];
        d = S.dd[i];
#pragma unroll
        for (int v = 0; v < NV; ++v) rhs[v] = HWA_(v, i);
    }
    if (FLOW) {
        // wait for the increments this lane reads from outside its run
        const int need = FWD ? 2 * epoch + 1 : 2 * epoch + 2;
        for (int it = 0;; ++it) {
            bool ready = true;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int nr = nb[k];
                const bool ext = live && nr >= 0 && (FWD ? nr < i : nr > i) && !(nr >= first && nr < first + cnt);
                if (ext) ready = ld_agent(S.flag + nr) >= need && ready;
            }
            if (!FWD && live) ready = ld_agent(S.flag + i) >= 2 * epoch + 1 && ready;
            if (ready) break;
            if (it >= kSpinMax || ld_agent(S.err) != 0) {
                if (it >= kSpinMax) flag_error(S.err, RC_ERR_DEVICE, i, K_FLOW);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        if (!FWD) {
#pragma unroll
            for (int v = 0; v < NV; ++v) rhs[v] = live ? ld_agent(&HWA_(v, i)) : 0.0;
        }
    }
    // the terms of the cell's three faces, in LDS: tl[(k*8+v)*bs + tid]
    // (face k, variable v); a face that does not take part holds +0.0
    const int bs = blockDim.x, tid = threadIdx.x;
    bool use[3];
    int qs[3];         // sweep step of a neighbour inside the run, else -1
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int nr = nb[k];
        use[k] = live && nr >= 0 && (FWD ? nr < i : nr > i);
        const bool inrun = use[k] && nr >= first && nr < first + cnt;
        qs[k] = inrun ? (FWD ? nr - first : first + cnt - 1 - nr) : -1;
        // faces that do not take part read cell 0's cache record, and so do
        // increments that come from the run: one shared line, not one per
        // lane, and never a record this launch writes
        const int j = use[k] ? nr : 0;
        const LuNb nbk = load_lunb(S, j);
        double dw[NV], ck[NV];
        if (FLOW && use[k] && !inrun) {
#pragma unroll
            for (int v = 0; v < NV; ++v) dw[v] = ld_agent(&HWA_(v, nr));
        } else {
            const RC_G double *hw = (use[k] && !inrun) ? &HWA_(0, nr) : &LC_(0, 0);
#pragma unroll
            for (int v = 0; v < NV; ++v) dw[v] = hw[v];
        }
        lusgs_term<NSP>(th, nbk, dw, nx[k], ny[k], af[k], ck);
#pragma unroll
        for (int v = 0; v < NV; ++v) tl[(k * 8 + v) * bs + tid] = use[k] ? ck[v] : 0.0;
    }
    const int mystep = FWD ? lane : cnt - 1 - lane;
    double hv[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) hv[v] = 0.0;
#pragma unroll
    for (int s = 0; s < R; ++s) {
        // the lane of step s: all its terms are ready (in-run neighbours were
        // swept at earlier steps); dflux adds the three faces' terms in face
        // order, as the reference's face loop does
        if (mystep == s) {
            double dflux[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) dflux[v] = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int v = 0; v < NV; ++v) dflux[v] = dflux[v] + tl[(k * 8 + v) * bs + tid];
#pragma unroll
            for (int v = 0; v < NV; ++v) hv[v] = FWD ? (rhs[v] - dflux[v]) / d : rhs[v] - dflux[v] / d;
        }
        if (s < R - 1) {
            // the increment of the cell swept at step s, to the whole group
            const int src = FWD ? s : (cnt - 1 - s > 0 ? cnt - 1 - s : 0);
            double dw[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) dw[v] = __shfl(hv[v], src, R);
            // every lane with that cell among its neighbours (one face at
            // most) forms the term of that face now, from the neighbour's
            // cache record (loaded again: it was read above)
            const int ks = qs[0] == s ? 0 : (qs[1] == s ? 1 : (qs[2] == s ? 2 : -1));
            if (ks >= 0) {
                const int nrs = ks == 0 ? nb[0] : (ks == 1 ? nb[1] : nb[2]);
                const LuNb nb_s = load_lunb(S, nrs);
                double cs[NV];
                lusgs_term<NSP>(th, nb_s, dw, sel3(ks, nx[0], nx[1], nx[2]), sel3(ks, ny[0], ny[1], ny[2]),
                                sel3(ks, af[0], af[1], af[2]), cs);
#pragma unroll
                for (int v = 0; v < NV; ++v) tl[(ks * 8 + v) * bs + tid] = cs[v];
            }
        }
    }
    if constexpr (TOP && FWD && !FLOW) {
        // the backward sweep of the run (Time_LUSGS_steady.f90:127-190):
        // faces to higher neighbours (all in the run) start at +0.0 and take
        // their term at the step of that neighbour, the others stay +0.0
        int qb[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int nr = nb[k];
            qb[k] = live && nr >= 0 && nr > i ? first + cnt - 1 - nr : -1;
#pragma unroll
            for (int v = 0; v < NV; ++v) tl[(k * 8 + v) * bs + tid] = 0.0;
        }
        const int bstep = cnt - 1 - lane;
        double hb[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) hb[v] = 0.0;
#pragma unroll
        for (int s = 0; s < R; ++s) {
            if (bstep == s) {
                double dflux[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) dflux[v] = 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k)
#pragma unroll
                    for (int v = 0; v < NV; ++v) dflux[v] = dflux[v] + tl[(k * 8 + v) * bs + tid];
#pragma unroll
                for (int v = 0; v < NV; ++v) hb[v] = hv[v] - dflux[v] / d;
            }
            if (s < R - 1) {
                const int src = cnt - 1 - s > 0 ? cnt - 1 - s : 0;
                double dw[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) dw[v] = __shfl(hb[v], src, R);
                const int ks = qb[0] == s ? 0 : (qb[1] == s ? 1 : (qb[2] == s ? 2 : -1));
                if (ks >= 0) {
                    const int nrs = ks == 0 ? nb[0] : (ks == 1 ? nb[1] : nb[2]);
                    const LuNb nb_s = load_lunb(S, nrs);
                    double cs[NV];
                    lusgs_term<NSP>(th, nb_s, dw, sel3(ks, nx[0], nx[1], nx[2]), sel3(ks, ny[0], ny[1], ny[2]),
                                    sel3(ks, af[0], af[1], af[2]), cs);
#pragma unroll
                    for (int v = 0; v < NV; ++v) tl[(ks * 8 + v) * bs + tid] = cs[v];
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) hv[v] = hb[v];
    }
    if (live) {
        if (FLOW) {
#pragma unroll
            for (int v = 0; v < NV; ++v) st_agent(&HWA_(v, i), hv[v]);
            // the increment's stores complete before the flag is raised
            __builtin_amdgcn_s_waitcnt(0x0f70);   // s_waitcnt vmcnt(0)
            st_agent(S.flag + i, FWD ? 2 * epoch + 1 : 2 * epoch + 2);
        } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) HWA_(v, i) = hv[v];
            // the record's unused slot too: whole lines go to memory
#pragma unroll
            for (int v = NV; v < kRecHW; ++v) HWA_(v, i) = 0.0;
        }
    }
}

// The residual (what k_residual forms, with the TimeStep cell part) of the
// cells of runs t .. tend-1, lane p of a run's group taking its cell p.
template <int NSP>
__device__ __forceinline__ void run_residual(const RC_G CaseDev &C, int t, int tend) {
    if (t >= tend) return;
    const GMesh M = gview(C.M);
    const GState S = gview(C.S);
    GThermo &th = *gp(C.th);
    GParams &PAR = *gp(C.par);
    const int q = 4 * t + (int)threadIdx.x % kRunMax, ss = M.ss;
    if ((int)threadIdx.x % kRunMax >= M.slot[SL_CNT * ss + q]) return;
    const int i = M.slot[SL_CELL * ss + q];
    int cfs[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) cfs[k] = M.slot[(SL_CF + k) * ss + q];
    double al[3];
    cell_step_dd(M, S, PAR, i, cfs, al);
    double h[NV];
    cell_residual<NSP>(M, S, th, PAR, i, cfs, h);
#pragma unroll
    for (int k = 0; k < NV; ++k) HWA_(k, i) = h[k];
#pragma unroll
    for (int k = NV; k < kRecHW; ++k) HWA_(k, i) = 0.0;
}

// One level of the sweep: kRunChunk runs of one case per thread block (the
// work item), or runs lo + ... below hi of the one case.
// NX (forward levels of the fused step): thread blocks nb1 and up form the
// residual of the next level's runs instead (items work2[b - nb1], or runs
// lo2 + ... below hi2), alongside the sweep: the forward sweep of a level
// reads the increments of lower levels only and writes those of its own
// cells, and the residual of a higher level's cell reads neither, so the
// next level sweeps with its residual, step and dd already in memory.
template <int NSP, bool FWD, bool RES, bool B, bool NX, bool TOP = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_lusgs(
        ITEM_ARGS, const int4 *__restrict__ work2, int lo2, int hi2, int nb1) {
    constexpr int R = kRunMax;
    if constexpr (NX) {
        if ((int)blockIdx.x >= nb1) {
            const int b = (int)blockIdx.x - nb1;
            if constexpr (B) {
                const int4 w_ = work2[b];
                run_residual<NSP>(gp(cases)[w_.x], w_.y + (int)threadIdx.x / R, w_.z);
            } else {
                run_residual<NSP>(*gp(cases), lo2 + (b * (int)blockDim.x + (int)threadIdx.x) / R, hi2);
            }
            return;
        }
    }
    const RC_G CaseDev *Cp_;
    int t, tend;
    if constexpr (B) {
        const int4 w_ = work[blockIdx.x];
        Cp_ = gp(cases) + w_.x;
        t = w_.y + (int)threadIdx.x / R;
        tend = w_.z;
    } else {
        Cp_ = gp(cases);
        t = lo + (int)(blockIdx.x * blockDim.x + threadIdx.x) / R;
        tend = hi;
    }
    extern __shared__ double lusgs_tl[];
    lusgs_runs<NSP, FWD, RES, false, TOP>(*Cp_, t, tend, 0, lusgs_tl);
}

// The dataflow sweep: forward sweep (with the residual) and backward sweep
// of every case of the launch set in one launch.  Each thread block (one
// wave, kRunChunk runs) takes items from its counter until its sub-list is
// done; the counter's next value is fetched while the current item runs.
template <int NSP>
__global__ void __launch_bounds__(kRunThreads) __attribute__((amdgpu_waves_per_eu(2))) k_lusgs_flow(const CaseDev *__restrict__ cases,
                                                            const int4 *__restrict__ items, FlowSpan span,
                                                            int *__restrict__ ctr) {
    extern __shared__ double lusgs_tl[];
    const int x = blockIdx.x % kXcdLists;
    const int beg = span.off[x], len = span.off[x + 1] - beg;
    RC_G int *cx = gp(ctr) + x;
    int k = __builtin_amdgcn_readfirstlane(
        threadIdx.x == 0 ? __hip_atomic_fetch_add(cx, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0);
    while (k < len) {
        const int nk = threadIdx.x == 0 ? __hip_atomic_fetch_add(cx, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                        : 0;
        const int4 w = items[beg + k];
        const RC_G CaseDev &C = gp(cases)[w.x];
        const int epoch = gp(C.S.err)[4];
        const int t = w.y + (int)threadIdx.x / kRunMax;
        if (w.w) lusgs_runs<NSP, true, true, true>(C, t, w.z, epoch, lusgs_tl);
        else lusgs_runs<NSP, false, false, true>(C, t, w.z, epoch, lusgs_tl);
        k = __builtin_amdgcn_readfirstlane(nk);
    }
}

// ---------------------------------------------------------------------------
// The persistent step (opt-in): the phases of the fused step -- the step
// counters, gradient + limiter, flux, each forward level (with its cells'
// residual and source), each backward level, update -- for nsteps steps in
// one cooperative launch, a grid barrier between consecutive phases.  Every
// phase calls the device functions of the kernels above on the same items,
// so the results are those of the step's kernels bit for bit (the residual
// is formed in each forward level's own lanes).  The barrier: each block
// writes its phase back (agent-scope release), counts itself in, waits for
// the count of the whole grid (a bounded wait: past kBarSpin polls it raises
// the abort flag and a device error, and every block leaves), then acquires
// (the L2 write-back and invalidate of the release and acquire fences are
// what makes one XCD's stores visible to the CUs of another).
// A cooperative launch has every block of the grid resident at once, so
// every wait ends.
constexpr int kBarSpin = 1 << 22;

__device__ bool persist_sync(RC_G unsigned *bar, unsigned &nbar, const RC_G CaseDev *cases) {
    __shared__ int ok_;
    // every wave's stores have reached the L2 before the block's write-back
    __builtin_amdgcn_s_waitcnt(0x0f70);   // s_waitcnt vmcnt(0)
    __syncthreads();
    if (threadIdx.x == 0) {
        ++nbar;
        const unsigned target = nbar * gridDim.x;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        for (int it = 0;; ++it) {
            if (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
            if (__hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                ok = 0;
                break;
            }
            if (it >= kBarSpin) {
                __hip_atomic_store(bar + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                flag_error(gp(cases[0].S.err), RC_ERR_DEVICE, 0, K_PERSIST);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ok_ = ok;
    }
    __syncthreads();
    return ok_ != 0;
}

template <int NSP>
__global__ void __launch_bounds__(kPersistThreads) k_step_persist(PersistSet a) {
    __shared__ DevThermo th_lds_;
    extern __shared__ double lusgs_tl[];
    const int tid = (int)threadIdx.x;
    const RC_G CaseDev *cases = gp(a.cases);
    RC_G unsigned *bar = gp(a.bar);
    unsigned nbar = 0;
    for (int st = 0; st < a.nsteps; ++st) {
        // the device step counters of the error reports (k_step_begin)
        if (blockIdx.x == 0)
            for (int c = tid; c < a.ncases; c += kPersistThreads) {
                RC_G int *err = gp(cases[c].S.err);
                err[4] = err[4] + 1;
                err[5] = err[4];
            }
        if (!persist_sync(bar, nbar, cases)) return;
        // gradient + limiter: one thread per cell
        for (int it = blockIdx.x; it < a.ncell; it += gridDim.x) {
            const int4 w = a.cell[it];
            const int i = w.y + tid;
            if (i < w.z) gradlim_cell<NSP, true, LIM_CASE>(cases[w.x], i);
        }
        if (!persist_sync(bar, nbar, cases)) return;
        // flux: two lanes per face, 128 faces per pass over an item
        int staged = -1;
        for (int it = blockIdx.x; it < a.fpass * a.nface; it += gridDim.x) {
            const int4 w = a.face[it / a.fpass];
            if (w.x != staged) {
                __syncthreads();
                stage_thermo(cases[w.x], th_lds_);
                __syncthreads();
                staged = w.x;
            }
            const int f = w.y + (it % a.fpass) * (kPersistThreads / 2) + (tid >> 1);
            if (f < w.z) flux2_face<NSP, true>(cases[w.x], f, tid & 1, th_lds_);
        }
        if (!persist_sync(bar, nbar, cases)) return;
        // the forward levels, each item of runs to one wave (16 runs a pass)
        constexpr int per = kPersistThreads / kRunThreads;
        for (int l = 0; l < a.nlev; ++l) {
            const int n = a.nlevitems[l];
            for (int it = blockIdx.x; it * per < n; it += gridDim.x) {
                const int u = it * per + tid / kRunThreads;
                if (u < n) {
                    const int4 w = a.lev[l][u];
                    for (int t0 = w.y; t0 < w.z; t0 += kRunChunk)
                        lusgs_runs<NSP, true, true, false>(cases[w.x], t0 + (tid % kRunThreads) / kRunMax, w.z, 0,
                                                           lusgs_tl);
                }
            }
            if (!persist_sync(bar, nbar, cases)) return;
        }
        // the backward levels
        for (int l = a.nlev - 1; l >= 0; --l) {
            const int n = a.nlevitems[l];
            for (int it = blockIdx.x; it * per < n; it += gridDim.x) {
                const int u = it * per + tid / kRunThreads;
                if (u < n) {
                    const int4 w = a.lev[l][u];
                    for (int t0 = w.y; t0 < w.z; t0 += kRunChunk)
                        lusgs_runs<NSP, false, false, false>(cases[w.x], t0 + (tid % kRunThreads) / kRunMax, w.z, 0,
                                                             lusgs_tl);
                }
            }
            if (!persist_sync(bar, nbar, cases)) return;
        }
        // update: one thread per cell
        staged = -1;
        for (int it = blockIdx.x; it < a.ncell; it += gridDim.x) {
            const int4 w = a.cell[it];
            if (w.x != staged) {
                __syncthreads();
                stage_thermo(cases[w.x], th_lds_);
                __syncthreads();
                staged = w.x;
            }
            const int i = w.y + tid;
            if (i < w.z) update_cell<NSP>(cases[w.x], i, th_lds_);
        }
        if (st + 1 < a.nsteps && !persist_sync(bar, nbar, cases)) return;
    }
}

// WA += dW (Time_LUSGS_steady.f90:193-197) and update_primitve_variables
// (update_flow_variables.f90:4-48), plus the NaN check of lu_sgs_stage;
// the new cell-average quantities go to the cache
// (four waves per SIMD: 127 VGPRs, no spill; three at the compiler's own choice)
// (the work of cell i of case C, the thermo table staged in LDS; k_update
// and the persistent step call it)
template <int NSP, bool SF = false>
__device__ __forceinline__ void update_cell(const RC_G CaseDev &C, int i, const DevThermo &th) {
    const GMesh M = gview(C.M);
    const GState S = gview(C.S);
    // SF: the last kernel of the fused step ends the step on the device
    // counter (see k_step_begin)
    if (SF && i == 0) S.err[5] = S.err[4];
    double w[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) { w[k] = WA_(k, i) + HWA_(k, i); WA_(k, i) = w[k]; }
    const double rho = w[0], u = w[1] / rho, v = w[2] / rho;
    double y[NSP], s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) y[k] = w[kNFlow + k] / rho;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) s = s + y[k];
    y[NSP - 1] = 1.0 - s;
    double t, cp, gam, r;
    bool ok = newton_temperature<NSP>(th, w[3] / rho - 0.5 * (u * u + v * v), y, PA_(3, i), t, cp, gam, r);
    const double p = rho * r * t;
    PA_(0, i) = p;
    PA_(1, i) = u;
    PA_(2, i) = v;
    PA_(3, i) = t;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) PA_(kNFlow + k, i) = y[k];
#pragma unroll
    for (int k = NV; k < kRecPA; ++k) PA_(k, i) = 0.0;     // whole record: whole lines go to memory
    ok = write_lucoef<NSP>(M, S, th, i, p, u, v, t, y) && ok;
    if (!ok) flag_error(S.err, RC_ERR_TEMPERATURE, i, K_UPDATE);
    if (i == 0 && isnan(w[0])) flag_error(S.err, RC_ERR_NAN, 0, K_UPDATE);
}

template <int NSP, bool SF, bool B>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_update(ITEM_ARGS) {
    __shared__ DevThermo th_lds_;
    ITEM_CASE_END(i, end_);
    stage_thermo(*Cp_, th_lds_);
    __syncthreads();
    if (i >= end_) return;
    update_cell<NSP, SF>(*Cp_, i, th_lds_);
}

// primitve_to_conserved_variables (update_flow_variables.f90:50-84), and
// the cell-average cache of the new state
template <int NSP, bool B>
__global__ void __launch_bounds__(256) k_prim_to_cons(ITEM_ARGS) {
    ITEM_PROLOGUE(i);
    double pa[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) pa[k] = PA_(k, i);
    if (!cell_prim_to_cons<NSP>(M, S, TH, i, pa)) flag_error(S.err, RC_ERR_TEMPERATURE, i, K_P2C);
}

// the derived cell fields of output_vtk (output.f90:26-40): density,
// Mach number, and the production rates of the current state as sourceterm
// forms them (sourceterm.f90:17-43; zero without combustion); out holds
// rho, Ma and NSP-1 rates per cell
template <int NSP, bool B>
__global__ void __launch_bounds__(256) k_derived(ITEM_ARGS, double *__restrict__ out) {
    ITEM_PROLOGUE(i);
    const double pre = PA_(0, i), u = PA_(1, i), v = PA_(2, i), tem = PA_(3, i);
    double y[NSP], s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) { y[k] = PA_(kNFlow + k, i); s = s + y[k]; }
    y[NSP - 1] = 1.0 - s;
    double cp, gama, rcp;
    if (!gas_param<NSP>(TH, y, tem, cp, gama, rcp)) flag_error(S.err, RC_ERR_TEMPERATURE, i, K_P2C);
    const double rho = pre / (rcp * tem);
    const double ac = sqrt(gama * rcp * tem);
    RC_G double *o = gp(out) + (size_t)i * (NSP + 1);
    o[0] = rho;
    o[1] = sqrt(u * u + v * v) / ac;
    double w[NSP];
#pragma unroll
    for (int k = 0; k < NSP; ++k) w[k] = 0.0;
    if (PAR.combustion) production_rate<NSP>(TH, tem, pre / (tem * rcp), y, w);
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) o[2 + k] = w[k];
}

// Residual (residual.f90:4-21): partial sums of vol*|hWA| and max(WA(1))
// per cell chunk (one work item per block, slot w.w of the case's partials),
// then one block per case folds its partials in a fixed order, so results
// are deterministic (the summation order is not the reference's serial one:
// the norms agree to rounding).
template <int NSP, bool B>
__global__ void __launch_bounds__(256) k_resnorm_partial(ITEM_ARGS) {
    __shared__ double sh[256][NV + 1];
    const RC_G CaseDev *Cp_;
    int i, iend, slot;
    if constexpr (B) {
        const int4 w_ = work[blockIdx.x];
        Cp_ = gp(cases) + w_.x;
        i = w_.y + (int)threadIdx.x;
        iend = w_.z;
        slot = w_.w;
    } else {
        Cp_ = gp(cases);
        i = lo + (int)(blockIdx.x * blockDim.x + threadIdx.x);
        iend = hi;
        slot = blockIdx.x;
    }
    const RC_G CaseDev &C = *Cp_;
    const GMesh M = gview(C.M);
    const GState S = gview(C.S);
    const int tid = threadIdx.x;
    double acc[NV + 1];
#pragma unroll
    for (int k = 0; k <= NV; ++k) acc[k] = 0.0;
    if (i < iend) {
        const double v = M.vol[i];
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] = v * fabs(HWA_(k, i));
        acc[NV] = WA_(0, i);
    } else {
        acc[NV] = -INFINITY;
    }
#pragma unroll
    for (int k = 0; k <= NV; ++k) sh[tid][k] = acc[k];
    __syncthreads();
    // a tree over the block's threads (64, 128 or 256: a power of two)
    for (int h = (int)blockDim.x / 2; h > 0; h >>= 1) {
        if (tid < h) {
#pragma unroll
            for (int k = 0; k < NV; ++k) sh[tid][k] += sh[tid + h][k];
            sh[tid][NV] = fmax(sh[tid][NV], sh[tid + h][NV]);
        }
        __syncthreads();
    }
    if (tid <= NV) gp(C.partial)[slot * (NV + 1) + tid] = sh[0][tid];
}

__global__ void __launch_bounds__(256) k_resnorm_final(const CaseDev *__restrict__ cases, double *out) {
    __shared__ double sh[256][NV + 1];
    const RC_G CaseDev &C = gp(cases)[blockIdx.x];
    const RC_G double *partial = gp(C.partial);
    const int tid = threadIdx.x;
    double acc[NV + 1];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    acc[NV] = -INFINITY;
    for (int b = tid; b < C.npartial; b += 256) {
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] += partial[b * (NV + 1) + k];
        acc[NV] = fmax(acc[NV], partial[b * (NV + 1) + NV]);
    }
#pragma unroll
    for (int k = 0; k <= NV; ++k) sh[tid][k] = acc[k];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) {
#pragma unroll
            for (int k = 0; k < NV; ++k) sh[tid][k] += sh[tid + h][k];
            sh[tid][NV] = fmax(sh[tid][NV], sh[tid + h][NV]);
        }
        __syncthreads();
    }
    if (tid < NV) out[blockIdx.x * (NV + 1) + tid] = sh[0][tid] * C.inv_volume;
    if (tid == NV) out[blockIdx.x * (NV + 1) + tid] = sh[0][NV];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

void launch_ts_face(const LaunchSet &L, hipStream_t st) { LAUNCH(L.face, kChunk, st, k_timestep_face, kNsp); }

void launch_ts_cell(const LaunchSet &L, hipStream_t st) { LAUNCH(L.cell, kChunk, st, k_timestep_cell, kNsp); }

void launch_ghost(const LaunchSet &L, hipStream_t st) { LAUNCH(L.ghost, kChunk, st, k_ghost, kNsp); }

// grad: gradient; limit: 0 none, 1 Barth limiter always, 2 when the
// case's limit_ID > 0
void launch_gradlim(const LaunchSet &L, bool grad, int limit, hipStream_t st) {
    if (grad && limit == 2) LAUNCH(L.cell, kChunk, st, k_gradlim, kNsp, true, LIM_CASE, false);
    else if (grad && limit == 1) LAUNCH(L.cell, kChunk, st, k_gradlim, kNsp, true, LIM_ALWAYS, false);
    else if (grad) LAUNCH(L.cell, kChunk, st, k_gradlim, kNsp, true, LIM_NONE, false);
    else if (limit == 1) LAUNCH(L.cell, kChunk, st, k_gradlim, kNsp, false, LIM_ALWAYS, false);
}
// reconstruction as the first kernel of the fused step: it also begins the
// step on the device counter (no k_step_begin launch)
void launch_gradlim_step(const LaunchSet &L, hipStream_t st) {
    LAUNCH(L.cell, kChunk, st, k_gradlim, kNsp, true, LIM_CASE, true);
}

// LDS of a sweep block: three faces' terms (8 doubles each) per thread
constexpr size_t lusgs_lds(int threads) { return (size_t)threads * 3 * 8 * sizeof(double); }

void launch_step_begin(const LaunchSet &L, hipStream_t st) {
    k_step_begin<<<(std::max(L.ncases, kXcdLists) + 63) / 64, 64, 0, st>>>(L.cases, L.ncases, L.flow.ctr);
}

// the whole LU-SGS sweep of the fused step (forward with the residual, then
// backward) as one launch; needs k_step_begin before it in the same stream
void launch_lusgs_flow(const LaunchSet &L, hipStream_t st) {
    if (L.flow.grid > 0)
        k_lusgs_flow<kNsp><<<L.flow.grid, kRunThreads, lusgs_lds(kRunThreads), st>>>(L.cases, L.flow.w, L.flow.span,
                                                                                    L.flow.ctr);
}

// blocks of the dataflow sweep resident at once on the device (its
// occupancy times the CU count), at least kXcdLists
int flow_max_blocks(int device) {
    static int cache[64] = {};
    if (device < 0 || device >= 64) return kXcdLists;
    if (cache[device]) return cache[device];
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_lusgs_flow<kNsp>, kRunThreads,
                                                     lusgs_lds(kRunThreads)) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return kXcdLists;
    cache[device] = std::max(kXcdLists, per_cu * cus / kXcdLists * kXcdLists);
    return cache[device];
}

// two lanes per face (k_flux2; RC_FLUX1=1: k_flux, one thread per face)
template <bool TS>
void launch_flux2(const LaunchSet &L, hipStream_t st) {
    const WorkList &wl = L.face;
    if (L.single) {
        const int n = wl.hi - wl.lo;
        if (n > 0) k_flux2<kNsp, TS, false><<<(n + kChunk - 1) / kChunk, 2 * kChunk, 0, st>>>(L.cases, nullptr, wl.lo, wl.hi);
    } else if (wl.n) {
        k_flux2<kNsp, TS, true><<<wl.n, 2 * wl.bs, 0, st>>>(L.cases, wl.w, 0, 0);
    }
}

bool flux_one_lane() {
    const char *e = std::getenv("RC_FLUX1");
    return e && e[0] == '1';
}

void launch_flux(const LaunchSet &L, bool ts, hipStream_t st) {
    if (!flux_one_lane()) {
        if (ts) launch_flux2<true>(L, st);
        else launch_flux2<false>(L, st);
        return;
    }
    if (ts) LAUNCH(L.face, kChunk, st, k_flux, kNsp, true);
    else LAUNCH(L.face, kChunk, st, k_flux, kNsp, false);
}

void launch_residual(const LaunchSet &L, bool ts, hipStream_t st) {
    if (ts) LAUNCH(L.cell, kChunk, st, k_residual, kNsp, true);
    else LAUNCH(L.cell, kChunk, st, k_residual, kNsp, false);
}

// res: forward levels that also form their cells' residual; nx: forward
// levels that also form the residual of the next level's cells (list nx)
template <bool FWD, bool RES, bool NX, bool TOP = false>
void launch_level(const LaunchSet &L, const WorkList &wl, const WorkList *nx, hipStream_t st) {
    if (L.single) {
        const int n = wl.hi - wl.lo, nb1 = (n + kRunChunk - 1) / kRunChunk;
        const int n2 = NX ? nx->hi - nx->lo : 0, nb2 = (n2 + kRunChunk - 1) / kRunChunk;
        if (nb1 + nb2 > 0)
            k_lusgs<kNsp, FWD, RES, false, NX, TOP><<<nb1 + nb2, kRunThreads, lusgs_lds(kRunThreads), st>>>(
                L.cases, nullptr, wl.lo, wl.hi, nullptr, NX ? nx->lo : 0, NX ? nx->hi : 0, nb1);
    } else {
        const int nb2 = NX ? nx->n : 0;
        if (wl.n + nb2 > 0)
            k_lusgs<kNsp, FWD, RES, true, NX, TOP><<<wl.n + nb2, wl.bs, lusgs_lds(wl.bs), st>>>(
                L.cases, wl.w, 0, 0, NX ? nx->w : nullptr, 0, 0, wl.n);
    }
}

// the forward level l of the fused step: it forms its own cells' residual
// unless the launch of level l-1 formed it, and the next level's when
// L.next_res[l] says so (the level lists then share one block size); the
// highest level also runs its backward sweep when L.top_fused
void launch_lusgs_fused_forward(const LaunchSet &L, int l, hipStream_t st) {
    const bool res = l == 0 || !L.next_res[l - 1];
    const bool nx = L.next_res[l] != 0;
    if (L.top_fused && l + 1 == (int)L.level.size()) {
        // the highest level: its backward sweep in the same launch
        if (res) launch_level<true, true, false, true>(L, L.level[l], nullptr, st);
        else launch_level<true, false, false, true>(L, L.level[l], nullptr, st);
        return;
    }
    const WorkList *n = nx ? &L.level[l + 1] : nullptr;
    if (res && nx) launch_level<true, true, true>(L, L.level[l], n, st);
    else if (res) launch_level<true, true, false>(L, L.level[l], nullptr, st);
    else if (nx) launch_level<true, false, true>(L, L.level[l], n, st);
    else launch_level<true, false, false>(L, L.level[l], nullptr, st);
}

// threads of the forward level kernel with the next level's residual blocks
// resident at once on the device (blocks of bs threads)
long lusgs_resident_threads(int device, int bs) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, long> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({device, bs});
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_lusgs<kNsp, true, true, true, true>, bs,
                                                     lusgs_lds(bs)) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 0;
    const long n = (long)per_cu * cus * bs;
    cache[{device, bs}] = n;
    return n;
}

void launch_lusgs_level(const LaunchSet &L, bool fwd, int l, bool res, hipStream_t st) {
    if (fwd && res) launch_level<true, true, false>(L, L.level[l], nullptr, st);
    else if (fwd) launch_level<true, false, false>(L, L.level[l], nullptr, st);
    else launch_level<false, false, false>(L, L.level[l], nullptr, st);
}

void launch_update(const LaunchSet &L, hipStream_t st) { LAUNCH(L.cell, kChunk, st, k_update, kNsp, false); }
// the update as the last kernel of the fused step: it also ends the step on
// the device counter
void launch_update_step(const LaunchSet &L, hipStream_t st) { LAUNCH(L.cell, kChunk, st, k_update, kNsp, true); }

void launch_prim_to_cons(const LaunchSet &L, hipStream_t st) { LAUNCH(L.cell, kChunk, st, k_prim_to_cons, kNsp); }

void launch_derived(const LaunchSet &L, double *out, hipStream_t st) {
    if (L.single) {
        const int n = L.cell.hi - L.cell.lo;
        if (n > 0)
            k_derived<kNsp, false><<<(n + kChunk - 1) / kChunk, kChunk, 0, st>>>(L.cases, nullptr, L.cell.lo, L.cell.hi,
                                                                                 out);
    } else if (L.cell.n) {
        k_derived<kNsp, true><<<L.cell.n, kChunk, 0, st>>>(L.cases, L.cell.w, 0, 0, out);
    }
}

// the error flags err[0..3] of every case of the set, out[4c + k]
__global__ void k_gather_err(const CaseDev *__restrict__ cases, int n, int *__restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) {
        const RC_G int *e = gp(gp(cases)[c].S.err);
#pragma unroll
        for (int k = 0; k < 4; ++k) out[4 * c + k] = e[k];
    }
}

void launch_gather_err(const LaunchSet &L, int *out, hipStream_t st) {
    if (L.ncases > 0) k_gather_err<<<(L.ncases + 255) / 256, 256, 0, st>>>(L.cases, L.ncases, out);
}

// ---------------------------------------------------------------------------
// the persistent step on the host side
// ---------------------------------------------------------------------------
constexpr size_t persist_lds() { return (size_t)kPersistThreads * 3 * 8 * sizeof(double); }

// whether a launch set can run as the persistent step: at most
// kPersistMaxLevels levels, cell and face items of at most kPersistThreads
// (a wave sweeps a level item kRunChunk runs at a time, whatever its size)
bool persist_supported(const LaunchSet &L) {
    if (L.level.size() > (size_t)kPersistMaxLevels || L.cell.bs > kPersistThreads ||
        L.face.bs > kPersistThreads || !L.cell.w || !L.face.w)
        return false;
    for (const WorkList &w : L.level)
        if (w.n > 0 && !w.w) return false;
    return true;
}

// blocks of the persistent step resident at once on the device
int persist_max_blocks(int device) {
    static std::mutex mu;
    static std::map<int, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(device);
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_step_persist<kNsp>, kPersistThreads, persist_lds()) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 0;
    cache[device] = per_cu * cus;
    return per_cu * cus;
}

// blocks a launch set's persistent step uses: as many as its widest phase
// can keep busy, no more than are resident at once (RC_PERSIST_GRID=n caps it)
int persist_grid(const LaunchSet &L, int device) {
    int need = std::max(L.cell.n, (L.face.bs + kPersistThreads / 2 - 1) / (kPersistThreads / 2) * L.face.n);
    for (const WorkList &w : L.level) need = std::max(need, (w.n + kPersistThreads / kRunThreads - 1) /
                                                                    (kPersistThreads / kRunThreads));
    int g = std::min(std::max(need, 1), persist_max_blocks(device));
    if (const char *e = std::getenv("RC_PERSIST_GRID")) {
        const int cap = std::atoi(e);
        if (cap > 0) g = std::min(g, cap);
    }
    return g;
}

// nsteps steps of the launch set as one cooperative launch; bar: two
// device words, cleared here
hipError_t launch_step_persist(const LaunchSet &L, unsigned *bar, int nsteps, int grid, hipStream_t st) {
    PersistSet a{};
    a.cases = L.cases;
    a.ncases = L.ncases;
    a.cell = L.cell.w;
    a.ncell = L.cell.n;
    a.face = L.face.w;
    a.nface = L.face.n;
    a.fpass = (L.face.bs + kPersistThreads / 2 - 1) / (kPersistThreads / 2);
    a.nlev = (int)L.level.size();
    for (int l = 0; l < a.nlev; ++l) {
        a.lev[l] = L.level[l].w;
        a.nlevitems[l] = L.level[l].n;
    }
    a.bar = bar;
    a.nsteps = nsteps;
    hipError_t e = hipMemsetAsync(bar, 0, 2 * sizeof(unsigned), st);
    if (e != hipSuccess) return e;
    void *args[] = {&a};
    return hipLaunchCooperativeKernel((const void *)k_step_persist<kNsp>, dim3(grid), dim3(kPersistThreads), args,
                                      persist_lds(), st);
}

void launch_resnorm(const LaunchSet &L, hipStream_t st) {
    LAUNCH(L.norm, kChunk, st, k_resnorm_partial, kNsp);
    k_resnorm_final<<<L.ncases, 256, 0, st>>>(L.cases, L.resout);
}

}  // namespace rc
