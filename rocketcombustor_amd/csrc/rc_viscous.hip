// FP64 HIP face kernel of the viscous path (IfViscous): TimeStep's viscous
// spectral radius (TimeStep.f90:73-85) and flux_viscous
// (flux_viscous.f90:4-350) with compute_transport_property
// (transport_property.f90:78-125) and compute_diffusion_velocity
// (diffusion_velocity.f90:55-117).
//
// One thread per face, gather-only like the inviscid flux: the face writes
// its record (alpa, Fv(2:Nvar), alpaVis) and the cells add the viscous
// fluxes after the inviscid ones, in face order, when they assemble their
// residual.  The pair constants of the Wilke rule and of the binary
// diffusion coefficients come precomputed in the case's transport table;
// both tables are staged in LDS by the block.
#include "rc_items.h"

namespace rc {

// compute_transport_property: mixture viscosity (Wilke) and conductivity;
// false above the table's range (the reference stops there)
template <int NSP, class TR, class TT>
__device__ __forceinline__ bool transport_property(TR &tr, TT &th, double tem2, const double *yy, double &vmul,
                                                   double &akmu) {
    const double tem = fmax(tr.t_low, tem2);
    double xx[NSP], s = 0.0;
#pragma unroll
    for (int i = 0; i < NSP; ++i) s = s + yy[i] / th.mw[i];
    const double ww = 1. / s;
#pragma unroll
    for (int i = 0; i < NSP; ++i) xx[i] = ww * yy[i] / th.mw[i];
    const bool low = tr.t_low <= tem && tem <= tr.t_crit;
    const bool ok = low || (tr.t_crit < tem && tem <= tr.t_high);
    const auto *mu = low ? tr.mu_low : tr.mu_high;
    const auto *kk = low ? tr.k_low : tr.k_high;
    const double lt = log(tem);
    double mul[NSP], kl[NSP];
#pragma unroll
    for (int i = 0; i < NSP; ++i) {
        mul[i] = exp(mu[i][0] * lt + mu[i][1] / tem + mu[i][2] / (tem * tem) + mu[i][3]);
        kl[i] = exp(kk[i][0] * lt + kk[i][1] / tem + kk[i][2] / (tem * tem) + kk[i][3]);
    }
#pragma unroll
    for (int i = 0; i < NSP; ++i) { mul[i] = mul[i] * 1.0e-7; kl[i] = kl[i] * 1.0e-4; }
    double v = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
    for (int i = 0; i < NSP; ++i) {
        double phil = 0.;
#pragma unroll
        for (int j = 0; j < NSP; ++j) {
            const double t = 1. + sqrt(mul[i] / mul[j]) * tr.wa[i][j];
            phil = phil + xx[j] * (t * t) / tr.wb[i][j];
        }
        phil = 0.25 * sqrt(2.) * phil;
        v = v + xx[i] * mul[i] / phil;
    }
#pragma unroll
    for (int i = 0; i < NSP; ++i) a1 = a1 + xx[i] * kl[i];
#pragma unroll
    for (int i = 0; i < NSP; ++i) a2 = a2 + xx[i] / kl[i];
    vmul = v;
    akmu = 0.5 * (a1 + 1. / a2);
    return ok;
}

// compute_diffusion_velocity: mixture-averaged coefficients from the binary
// ones (Chapman-Enskog, Neufeld collision-integral fit); D_ij = D_ji, so each
// pair is formed once
template <int NSP, class TR, class TT>
__device__ __forceinline__ void diffusion_coef(TR &tr, TT &th, double pp, double tem, const double *yy,
                                               double *dk) {
    constexpr double A = 1.06036, B = 0.15610, C = 0.19300, D = 0.47635, E = 1.03587, F = 1.52996,
                     G = 1.76474, H = 3.89411;
    double xx[NSP], s = 0.0;
#pragma unroll
    for (int i = 0; i < NSP; ++i) s = s + yy[i] / th.mw[i];
    const double ww = 1. / s;
#pragma unroll
    for (int i = 0; i < NSP; ++i) xx[i] = ww * yy[i] / th.mw[i];
    const double t32 = sqrt(tem * tem * tem);
    double dij[NSP][NSP];
#pragma unroll
    for (int i = 0; i < NSP; ++i)
#pragma unroll
        for (int j = i + 1; j < NSP; ++j) {
            const double tstar = tem / tr.sqeps[i][j];
            const double om = A * pow(tstar, -B) + C * exp(-D * tstar) + E * exp(-F * tstar) + G * exp(-H * tstar);
            dij[i][j] = 0.0266 * t32 / (pp * tr.sqmw[i][j] * tr.cd2[i][j] * om);
            dij[j][i] = dij[i][j];
        }
#pragma unroll
    for (int i = 0; i < NSP; ++i) {
        double xd = 0.;
#pragma unroll
        for (int j = 0; j < NSP; ++j)
            if (j != i) xd = xd + xx[j] / dij[i][j];
        dk[i] = (1. - xx[i]) / (xd + 1.0e-16);
    }
}

__device__ __forceinline__ void removedirection(double &vx, double &vy, double nx, double ny) {
    const double vn = vx * nx + vy * ny;
    vx = vx - vn * nx;
    vy = vy - vn * ny;
}

// TS: the viscous part of TimeStep's face loop (alpaVis; alpaF = alpa +
// alpaVis/vol(NL), alpa as the inviscid face kernel left it in the face
// record).  FLUX: the viscous flux of the face.  Two waves per SIMD (one
// at the compiler's own choice of 256 VGPRs), with a 108-byte spill.
template <int NSP, bool TS, bool FLUX, bool B>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_viscous(ITEM_ARGS) {
    __shared__ DevThermo th_lds_;
    __shared__ DevTransport tr_lds_;
    const RC_G CaseDev *Cp_;
    int f, end_;
    if constexpr (B) {
        const int4 w_ = work[blockIdx.x];
        Cp_ = gp(cases) + w_.x;
        f = w_.y + (int)threadIdx.x;
        end_ = w_.z;
    } else {
        Cp_ = gp(cases);
        f = lo + (int)(blockIdx.x * blockDim.x + threadIdx.x);
        end_ = hi;
    }
    const RC_G CaseDev &C = *Cp_;
    // a case without transport tables is inviscid: its whole work item ends
    if (C.tr == nullptr) return;
    {
        const RC_G double *src_ = (const RC_G double *)gp(C.th);
        double *dst_ = reinterpret_cast<double *>(&th_lds_);
        for (int k_ = threadIdx.x; k_ < (int)(sizeof(DevThermo) / 8); k_ += blockDim.x) dst_[k_] = src_[k_];
        const RC_G double *src2_ = (const RC_G double *)gp(C.tr);
        double *dst2_ = reinterpret_cast<double *>(&tr_lds_);
        for (int k_ = threadIdx.x; k_ < (int)(sizeof(DevTransport) / 8); k_ += blockDim.x) dst2_[k_] = src2_[k_];
    }
    __syncthreads();
    if (f >= end_) return;
    const GMesh M = gview(C.M);
    const GState S = gview(C.S);
    const DevThermo &th = th_lds_;
    const DevTransport &tr = tr_lds_;
    GParams &P = *gp(C.par);
    const int nl = M.fnl[f], nr = M.fnr[f];
    const double dx = M.vecx[f], dy = M.vecx[M.sf + f];

    if (TS) {
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
        (void)uu; (void)vv;
        double cp, gam, r, vmul, akmu;
        bool ok = gas_param<NSP>(th, y, tt, cp, gam, r);
        ok = transport_property<NSP>(tr, th, tt, y, vmul, akmu) && ok;
        if (!ok) flag_error(S.err, RC_ERR_TEMPERATURE, f, K_VISC);
        const double rr = pp / (r * tt);
        const double prl = vmul * cp / akmu;
        const double alpavis = 2. * (dx * dx + dy * dy) * fmax(4. / 3., gam / prl) * vmul / rr;
        const double alpa = ALPHA_(f);
        FV_(0, f) = alpa;
        FV_(kRecF - 1, f) = alpavis;
        ALPHA_(f) = alpa + alpavis / M.vol[nl];
    }
    if (!FLUX) return;

    constexpr double yita = 4. / 3;
    const double sav1 = dx, sav2 = dy;
    const double sav = fmax(sqrt(dx * dx + dy * dy), kSmall);
    const double s1n = sav1 / sav, s2n = sav2 / sav;
    const double v0 = M.fvl[f], v1 = M.fvl[M.sf + f];
    double ql[NV], qr[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) ql[k] = PA_(k, nl) + (GR_(k, 0, nl) * v0 + GR_(k, 1, nl) * v1);
    double pl = ql[0], ul = ql[1], vl = ql[2], tl = ql[3];
    double yl[NSP], yr[NSP], s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) { yl[k] = ql[kNFlow + k]; s = s + yl[k]; }
    yl[NSP - 1] = 1. - s;
    double cpl, gaml, rcl;
    bool ok = gas_param<NSP>(th, yl, tl, cpl, gaml, rcl);
    double rl = pl / (rcl * tl);
    double pr, ur, vr, tr_, rr;
    if (nr >= 0) {
        const double w0 = M.fvr[f], w1 = M.fvr[M.sf + f];
#pragma unroll
        for (int k = 0; k < NV; ++k) qr[k] = PA_(k, nr) + (GR_(k, 0, nr) * w0 + GR_(k, 1, nr) * w1);
        pr = qr[0]; ur = qr[1]; vr = qr[2]; tr_ = qr[3];
        s = 0.0;
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k) { yr[k] = qr[kNFlow + k]; s = s + yr[k]; }
        yr[NSP - 1] = 1. - s;
        double cpr, gamr, rcr;
        ok = gas_param<NSP>(th, yr, tr_, cpr, gamr, rcr) && ok;
        rr = pr / (rcr * tr_);
    } else {
#pragma unroll
        for (int k = 0; k < NV; ++k) qr[k] = 0.0;
#pragma unroll
        for (int k = 0; k < NSP; ++k) yr[k] = yl[k];
        const int type = -nr;
        if (type == 1) {
            rr = rl; pr = pl; tr_ = tl;
            ul = 0.; vl = 0.;
            ur = 0.; vr = 0.;
        } else if (type == 2 || type == 6) {
            removedirection(ul, vl, s1n, s2n);
            rr = rl; pr = pl; tr_ = tl;
            ur = ul; vr = vl;
        } else if (type == 3 || type == 4) {
            const RC_G double *yin = type == 3 ? P.ox_y : P.fu_y;
#pragma unroll
            for (int k = 0; k < NSP; ++k) yr[k] = yin[k];
            tr_ = type == 3 ? P.ox_t : P.fu_t;
            double cpr, gamr, rcr;
            ok = gas_param<NSP>(th, yr, tr_, cpr, gamr, rcr) && ok;
            pr = pl;
            rr = pr / (rcr * tr_);
            ur = (type == 3 ? P.ox_g : P.fu_g) / rr;
            vr = 0.0;
            pl = pr; ul = ur; vl = vr; tl = tr_; rl = rr;
#pragma unroll
            for (int k = 0; k < NSP; ++k) yl[k] = yr[k];
        } else if (type == 5) {
            const double ac = sqrt(gaml * pl / rl);
            const double vn = ul * s1n + vl * s2n;
            if (fabs(vn) > ac) {
                rr = rl; ur = ul; vr = vl; pr = pl; tr_ = tl;
            } else {
                pr = P.p_out;
                ur = ul + s1n * (pl - pr) / (rl * ac);
                vr = vl + s2n * (pl - pr) / (rl * ac);
                rr = rl + (pr - pl) / (ac * ac);
                tr_ = pr / (rcl * rr);
                pl = pr; ul = ur; vl = vr; tl = tr_; rl = rr;
            }
        } else {
            rr = rl; pr = pl; tr_ = tl;
            ur = ul; vr = vl;
        }
    }
    (void)rr;
    const double pp = 0.5 * (pl + pr), uu = 0.5 * (ul + ur), vv = 0.5 * (vl + vr), tem = 0.5 * (tl + tr_);
    double yy[NSP];
#pragma unroll
    for (int k = 0; k < NSP; ++k) yy[k] = 0.5 * (yl[k] + yr[k]);
    double cp, gam, rc;
    ok = gas_param<NSP>(th, yy, tem, cp, gam, rc) && ok;
    const double rho = pp / (rc * tem);
    // gradients at the face: fbaseGrad_dxdy scales the scaled-coordinate
    // gradients by 1/reflen; interior faces average both cells and add the
    // face-normal jump term
    double gx[NV], gy[NV];
    const double ixl = M.vgeo[f], iyl = M.vgeo[M.sf + f];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        gx[k] = GR_(k, 0, nl) * ixl + GR_(k, 1, nl) * 0.;
        gy[k] = GR_(k, 0, nl) * 0. + GR_(k, 1, nl) * iyl;
    }
    if (nr >= 0) {
        const double ixr = M.vgeo[2 * M.sf + f], iyr = M.vgeo[3 * M.sf + f];
        const double dlrx = M.vgeo[4 * M.sf + f], dlry = M.vgeo[5 * M.sf + f];
        const double dxk = fabs(dlrx * s1n + dlry * s2n);
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            gx[k] = 0.5 * (gx[k] + (GR_(k, 0, nr) * ixr + GR_(k, 1, nr) * 0.)) + yita * s1n * (qr[k] - ql[k]) / dxk;
            gy[k] = 0.5 * (gy[k] + (GR_(k, 0, nr) * 0. + GR_(k, 1, nr) * iyr)) + yita * s2n * (qr[k] - ql[k]) / dxk;
        }
    }
    double dudx = gx[1], dudy = gy[1], dvdx = gx[2], dvdy = gy[2], dtdx = gx[3], dtdy = gy[3];
    double dydx[NSP], dydy[NSP];
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) { dydx[k] = gx[kNFlow + k]; dydy[k] = gy[kNFlow + k]; }
    if (nr == -1 || nr == -6) removedirection(dtdx, dtdy, s1n, s2n);
    if (nr == -2) {
        const double snx = s1n, sny = s2n, stx = -sny, sty = snx;
        removedirection(dtdx, dtdy, snx, sny);
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k) removedirection(dydx[k], dydy[k], snx, sny);
        double dutdx = dudx * stx + dvdx * sty, dutdy = dudy * stx + dvdy * sty;
        double dundx = dudx * snx + dvdx * sny, dundy = dudy * snx + dvdy * sny;
        removedirection(dutdx, dutdy, snx, sny);
        removedirection(dundx, dundy, stx, sty);
        dudx = dutdx * stx + dundx * snx;
        dudy = dutdy * stx + dundy * snx;
        dvdx = dutdx * sty + dundx * sny;
        dvdy = dutdy * sty + dundy * sny;
    }
    s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) s = s + dydx[k];
    dydx[NSP - 1] = -s;
    s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) s = s + dydy[k];
    dydy[NSP - 1] = -s;
    double entha[NSP], vmulc, akmu, dk[NSP];
    ok = enthalpy<NSP>(th, tem, entha) && ok;
    ok = transport_property<NSP>(tr, th, tem, yy, vmulc, akmu) && ok;
    if (!ok) flag_error(S.err, RC_ERR_TEMPERATURE, f, K_VISC);
    diffusion_coef<NSP>(tr, th, pp, tem, yy, dk);
    double uicx = 0.0, uicy = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) uicx = uicx + dk[k] * dydx[k];
#pragma unroll
    for (int k = 0; k < NSP; ++k) uicy = uicy + dk[k] * dydy[k];
    const double vlac = -2. / 3. * vmulc;
    const double div = dudx + dvdy;
    const double txx = 2. * vmulc * dudx + vlac * div;
    const double tyy = 2. * vmulc * dvdy + vlac * div;
    const double txy = vmulc * (dudy + dvdx);
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) sx = sx + (dk[k] * dydx[k] - yy[k] * uicx) * entha[k];
#pragma unroll
    for (int k = 0; k < NSP; ++k) sy = sy + (dk[k] * dydy[k] - yy[k] * uicy) * entha[k];
    const double qx = -akmu * dtdx - rho * sx;
    const double qy = -akmu * dtdy - rho * sy;
    const double btx = uu * txx + vv * txy - qx;
    const double bty = uu * txy + vv * tyy - qy;
    FV_(1, f) = -(sav1 * txx + sav2 * txy);
    FV_(2, f) = -(sav1 * txy + sav2 * tyy);
    FV_(3, f) = -(sav1 * btx + sav2 * bty);
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) {
        const double spx = rho * (dk[k] * dydx[k] - yy[k] * uicx);
        const double spy = rho * (dk[k] * dydy[k] - yy[k] * uicy);
        FV_(kNFlow + k, f) = -(sav1 * spx + sav2 * spy);
    }
}

// ts: the TimeStep part; flux: the viscous flux.  Launched for viscous
// cases only; in a batch, inviscid members' work items return at once.
void launch_viscous(const LaunchSet &L, bool ts, bool flux, hipStream_t st) {
    if (ts && flux) LAUNCH(L.face, kChunk, st, k_viscous, kNsp, true, true);
    else if (ts) LAUNCH(L.face, kChunk, st, k_viscous, kNsp, true, false);
    else if (flux) LAUNCH(L.face, kChunk, st, k_viscous, kNsp, false, true);
}

}  // namespace rc
