// Item macros and leaf functions shared by the kernel files: record
// accessors of the device state, the work-item prologue of every per-item
// kernel (one case over a range, or a batch work list), the launcher macro,
// and the cell-average cache writer.
#pragma once

#include "rc_device.h"

namespace rc {

// gathered arrays are records per cell / face (one or two cache lines per
// gather); WA, only ever read by its own cell, stays one array per variable
#define PA_(k, i) S.pa[(i) * kRecPA + (k)]
#define WA_(k, i) S.wa[(k) * M.sc + (i)]
#define HWA_(k, i) S.hwa[(i) * kRecHW + (k)]
#define GR_(k, r, i) S.gradc[(i) * kRecGR + (k) * 2 + (r)]
#define GPA_(k, j) S.gpa[(j) * kRecPA + (k)]
#define LC_(c, i) S.lucoef[(i) * kRecLC + (c)]
#define F_(k, f) S.fflux[(f) * kRecF + (k)]
#define ALPHA_(f) S.fflux[(f) * kRecF + kNvar]
// viscous face record: alpa (inviscid spectral radius), Fv(2:Nvar), alpaVis
#define FV_(k, f) S.fvis[(f) * kRecF + (k)]

constexpr int NV = kNvar;

// Every per-item kernel has two forms.  B (batch): thread block b takes
// work item work[b] = (case, first, end, slot) of the case table.  !B (one
// case): cases[0] is the case and block b takes items lo + b*blockDim ...
// below hi, so no table load precedes the case's fields.
#define ITEM_ARGS const CaseDev *__restrict__ cases, const int4 *__restrict__ work, int lo, int hi
#define ITEM_PROLOGUE(idx)                                              \
    const RC_G CaseDev *Cp_;                                            \
    int idx;                                                            \
    if constexpr (B) {                                                  \
        const int4 w_ = work[blockIdx.x];                           \
        Cp_ = gp(cases) + w_.x;                                         \
        idx = w_.y + (int)threadIdx.x;                                  \
        if (idx >= w_.z) return;                                        \
    } else {                                                            \
        Cp_ = gp(cases);                                                \
        idx = lo + (int)(blockIdx.x * blockDim.x + threadIdx.x);        \
        if (idx >= hi) return;                                          \
    }                                                                   \
    const RC_G CaseDev &C = *Cp_;                                       \
    const GMesh M = gview(C.M);                                         \
    const GState S = gview(C.S);                                        \
    GThermo &TH = *gp(C.th);                                            \
    GParams &PAR = *gp(C.par);

// the case and the item of this thread only (for kernels whose work is a
// device function of (case, item)): Cp_ and idx, returning past the end
#define ITEM_CASE(idx)                                                  \
    const RC_G CaseDev *Cp_;                                            \
    int idx;                                                            \
    if constexpr (B) {                                                  \
        const int4 w_ = work[blockIdx.x];                               \
        Cp_ = gp(cases) + w_.x;                                         \
        idx = w_.y + (int)threadIdx.x;                                  \
        if (idx >= w_.z) return;                                        \
    } else {                                                            \
        Cp_ = gp(cases);                                                \
        idx = lo + (int)(blockIdx.x * blockDim.x + threadIdx.x);        \
        if (idx >= hi) return;                                          \
    }
// the same without returning (the block stages LDS first): end_ is the end
#define ITEM_CASE_END(idx, end_)                                        \
    const RC_G CaseDev *Cp_;                                            \
    int idx, end_;                                                      \
    if constexpr (B) {                                                  \
        const int4 w_ = work[blockIdx.x];                               \
        Cp_ = gp(cases) + w_.x;                                         \
        idx = w_.y + (int)threadIdx.x;                                  \
        end_ = w_.z;                                                    \
    } else {                                                            \
        Cp_ = gp(cases);                                                \
        idx = lo + (int)(blockIdx.x * blockDim.x + threadIdx.x);        \
        end_ = hi;                                                      \
    }

// the case's thermo table copied into LDS by every thread of the block (the
// caller synchronises the block before reading it)
__device__ __forceinline__ void stage_thermo(const RC_G CaseDev &C, DevThermo &dst) {
    const RC_G double *src_ = (const RC_G double *)gp(C.th);
    double *dst_ = reinterpret_cast<double *>(&dst);
    for (int k_ = threadIdx.x; k_ < (int)(sizeof(DevThermo) / 8); k_ += blockDim.x) dst_[k_] = src_[k_];
}

// the same with the case's thermo table staged in LDS by the whole block
// first (the thermo functions then read their coefficients with LDS
// latency); TH refers to that copy
#define ITEM_PROLOGUE_LDS(idx)                                          \
    __shared__ DevThermo th_lds_;                                       \
    const RC_G CaseDev *Cp_;                                            \
    int idx, end_;                                                      \
    if constexpr (B) {                                                  \
        const int4 w_ = work[blockIdx.x];                               \
        Cp_ = gp(cases) + w_.x;                                         \
        idx = w_.y + (int)threadIdx.x;                                  \
        end_ = w_.z;                                                    \
    } else {                                                            \
        Cp_ = gp(cases);                                                \
        idx = lo + (int)(blockIdx.x * blockDim.x + threadIdx.x);        \
        end_ = hi;                                                      \
    }                                                                   \
    const RC_G CaseDev &C = *Cp_;                                       \
    {                                                                   \
        const RC_G double *src_ = (const RC_G double *)gp(C.th);        \
        double *dst_ = reinterpret_cast<double *>(&th_lds_);            \
        for (int k_ = threadIdx.x; k_ < (int)(sizeof(DevThermo) / 8); k_ += blockDim.x) dst_[k_] = src_[k_]; \
    }                                                                   \
    __syncthreads();                                                    \
    if (idx >= end_) return;                                            \
    const GMesh M = gview(C.M);                                         \
    const GState S = gview(C.S);                                        \
    const DevThermo &TH = th_lds_;                                      \
    GParams &PAR = *gp(C.par);
static_assert(sizeof(DevThermo) % 8 == 0, "thermo table is copied as doubles");

// the cell-average quantities of Time_LUSGS_steady.f90:72-87 (and of the
// source term and ghost update) from one cell's primitive state, written to
// the per-cell cache; the arithmetic is that of the reference's neighbour
// term, so the cached values equal recomputing them there
template <int NSP, class TT>
__device__ __forceinline__ bool write_lucoef(const GMesh &M, const GState &S, TT &th, int i,
                                             double pp, double uu, double vv, double tt, const double *yy) {
    double cp, gama, rcp, h[NSP];
    bool ok = gas_param<NSP>(th, yy, tt, cp, gama, rcp);
    ok = enthalpy<NSP>(th, tt, h) && ok;
    double hh = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) hh = hh + yy[k] * h[k];
    const double rr = pp / (rcp * tt);
    const double roe = rr * (hh + 0.5 * (uu * uu + vv * vv)) - pp;
    LC_(LC_RR, i) = rr;
    LC_(LC_ROU, i) = rr * uu;
    LC_(LC_ROV, i) = rr * vv;
    LC_(LC_ROEP, i) = roe + pp;
    LC_(LC_U, i) = uu;
    LC_(LC_V, i) = vv;
    LC_(LC_T, i) = tt;
    LC_(LC_GAM, i) = gama;
#pragma unroll
    for (int k = 0; k < NSP; ++k) LC_(LC_H0 + k, i) = h[k];
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) LC_(LC_RY0 + k, i) = rr * yy[k];
#pragma unroll
    for (int k = LC_N; k < kRecLC; ++k) LC_(k, i) = 0.0;   // whole record: whole lines go to memory
    return ok;
}


// primitve_to_conserved_variables (update_flow_variables.f90:50-84) of
// cell i from its primitive state pa, and the cell-average cache of that
// state; false if the temperature is out of the polynomial range
template <int NSP, class TT>
__device__ __forceinline__ bool cell_prim_to_cons(const GMesh &M, const GState &S, TT &th, int i,
                                                  const double *pa) {
    const double pp = pa[0], uu = pa[1], vv = pa[2], tt = pa[3];
    double yy[NSP], s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) { yy[k] = pa[kNFlow + k]; s = s + yy[k]; }
    yy[NSP - 1] = 1.0 - s;
    double cp, gam, r, h[NSP];
    bool ok = gas_param<NSP>(th, yy, tt, cp, gam, r);
    const double rr = pp / (r * tt);
    ok = enthalpy<NSP>(th, tt, h) && ok;
    ok = write_lucoef<NSP>(M, S, th, i, pp, uu, vv, tt, yy) && ok;
    double hh = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) hh = hh + yy[k] * h[k];
    WA_(0, i) = rr;
    WA_(1, i) = rr * uu;
    WA_(2, i) = rr * vv;
    WA_(3, i) = rr * (hh + 0.5 * (uu * uu + vv * vv)) - pp;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) WA_(kNFlow + k, i) = rr * yy[k];
    return ok;
}

// the one-case form over [lo, hi) (per items per block of the list's size)
// or the batch form over the work list (one block of the list's size per
// work item)
#define LAUNCH(list, per, st, KERN, ...)                                                              \
    do {                                                                                              \
        if (L.single) {                                                                               \
            const int n_ = (list).hi - (list).lo;                                                     \
            if (n_ > 0)                                                                               \
                KERN<__VA_ARGS__, false><<<(n_ + (per) - 1) / (per), (list).bs, 0, st>>>(L.cases, nullptr, \
                                                                                     (list).lo, (list).hi); \
        } else if ((list).n) {                                                                        \
            KERN<__VA_ARGS__, true><<<(list).n, (list).bs, 0, st>>>(L.cases, (list).w, 0, 0);          \
        }                                                                                             \
    } while (0)

}  // namespace rc
