// Device-side views and FP64 leaf functions shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "rc_internal.h"

namespace rc {

// Device views of one case.  Arrays a cell reads only for itself are one
// array per value with stride sc (faces sf, ghosts sg):
//   stencil[s*sc+i], amls[(s*2+r)*sc+i], lvv[(kf*2+d)*sc+i], cface[kf*sc+i],
//   vecx[d*sf+f]; gathered arrays are records (see kRec* below).
struct DevMesh {
    int ne, nf, ng;
    int sc, sf, sg;
    int ss;                              // slot stride (4 lane slots per LU-SGS run)
    const int *__restrict__ fnl;
    const int *__restrict__ fnr;
    const int *__restrict__ stencil;
    const int *__restrict__ cface;
    const int *__restrict__ tsx;         // periodic meshes: faces of the TimeStep sum, tsx[k*sc+i] (-1 pad)
    const int *__restrict__ ghost_host;
    const int *__restrict__ ghost_type;
    const int *__restrict__ runs;        // LU-SGS runs (first cell, count), by level
    const int *__restrict__ pi_cells;    // point-implicit sweep: cells by level
    const double *__restrict__ vecx;
    const double *__restrict__ fvl;
    const double *__restrict__ fvr;
    const double *__restrict__ vol;
    const double *__restrict__ wdis;
    const double *__restrict__ amls;
    const double *__restrict__ lvv;
    const double *__restrict__ gnrm;
    const double *__restrict__ vgeo;     // viscous face operands, vgeo[k*sf+f], k < 6
    // LU-SGS face slots of each cell, in NeighF order: the neighbour across
    // the face (-1 on a boundary) and the face normal oriented out of the
    // cell (Time_LUSGS_steady.f90:60-68): lu_nb[k*sc+i], lu_n[(k*2+d)*sc+i]
    const int *__restrict__ lu_nb;
    const double *__restrict__ lu_n;
    // the same indices in lane-slot order of the LU-SGS sweep kernels: slot
    // 4t+p serves lane p of run t, so a lane's indices depend on its thread
    // index alone; slot[k*ss+s]: k = SL_CELL (the cell, or cell 0 for a lane
    // past the run's end), SL_FIRST, SL_CNT (the run), SL_CF+kf (cface),
    // SL_NB+kf (lu_nb); slot_n[(kf*2+d)*ss+s] (lu_n)
    const int *__restrict__ slot;
    const double *__restrict__ slot_n;
};
enum SlotField { SL_CELL = 0, SL_FIRST, SL_CNT, SL_CF, SL_NB = SL_CF + 3, SL_N = SL_NB + 3 };

// per-cell quantities of the cell-average state that the LU-SGS neighbour
// term (Time_LUSGS_steady.f90:72-87), the source term and the ghost update
// evaluate from PA alone; formed whenever PA is written
enum LuCoef { LC_RR = 0, LC_ROU, LC_ROV, LC_ROEP, LC_U, LC_V, LC_T, LC_GAM, LC_H0, LC_RY0 = LC_H0 + kNsp,
              LC_N = LC_RY0 + kNsp - 1 };

// record widths (doubles) of the arrays that kernels gather by neighbour
// index: one 64-byte or 128-byte record per cell or face, so a gather reads
// one or two cache lines instead of one line per variable
constexpr int kRecPA = 8;    // PA(7) and GPA(7)
constexpr int kRecHW = 8;    // hWA(7)
constexpr int kRecGR = 16;   // GradC(2,7)
constexpr int kRecLC = 16;   // LU cache (LC_N)
constexpr int kRecF = 8;     // face flux F(7) and alpaF
static_assert(LC_N <= kRecLC, "LU cache record too small");

struct DevState {
    double *pa;      // kRecPA per cell
    double *wa;      // Nvar arrays of sc
    double *hwa;     // kRecHW per cell
    double *gradc;   // kRecGR per cell
    double *gpa;     // kRecPA per ghost
    double *step, *dd;
    double *fflux;   // kRecF per face: F(1:7), alpaF
    double *lucoef;  // kRecLC per cell
    double *lu_af;   // alpaF of the cell's 3 face slots, 3 x sc
    double *fixed;   // backward Euler: vol/dt*WA at the first inner iteration, Nvar arrays of sc
    double *fvis;    // viscous path, kRecF per face: alpa, Fv(2:Nvar), alpaVis
    double *src;     // learned-chemistry source: omega_1..omega_{Nsp-1}, arrays of sc (null when exact)
    int *err;        // [0] code, [1] index, [2] kernel, [3] step of the error, [4] current step,
                     // [5] the step err[4] had when the step began (err[4] - 1 from the fused
                     // step's gradient kernel to its update, err[4] otherwise)
    int *flag;       // per cell: sweep progress of the dataflow sweep, 2e+1 after the forward
                     // and 2e+2 after the backward sweep of step e (step counter err[4])
};

// one case as the kernels see it; an array of these (one per resident case)
// is what every launch reads, so one launch serves any set of cases
struct CaseDev {
    DevMesh M;
    DevState S;
    const DevThermo *th;
    const DevParams *par;
    const DevTransport *tr;   // viscous path: transport tables (null when inviscid)
    double *partial;     // residual partial sums, (kNvar+1) per cell chunk
    int npartial;
    double inv_volume;   // 1 / total volume
};

// cells / faces / ghosts per thread block of the per-item kernels (the
// face kernel runs two lanes per face).  Measured at 12k / 400k / 1.6M
// cells: 128 against 256, 83.1 -> 80.1 us, 0.2967 -> 0.2911 ms, 1.042 ->
// 1.033 ms per step (the face kernel gains most: smaller blocks spread over
// more CUs); 64 is no better.
constexpr int kChunkDefault = 128;

// work of one launch: one int4 per thread block, (case index, first item,
// end item, partial-sum slot); items are cells, faces, ghosts or LU-SGS runs.
// [lo, hi) is the same range for a set of one case, launched without the
// table.
struct WorkList {
    const int4 *w = nullptr;
    int n = 0;
    int lo = 0, hi = 0;
    int bs = kChunkDefault;        // threads per block of the table form
};

// The dataflow sweep: every LU-SGS item of a step (forward levels in order,
// then backward levels in reverse order) in one launch.  Items are split into
// kXcdLists sub-lists, each in increasing order of the whole list (a
// topological order of the sweep); thread block b takes the next item of
// sub-list b % kXcdLists from an atomic counter, so a block waits only for
// items handed out before its own, and the blocks that serve every counter
// are the first ones dispatched: the launch cannot deadlock, whatever number
// of blocks is resident.
constexpr int kXcdLists = 8;
struct FlowSpan {
    int off[kXcdLists + 1];        // sub-list x: items[off[x] .. off[x+1])
};
struct FlowList {
    const int4 *w = nullptr;       // (case, first run, end run, 1 forward / 0 backward)
    FlowSpan span{};
    int *ctr = nullptr;            // kXcdLists counters, zeroed by k_step_begin
    int grid = 0;                  // thread blocks of the launch (0: no dataflow sweep)
};

// The persistent step: every phase of nsteps steps of a set of cases in one
// cooperative launch, grid barriers between the phases.  The lists are the
// step's own (cells and faces in items of at most 256, LU-SGS runs in items
// of at most 16 per level); bar: [0] barrier arrivals, [1] abort.
constexpr int kPersistThreads = 256;
constexpr int kPersistMaxLevels = 32;
struct PersistSet {
    const CaseDev *cases;
    int ncases;
    const int4 *cell;
    int ncell;
    const int4 *face;
    int nface;
    int fpass;                     // passes of kPersistThreads/2 faces per face item
    const int4 *lev[kPersistMaxLevels];
    int nlevitems[kPersistMaxLevels];
    int nlev;
    unsigned *bar;
    int nsteps;
};

// everything one step launches over a set of cases
struct LaunchSet {
    const CaseDev *cases = nullptr;
    int ncases = 0;
    bool single = false;           // one case: cases[0], launched over [lo, hi)
    WorkList cell, face, ghost;
    WorkList norm;                 // cell chunks owning the residual partial slots
    std::vector<WorkList> level;   // LU-SGS runs of each level
    std::vector<WorkList> pi_level;   // point-implicit sweep: cells (positions in pi_cells) of each level
    FlowList flow;                 // the whole LU-SGS sweep of the fused step as one launch
    std::vector<char> next_res;    // forward level l also forms level l+1's residual
    bool top_fused = false;        // the highest level's forward launch also runs its backward sweep
    double *resout = nullptr;      // ncases x (kNvar+1) residual norms
};

// Device-side views of a case whose pointers are typed as global memory.
// The host-side structs above hold plain pointers; a kernel reads them from
// the case table and casts them here, so every access compiles to a global
// (not flat) load or store.
#define RC_G __attribute__((address_space(1)))
struct GMesh {
    int ne, nf, ng;
    int sc, sf, sg, ss;
    const RC_G int *fnl, *fnr, *stencil, *cface, *tsx, *ghost_host, *ghost_type, *runs, *lu_nb, *pi_cells;
    const RC_G double *vecx, *fvl, *fvr, *vol, *wdis, *amls, *lvv, *gnrm, *lu_n, *vgeo;
    const RC_G int *slot;
    const RC_G double *slot_n;
};
struct GState {
    RC_G double *pa, *wa, *hwa, *gradc, *gpa, *step, *dd, *fflux, *lucoef, *lu_af, *fixed, *fvis, *src;
    RC_G int *err, *flag;
};
typedef const RC_G DevThermo GThermo;
typedef const RC_G DevParams GParams;

template <class T>
__device__ __forceinline__ RC_G T *gp(T *p) {
    return (RC_G T *)p;
}

__device__ __forceinline__ GMesh gview(const RC_G DevMesh &m) {
    GMesh g;
    g.ne = m.ne; g.nf = m.nf; g.ng = m.ng; g.sc = m.sc; g.sf = m.sf; g.sg = m.sg; g.ss = m.ss;
    g.slot = gp(m.slot); g.slot_n = gp(m.slot_n);
    g.fnl = gp(m.fnl); g.fnr = gp(m.fnr); g.stencil = gp(m.stencil); g.cface = gp(m.cface); g.tsx = gp(m.tsx);
    g.ghost_host = gp(m.ghost_host); g.ghost_type = gp(m.ghost_type); g.runs = gp(m.runs); g.lu_nb = gp(m.lu_nb);
    g.pi_cells = gp(m.pi_cells);
    g.vecx = gp(m.vecx); g.fvl = gp(m.fvl); g.fvr = gp(m.fvr); g.vol = gp(m.vol); g.wdis = gp(m.wdis);
    g.amls = gp(m.amls); g.lvv = gp(m.lvv); g.gnrm = gp(m.gnrm); g.lu_n = gp(m.lu_n); g.vgeo = gp(m.vgeo);
    return g;
}

__device__ __forceinline__ GState gview(const RC_G DevState &s) {
    GState g;
    g.pa = gp(s.pa); g.wa = gp(s.wa); g.hwa = gp(s.hwa); g.gradc = gp(s.gradc); g.gpa = gp(s.gpa);
    g.step = gp(s.step); g.dd = gp(s.dd); g.fflux = gp(s.fflux); g.lucoef = gp(s.lucoef); g.lu_af = gp(s.lu_af);
    g.fixed = gp(s.fixed);
    g.fvis = gp(s.fvis);
    g.src = gp(s.src);
    g.err = gp(s.err);
    g.flag = gp(s.flag);
    return g;
}

constexpr int kChunk = kChunkDefault;  // cells / faces / ghosts per thread block
constexpr int kRunThreads = 64;       // threads per block of the LU-SGS level kernels (default shape)
constexpr int kRunChunk = kRunThreads / kRunMax;   // LU-SGS runs per thread block

enum KernelId { K_TS_FACE = 1, K_TS_CELL, K_GHOST, K_GRADLIM, K_FLUX, K_RES, K_FWD, K_BWD, K_UPDATE, K_P2C,
                K_BE_FWD, K_BE_BWD, K_BE_UPDATE, K_VISC, K_FLOW, K_PERSIST };

// first error wins; the step number comes from the device-side step
// counter err[4] (advanced by k_step_begin), so captured graphs report it too
__device__ inline void flag_error(RC_G int *err, int code, int index, int kernel) {
    if (atomicCAS((int *)err, 0, code) == 0) {
        err[1] = index;
        err[2] = kernel;
        err[3] = __atomic_load_n(err + 4, __ATOMIC_RELAXED);
    }
}
// the same with the step given (a kernel that advances err[4] itself)
__device__ inline void flag_error_at(RC_G int *err, int code, int index, int kernel, int step) {
    if (atomicCAS((int *)err, 0, code) == 0) {
        err[1] = index;
        err[2] = kernel;
        err[3] = step;
    }
}

// real**integer as compiler-rt's __powidf2 evaluates it (the reference's
// `tem**(j-3)`); with a constant exponent the loop unrolls to the same
// multiply chain.
__device__ __forceinline__ double powi_d(double a, int b) {
    const bool recip = b < 0;
    double r = 1.0;
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

// ComputeGasParameter, thermo_property.f90:78-112.  Returns false when T is
// outside [t_low, t_high] (the reference stops there); the high-range
// coefficients are then used so that the arithmetic stays finite.
template <int NSP, class TT>
__device__ __forceinline__ bool gas_param(TT &th, const double *y, double t,
                                          double &cp, double &gam, double &r) {
    double basis[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) basis[j] = powi_d(t, j - 2);
    bool ok = true;
    const bool low = (th.t_low <= t && t <= th.t_crit);
    if (!low && !(th.t_crit < t && t <= th.t_high)) ok = false;
    const auto *coef = low ? th.low : th.high;
    double c = 0.0, rr = 0.0;
    double cps[NSP];
#pragma unroll
    for (int i = 0; i < NSP; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < 7; ++j) s = s + coef[i][j] * basis[j];
        cps[i] = s * th.rsp[i];
    }
#pragma unroll
    for (int i = 0; i < NSP; ++i) c = c + cps[i] * y[i];
#pragma unroll
    for (int i = 0; i < NSP; ++i) rr = rr + th.rsp[i] * y[i];
    cp = c;
    gam = c / (c - rr);
    r = rr;
    return ok;
}

// the mixture gas constant of ComputeGasParameter alone (its sum over the
// species in the same order, so the value is gas_param's r bit for bit)
template <int NSP, class TT>
__device__ __forceinline__ double gas_rconst(TT &th, const double *y) {
    double rr = 0.0;
#pragma unroll
    for (int i = 0; i < NSP; ++i) rr = rr + th.rsp[i] * y[i];
    return rr;
}

// compute_enthalpy, thermo_property.f90:114-144
template <int NSP, class TT>
__device__ __forceinline__ bool enthalpy(TT &th, double t, double *h) {
    double basis[7];
    basis[0] = -powi_d(t, -2);
    basis[1] = log(t) / t;
#pragma unroll
    for (int j = 2; j < 7; ++j) basis[j] = powi_d(t, j - 2) / (double)(j - 1);
    bool ok = true;
    const bool low = (th.t_low <= t && t <= th.t_crit);
    if (!low && !(th.t_crit < t && t <= th.t_high)) ok = false;
    const auto *coef = low ? th.low : th.high;
    const auto *b1 = low ? th.low_b1 : th.high_b1;
#pragma unroll
    for (int i = 0; i < NSP; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < 7; ++j) s = s + coef[i][j] * basis[j];
        h[i] = th.rsp[i] * (t * s + b1[i]);
    }
    return ok;
}

// Newton_temperature, thermo_property.f90:146-180
template <int NSP, class TT>
__device__ __forceinline__ bool newton_temperature(TT &th, double et, const double *y, double t0,
                                                   double &t, double &cp, double &gam, double &r) {
    t = t0;
    bool ok = true;
    double h[NSP];
    for (int it = 0; it < 200; ++it) {
        ok = enthalpy<NSP>(th, t, h) && ok;
        ok = gas_param<NSP>(th, y, t, cp, gam, r) && ok;
        if (!ok) break;
        double hh = 0.0;
#pragma unroll
        for (int i = 0; i < NSP; ++i) hh = hh + y[i] * h[i];
        const double ft = hh - r * t - et;
        const double dft = cp - r;
        const double dtt = -ft / dft;
        t = t + dtt;
        t = fmax(t, th.t_low);
        t = fmin(t, th.t_high);
        if (fabs(dtt) < 1.0e-5) break;
    }
    return ok;
}

// compute_production_rate, chemical_kinetics.f90:109-165 (one reaction;
// model 0: A exp(-E/RT) [H2]^0.2 [O2]^1.3, model 1: 1e4 [H2] above 710 K)
template <int NSP, class TT>
__device__ __forceinline__ void production_rate(TT &th, double t, double rho, const double *y2, double *w) {
    double y[NSP];
#pragma unroll
    for (int k = 0; k < NSP; ++k) y[k] = fmin(1.0, fmax(0.0, y2[k]));
    double kf;
    if (th.model == 0) {
        kf = th.a_f * exp(-th.e_act / (th.rconstant * t));
        kf = kf * pow(y[0] * rho / (th.mw[0] * 1000.0), 0.2);
        kf = kf * pow(y[1] * rho / (th.mw[1] * 1000.0), 1.3);
    } else {
        kf = 0.0;
        if (t > 710.0) kf = 1e4;
        kf = kf * (y[0] * rho / (th.mw[0] * 1000.0));   // x**1.0 == x
        kf = kf * 1.0;                                   // x**0.0 == 1
    }
#pragma unroll
    for (int k = 0; k < NSP; ++k) w[k] = 1000.0 * th.mw[k] * th.dnu[k] * kf;
}

}  // namespace rc
