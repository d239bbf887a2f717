// C-ABI of the library (include/rocketcombustor.h): case set-up, resident
// device state, the per-operator entry points, the steady step loop, and
// batches of resident cases stepped by the same launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "rc_device.h"

namespace rc {
void launch_ts_face(const LaunchSet &, hipStream_t);
void launch_ts_cell(const LaunchSet &, hipStream_t);
void launch_ghost(const LaunchSet &, hipStream_t);
void launch_gradlim(const LaunchSet &, bool, int, hipStream_t);
void launch_gradlim_step(const LaunchSet &, hipStream_t);
void launch_update_step(const LaunchSet &, hipStream_t);
void launch_flux(const LaunchSet &, bool, hipStream_t);
void launch_residual(const LaunchSet &, bool, hipStream_t);
void launch_step_begin(const LaunchSet &, hipStream_t);
void launch_lusgs_level(const LaunchSet &, bool, int, bool, hipStream_t);
void launch_lusgs_flow(const LaunchSet &, hipStream_t);
void launch_lusgs_fused_forward(const LaunchSet &, int, hipStream_t);
long lusgs_resident_threads(int, int);
int flow_max_blocks(int);
void launch_update(const LaunchSet &, hipStream_t);
void launch_prim_to_cons(const LaunchSet &, hipStream_t);
void launch_resnorm(const LaunchSet &, hipStream_t);
void launch_gather_err(const LaunchSet &, int *, hipStream_t);
bool persist_supported(const LaunchSet &);
int persist_grid(const LaunchSet &, int);
hipError_t launch_step_persist(const LaunchSet &, unsigned *, int, int, hipStream_t);
void launch_derived(const LaunchSet &, double *, hipStream_t);
void launch_be_prepare(const LaunchSet &, double, bool, const double *, const double *, double, hipStream_t);
void launch_be_factor(const LaunchSet &, double *, int *, double *, hipStream_t);
void launch_be_level(const LaunchSet &, bool, int, const double *, const int *, const double *, hipStream_t);
size_t be_factor_doubles();
size_t be_factor_ints();
void launch_be_update(const LaunchSet &, hipStream_t);
int pi_chunk();
void launch_viscous(const LaunchSet &, bool, bool, hipStream_t);
size_t be_cache_doubles();
struct Surrogate;
int surrogate_xform_len();
int surrogate_create(rc_mlp *, int, const double *, double, double, Surrogate **, std::string &);
void surrogate_destroy(Surrogate *);
void surrogate_enqueue(Surrogate *, const CaseDev *, int, hipStream_t);
int surrogate_accepted(Surrogate *, hipStream_t, long long &);
int surrogate_set_trust(Surrogate *, const int *, const double *, const double *, const unsigned char *,
                        std::string &);
}  // namespace rc

using namespace rc;

static thread_local std::string g_error;
static constexpr int kTimeSlots = 7;
static constexpr int kGraphSteps = 8;   // most steps in one graph (graph_steps())
static constexpr int kGraphPows = 6;    // graphs of 2^1 .. 2^6 steps at most

struct rc_case {
    int device = 0;
    hipStream_t stream = nullptr;
    HostMesh hm;
    rc_mechanism mech{};
    rc_params par{};
    int ne = 0, nf = 0, ng = 0, nvar = kNvar, nlev = 0;
    int npl = 0;                   // point-implicit sweep levels
    DevMesh M{};
    DevState S{};
    DevThermo *d_th = nullptr;
    DevParams *d_par = nullptr;
    DevTransport *d_tr = nullptr;  // viscous path tables (uploaded by rc_set_transport)
    double *be_fac = nullptr;      // backward Euler: per-cell block factors (allocated at first use)
    int *be_piv = nullptr;
    double *be_nbc = nullptr;      // backward Euler: per-cell thermo records of the neighbour terms
    double *bdf_wp = nullptr;      // BDF2: WP and WP2, the states at the start of the last two steps
    double *bdf_wp2 = nullptr;     //       (Nvar arrays of sc each; allocated at the first BDF2 call)
    Surrogate *sur = nullptr;      // learned-chemistry source (null: exact kinetics)
    double *sur_src = nullptr;     // its per-cell rates, (Nsp-1) arrays of sc
    long long sur_evals = 0;       // cell evaluations since the surrogate was set
    bool viscous = false;
    CaseDev hc{};                  // host copy of the case's descriptor
    CaseDev *d_case = nullptr;
    // work lists of the case alone (case index 0), kept on the host too so
    // that batches can concatenate them
    std::vector<int4> w_cell, w_face, w_ghost;
    std::vector<std::vector<int4>> w_level;
    std::vector<std::vector<int4>> w_pi;   // point-implicit sweep levels
    // the cell and face segments the lists band over the XCDs (the colours
    // of a block-multicolour order), for batches that re-cut the lists
    std::vector<std::pair<int, int>> seg_cell, seg_face;
    LaunchSet L;                   // the case alone
    rc_batch *batch = nullptr;     // batch the case belongs to, if any
    void *arena = nullptr;
    size_t arena_bytes = 0;
    int *h_err = nullptr;        // pinned
    double *h_res = nullptr;     // pinned
    int steps_done = 0;
    bool timing = false;
    // timing slots: 0 gradient+limiter, 1 flux faces, 2 residual+source,
    // 3 LU forward below the highest level, 4 the highest level, 5 LU
    // backward, 6 update
    // events are recorded without synchronising; each step uses kTimeSlots+1
    // consecutive pool entries, drained into tms[] at the next sync point
    std::vector<hipEvent_t> evpool;
    size_t ev_used = 0;
    double tms[kTimeSlots] = {};
    long timed_steps = 0;
    bool use_graph = true;
    hipGraphExec_t graph_exec = nullptr;
    hipGraphExec_t graph_pow[kGraphPows] = {};   // 2, 4, 8, ... steps in one graph
    int graph_pow_top = 0;                       // the largest of them (0: none captured)
    int persist = 0;               // steps per persistent launch (0: graphs)
    unsigned *d_bar = nullptr;     // the persistent step's barrier words
    std::string err;
};

// resident cases stepped together: every launch covers all members
struct rc_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<rc_case *> members;
    void *arena = nullptr;         // CaseDev array, work lists, residual output
    LaunchSet L;
    int nlev = 0;
    double *h_res = nullptr;       // pinned, members x (kNvar+1)
    size_t h_res_n = 0;
    int *d_err = nullptr;          // the members' error flags gathered, members x 4 (in the arena)
    int *h_err = nullptr;          // pinned copy of them
    size_t h_err_n = 0;
    bool dirty = true;             // membership or launch shape changed since the last build
    bool use_graph = true;
    hipGraphExec_t graph_exec = nullptr;
    hipGraphExec_t graph_pow[kGraphPows] = {};   // 2, 4, 8, ... steps in one graph
    int graph_pow_top = 0;
    int persist = 0;               // steps per persistent launch (0: graphs)
    unsigned *d_bar = nullptr;     // the persistent step's barrier words
    // launch shape: threads per block of the cell / face / ghost kernels and
    // of the LU-SGS level kernels (kRunMax lanes per run), XCD banding of the
    // work lists (-1: as the members were laid out)
    int cell_bs = 256, run_bs = kRunThreads, xcd = -1;
    std::string err;
};

namespace {

int set_err(rc_case *h, int code, const std::string &msg) {
    if (h) h->err = msg;
    g_error = msg;
    return code;
}

#define HIPCHK(h, x)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return set_err(h, RC_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
    } while (0)

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
int pad64(int n) { return (int)align_up((size_t)std::max(n, 1), 64); }

const char *kernel_name(int k) {
    switch (k) {
    case K_TS_FACE: return "timestep (face)";
    case K_TS_CELL: return "timestep (cell)";
    case K_GHOST: return "ghost-cell update";
    case K_GRADLIM: return "gradient/limiter";
    case K_FLUX: return "flux_lax";
    case K_RES: return "sourceterm";
    case K_FWD: return "LU-SGS forward sweep";
    case K_BWD: return "LU-SGS backward sweep";
    case K_UPDATE: return "update_primitve_variables";
    case K_P2C: return "primitve_to_conserved_variables";
    case K_BE_FWD: return "backward Euler forward sweep";
    case K_BE_BWD: return "backward Euler backward sweep";
    case K_BE_UPDATE: return "backward Euler update";
    case K_VISC: return "flux_viscous";
    case K_FLOW: return "LU-SGS dataflow sweep (a dependency wait timed out)";
    default: return "?";
    }
}

// reads the device error flag (stream must be synchronised)
int check_device_error(rc_case *h, bool copied = false) {
    if (!copied) {
        HIPCHK(h, hipMemcpyAsync(h->h_err, h->S.err, 4 * sizeof(int), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    if (h->h_err[0] == 0) return RC_OK;
    const int code = h->h_err[0];
    char buf[256];
    const char *what = code == RC_ERR_TEMPERATURE ? "temperature outside the thermo polynomial range"
                     : code == RC_ERR_NEGATIVE ? "negative density or pressure"
                     : code == RC_ERR_NAN ? "NaN in WA(1,1)"
                     : code == RC_ERR_SINGULAR ? "error in routine dgesvx (singular block)" : "error";
    std::snprintf(buf, sizeof buf, "%s in %s (index %d, step %d)", what, kernel_name(h->h_err[2]), h->h_err[1] + 1,
                  h->h_err[3]);
    HIPCHK(h, hipMemsetAsync(h->S.err, 0, 4 * sizeof(int), h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return set_err(h, code, buf);
}

int drain_timing(rc_case *h);

// the error flags copied behind the stream's work, one synchronisation
int sync_check(rc_case *h) {
    HIPCHK(h, hipMemcpyAsync(h->h_err, h->S.err, 4 * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    int rc = drain_timing(h);
    if (rc) return rc;
    return check_device_error(h, true);
}

DevThermo make_thermo(const rc_mechanism &m) {
    DevThermo t{};
    t.nsp = m.nsp;
    t.model = m.model;
    for (int k = 0; k < m.nsp; ++k) {
        t.rsp[k] = m.rconstant / m.mole_weight[k];
        t.mw[k] = m.mole_weight[k];
        t.low_b1[k] = m.low_b1[k];
        t.high_b1[k] = m.high_b1[k];
        t.dnu[k] = m.coef_r[k] - m.coef_f[k];
        for (int j = 0; j < 7; ++j) {
            t.low[k][j] = m.low_coef[k][j];
            t.high[k][j] = m.high_coef[k][j];
        }
    }
    t.t_low = m.t_low;
    t.t_crit = m.t_crit;
    t.t_high = m.t_high;
    t.a_f = m.a_f;
    t.e_act = m.e_act;
    t.rconstant = m.rconstant;
    return t;
}

DevParams make_params(const rc_params &p, int nsp, bool periodic) {
    DevParams d{};
    d.periodic = periodic ? 1 : 0;
    d.cfl = p.cfl;
    d.p_out = p.p_out;
    for (int k = 0; k < nsp; ++k) {
        d.ox_y[k] = p.ox_y[k];
        d.fu_y[k] = p.fu_y[k];
    }
    d.ox_t = p.ox_t;
    d.ox_g = p.ox_mass_flux;
    d.fu_t = p.fu_t;
    d.fu_g = p.fu_mass_flux;
    d.limit_id = p.limit_id;
    d.combustion = p.combustion;
    return d;
}

// carves typed sub-arrays out of one device allocation
struct Carver {
    size_t off = 0;
    char *base = nullptr;
    template <class T>
    T *take(size_t n) {
        off = align_up(off, 256);
        T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off += n * sizeof(T);
        return p;
    }
};

template <class T>
std::vector<T> soa(const std::vector<T> &aos, int n, int w, int stride) {
    // aos[i*w + c] -> soa[c*stride + i]
    std::vector<T> out((size_t)w * stride, T(0));
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < w; ++c) out[(size_t)c * stride + i] = aos[(size_t)i * w + c];
    return out;
}

// thread-block work items over [lo, hi) in chunks, case index c, partial
// slots counted from slot0
void chunk_list(std::vector<int4> &out, int c, int lo, int hi, int chunk, int slot0 = 0) {
    int slot = slot0;
    for (int a = lo; a < hi; a += chunk) out.push_back(make_int4(c, a, std::min(hi, a + chunk), slot++));
}

// Work items laid out for the 8 XCDs (thread blocks are dispatched to the
// XCDs round-robin: block b of a launch to XCD (b + s) % 8, s carried over
// from the blocks dispatched before it, so the blocks b = x mod 8 of one
// launch share an XCD): XCD x takes the x-th eighth of every segment, in
// chunks; item p of the list goes to XCD group p % 8, shorter lists padded
// with empty items.  With segments that are the colours of a
// block-multicolour order (each sorted along the mesh), an XCD sweeps one
// band of the domain in every colour, so the neighbours it gathers are
// mostly in its own L2.
constexpr int kXcd = 8;
void xcd_list(std::vector<int4> &out, int c, const std::vector<std::pair<int, int>> &segs, int chunk) {
    std::vector<std::vector<int2>> per(kXcd);
    for (const auto &sg : segs) {
        const long a = sg.first, len = (long)sg.second - sg.first;
        for (int x = 0; x < kXcd; ++x) {
            const int lo = (int)(a + len * x / kXcd), hi = (int)(a + len * (x + 1) / kXcd);
            for (int q = lo; q < hi; q += chunk) per[x].push_back(make_int2(q, std::min(hi, q + chunk)));
        }
    }
    size_t n = 0;
    for (auto &v : per) n = std::max(n, v.size());
    int slot = 0;
    for (size_t j = 0; j < n; ++j)
        for (int x = 0; x < kXcd; ++x) {
            const int2 r = j < per[x].size() ? per[x][j] : make_int2(0, 0);
            out.push_back(make_int4(c, r.x, r.y, slot++));
        }
}

// empty items up to a multiple of kXcd, so that the next XCD-banded list
// starts on XCD 0
void pad_xcd(std::vector<int4> &out, int c) {
    while (out.size() % kXcd) out.push_back(make_int4(c, 0, 0, 0));
}

// The items of the dataflow sweep from the per-level work lists of the
// fused step (items (case, first run, end run, slot) in launch order, any
// number of runs each): the forward levels in order, then the backward
// levels in reverse order, each item cut into chunks of kRunChunk runs (one
// wave) that keep its XCD (item p of a level list runs on XCD p % 8, the
// order in which a level launch dispatches its blocks).  Sub-list x holds
// the chunks of XCD x in that global order, which is a topological order of
// the sweep: every run a chunk waits for comes earlier.
void build_flow(const std::vector<std::vector<int4>> &levels, std::vector<int4> &items, FlowSpan &span) {
    std::vector<std::vector<int4>> per(kXcdLists);
    const int nlev = (int)levels.size();
    for (int pass = 0; pass < 2; ++pass)
        for (int q = 0; q < nlev; ++q) {
            const int l = pass == 0 ? q : nlev - 1 - q;
            const std::vector<int4> &lv = levels[l];
            for (size_t p = 0; p < lv.size(); ++p)
                for (int a = lv[p].y; a < lv[p].z; a += kRunChunk)
                    per[p % kXcdLists].push_back(make_int4(lv[p].x, a, std::min(lv[p].z, a + kRunChunk),
                                                           pass == 0 ? 1 : 0));
        }
    items.clear();
    for (int x = 0; x < kXcdLists; ++x) {
        span.off[x] = (int)items.size();
        items.insert(items.end(), per[x].begin(), per[x].end());
    }
    span.off[kXcdLists] = (int)items.size();
}

// thread blocks of a dataflow sweep over these items: as many as the
// device holds at once, no more than the longest sub-list needs, a multiple
// of kXcdLists; 0 is one launch per level and direction instead.  The
// dataflow sweep is opt-in (RC_FLOW=1): it gives the same results, but
// measured on the 400k-cell step it takes 0.30 ms against 0.18 ms for the
// level launches (its increments and flags bypass the L2 as agent-scope
// accesses, and a wait notices its producer one poll late), so the level
// launches are the default.
int flow_grid(int device, const FlowSpan &span) {
    const char *e = std::getenv("RC_FLOW");
    if (!(e && e[0] == '1')) return 0;
    int longest = 0;
    for (int x = 0; x < kXcdLists; ++x) longest = std::max(longest, span.off[x + 1] - span.off[x]);
    if (longest == 0) return 0;
    return std::min(flow_max_blocks(device), longest * kXcdLists);
}

// XCD banding pays while a band's neighbour records stay in its XCD's L2:
// measured on the steady step, 2.3% faster at 400k cells, 0.5% slower at
// 784k, 2.8% slower at 1.6M (with 128-cell blocks and the fused highest
// level: 0.9% faster at 400k, 2.0% slower at 1.6M).  RC_XCD=0 / 1 forces it
// off / on.
constexpr int kXcdMaxCells = 600000;
bool use_xcd_lists(int ne) {
    const char *e = std::getenv("RC_XCD");
    if (e && e[0] == '0') return false;
    if (e && e[0] == '1') return true;
    return ne <= kXcdMaxCells;
}

// Which forward levels of the fused step also form the next level's
// residual, in extra thread blocks of their own launch: those whose threads
// and the next level's fit on the device at once with a quarter to spare,
// so the residual blocks run beside the sweep and take the residual's chain
// off the next level's.  Measured: the 12k-cell step 102.6 -> 93.4 us; on
// the 400k step, whose levels each fill most of the device, doing it at
// every level costs 1.3%.  The level lists must share one block size.
// RC_NEXTRES=0 / 1 turns it off / on at every level (same results).
std::vector<char> next_res_levels(const LaunchSet &L, int device) {
    const int nlev = (int)L.level.size();
    std::vector<char> nx(nlev, 0);
    const char *e = std::getenv("RC_NEXTRES");
    if (e && e[0] == '0') return nx;
    for (const WorkList &w : L.level)
        if (w.bs != L.level.front().bs) return nx;
    const bool force = e && e[0] == '1';
    const long cap = force || nlev < 2 ? 0 : lusgs_resident_threads(device, L.single ? kRunThreads : L.level[0].bs);
    auto threads = [&](const WorkList &w) -> long {
        return L.single ? (long)((w.hi - w.lo + kRunChunk - 1) / kRunChunk) * kRunThreads : (long)w.n * w.bs;
    };
    for (int l = 0; l + 1 < nlev; ++l)
        nx[l] = force || 4 * (threads(L.level[l]) + threads(L.level[l + 1])) <= 3 * cap;
    return nx;
}

// the fused step's sweeps: forward levels (the highest one also running
// its backward sweep when L.top_fused), then the other backward levels
void enqueue_sweeps(const LaunchSet &L, int nlev, hipStream_t st) {
    for (int l = 0; l < nlev; ++l) launch_lusgs_fused_forward(L, l, st);
}
void enqueue_backward_levels(const LaunchSet &L, int nlev, hipStream_t st) {
    for (int l = nlev - 1 - (L.top_fused ? 1 : 0); l >= 0; --l) launch_lusgs_level(L, false, l, false, st);
}

// RC_TOPFUSE=0: the highest level's backward sweep as a launch of its own
// (read when a launch set is built; default: fused into its forward launch)
bool top_fused_env() {
    const char *e = std::getenv("RC_TOPFUSE");
    return !(e && e[0] == '0');
}

// cell index ranges of the sweep levels when every level's runs are one
// contiguous range (a colour of a block-multicolour order); else the whole
// range
std::vector<std::pair<int, int>> level_cell_segments(const HostMesh &m) {
    const int nlev = (int)m.lvl_ptr.size() - 1;
    std::vector<std::pair<int, int>> segs;
    long covered = 0;
    for (int l = 0; l < nlev; ++l) {
        int a = m.ne, b = 0;
        long cells = 0;
        for (int r = m.lvl_ptr[l]; r < m.lvl_ptr[l + 1]; ++r) {
            a = std::min(a, m.runs[2 * r]);
            b = std::max(b, m.runs[2 * r] + m.runs[2 * r + 1]);
            cells += m.runs[2 * r + 1];
        }
        if (cells != (long)b - a) return {{0, m.ne}};
        segs.push_back({a, b});
        covered += cells;
    }
    std::sort(segs.begin(), segs.end());
    for (size_t k = 1; k < segs.size(); ++k)
        if (segs[k].first != segs[k - 1].second) return {{0, m.ne}};
    if (covered != m.ne) return {{0, m.ne}};
    return segs;
}

int layout_and_upload(rc_case *h) {
    const HostMesh &m = h->hm;
    const int sc = pad64(m.ne), sf = pad64(m.nf), sg = pad64(m.ng);
    h->w_cell.clear(); h->w_face.clear(); h->w_ghost.clear();
    const int nlev = (int)m.lvl_ptr.size() - 1;
    h->w_level.assign(nlev, {});
    const bool xcd = use_xcd_lists(m.ne);
    if (xcd) {
        auto cs = level_cell_segments(m);
        if (!std::is_sorted(m.fnl.begin(), m.fnl.end())) cs = {{0, m.ne}};
        // faces are numbered in order of first appearance over the cells, so
        // each cell segment owns one contiguous face range
        std::vector<std::pair<int, int>> fs;
        for (const auto &c : cs) {
            const int fa = (int)(std::lower_bound(m.fnl.begin(), m.fnl.end(), c.first) - m.fnl.begin());
            const int fb = (int)(std::lower_bound(m.fnl.begin(), m.fnl.end(), c.second) - m.fnl.begin());
            fs.push_back({fa, fb});
        }
        xcd_list(h->w_cell, 0, cs, kChunk);
        xcd_list(h->w_face, 0, fs, kChunk);
        h->seg_cell = cs;
        h->seg_face = fs;
        chunk_list(h->w_ghost, 0, 0, m.ng, kChunk);
        for (int l = 0; l < nlev; ++l) xcd_list(h->w_level[l], 0, {{m.lvl_ptr[l], m.lvl_ptr[l + 1]}}, kRunChunk);
    } else {
        chunk_list(h->w_cell, 0, 0, m.ne, kChunk);
        chunk_list(h->w_face, 0, 0, m.nf, kChunk);
        chunk_list(h->w_ghost, 0, 0, m.ng, kChunk);
        for (int l = 0; l < nlev; ++l) chunk_list(h->w_level[l], 0, m.lvl_ptr[l], m.lvl_ptr[l + 1], kRunChunk);
        h->seg_cell = {{0, m.ne}};
        h->seg_face = {{0, m.nf}};
    }
    const int npl = (int)m.pi_lvl_ptr.size() - 1;
    h->w_pi.assign(npl, {});
    for (int l = 0; l < npl; ++l) chunk_list(h->w_pi[l], 0, m.pi_lvl_ptr[l], m.pi_lvl_ptr[l + 1], pi_chunk());
    size_t nlw = 0;
    for (int l = 0; l < nlev; ++l) nlw += h->w_level[l].size();
    size_t npw = 0;
    for (int l = 0; l < npl; ++l) npw += h->w_pi[l].size();
    const int npartial = (int)h->w_cell.size();
    std::vector<int4> fitems;
    FlowSpan fspan{};
    build_flow(h->w_level, fitems, fspan);
    int4 *d_flow = nullptr;
    int *d_ctr = nullptr;
    int4 *d_wc = nullptr, *d_wf = nullptr, *d_wg = nullptr, *d_wl = nullptr, *d_wp = nullptr;
    double *d_partial = nullptr, *d_resout = nullptr;
    const int nruns = (int)m.runs.size() / 2;
    const int ss = (int)align_up((size_t)std::max(4 * nruns, 1), 64);
    auto plan = [&](Carver &c, DevMesh &M, DevState &S) {
        M.ne = m.ne; M.nf = m.nf; M.ng = m.ng; M.sc = sc; M.sf = sf; M.sg = sg; M.ss = ss;
        S.pa = c.take<double>((size_t)kRecPA * sc);
        S.wa = c.take<double>((size_t)kNvar * sc);
        S.hwa = c.take<double>((size_t)kRecHW * sc);
        S.gradc = c.take<double>((size_t)kRecGR * sc);
        S.step = c.take<double>(sc);
        S.dd = c.take<double>(sc);
        S.fflux = c.take<double>((size_t)kRecF * sf);
        S.gpa = c.take<double>((size_t)kRecPA * sg);
        S.err = c.take<int>(8);
        S.flag = c.take<int>(sc);
        S.lucoef = c.take<double>((size_t)kRecLC * sc);
        S.lu_af = c.take<double>((size_t)3 * sc);
        S.fixed = c.take<double>((size_t)kNvar * sc);
        S.fvis = c.take<double>((size_t)kRecF * sf);
        M.lu_nb = c.take<int>((size_t)3 * sc);
        M.lu_n = c.take<double>((size_t)6 * sc);
        M.vol = c.take<double>(sc);
        M.wdis = c.take<double>((size_t)3 * sc);
        M.amls = c.take<double>((size_t)6 * sc);
        M.lvv = c.take<double>((size_t)6 * sc);
        M.stencil = c.take<int>((size_t)3 * sc);
        M.cface = c.take<int>((size_t)3 * sc);
        M.tsx = m.tsx.empty() ? nullptr : c.take<int>((size_t)kTsFaces * sc);
        M.runs = c.take<int>(m.runs.size());
        M.slot = c.take<int>((size_t)SL_N * ss);
        M.slot_n = c.take<double>((size_t)6 * ss);
        M.pi_cells = c.take<int>(m.pi_cells.size());
        M.vecx = c.take<double>((size_t)2 * sf);
        M.fvl = c.take<double>((size_t)2 * sf);
        M.fvr = c.take<double>((size_t)2 * sf);
        M.fnl = c.take<int>(sf);
        M.fnr = c.take<int>(sf);
        M.vgeo = c.take<double>((size_t)6 * sf);
        M.gnrm = c.take<double>((size_t)2 * sg);
        M.ghost_host = c.take<int>(sg);
        M.ghost_type = c.take<int>(sg);
        h->d_th = c.take<DevThermo>(1);
        h->d_par = c.take<DevParams>(1);
        h->d_tr = c.take<DevTransport>(1);
        h->d_case = c.take<CaseDev>(1);
        d_partial = c.take<double>((size_t)npartial * (kNvar + 1));
        d_resout = c.take<double>(kNvar + 1);
        d_wc = c.take<int4>(h->w_cell.size());
        d_wf = c.take<int4>(h->w_face.size());
        d_wg = c.take<int4>(h->w_ghost.size());
        d_wl = c.take<int4>(nlw);
        d_wp = c.take<int4>(npw);
        d_flow = c.take<int4>(fitems.size());
        d_ctr = c.take<int>(kXcdLists);
    };
    Carver dry;
    DevMesh M{};
    DevState S{};
    plan(dry, M, S);
    h->arena_bytes = align_up(dry.off, 256);
    HIPCHK(h, hipMalloc(&h->arena, h->arena_bytes));
    HIPCHK(h, hipMemsetAsync(h->arena, 0, h->arena_bytes, h->stream));
    Carver real;
    real.base = static_cast<char *>(h->arena);
    plan(real, M, S);
    h->M = M;
    h->S = S;
    auto up = [&](const void *dst, const void *src, size_t bytes) -> int {
        HIPCHK(h, hipMemcpyAsync(const_cast<void *>(dst), src, bytes, hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        return RC_OK;
    };
    int rc;
    std::vector<double> vol(m.vol);
    vol.resize(sc, 0.0);
    if ((rc = up(M.vol, vol.data(), sizeof(double) * sc))) return rc;
    auto wd = soa(m.wdis, m.ne, 3, sc);
    if ((rc = up(M.wdis, wd.data(), sizeof(double) * wd.size()))) return rc;
    auto am = soa(m.amls, m.ne, 6, sc);
    if ((rc = up(M.amls, am.data(), sizeof(double) * am.size()))) return rc;
    auto lv = soa(m.lvv, m.ne, 6, sc);
    if ((rc = up(M.lvv, lv.data(), sizeof(double) * lv.size()))) return rc;
    auto lnb = soa(m.lu_nb, m.ne, 3, sc);
    if ((rc = up(M.lu_nb, lnb.data(), sizeof(int) * lnb.size()))) return rc;
    auto lnn = soa(m.lu_n, m.ne, 6, sc);
    if ((rc = up(M.lu_n, lnn.data(), sizeof(double) * lnn.size()))) return rc;
    auto stn = soa(m.stencil, m.ne, 3, sc);
    if ((rc = up(M.stencil, stn.data(), sizeof(int) * stn.size()))) return rc;
    auto cf = soa(m.cface, m.ne, 3, sc);
    if ((rc = up(M.cface, cf.data(), sizeof(int) * cf.size()))) return rc;
    if (M.tsx) {
        auto tx = soa(m.tsx, m.ne, kTsFaces, sc);
        if ((rc = up(M.tsx, tx.data(), sizeof(int) * tx.size()))) return rc;
    }
    if ((rc = up(M.runs, m.runs.data(), sizeof(int) * m.runs.size()))) return rc;
    {
        // lane slots of the sweep kernels (lane p of run t: slot 4t+p)
        std::vector<int> sl((size_t)SL_N * ss, 0);
        std::vector<double> sn((size_t)6 * ss, 0.0);
        for (int t = 0; t < nruns; ++t)
            for (int p = 0; p < 4; ++p) {
                const size_t q = 4 * (size_t)t + p;
                const int first = m.runs[2 * t], cnt = m.runs[2 * t + 1];
                const int i = p < cnt ? first + p : 0;
                sl[SL_CELL * (size_t)ss + q] = i;
                sl[SL_FIRST * (size_t)ss + q] = first;
                sl[SL_CNT * (size_t)ss + q] = cnt;
                for (int k = 0; k < 3; ++k) {
                    sl[(SL_CF + k) * (size_t)ss + q] = m.cface[3 * (size_t)i + k];
                    sl[(SL_NB + k) * (size_t)ss + q] = m.lu_nb[3 * (size_t)i + k];
                }
                for (int k = 0; k < 6; ++k) sn[k * (size_t)ss + q] = m.lu_n[6 * (size_t)i + k];
            }
        if ((rc = up(M.slot, sl.data(), sizeof(int) * sl.size()))) return rc;
        if ((rc = up(M.slot_n, sn.data(), sizeof(double) * sn.size()))) return rc;
    }
    if ((rc = up(M.pi_cells, m.pi_cells.data(), sizeof(int) * m.pi_cells.size()))) return rc;
    auto vx = soa(m.vecx, m.nf, 2, sf);
    if ((rc = up(M.vecx, vx.data(), sizeof(double) * vx.size()))) return rc;
    auto fl = soa(m.fvl, m.nf, 2, sf);
    if ((rc = up(M.fvl, fl.data(), sizeof(double) * fl.size()))) return rc;
    auto fr = soa(m.fvr, m.nf, 2, sf);
    if ((rc = up(M.fvr, fr.data(), sizeof(double) * fr.size()))) return rc;
    if ((rc = up(M.fnl, m.fnl.data(), sizeof(int) * m.nf))) return rc;
    if ((rc = up(M.fnr, m.fnr.data(), sizeof(int) * m.nf))) return rc;
    auto vg = soa(m.vgeo, m.nf, 6, sf);
    if ((rc = up(M.vgeo, vg.data(), sizeof(double) * vg.size()))) return rc;
    if (m.ng) {
        auto gn = soa(m.gnrm, m.ng, 2, sg);
        if ((rc = up(M.gnrm, gn.data(), sizeof(double) * gn.size()))) return rc;
        if ((rc = up(M.ghost_host, m.ghost_host.data(), sizeof(int) * m.ng))) return rc;
        if ((rc = up(M.ghost_type, m.ghost_type.data(), sizeof(int) * m.ng))) return rc;
    }
    DevThermo th = make_thermo(h->mech);
    DevParams dp = make_params(h->par, h->mech.nsp, !h->hm.tsx.empty());
    if ((rc = up(h->d_th, &th, sizeof th))) return rc;
    if ((rc = up(h->d_par, &dp, sizeof dp))) return rc;
    CaseDev &hc = h->hc;
    hc.M = M;
    hc.S = S;
    hc.th = h->d_th;
    hc.par = h->d_par;
    hc.tr = nullptr;
    hc.partial = d_partial;
    hc.npartial = npartial;
    hc.inv_volume = 1.0 / m.total_volume;
    if ((rc = up(h->d_case, &hc, sizeof hc))) return rc;
    if (!h->w_cell.empty() && (rc = up(d_wc, h->w_cell.data(), sizeof(int4) * h->w_cell.size()))) return rc;
    if (!h->w_face.empty() && (rc = up(d_wf, h->w_face.data(), sizeof(int4) * h->w_face.size()))) return rc;
    if (!h->w_ghost.empty() && (rc = up(d_wg, h->w_ghost.data(), sizeof(int4) * h->w_ghost.size()))) return rc;
    LaunchSet &L = h->L;
    L.cases = h->d_case;
    L.ncases = 1;
    L.single = !xcd;
    L.cell = {d_wc, (int)h->w_cell.size(), 0, m.ne};
    L.norm = L.cell;
    L.face = {d_wf, (int)h->w_face.size(), 0, m.nf};
    L.ghost = {d_wg, (int)h->w_ghost.size(), 0, m.ng};
    L.level.assign(nlev, {});
    size_t off = 0;
    for (int l = 0; l < nlev; ++l) {
        const size_t n = h->w_level[l].size();
        if ((rc = up(d_wl + off, h->w_level[l].data(), sizeof(int4) * n))) return rc;
        L.level[l] = {d_wl + off, (int)n, m.lvl_ptr[l], m.lvl_ptr[l + 1], kRunThreads};
        off += n;
    }
    L.pi_level.assign(npl, {});
    off = 0;
    for (int l = 0; l < npl; ++l) {
        const size_t n = h->w_pi[l].size();
        if ((rc = up(d_wp + off, h->w_pi[l].data(), sizeof(int4) * n))) return rc;
        L.pi_level[l] = {d_wp + off, (int)n, m.pi_lvl_ptr[l], m.pi_lvl_ptr[l + 1]};
        off += n;
    }
    if (!fitems.empty() && (rc = up(d_flow, fitems.data(), sizeof(int4) * fitems.size()))) return rc;
    L.flow = {d_flow, fspan, d_ctr, flow_grid(h->device, fspan)};
    L.next_res = next_res_levels(L, h->device);
    L.top_fused = top_fused_env();
    L.resout = d_resout;
    return RC_OK;
}

// the stages of one step; with timing on, events bracket every kernel
int ensure_events(rc_case *h, size_t n) {
    while (h->evpool.size() < n) {
        hipEvent_t e;
        HIPCHK(h, hipEventCreate(&e));
        h->evpool.push_back(e);
    }
    return RC_OK;
}
void mark(rc_case *h, int k) {
    if (!h->timing) return;
    if (ensure_events(h, h->ev_used + kTimeSlots + 1)) return;
    (void)hipEventRecord(h->evpool[h->ev_used + k], h->stream);
}
// accumulates the recorded steps' kernel times (after the stream has drained)
int drain_timing(rc_case *h) {
    if (h->ev_used == 0) return RC_OK;
    HIPCHK(h, hipEventSynchronize(h->evpool[h->ev_used - 1]));
    for (size_t b = 0; b + kTimeSlots < h->ev_used; b += kTimeSlots + 1) {
        for (int k = 0; k < kTimeSlots; ++k) {
            float ms = 0.f;
            HIPCHK(h, hipEventElapsedTime(&ms, h->evpool[b + k], h->evpool[b + k + 1]));
            h->tms[k] += ms;
        }
        h->timed_steps++;
    }
    h->ev_used = 0;
    return RC_OK;
}
// one steady step as the fused kernel pipeline:
//   step counter, gradient+limiter (ghost states formed in-kernel), flux
//   (+ TimeStep face part), LU-SGS forward levels that first form their
//   cells' residual + source (+ TimeStep cell part), backward levels, update.
void enqueue_step_kernels(rc_case *h) {
    const LaunchSet &L = h->L;
    // the dataflow sweep needs k_step_begin (its item counters); otherwise
    // the gradient and update kernels begin and end the step on the device
    // counter themselves
    const bool flow = L.flow.grid > 0;
    if (flow) launch_step_begin(L, h->stream);
    mark(h, 0);
    if (flow) launch_gradlim(L, true, 2, h->stream);
    else launch_gradlim_step(L, h->stream);
    mark(h, 1);
    launch_flux(L, true, h->stream);
    if (h->viscous) launch_viscous(L, true, true, h->stream);
    if (h->sur) surrogate_enqueue(h->sur, h->d_case, h->ne, h->stream);
    mark(h, 2);
    mark(h, 3);                 // the residual is formed in the forward sweep
    if (L.flow.grid > 0) {
        // both sweeps in one launch (timed in the forward slot)
        launch_lusgs_flow(L, h->stream);
        mark(h, 4);
        mark(h, 5);
    } else {
        // forward levels below the highest; the highest level (with its
        // backward sweep when fused); the other backward levels
        for (int l = 0; l + 1 < h->nlev; ++l) launch_lusgs_fused_forward(L, l, h->stream);
        mark(h, 4);
        if (h->nlev > 0) launch_lusgs_fused_forward(L, h->nlev - 1, h->stream);
        mark(h, 5);
        enqueue_backward_levels(L, h->nlev, h->stream);
    }
    mark(h, 6);
    if (flow) launch_update(L, h->stream);
    else launch_update_step(L, h->stream);
}

// the step captured once into a HIP graph and replayed (no memset nodes:
// every node is a kernel)
int ensure_graph(rc_case *h) {
    if (h->graph_exec) return RC_OK;
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step_kernels(h);
    hipGraph_t g = nullptr;
    HIPCHK(h, hipStreamEndCapture(h->stream, &g));
    hipError_t e = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        h->graph_exec = nullptr;
        return set_err(h, RC_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    // its one-time set-up now, not at its first replay
    HIPCHK(h, hipGraphUpload(h->graph_exec, h->stream));
    return RC_OK;
}

// Most steps in one graph: a run of steps between residual reports is
// replayed as graphs of several steps instead of one graph per step (a
// graph holds the same kernels, step after step, so results are the same
// bit for bit); RC_GRAPH_STEPS=n sets it (rounded down to a power of two;
// 1: one graph per step only).
int graph_steps() {
    const char *e = std::getenv("RC_GRAPH_STEPS");
    const int v = e ? std::atoi(e) : kGraphSteps;
    return v < 1 ? 1 : (v > 64 ? 64 : v);
}

// the largest power of two not above graph_steps()
int graph_top() {
    int p = 1;
    while (2 * p <= graph_steps()) p *= 2;
    return p;
}
// RC_PERSIST=n (1 <= n <= 64): steps run as the persistent step, n steps
// per cooperative launch at most (read when a case or batch is created;
// unset or 0: one graph launch per step or per run of steps)
int persist_env() {
    const char *e = std::getenv("RC_PERSIST");
    const int n = e ? std::atoi(e) : 0;
    return std::max(0, std::min(n, 64));
}

// slot of the graph of p = 2, 4, 8, ... steps
int pow_index(int p) {
    int j = -1;
    while (p > 1) {
        p >>= 1;
        ++j;
    }
    return j;
}

void drop_graphs(rc_case *h) {
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    h->graph_exec = nullptr;
    for (auto &g : h->graph_pow) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    h->graph_pow_top = 0;
}

int capture_steps(rc_case *h, int nsteps, hipGraphExec_t &out) {
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < nsteps; ++k) enqueue_step_kernels(h);
    hipGraph_t g = nullptr;
    HIPCHK(h, hipStreamEndCapture(h->stream, &g));
    hipError_t e = hipGraphInstantiate(&out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        out = nullptr;
        return set_err(h, RC_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    HIPCHK(h, hipGraphUpload(out, h->stream));
    return RC_OK;
}

// the one-step graph and the graphs of 2, 4, ... graph_top() steps, all at
// the first run of more than one step (so that no later run captures inside
// its steps; a case only ever stepped one step at a time captures one)
int ensure_graphs(rc_case *h) {
    int rc = ensure_graph(h);
    if (rc) return rc;
    const int top = graph_top();
    if (h->graph_pow_top == top) return RC_OK;
    for (auto &g : h->graph_pow) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    h->graph_pow_top = 0;
    for (int p = 2; p <= top; p *= 2)
        if ((rc = capture_steps(h, p, h->graph_pow[pow_index(p)]))) return rc;
    h->graph_pow_top = top;
    return RC_OK;
}

int enqueue_stage(rc_case *h, bool allow_graph);

// k steps as graphs of graph_top() steps, then the rest by its binary
// digits (at most k/top + log2(top) graph launches); each graph holds the
// same kernels step after step, so the results are those of one graph per
// step
int enqueue_stages(rc_case *h, int k) {
    if (h->persist > 0 && !h->timing && !h->viscous && !h->sur && persist_supported(h->L)) {
        if (!h->d_bar) HIPCHK(h, hipMalloc(&h->d_bar, 2 * sizeof(unsigned)));
        const int grid = persist_grid(h->L, h->device);
        if (grid <= 0) return set_err(h, RC_ERR_DEVICE, "persistent step: no resident blocks");
        while (k > 0) {
            const int run = std::min(k, h->persist);
            HIPCHK(h, launch_step_persist(h->L, h->d_bar, run, grid, h->stream));
            h->steps_done += run;
            k -= run;
        }
        return RC_OK;
    }
    if (h->use_graph && !h->timing && k > 1) {
        int rc = ensure_graphs(h);
        if (rc) return rc;
        for (int p = h->graph_pow_top; p >= 2; p /= 2) {
            while (k >= p) {
                HIPCHK(h, hipGraphLaunch(h->graph_pow[pow_index(p)], h->stream));
                h->steps_done += p;
                if (h->sur) h->sur_evals += (long long)p * h->ne;
                k -= p;
            }
        }
    }
    for (; k > 0; --k) {
        int rc = enqueue_stage(h, true);
        if (rc) return rc;
    }
    return RC_OK;
}

int enqueue_stage(rc_case *h, bool allow_graph) {
    if (h->timing) {
        if (h->ev_used + kTimeSlots + 1 > 4096) {
            int rc = drain_timing(h);
            if (rc) return rc;
        }
        enqueue_step_kernels(h);
        int rc = ensure_events(h, h->ev_used + kTimeSlots + 1);
        if (rc) return rc;
        HIPCHK(h, hipEventRecord(h->evpool[h->ev_used + kTimeSlots], h->stream));
        h->ev_used += kTimeSlots + 1;
    } else if (allow_graph && h->use_graph) {
        int rc = ensure_graph(h);
        if (rc) return rc;
        HIPCHK(h, hipGraphLaunch(h->graph_exec, h->stream));
    } else {
        enqueue_step_kernels(h);
    }
    h->steps_done++;
    if (h->sur) h->sur_evals += h->ne;
    return RC_OK;
}

int residual_async(rc_case *h, double *host_out) {
    launch_resnorm(h->L, h->stream);
    HIPCHK(h, hipMemcpyAsync(host_out, h->L.resout, sizeof(double) * (kNvar + 1), hipMemcpyDeviceToHost, h->stream));
    return RC_OK;
}

// the per-cell block factors of the backward-Euler sweeps (allocated once,
// at the first backward-Euler call of the case)
int ensure_be_storage(rc_case *h) {
    if (h->be_fac) return RC_OK;
    HIPCHK(h, hipMalloc(&h->be_fac, sizeof(double) * be_factor_doubles() * h->M.sc));
    HIPCHK(h, hipMalloc(&h->be_piv, sizeof(int) * be_factor_ints() * h->M.sc));
    HIPCHK(h, hipMalloc(&h->be_nbc, sizeof(double) * be_cache_doubles() * h->M.sc));
    return RC_OK;
}

// the BDF2 states (allocated once, at the first BDF2 call of the case)
int ensure_bdf2_storage(rc_case *h) {
    if (h->bdf_wp) return RC_OK;
    HIPCHK(h, hipMalloc(&h->bdf_wp, sizeof(double) * kNvar * h->M.sc));
    HIPCHK(h, hipMalloc(&h->bdf_wp2, sizeof(double) * kNvar * h->M.sc));
    return RC_OK;
}

// the prologue of BDF2_stage (Time_BDF2.f90:12-20): at the first step WP =
// WA, WP2 = WP, phi = 0; later WP2 = WP, WP = WA, phi = 0.5
int bdf2_begin_step(rc_case *h, bool first, double *phi) {
    const size_t bytes = sizeof(double) * kNvar * h->M.sc;
    if (first) {
        HIPCHK(h, hipMemcpyAsync(h->bdf_wp, h->S.wa, bytes, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->bdf_wp2, h->bdf_wp, bytes, hipMemcpyDeviceToDevice, h->stream));
        *phi = 0.0;
    } else {
        HIPCHK(h, hipMemcpyAsync(h->bdf_wp2, h->bdf_wp, bytes, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->bdf_wp, h->S.wa, bytes, hipMemcpyDeviceToDevice, h->stream));
        *phi = 0.5;
    }
    return RC_OK;
}

// backwardEuler_LU_SGS: dual-time right-hand side and diagonal, every
// cell's block factored, the point-implicit sweeps level by level, PA +=
// dPA and WA = U(PA); with bdf2 the right-hand side and diagonal of
// BDF2_LU_SGS(Inner_rk, phi) (the sweeps are the same)
void enqueue_be_lusgs(rc_case *h, double dt, bool first, bool bdf2 = false, double phi = 0.0) {
    const LaunchSet &L = h->L;
    launch_be_prepare(L, dt, first, bdf2 ? h->bdf_wp : nullptr, bdf2 ? h->bdf_wp2 : nullptr, phi, h->stream);
    launch_be_factor(L, h->be_fac, h->be_piv, h->be_nbc, h->stream);
    for (int l = 0; l < h->npl; ++l) launch_be_level(L, true, l, h->be_fac, h->be_piv, h->be_nbc, h->stream);
    for (int l = h->npl - 1; l >= 0; --l)
        launch_be_level(L, false, l, h->be_fac, h->be_piv, h->be_nbc, h->stream);
    launch_be_update(L, h->stream);
}

// copies n items of `width` values out of a device array laid out either as
// one array per value (stride = array length) or as records (rec > 0 values
// per item, the first `width` used), into out[i*width + c]
int get_items(rc_case *h, double *out, const double *dev, int n, int width, int stride, int rec) {
    if (!out || n == 0) return RC_OK;
    const size_t len = rec > 0 ? (size_t)rec * n : (size_t)width * stride;
    std::vector<double> tmp(len);
    HIPCHK(h, hipMemcpyAsync(tmp.data(), dev, sizeof(double) * len, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < width; ++c)
            out[(size_t)i * width + c] = rec > 0 ? tmp[(size_t)i * rec + c] : tmp[(size_t)c * stride + i];
    return RC_OK;
}

int bset_err(rc_batch *b, int code, const std::string &msg) {
    if (b) b->err = msg;
    g_error = msg;
    return code;
}

#define BHIPCHK(b, x)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return bset_err(b, RC_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
    } while (0)

void batch_drop_graphs(rc_batch *b) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    b->graph_exec = nullptr;
    for (auto &g : b->graph_pow) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    b->graph_pow_top = 0;
}

void batch_release(rc_batch *b) {
    batch_drop_graphs(b);
    if (b->arena) {
        (void)hipStreamSynchronize(b->stream);
        (void)hipFree(b->arena);
    }
    b->arena = nullptr;
    b->d_err = nullptr;
}

// concatenates the members' descriptors and work lists (case index =
// position in the batch) into one device allocation
int batch_build(rc_batch *b) {
    if (!b->dirty) return RC_OK;
    batch_release(b);
    const int n = (int)b->members.size();
    std::vector<CaseDev> hcs(n);
    std::vector<int4> wc, wf, wg, wn;
    std::vector<std::vector<int4>> wl;
    const int rchunk = b->run_bs / kRunMax;
    for (int c = 0; c < n; ++c) {
        const rc_case *h = b->members[c];
        hcs[c] = h->hc;
        const bool xcd = b->xcd < 0 ? h->L.single == false : b->xcd == 1;
        // residual partial sums: the member's own chunks own its slots
        for (int4 w : h->w_cell) wn.push_back(make_int4(c, w.y, w.z, w.w));
        // the step's lists, cut to the batch's launch shape (item lists of
        // XCD-banded members stay multiples of 8 long, so the banding holds
        // in the concatenation)
        if (xcd) {
            pad_xcd(wc, c);
            pad_xcd(wf, c);
            xcd_list(wc, c, h->seg_cell, b->cell_bs);
            xcd_list(wf, c, h->seg_face, b->cell_bs);
        } else {
            chunk_list(wc, c, 0, h->ne, b->cell_bs);
            chunk_list(wf, c, 0, h->nf, b->cell_bs);
        }
        chunk_list(wg, c, 0, h->ng, b->cell_bs);
        const int nl = (int)h->hm.lvl_ptr.size() - 1;
        if (nl > (int)wl.size()) wl.resize(nl);
        for (int l = 0; l < nl; ++l) {
            const int a = h->hm.lvl_ptr[l], e = h->hm.lvl_ptr[l + 1];
            if (xcd) {
                pad_xcd(wl[l], c);
                xcd_list(wl[l], c, {{a, e}}, rchunk);
            } else {
                chunk_list(wl[l], c, a, e, rchunk);
            }
        }
    }
    b->nlev = (int)wl.size();
    size_t nlw = 0;
    for (auto &v : wl) nlw += v.size();
    std::vector<int4> fitems;
    FlowSpan fspan{};
    build_flow(wl, fitems, fspan);
    Carver dry, real;
    int4 *dwn = nullptr, *dflow = nullptr;
    int *dctr = nullptr, *derr = nullptr;
    auto plan = [&](Carver &c, CaseDev *&dc, int4 *&dwc, int4 *&dwf, int4 *&dwg, int4 *&dwl, double *&dres) {
        dc = c.take<CaseDev>(std::max(n, 1));
        derr = c.take<int>(4 * (size_t)std::max(n, 1));
        dflow = c.take<int4>(fitems.size());
        dctr = c.take<int>(kXcdLists);
        dwn = c.take<int4>(wn.size());
        dwc = c.take<int4>(wc.size());
        dwf = c.take<int4>(wf.size());
        dwg = c.take<int4>(wg.size());
        dwl = c.take<int4>(nlw);
        dres = c.take<double>((size_t)std::max(n, 1) * (kNvar + 1));
    };
    CaseDev *dc;
    int4 *dwc, *dwf, *dwg, *dwl;
    double *dres;
    plan(dry, dc, dwc, dwf, dwg, dwl, dres);
    BHIPCHK(b, hipSetDevice(b->device));
    BHIPCHK(b, hipMalloc(&b->arena, align_up(dry.off, 256)));
    real.base = static_cast<char *>(b->arena);
    plan(real, dc, dwc, dwf, dwg, dwl, dres);
    auto up = [&](void *dst, const void *src, size_t bytes) -> int {
        if (bytes) BHIPCHK(b, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b->stream));
        return RC_OK;
    };
    int rc;
    if ((rc = up(dc, hcs.data(), sizeof(CaseDev) * n))) return rc;
    if ((rc = up(dwc, wc.data(), sizeof(int4) * wc.size()))) return rc;
    if ((rc = up(dwf, wf.data(), sizeof(int4) * wf.size()))) return rc;
    if ((rc = up(dwg, wg.data(), sizeof(int4) * wg.size()))) return rc;
    if ((rc = up(dwn, wn.data(), sizeof(int4) * wn.size()))) return rc;
    if ((rc = up(dflow, fitems.data(), sizeof(int4) * fitems.size()))) return rc;
    BHIPCHK(b, hipMemsetAsync(dctr, 0, sizeof(int) * kXcdLists, b->stream));
    LaunchSet &L = b->L;
    L.cases = dc;
    L.ncases = n;
    L.single = false;
    L.cell = {dwc, (int)wc.size(), 0, 0, b->cell_bs};
    L.face = {dwf, (int)wf.size(), 0, 0, b->cell_bs};
    L.ghost = {dwg, (int)wg.size(), 0, 0, b->cell_bs};
    L.norm = {dwn, (int)wn.size(), 0, 0, kChunk};
    L.level.assign(b->nlev, {});
    size_t off = 0;
    for (int l = 0; l < b->nlev; ++l) {
        if ((rc = up(dwl + off, wl[l].data(), sizeof(int4) * wl[l].size()))) return rc;
        L.level[l] = {dwl + off, (int)wl[l].size(), 0, 0, b->run_bs};
        off += wl[l].size();
    }
    L.flow = {dflow, fspan, dctr, flow_grid(b->device, fspan)};
    L.next_res = next_res_levels(L, b->device);
    L.top_fused = top_fused_env();
    L.resout = dres;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    const size_t need = (size_t)std::max(n, 1) * (kNvar + 1);
    if (b->h_res_n < need) {
        if (b->h_res) (void)hipHostFree(b->h_res);
        b->h_res = nullptr;
        b->h_res_n = 0;
        BHIPCHK(b, hipHostMalloc(&b->h_res, need * sizeof(double)));
        b->h_res_n = need;
    }
    b->d_err = derr;
    const size_t nerr = 4 * (size_t)std::max(n, 1);
    if (b->h_err_n < nerr) {
        if (b->h_err) (void)hipHostFree(b->h_err);
        b->h_err = nullptr;
        b->h_err_n = 0;
        BHIPCHK(b, hipHostMalloc(&b->h_err, nerr * sizeof(int)));
        b->h_err_n = nerr;
    }
    b->dirty = false;
    return RC_OK;
}

// one steady step of every member: the same kernels as a single case's step
void enqueue_batch_step(rc_batch *b) {
    const LaunchSet &L = b->L;
    const bool flow = L.flow.grid > 0;
    if (flow) launch_step_begin(L, b->stream);
    if (flow) launch_gradlim(L, true, 2, b->stream);
    else launch_gradlim_step(L, b->stream);
    launch_flux(L, true, b->stream);
    bool visc = false;
    for (const rc_case *m : b->members) visc = visc || m->viscous;
    if (visc) launch_viscous(L, true, true, b->stream);
    if (L.flow.grid > 0) {
        launch_lusgs_flow(L, b->stream);
    } else {
        enqueue_sweeps(L, b->nlev, b->stream);
        enqueue_backward_levels(L, b->nlev, b->stream);
    }
    if (flow) launch_update(L, b->stream);
    else launch_update_step(L, b->stream);
}

// the batch step captured as a graph of nsteps steps
int batch_capture(rc_batch *b, int nsteps, hipGraphExec_t &out) {
    BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < nsteps; ++k) enqueue_batch_step(b);
    hipGraph_t g = nullptr;
    BHIPCHK(b, hipStreamEndCapture(b->stream, &g));
    hipError_t e = hipGraphInstantiate(&out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        out = nullptr;
        return bset_err(b, RC_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    BHIPCHK(b, hipGraphUpload(out, b->stream));
    return RC_OK;
}

// the one-step graph and the graphs of 2, 4, ... graph_top() steps, all at
// the first replay of a build
int batch_graphs(rc_batch *b) {
    int rc;
    if (!b->graph_exec && (rc = batch_capture(b, 1, b->graph_exec))) return rc;
    const int top = graph_top();
    if (b->graph_pow_top == top) return RC_OK;
    for (auto &g : b->graph_pow) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    b->graph_pow_top = 0;
    for (int p = 2; p <= top; p *= 2)
        if ((rc = batch_capture(b, p, b->graph_pow[pow_index(p)]))) return rc;
    b->graph_pow_top = top;
    return RC_OK;
}

// device error flags of every member (gathered by one kernel behind the
// batch's work, one copy, one synchronisation); the first failing member's
// code
int batch_check(rc_batch *b, int *err_case) {
    launch_gather_err(b->L, b->d_err, b->stream);
    BHIPCHK(b, hipMemcpyAsync(b->h_err, b->d_err, sizeof(int) * 4 * b->members.size(), hipMemcpyDeviceToHost,
                              b->stream));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    for (size_t c = 0; c < b->members.size(); ++c) {
        rc_case *h = b->members[c];
        if (b->h_err[4 * c] == 0) continue;
        for (int k = 0; k < 4; ++k) h->h_err[k] = b->h_err[4 * c + k];
        const int rc = check_device_error(h, true);
        if (rc) {
            if (err_case) *err_case = (int)c;
            return bset_err(b, rc, "case " + std::to_string(c) + ": " + h->err);
        }
    }
    return RC_OK;
}

}  // namespace

extern "C" {

const char *rc_last_error(const rc_case *h) { return h ? h->err.c_str() : g_error.c_str(); }

int rc_case_create(const rc_mechanism *mech, const rc_params *par, int nnt, const double *coor, int ne,
                   const int *cellnodes, int nb, const int *bnodes, const int *btype, int device, rc_case **out) {
    return rc_case_create_periodic(mech, par, nnt, coor, ne, cellnodes, nb, bnodes, btype, 0, nullptr, nullptr,
                                   device, out);
}

int rc_case_create_periodic(const rc_mechanism *mech, const rc_params *par, int nnt, const double *coor, int ne,
                            const int *cellnodes, int nb, const int *bnodes, const int *btype, int ncon,
                            const int *con, const double *dxyz, int device, rc_case **out) {
    if (!out) return set_err(nullptr, RC_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!mech || !par || !coor || !cellnodes || (nb > 0 && (!bnodes || !btype)) || ncon < 0 ||
        (ncon > 0 && (!con || !dxyz)))
        return set_err(nullptr, RC_ERR_ARGUMENT, "NULL argument");
    if (mech->nsp != kNsp)
        return set_err(nullptr, RC_ERR_ARGUMENT, "the kernels are built for 4 species (Nvar = 7)");
    if (mech->model != 0 && mech->model != 1)
        return set_err(nullptr, RC_ERR_ARGUMENT, "No such code_chemical_model");
    rc_case *h = new rc_case();
    h->mech = *mech;
    h->par = *par;
    h->device = device;
    std::vector<int> c0(3 * (size_t)ne), b0(2 * (size_t)std::max(nb, 0));
    for (size_t k = 0; k < c0.size(); ++k) c0[k] = cellnodes[k] - 1;
    for (size_t k = 0; k < b0.size(); ++k) b0[k] = bnodes[k] - 1;
    std::string msg;
    int rc = build_mesh(nnt, coor, ne, c0.data(), nb, b0.data(), btype, ncon, con, dxyz, h->hm, msg);
    if (rc) {
        delete h;
        return set_err(nullptr, rc, msg);
    }
    h->ne = h->hm.ne;
    h->nf = h->hm.nf;
    h->ng = h->hm.ng;
    h->nlev = (int)h->hm.lvl_ptr.size() - 1;
    h->npl = (int)h->hm.pi_lvl_ptr.size() - 1;
    auto bail = [&](int code) {
        g_error = h->err;
        rc_case_destroy(h);
        return code;
    };
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return bail(set_err(h, RC_ERR_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e)));
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(set_err(h, RC_ERR_DEVICE, std::string("stream: ") + hipGetErrorString(e)));
    if (hipHostMalloc(&h->h_err, 4 * sizeof(int)) != hipSuccess ||
        hipHostMalloc(&h->h_res, (kNvar + 1) * sizeof(double)) != hipSuccess)
        return bail(set_err(h, RC_ERR_MEMORY, "pinned host allocation failed"));
    rc = layout_and_upload(h);
    if (rc) return bail(rc);
    const char *ng = std::getenv("RC_NO_GRAPH");
    h->use_graph = !(ng && ng[0] == '1');
    h->persist = persist_env();
    *out = h;
    return RC_OK;
}

void rc_case_destroy(rc_case *h) {
    if (!h) return;
    if (h->batch) rc_batch_remove(h->batch, h);
    if (h->stream) {
        (void)hipSetDevice(h->device);
        (void)hipStreamSynchronize(h->stream);
    }
    drop_graphs(h);
    if (h->arena) (void)hipFree(h->arena);
    if (h->be_fac) (void)hipFree(h->be_fac);
    if (h->be_piv) (void)hipFree(h->be_piv);
    if (h->be_nbc) (void)hipFree(h->be_nbc);
    if (h->bdf_wp) (void)hipFree(h->bdf_wp);
    if (h->bdf_wp2) (void)hipFree(h->bdf_wp2);
    if (h->sur) surrogate_destroy(h->sur);
    if (h->sur_src) (void)hipFree(h->sur_src);
    if (h->d_bar) (void)hipFree(h->d_bar);
    for (auto &ev : h->evpool) (void)hipEventDestroy(ev);
    if (h->h_err) (void)hipHostFree(h->h_err);
    if (h->h_res) (void)hipHostFree(h->h_res);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int rc_case_info(const rc_case *h, long long *s) {
    if (!h || !s) return set_err(nullptr, RC_ERR_ARGUMENT, "NULL argument");
    s[0] = h->ne; s[1] = h->nf; s[2] = h->ng; s[3] = h->nvar; s[4] = h->nlev; s[5] = (long long)h->arena_bytes;
    return RC_OK;
}

int rc_case_levels(const rc_case *h, int *cells, int n) {
    if (!h || !cells) return set_err(nullptr, RC_ERR_ARGUMENT, "NULL argument");
    if (n < h->nlev) return set_err(nullptr, RC_ERR_ARGUMENT, "fewer slots than levels");
    const HostMesh &m = h->hm;
    for (int l = 0; l < n; ++l) cells[l] = 0;
    for (int l = 0; l < h->nlev; ++l)
        for (int r = m.lvl_ptr[l]; r < m.lvl_ptr[l + 1]; ++r) cells[l] += m.runs[2 * r + 1];
    return RC_OK;
}

int rc_set_primitive(rc_case *h, const double *pa) {
    if (!h || !pa) return set_err(h, RC_ERR_ARGUMENT, "NULL argument");
    HIPCHK(h, hipSetDevice(h->device));
    std::vector<double> tmp((size_t)kRecPA * h->M.sc, 0.0);
    for (int i = 0; i < h->ne; ++i)
        for (int k = 0; k < kNvar; ++k) tmp[(size_t)i * kRecPA + k] = pa[(size_t)i * kNvar + k];
    HIPCHK(h, hipMemcpyAsync(h->S.pa, tmp.data(), sizeof(double) * tmp.size(), hipMemcpyHostToDevice, h->stream));
    launch_prim_to_cons(h->L, h->stream);
    return sync_check(h);
}

int rc_timestep(rc_case *h) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    HIPCHK(h, hipSetDevice(h->device));
    launch_ts_face(h->L, h->stream);
    if (h->viscous) launch_viscous(h->L, true, false, h->stream);
    launch_ts_cell(h->L, h->stream);
    return sync_check(h);
}

int rc_gradient(rc_case *h) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    HIPCHK(h, hipSetDevice(h->device));
    launch_gradlim(h->L, true, 0, h->stream);
    return sync_check(h);
}

int rc_limiter_barth(rc_case *h) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    HIPCHK(h, hipSetDevice(h->device));
    launch_gradlim(h->L, false, 1, h->stream);
    return sync_check(h);
}

int rc_reconstruction(rc_case *h) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    HIPCHK(h, hipSetDevice(h->device));
    launch_gradlim(h->L, true, 2, h->stream);
    return sync_check(h);
}

int rc_flux_sourceterm(rc_case *h) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    HIPCHK(h, hipSetDevice(h->device));
    launch_flux(h->L, false, h->stream);
    if (h->viscous) launch_viscous(h->L, false, true, h->stream);
    if (h->sur) {
        surrogate_enqueue(h->sur, h->d_case, h->ne, h->stream);
        h->sur_evals += h->ne;
    }
    launch_residual(h->L, false, h->stream);
    return sync_check(h);
}

int rc_lusgs_steady(rc_case *h) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    HIPCHK(h, hipSetDevice(h->device));
    for (int l = 0; l < h->nlev; ++l) launch_lusgs_level(h->L, true, l, false, h->stream);
    for (int l = h->nlev - 1; l >= 0; --l) launch_lusgs_level(h->L, false, l, false, h->stream);
    launch_update(h->L, h->stream);
    return sync_check(h);
}

int rc_lusgs_stage(rc_case *h) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    HIPCHK(h, hipSetDevice(h->device));
    int rc = enqueue_stage(h, false);
    if (rc) return rc;
    return sync_check(h);
}

int rc_run_steady(rc_case *h, int nsteps, int every, double *res_out, int *done) {
    if (!h || nsteps < 0) return set_err(h, RC_ERR_ARGUMENT, "bad argument");
    HIPCHK(h, hipSetDevice(h->device));
    if (every <= 0) every = nsteps + 1;
    int nrep = 0;
    // reports (residual rows, device error check): at step 1 when rows are
    // asked for, and every `every` steps
    const bool first = res_out != nullptr;
    for (int s = 0; s < nsteps;) {
        const int next = std::min(nsteps, s == 0 && first ? 1 : (s / every + 1) * every);
        int rc = enqueue_stages(h, next - s);
        if (rc) return rc;
        s = next;
        if ((s == 1 && first) || s % every == 0) {
            if (res_out) {
                rc = residual_async(h, h->h_res);
                if (rc) return rc;
            }
            rc = sync_check(h);
            if (rc) {
                if (done) *done = s;
                return rc;
            }
            if (res_out) std::memcpy(res_out + (size_t)nrep * (kNvar + 1), h->h_res, sizeof(double) * (kNvar + 1));
            ++nrep;
        }
    }
    int rc = sync_check(h);
    if (done) *done = nsteps;
    return rc;
}

int rc_set_transport(rc_case *h, const rc_transport *t) {
    if (!h || !t) return set_err(h, RC_ERR_ARGUMENT, "NULL argument");
    const rc_mechanism &m = h->mech;
    if (!(t->t_low < t->t_crit && t->t_crit < t->t_high)) return set_err(h, RC_ERR_ARGUMENT, "bad transport ranges");
    DevTransport d{};
    d.t_low = t->t_low;
    d.t_crit = t->t_crit;
    d.t_high = t->t_high;
    for (int i = 0; i < kNsp; ++i) {
        for (int j = 0; j < 4; ++j) {
            d.mu_low[i][j] = t->mu_low[i][j];
            d.mu_high[i][j] = t->mu_high[i][j];
            d.k_low[i][j] = t->k_low[i][j];
            d.k_high[i][j] = t->k_high[i][j];
        }
        // the per-pair constants of compute_transport_property (:113-116)
        // and compute_diffusion_velocity (:72-80)
        for (int j = 0; j < kNsp; ++j) {
            const double mi = m.mole_weight[i], mj = m.mole_weight[j];
            d.wa[i][j] = std::pow(mj / mi, 0.25);
            d.wb[i][j] = std::sqrt(1. + mi / mj);
            d.sqmw[i][j] = std::sqrt(2. / (1. / mi + 1. / mj) * 1000);
            const double cd = 0.5 * (t->sigma[i] + t->sigma[j]);
            d.cd2[i][j] = cd * cd;
            d.sqeps[i][j] = std::sqrt(t->eps[i] * t->eps[j]);
        }
    }
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_tr, &d, sizeof d, hipMemcpyHostToDevice, h->stream));
    h->hc.tr = h->d_tr;
    HIPCHK(h, hipMemcpyAsync(h->d_case, &h->hc, sizeof h->hc, hipMemcpyHostToDevice, h->stream));
    DevParams dp = make_params(h->par, h->mech.nsp, !h->hm.tsx.empty());
    dp.viscous = 1;
    dp.surrogate = h->sur ? 1 : 0;
    HIPCHK(h, hipMemcpyAsync(h->d_par, &dp, sizeof dp, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->viscous = true;
    // the captured step and any batch holding the case change shape
    drop_graphs(h);
    if (h->batch) {
        h->batch->dirty = true;
        batch_drop_graphs(h->batch);
    }
    return RC_OK;
}

int rc_set_surrogate(rc_case *h, rc_mlp *m, const double *xform, double tol, double rtol) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    if (m && !xform) return set_err(h, RC_ERR_ARGUMENT, "NULL transform");
    if (h->batch) return set_err(h, RC_ERR_ARGUMENT, "a case with the learned source term steps alone");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (h->sur) surrogate_destroy(h->sur);
    h->sur = nullptr;
    h->sur_evals = 0;
    if (m) {
        std::string msg;
        int rc = surrogate_create(m, h->ne, xform, tol, rtol, &h->sur, msg);
        if (rc) return set_err(h, rc, msg);
        if (!h->sur_src) HIPCHK(h, hipMalloc(&h->sur_src, sizeof(double) * (kNsp - 1) * h->M.sc));
    }
    h->hc.S.src = m ? h->sur_src : nullptr;
    h->S.src = h->hc.S.src;
    HIPCHK(h, hipMemcpyAsync(h->d_case, &h->hc, sizeof h->hc, hipMemcpyHostToDevice, h->stream));
    DevParams dp = make_params(h->par, h->mech.nsp, !h->hm.tsx.empty());
    dp.viscous = h->viscous ? 1 : 0;
    dp.surrogate = m ? 1 : 0;
    HIPCHK(h, hipMemcpyAsync(h->d_par, &dp, sizeof dp, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drop_graphs(h);
    return RC_OK;
}

int rc_set_surrogate_trust(rc_case *h, const int *dims3, const double *lo3, const double *hi3,
                           const unsigned char *accept) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    if (!h->sur) return set_err(h, RC_ERR_ARGUMENT, "no learned source set (rc_set_surrogate first)");
    if (accept && (!dims3 || !lo3 || !hi3)) return set_err(h, RC_ERR_ARGUMENT, "NULL argument");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    std::string msg;
    const int rc = surrogate_set_trust(h->sur, dims3, lo3, hi3, accept, msg);
    if (rc) return set_err(h, rc, msg);
    // the gate's table is a kernel argument: capture the step again
    drop_graphs(h);
    return RC_OK;
}

int rc_get_surrogate_stats(rc_case *h, long long *acc_eval) {
    if (!h || !acc_eval) return set_err(h, RC_ERR_ARGUMENT, "NULL argument");
    acc_eval[0] = acc_eval[1] = 0;
    if (!h->sur) return RC_OK;
    HIPCHK(h, hipSetDevice(h->device));
    long long acc = 0;
    if (surrogate_accepted(h->sur, h->stream, acc)) return set_err(h, RC_ERR_DEVICE, "reading the counters failed");
    acc_eval[0] = acc;
    acc_eval[1] = h->sur_evals;
    return RC_OK;
}

int rc_set_outlet_pressure(rc_case *h, double p_out) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    HIPCHK(h, hipSetDevice(h->device));
    h->par.p_out = p_out;
    HIPCHK(h, hipMemcpyAsync(reinterpret_cast<char *>(h->d_par) + offsetof(DevParams, p_out), &h->par.p_out,
                             sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return RC_OK;
}

int rc_backward_euler_lusgs(rc_case *h, double dt, int inner) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    if (!(dt > 0.0) || inner < 1) return set_err(h, RC_ERR_ARGUMENT, "dt must be > 0 and inner >= 1");
    HIPCHK(h, hipSetDevice(h->device));
    int rc = ensure_be_storage(h);
    if (rc) return rc;
    enqueue_be_lusgs(h, dt, inner == 1);
    return sync_check(h);
}

}  // extern "C"

namespace {
// backwardEuler_stage (Time_BackwardEuler.f90:4-56), or with bdf2 BDF2_stage
// (Time_BDF2.f90:4-69): the same inner loop around the two LU-SGS variants
int dual_time_stage(rc_case *h, double dt, int innerloop, double reltol, int *ninner, double *res_out, bool bdf2,
                    double phi) {
    const LaunchSet &L = h->L;
    launch_step_begin(L, h->stream);
    double ro1 = 0.0;
    int inner = 1;
    for (;; ++inner) {
        launch_ts_face(L, h->stream);
        if (h->viscous) launch_viscous(L, true, false, h->stream);
        launch_ts_cell(L, h->stream);
        launch_gradlim(L, true, 2, h->stream);
        launch_flux(L, false, h->stream);
        if (h->viscous) launch_viscous(L, false, true, h->stream);
        if (h->sur) {
            surrogate_enqueue(h->sur, h->d_case, h->ne, h->stream);
            h->sur_evals += h->ne;
        }
        launch_residual(L, false, h->stream);
        enqueue_be_lusgs(h, dt, inner == 1, bdf2, phi);
        int rc = residual_async(h, h->h_res);
        if (rc) return rc;
        rc = sync_check(h);
        if (ninner) *ninner = inner;
        if (rc) return rc;
        if (res_out) std::memcpy(res_out + (size_t)(inner - 1) * (kNvar + 1), h->h_res, sizeof(double) * (kNvar + 1));
        double ro = 0.0;
        for (int k = kNFlow; k < kNvar; ++k) ro = ro + h->h_res[k];
        if (inner == 1) ro1 = ro;
        if (ro < ro1 * reltol || inner == innerloop) {
            // the closing reconstruction + flux_sourceterm of the converged
            // inner loop (Time_BackwardEuler.f90:36-38): hWA and GradC then
            // hold the residual and gradient of the new state, which the
            // main loop's step_residual reads (fvm2d.f90:137)
            launch_gradlim(L, true, 2, h->stream);
            launch_flux(L, false, h->stream);
            if (h->viscous) launch_viscous(L, false, true, h->stream);
            if (h->sur) {
                surrogate_enqueue(h->sur, h->d_case, h->ne, h->stream);
                h->sur_evals += h->ne;
            }
            launch_residual(L, false, h->stream);
            rc = sync_check(h);
            if (rc) return rc;
            break;
        }
        // the NaN check of backwardEuler_stage (Time_BackwardEuler.f90:44-47)
        double wa11 = 0.0;
        HIPCHK(h, hipMemcpyAsync(&wa11, h->S.wa, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        if (std::isnan(wa11)) return set_err(h, RC_ERR_NAN, "there is NAN in WA(1,1) (inner iteration " +
                                                                std::to_string(inner) + ")");
    }
    h->steps_done++;
    return RC_OK;
}
}  // namespace

extern "C" {

int rc_backward_euler_stage(rc_case *h, double dt, int innerloop, double reltol, int *ninner, double *res_out) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    if (!(dt > 0.0) || innerloop < 1) return set_err(h, RC_ERR_ARGUMENT, "dt must be > 0 and innerloop >= 1");
    HIPCHK(h, hipSetDevice(h->device));
    int rc0 = ensure_be_storage(h);
    if (rc0) return rc0;
    return dual_time_stage(h, dt, innerloop, reltol, ninner, res_out, false, 0.0);
}

int rc_bdf2_begin_step(rc_case *h, int first, double *phi) {
    if (!h || !phi) return set_err(h, RC_ERR_ARGUMENT, "NULL argument");
    HIPCHK(h, hipSetDevice(h->device));
    int rc = ensure_bdf2_storage(h);
    if (rc) return rc;
    rc = bdf2_begin_step(h, first != 0, phi);
    if (rc) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return RC_OK;
}

int rc_bdf2_lusgs(rc_case *h, double dt, int inner, double phi) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    if (!(dt > 0.0) || inner < 1) return set_err(h, RC_ERR_ARGUMENT, "dt must be > 0 and inner >= 1");
    HIPCHK(h, hipSetDevice(h->device));
    if (!h->bdf_wp) return set_err(h, RC_ERR_ARGUMENT, "rc_bdf2_begin_step first");
    int rc = ensure_be_storage(h);
    if (rc) return rc;
    enqueue_be_lusgs(h, dt, inner == 1, true, phi);
    return sync_check(h);
}

int rc_bdf2_stage(rc_case *h, double dt, int innerloop, double reltol, int first, int *ninner, double *res_out) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    if (!(dt > 0.0) || innerloop < 1) return set_err(h, RC_ERR_ARGUMENT, "dt must be > 0 and innerloop >= 1");
    HIPCHK(h, hipSetDevice(h->device));
    int rc = ensure_be_storage(h);
    if (!rc) rc = ensure_bdf2_storage(h);
    if (rc) return rc;
    double phi = 0.0;
    rc = bdf2_begin_step(h, first != 0, &phi);
    if (rc) return rc;
    return dual_time_stage(h, dt, innerloop, reltol, ninner, res_out, true, phi);
}

int rc_residual(rc_case *h, double *res) {
    if (!h || !res) return set_err(h, RC_ERR_ARGUMENT, "NULL argument");
    HIPCHK(h, hipSetDevice(h->device));
    int rc = residual_async(h, h->h_res);
    if (rc) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    std::memcpy(res, h->h_res, sizeof(double) * (kNvar + 1));
    return RC_OK;
}

int rc_get_primitive(rc_case *h, double *pa) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    return get_items(h, pa, h->S.pa, h->ne, kNvar, 0, kRecPA);
}
int rc_get_conserved(rc_case *h, double *wa) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    return get_items(h, wa, h->S.wa, h->ne, kNvar, h->M.sc, 0);
}
int rc_get_increment(rc_case *h, double *hwa) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    return get_items(h, hwa, h->S.hwa, h->ne, kNvar, 0, kRecHW);
}
int rc_get_gradient(rc_case *h, double *g) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    return get_items(h, g, h->S.gradc, h->ne, 2 * kNvar, 0, kRecGR);
}
int rc_get_timestep(rc_case *h, double *step, double *alpaf) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    int rc = get_items(h, step, h->S.step, h->ne, 1, h->M.sc, 0);
    if (rc) return rc;
    if (!alpaf) return RC_OK;
    std::vector<double> rec((size_t)h->nf * kRecF);
    rc = get_items(h, rec.data(), h->S.fflux, h->nf, kRecF, 0, kRecF);
    if (rc) return rc;
    for (int f = 0; f < h->nf; ++f) alpaf[f] = rec[(size_t)f * kRecF + kNvar];
    return RC_OK;
}
int rc_get_ghost(rc_case *h, double *gpa) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    if (h->ng == 0) return RC_OK;
    return get_items(h, gpa, h->S.gpa, h->ng, kNvar, 0, kRecPA);
}

int rc_get_derived(rc_case *h, double *rho, double *mach, double *prod) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    HIPCHK(h, hipSetDevice(h->device));
    const int w = kNsp + 1;
    double *d = nullptr;
    HIPCHK(h, hipMalloc(&d, sizeof(double) * w * std::max(h->ne, 1)));
    launch_derived(h->L, d, h->stream);
    std::vector<double> tmp((size_t)w * h->ne);
    hipError_t e = hipMemcpyAsync(tmp.data(), d, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return set_err(h, RC_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e));
    int rc = check_device_error(h);
    if (rc) return rc;
    for (int i = 0; i < h->ne; ++i) {
        if (rho) rho[i] = tmp[(size_t)i * w];
        if (mach) mach[i] = tmp[(size_t)i * w + 1];
        if (prod)
            for (int k = 0; k < kNsp - 1; ++k) prod[(size_t)i * (kNsp - 1) + k] = tmp[(size_t)i * w + 2 + k];
    }
    return RC_OK;
}

}  // extern "C"

// The derived fields without any allocation (no reference counterpart): the
// same kernel as rc_get_derived stages its output in `scratch`, a
// file-scope device array of kNoallocCells x (Nsp+1) doubles (21 MB) that
// every GPU the library runs on holds until the process ends.  Host code
// reaches a device's copy only through the address hipGetSymbolAddress
// returns with that device current, taken at the entry point's first call on
// the device and kept per device; the array's name itself is no device
// address.  Calls of one process are serialised by a lock (they share each
// device's array); every element a call copies out is written by that
// call's kernel first, so nothing needs clearing.
constexpr int kNoallocCells = 1 << 19;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[(size_t)kNoallocCells * (kNsp + 1)];
namespace {
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// the device's copy of scratch (the lock held, the device current)
int scratch_address(rc_case *h, double **out) {
    if (!g_scratch_dev[h->device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(h, RC_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[h->device] = static_cast<double *>(p);
    }
    *out = g_scratch_dev[h->device];
    return RC_OK;
}
}  // namespace

extern "C" {

int rc_get_derived_noalloc(rc_case *h, double *rho, double *mach, double *prod) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    if (h->ne > kNoallocCells)
        return set_err(h, RC_ERR_ARGUMENT, "rc_get_derived_noalloc: the case has " + std::to_string(h->ne) +
                                               " cells, the static staging array holds " +
                                               std::to_string(kNoallocCells) + " (use rc_get_derived)");
    if (h->device < 0 || h->device >= kMaxDevices) return set_err(h, RC_ERR_ARGUMENT, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    HIPCHK(h, hipSetDevice(h->device));
    double *d = nullptr;
    int rc = scratch_address(h, &d);
    if (rc) return rc;
    launch_derived(h->L, d, h->stream);
    // strided copies straight into the caller's arrays (no host staging)
    const size_t w = kNsp + 1, n = (size_t)h->ne;
    if (n > 0) {
        if (rho)
            HIPCHK(h, hipMemcpy2DAsync(rho, sizeof(double), d, w * sizeof(double), sizeof(double), n,
                                       hipMemcpyDeviceToHost, h->stream));
        if (mach)
            HIPCHK(h, hipMemcpy2DAsync(mach, sizeof(double), d + 1, w * sizeof(double), sizeof(double), n,
                                       hipMemcpyDeviceToHost, h->stream));
        if (prod)
            HIPCHK(h, hipMemcpy2DAsync(prod, (kNsp - 1) * sizeof(double), d + 2, w * sizeof(double),
                                       (kNsp - 1) * sizeof(double), n, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return check_device_error(h);
}

int rc_set_timing(rc_case *h, int enable) {
    if (!h) return set_err(h, RC_ERR_ARGUMENT, "NULL handle");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->timing = enable != 0;
    h->ev_used = 0;
    for (double &t : h->tms) t = 0.0;
    h->timed_steps = 0;
    return RC_OK;
}

int rc_get_timing(rc_case *h, double *t8) {
    if (!h || !t8) return set_err(h, RC_ERR_ARGUMENT, "NULL argument");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    int rc = drain_timing(h);
    if (rc) return rc;
    for (int k = 0; k < kTimeSlots; ++k) t8[k] = h->tms[k];
    t8[kTimeSlots] = (double)h->timed_steps;
    return RC_OK;
}

int rc_mesh_order_multicolor(int ne, const int *cellnodes, int block, int *perm, int *ncolors) {
    if (ne <= 0 || !cellnodes || !perm || block < 1) return set_err(nullptr, RC_ERR_ARGUMENT, "bad argument");
    std::vector<int> c0(3 * (size_t)ne);
    for (size_t k = 0; k < c0.size(); ++k) c0[k] = cellnodes[k] - 1;
    int rc = order_multicolor(ne, c0.data(), block, perm, ncolors);
    if (rc) return set_err(nullptr, rc, "mesh is not a manifold triangle mesh");
    return RC_OK;
}

int rc_batch_create(int device, rc_batch **out) {
    if (!out) return set_err(nullptr, RC_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    rc_batch *b = new rc_batch();
    b->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return set_err(nullptr, RC_ERR_DEVICE, std::string("batch stream: ") + hipGetErrorString(e));
    }
    const char *ng = std::getenv("RC_NO_GRAPH");
    b->use_graph = !(ng && ng[0] == '1');
    b->persist = persist_env();
    *out = b;
    return RC_OK;
}

void rc_batch_destroy(rc_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    for (rc_case *h : b->members) h->batch = nullptr;
    batch_release(b);
    if (b->d_bar) (void)hipFree(b->d_bar);
    if (b->h_res) (void)hipHostFree(b->h_res);
    if (b->h_err) (void)hipHostFree(b->h_err);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int rc_batch_add(rc_batch *b, rc_case *h) {
    if (!b || !h) return set_err(nullptr, RC_ERR_ARGUMENT, "NULL argument");
    if (h->batch) return bset_err(b, RC_ERR_ARGUMENT, "the case already belongs to a batch");
    if (h->sur) return bset_err(b, RC_ERR_ARGUMENT, "a case with the learned source term steps alone");
    if (h->device != b->device) return bset_err(b, RC_ERR_ARGUMENT, "the case is on another device");
    BHIPCHK(b, hipSetDevice(b->device));
    BHIPCHK(b, hipStreamSynchronize(h->stream));
    b->members.push_back(h);
    h->batch = b;
    b->dirty = true;
    return RC_OK;
}

int rc_batch_remove(rc_batch *b, rc_case *h) {
    if (!b || !h) return set_err(nullptr, RC_ERR_ARGUMENT, "NULL argument");
    auto it = std::find(b->members.begin(), b->members.end(), h);
    if (it == b->members.end()) return bset_err(b, RC_ERR_ARGUMENT, "the case is not in the batch");
    BHIPCHK(b, hipSetDevice(b->device));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    b->members.erase(it);
    h->batch = nullptr;
    b->dirty = true;
    return RC_OK;
}

int rc_batch_info(const rc_batch *b, long long *info3) {
    if (!b || !info3) return set_err(nullptr, RC_ERR_ARGUMENT, "NULL argument");
    long long cells = 0;
    int nlev = 0;
    for (const rc_case *h : b->members) {
        cells += h->ne;
        nlev = std::max(nlev, h->nlev);
    }
    info3[0] = (long long)b->members.size();
    info3[1] = cells;
    info3[2] = nlev;
    return RC_OK;
}

int rc_batch_set_launch(rc_batch *b, int cell_threads, int run_threads, int xcd) {
    if (!b) return set_err(nullptr, RC_ERR_ARGUMENT, "NULL batch");
    auto ok = [](int t) { return t == 64 || t == 128 || t == 256; };
    if (!ok(cell_threads) || !ok(run_threads) || xcd < -1 || xcd > 1)
        return bset_err(b, RC_ERR_ARGUMENT, "threads per block must be 64, 128 or 256 and xcd -1, 0 or 1");
    if (b->cell_bs == cell_threads && b->run_bs == run_threads && b->xcd == xcd) return RC_OK;
    b->cell_bs = cell_threads;
    b->run_bs = run_threads;
    b->xcd = xcd;
    b->dirty = true;
    return RC_OK;
}

int rc_batch_run_steady(rc_batch *b, int nsteps, int every, double *res_out, int *done, int *err_case) {
    if (done) *done = 0;
    if (err_case) *err_case = -1;
    if (!b || nsteps < 0) return set_err(nullptr, RC_ERR_ARGUMENT, "bad argument");
    if (b->members.empty()) return RC_OK;
    BHIPCHK(b, hipSetDevice(b->device));
    int rc = batch_build(b);
    if (rc) return rc;
    const size_t n = b->members.size();
    if (every <= 0) every = nsteps + 1;
    int nrep = 0;
    // reports (residual rows, device error checks): at step 1 when rows are
    // asked for, and every `every` steps; between them whole multi-step
    // graphs first, then single steps
    const bool first = res_out != nullptr;
    bool visc = false;
    for (const rc_case *m : b->members) visc = visc || m->viscous;
    const bool persist = b->persist > 0 && !visc && persist_supported(b->L);
    const int pgrid = persist ? persist_grid(b->L, b->device) : 0;
    if (persist && pgrid <= 0) return bset_err(b, RC_ERR_DEVICE, "persistent step: no resident blocks");
    if (persist && !b->d_bar) BHIPCHK(b, hipMalloc(&b->d_bar, 2 * sizeof(unsigned)));
    for (int s = 0; s < nsteps;) {
        const int next = std::min(nsteps, s == 0 && first ? 1 : (s / every + 1) * every);
        for (int k = next - s; k > 0;) {
            int run = 1;
            if (persist) {
                run = std::min(k, b->persist);
                BHIPCHK(b, launch_step_persist(b->L, b->d_bar, run, pgrid, b->stream));
            } else if (b->use_graph) {
                if ((rc = batch_graphs(b))) return rc;
                // the largest graph that fits the rest (top, then its binary digits)
                for (int p = b->graph_pow_top; p >= 2 && run == 1; p /= 2)
                    if (k >= p) run = p;
                BHIPCHK(b, hipGraphLaunch(run == 1 ? b->graph_exec : b->graph_pow[pow_index(run)], b->stream));
            } else {
                enqueue_batch_step(b);
            }
            for (rc_case *h : b->members) h->steps_done += run;
            k -= run;
        }
        s = next;
        if ((s == 1 && first) || s % every == 0) {
            if (res_out) {
                launch_resnorm(b->L, b->stream);
                BHIPCHK(b, hipMemcpyAsync(b->h_res, b->L.resout, sizeof(double) * n * (kNvar + 1),
                                          hipMemcpyDeviceToHost, b->stream));
            }
            if ((rc = batch_check(b, err_case))) {
                if (done) *done = s;
                return rc;
            }
            if (res_out)
                std::memcpy(res_out + (size_t)nrep * n * (kNvar + 1), b->h_res, sizeof(double) * n * (kNvar + 1));
            ++nrep;
        }
    }
    rc = batch_check(b, err_case);
    if (done) *done = nsteps;
    return rc;
}

}  // extern "C"
