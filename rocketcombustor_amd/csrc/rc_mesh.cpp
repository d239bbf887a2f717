// Host-side mesh set-up of the library: faces, boundary tags, cell volumes
// and centroids, reconstruction stencils, mirrored ghost cells and the
// least-squares matrices, computed with the reference's numbering
// conventions so that every per-cell sum runs in the reference's order.
//
//   faces in order of first appearance over cells     src/mesh.f90:120-199
//   boundary tags Neighbor(2,f) = -type                src/mesh.f90:76-93
//   13-point triangle rule for vol / centroid          src/gaussInteg4.f90:4-50, mesh.f90:206-255
//   NeighF ascending, NeighC, Neigh                    src/reconstruction_initialization.f90:11-50
//   mirrored ghost cells                               src/ghostcell.f90:19-195
//   wdis = 1/|dc|, A = wdis*(<x_j>-<x_i>)/reflen       src/reconstruction_initialization.f90:93-150
//   pinv(A) (dgesvd in the reference, la_math.f90:1-62; here a thin QR)
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <queue>
#include <tuple>
#include <unordered_map>

#include "rc_internal.h"

namespace rc {
namespace {

// real**integer as compiler-rt's __powidf2 evaluates it
double powi(double a, int b) {
    const bool recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

// integrand selector of the quadrature calls: 1, x, y, or (x-cx)/rx, (y-cy)/ry
struct Integrand {
    int kind;
    double c, r;
    double operator()(const double *v) const {
        switch (kind) {
        case 0: return 1.0;
        case 1: return v[0];
        case 2: return v[1];
        case 3: return (v[0] - c) / r;
        default: return (v[1] - c) / r;
        }
    }
};

// 13-point, 7th-order triangle rule (weights and points of gaussInteg4.f90:14-25)
double tri_rule(const Integrand &f, const double *p1, const double *p2, const double *p3) {
    const double x1 = p1[0], y1 = p1[1], x2 = p2[0], y2 = p2[1], x3 = p3[0], y3 = p3[1];
    const double area = 0.5 * std::fabs((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1));
    double q[13][2];
    for (int d = 0; d < 2; ++d) {
        q[0][d] = (p1[d] + p2[d] + p3[d]) / 3.0;
        double a = 0.479308067841920, b = 0.260345966079040;
        q[1][d] = a * p1[d] + b * (p2[d] + p3[d]);
        q[2][d] = a * p2[d] + b * (p1[d] + p3[d]);
        q[3][d] = a * p3[d] + b * (p1[d] + p2[d]);
        a = 0.869739794195568; b = 0.065130102902216;
        q[4][d] = a * p1[d] + b * (p2[d] + p3[d]);
        q[5][d] = a * p2[d] + b * (p1[d] + p3[d]);
        q[6][d] = a * p3[d] + b * (p1[d] + p2[d]);
        a = 0.048690315425316; b = 0.312865496004874;
        const double w = 1 - a - b;
        q[7][d] = a * p1[d] + b * p2[d] + w * p3[d];
        q[8][d] = a * p1[d] + b * p3[d] + w * p2[d];
        q[9][d] = a * p2[d] + b * p1[d] + w * p3[d];
        q[10][d] = a * p2[d] + b * p3[d] + w * p1[d];
        q[11][d] = a * p3[d] + b * p1[d] + w * p2[d];
        q[12][d] = a * p3[d] + b * p2[d] + w * p1[d];
    }
    const double s = -0.149570044467682 * f(q[0])
        + 0.175615257433208 * (f(q[1]) + f(q[2]) + f(q[3]))
        + 0.053347235608838 * (f(q[4]) + f(q[5]) + f(q[6]))
        + 0.077113760890257 * (f(q[7]) + f(q[8]) + f(q[9]) + f(q[10]) + f(q[11]) + f(q[12]));
    return s * area;
}

// quadrilateral (v1,v2,v3,v4) as two triangles; a triangle repeats v3
double quad_rule(const Integrand &f, const double *a, const double *b, const double *c, const double *d) {
    return tri_rule(f, a, b, c) + tri_rule(f, a, c, d);
}

// pinv of a 3x2 matrix through a thin QR (Gram-Schmidt, one
// re-orthogonalisation pass): pinv = R^-1 Q^T.
void pinv32(const double a[3][2], double b[2][3]) {
    double q0[3], v[3];
    const double n0 = std::sqrt(a[0][0] * a[0][0] + a[1][0] * a[1][0] + a[2][0] * a[2][0]);
    for (int i = 0; i < 3; ++i) q0[i] = a[i][0] / n0;
    for (int i = 0; i < 3; ++i) v[i] = a[i][1];
    double r01 = q0[0] * v[0] + q0[1] * v[1] + q0[2] * v[2];
    for (int i = 0; i < 3; ++i) v[i] = v[i] - r01 * q0[i];
    const double s = q0[0] * v[0] + q0[1] * v[1] + q0[2] * v[2];
    for (int i = 0; i < 3; ++i) v[i] = v[i] - s * q0[i];
    r01 = r01 + s;
    const double n1 = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int i = 0; i < 3; ++i) {
        const double y1 = (v[i] / n1) / n1;
        const double y0 = (q0[i] - r01 * y1) / n0;
        b[0][i] = y0;
        b[1][i] = y1;
    }
}

inline uint64_t edge_key(int a, int b) {
    const uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
    return ((uint64_t)hi << 32) | lo;
}

int fail(std::string &msg, const char *fmt, int a = 0, int b = 0) {
    char buf[256];
    std::snprintf(buf, sizeof buf, fmt, a, b);
    msg = buf;
    return RC_ERR_MESH;
}

}  // namespace

int build_mesh(int nnt, const double *coor, int ne, const int *cells0, int nb, const int *bedges0,
               const int *btype, int ncon, const int *con, const double *dxyz, HostMesh &m, std::string &msg) {
    if (nnt <= 0 || ne <= 0 || nb < 0) return fail(msg, "empty mesh");
    m.nnt = nnt;
    m.ne = ne;
    m.coor.assign(coor, coor + 2 * (size_t)nnt);
    m.cn.assign(cells0, cells0 + 3 * (size_t)ne);
    for (size_t k = 0; k < m.cn.size(); ++k)
        if (m.cn[k] < 0 || m.cn[k] >= nnt) return fail(msg, "cell node %d out of range (cell %d)", m.cn[k], (int)(k / 3));
    const double *X = m.coor.data();

    // ---- faces in order of first appearance (mesh.f90:120-199)
    std::unordered_map<uint64_t, int> fidx;
    fidx.reserve((size_t)ne * 2);
    m.fnode.clear(); m.fnl.clear(); m.fnr.clear(); m.vecx.clear();
    for (int j = 0; j < ne; ++j) {
        for (int kf = 0; kf < 3; ++kf) {
            const int n1 = m.cn[3 * j + kf], n2 = m.cn[3 * j + (kf + 1) % 3];
            if (n1 == n2) return fail(msg, "degenerate cell %d", j);
            auto it = fidx.find(edge_key(n1, n2));
            if (it == fidx.end()) {
                const int f = (int)m.fnl.size();
                fidx.emplace(edge_key(n1, n2), f);
                m.fnl.push_back(j);
                m.fnr.push_back(-100);
                m.fnode.push_back(n1);
                m.fnode.push_back(n2);
                m.vecx.push_back(X[2 * n2 + 1] - X[2 * n1 + 1]);
                m.vecx.push_back(-(X[2 * n2] - X[2 * n1]));
            } else {
                if (m.fnr[it->second] != -100) return fail(msg, "face shared by more than two cells (cell %d)", j);
                m.fnr[it->second] = j;
            }
        }
    }
    m.nf = (int)m.fnl.size();

    // ---- periodic connections (CreatePeriodConnect, period.f90:15-104,
    // after the face creation as mesh.f90:69-73 calls it).  Connection c
    // takes every face of boundary con[2c], in its given order, to the first
    // face of boundary con[2c+1] whose midpoint is the source midpoint moved
    // by dxyz: the midpoints and their difference in single precision, the
    // test |ddd - dxyz|_1 < 1e-7 in double (period.f90:19,76-87).  The face's
    // neighbour becomes the owner of that face, FProperty 100, SideOfs dxyz.
    auto labelled = [&](int t) {
        for (int c = 0; c < ncon; ++c)
            if (con[2 * c] == t || con[2 * c + 1] == t) return true;
        return false;
    };
    m.sideofs.clear();
    m.periodic.clear();
    if (ncon > 0) {
        m.sideofs.assign(2 * (size_t)m.nf, 0.0);
        m.periodic.assign(m.nf, 0);
        auto bface = [&](int b) {
            auto it = fidx.find(edge_key(bedges0[2 * b], bedges0[2 * b + 1]));
            return it == fidx.end() ? -1 : it->second;
        };
        auto mid = [&](int b, int d) {
            return (float)((X[2 * bedges0[2 * b] + d] + X[2 * bedges0[2 * b + 1] + d]) / 2.);
        };
        for (int c = 0; c < ncon; ++c) {
            if (con[2 * c] < 100 || con[2 * c + 1] < 100)
                return fail(msg, "periodic connection %d names a boundary label below 100", c + 1);
            for (int b = 0; b < nb; ++b) {
                if (btype[b] != con[2 * c]) continue;
                const int fs = bface(b);
                if (fs < 0) return fail(msg, "not find boundary face (boundary face %d)", b + 1);
                const float xs0 = mid(b, 0), xs1 = mid(b, 1);
                int hit = -1;
                for (int b2 = 0; b2 < nb && hit < 0; ++b2) {
                    if (btype[b2] != con[2 * c + 1]) continue;
                    const float d0 = mid(b2, 0) - xs0, d1 = mid(b2, 1) - xs1;
                    if (std::fabs((double)d0 - dxyz[2 * c]) + std::fabs((double)d1 - dxyz[2 * c + 1]) < (double)1.e-7f)
                        hit = b2;
                }
                if (hit < 0) return fail(msg, "not find target (periodic connection %d, boundary face %d)", c + 1, b + 1);
                const int ft = bface(hit);
                if (ft < 0) return fail(msg, "not find boundary face (boundary face %d)", hit + 1);
                m.fnr[fs] = m.fnl[ft];
                m.periodic[fs] = 1;
                m.sideofs[2 * (size_t)fs] = dxyz[2 * c];
                m.sideofs[2 * (size_t)fs + 1] = dxyz[2 * c + 1];
            }
        }
    }
    auto ofs = [&](int f, int d) { return m.sideofs.empty() ? 0.0 : m.sideofs[2 * (size_t)f + d]; };
    auto is_periodic = [&](int f) { return !m.periodic.empty() && m.periodic[f] != 0; };

    // ---- boundary tags (mesh.f90:76-93); the boundaries a periodic
    // connection names take type 100 (mesh.f90:62-63)
    std::unordered_map<uint64_t, int> btag;
    btag.reserve((size_t)nb * 2);
    for (int b = 0; b < nb; ++b) {
        int t = btype[b];
        if (t >= 100 && labelled(t)) t = 100;
        // flux_lax knows types 1..6 and stops on any other face
        // (flux_lax.f90:232-235, "no face of this type"), every step before
        // flux_viscous, so flux_viscous's type 7 branch (:170-175) never runs
        else if (t < 1 || t > 6)
            return fail(msg, "unsupported boundary type %d (flux_lax handles types 1-6 and stops on others)", t);
        btag.emplace(edge_key(bedges0[2 * b], bedges0[2 * b + 1]), t);
    }
    for (int f = 0; f < m.nf; ++f) {
        if (m.fnr[f] < 0) {
            auto it = btag.find(edge_key(m.fnode[2 * f], m.fnode[2 * f + 1]));
            if (it == btag.end()) return fail(msg, "not find boundary of this type (face %d)", f + 1);
            // a face of a periodic boundary that no connection paired: type
            // 100, on which flux_lax stops
            if (it->second == 100)
                return fail(msg, "boundary face %d of a periodic boundary has no partner (flux_lax stops on type 100)",
                            f + 1);
            m.fnr[f] = -it->second;
        }
    }

    // ---- volumes, centroids, reflen (mesh.f90:206-255)
    m.vol.assign(ne, 0.0);
    m.cellxy.assign(2 * (size_t)ne, 0.0);
    m.reflen.assign(2 * (size_t)ne, 0.0);
    const Integrand one{0, 0, 0}, fx{1, 0, 0}, fy{2, 0, 0};
    for (int i = 0; i < ne; ++i) {
        const double *p1 = X + 2 * m.cn[3 * i], *p2 = X + 2 * m.cn[3 * i + 1], *p3 = X + 2 * m.cn[3 * i + 2];
        const double v = quad_rule(one, p1, p2, p3, p3);
        m.vol[i] = v;
        m.cellxy[2 * i] = quad_rule(fx, p1, p2, p3, p3) / v;
        m.cellxy[2 * i + 1] = quad_rule(fy, p1, p2, p3, p3) / v;
        if (v < 0.0) return fail(msg, "negative volume %d", i + 1);
        // a clockwise triangle gives owner-inward VECX; the reference assumes CCW
        const double cr = (p2[0] - p1[0]) * (p3[1] - p1[1]) - (p3[0] - p1[0]) * (p2[1] - p1[1]);
        if (!(cr > 0.0)) return fail(msg, "cell %d is not counter-clockwise", i + 1);
        for (int a = 0; a < 2; ++a)
            for (int b = a + 1; b < 3; ++b)
                for (int d = 0; d < 2; ++d)
                    m.reflen[2 * i + d] = std::max(0.5 * std::fabs(X[2 * m.cn[3 * i + a] + d] - X[2 * m.cn[3 * i + b] + d]),
                                                   m.reflen[2 * i + d]);
    }
    double sv = 0.0;
    for (int i = 0; i < ne; ++i) sv = sv + m.vol[i];
    m.total_volume = sv;

    // ---- NeighF / NeighC (reconstruction_initialization.f90:11-50)
    std::vector<int> cnt(ne, 0);
    m.neighf.assign(3 * (size_t)ne, -1);
    for (int f = 0; f < m.nf; ++f) {
        const int nl = m.fnl[f], nr = m.fnr[f];
        if (cnt[nl] >= 3) return fail(msg, "cell %d has more than three faces", nl + 1);
        m.neighf[3 * nl + cnt[nl]++] = f;
        // a periodic face is in its owner's list only (:20)
        if (nr >= 0 && !is_periodic(f)) {
            if (cnt[nr] >= 3) return fail(msg, "cell %d has more than three faces", nr + 1);
            m.neighf[3 * nr + cnt[nr]++] = f;
        }
    }
    m.neigh.assign(ne, 0);
    m.neighc.assign(3 * (size_t)ne, -1);
    m.nofs.assign(6 * (size_t)ne, 0.0);
    m.cface.assign(3 * (size_t)ne, 0);
    m.lu_nb.assign(3 * (size_t)ne, -1);
    m.lu_n.assign(6 * (size_t)ne, 0.0);
    for (int i = 0; i < ne; ++i) {
        if (cnt[i] != 3) return fail(msg, "cell %d has %d faces", i + 1, cnt[i]);
        for (int k = 0; k < 3; ++k) {
            const int f = m.neighf[3 * i + k];
            const int sides[2] = {m.fnl[f], m.fnr[f]};
            for (int s = 0; s < 2; ++s) {
                const int ic = sides[s];
                if (ic != i && ic >= 0) {
                    for (int d = 0; d < 2; ++d) m.nofs[6 * (size_t)i + 2 * m.neigh[i] + d] = ofs(f, d);   // :40
                    m.neighc[3 * i + m.neigh[i]++] = ic;
                }
            }
            m.cface[3 * i + k] = (m.fnl[f] == i) ? f : ~f;
            // LU-SGS face slot (Time_LUSGS_steady.f90:60-68): the other cell
            // and the normal VECX, negated when the cell is not the owner
            int nr = m.fnr[f];
            double dx = m.vecx[2 * f], dy = m.vecx[2 * f + 1];
            if (nr == i) { nr = m.fnl[f]; dx = -dx; dy = -dy; }
            m.lu_nb[3 * i + k] = nr >= 0 ? nr : -1;
            m.lu_n[6 * (size_t)i + 2 * k] = dx;
            m.lu_n[6 * (size_t)i + 2 * k + 1] = dy;
        }
    }

    // ---- ghost cells (ghostcell.f90:19-195)
    int ng = 0;
    for (int i = 0; i < ne; ++i) ng += 3 - m.neigh[i];
    m.ng = ng;
    m.ghost_host.assign(ng, 0);
    m.ghost_face.assign(ng, 0);
    m.ghost_type.assign(ng, 0);
    m.ghost_nodes.assign(8 * (size_t)ng, 0.0);
    m.ghost_vol.assign(ng, 0.0);
    m.ghost_xy.assign(2 * (size_t)ng, 0.0);
    m.gnrm.assign(2 * (size_t)ng, 0.0);
    m.num_ghost.assign(ne, 0);
    m.ghostcell.assign(3 * (size_t)ne, -1);
    int j = 0;
    for (int i = 0; i < ne; ++i) {
        if (m.neigh[i] < 3) {
            for (int kf = 0; kf < 3; ++kf) {
                const int f = m.neighf[3 * i + kf];
                if (m.fnr[f] >= 0) continue;
                m.ghost_host[j] = i;
                m.ghost_face[j] = f;
                m.ghost_type[j] = -m.fnr[f];
                m.ghostcell[3 * i + m.num_ghost[i]++] = j;
                const int n1 = m.fnode[2 * f], n2 = m.fnode[2 * f + 1];
                int k1 = 0;
                while (k1 < 3 && m.cn[3 * i + k1] != n1) ++k1;
                const int k2 = (k1 + 1) % 3;
                if (k1 == 3 || m.cn[3 * i + k2] != n2) return fail(msg, "ghost cell node order (cell %d)", i + 1);
                const int k3 = (k2 + 1) % 3, k4 = (k3 + 1) % 3;
                const int n3 = m.cn[3 * i + k3], n4 = m.cn[3 * i + k4];
                const double sav1 = m.vecx[2 * f], sav2 = m.vecx[2 * f + 1];
                const double sav = std::max(std::sqrt(sav1 * sav1 + sav2 * sav2), kSmall);
                const double s1n = sav1 / sav, s2n = sav2 / sav;
                double *g = &m.ghost_nodes[8 * (size_t)j];
                g[0] = X[2 * n2]; g[1] = X[2 * n2 + 1];
                g[2] = X[2 * n1]; g[3] = X[2 * n1 + 1];
                const double d1 = (X[2 * n1] - X[2 * n4]) * s1n + (X[2 * n1 + 1] - X[2 * n4 + 1]) * s2n;
                if (d1 < 0) return fail(msg, "ghost distance (n1,n4) negative (cell %d)", i + 1);
                g[4] = X[2 * n4] + 2 * d1 * s1n;
                g[5] = X[2 * n4 + 1] + 2 * d1 * s2n;
                const double d2 = (X[2 * n2] - X[2 * n3]) * s1n + (X[2 * n2 + 1] - X[2 * n3 + 1]) * s2n;
                if (d2 < 0) return fail(msg, "ghost distance (n2,n3) negative (cell %d)", i + 1);
                g[6] = X[2 * n3] + 2 * d2 * s1n;
                g[7] = X[2 * n3 + 1] + 2 * d2 * s2n;
                const double gv = quad_rule(one, g, g + 2, g + 4, g + 6);
                m.ghost_vol[j] = gv;
                m.ghost_xy[2 * j] = quad_rule(fx, g, g + 2, g + 4, g + 6) / gv;
                m.ghost_xy[2 * j + 1] = quad_rule(fy, g, g + 2, g + 4, g + 6) / gv;
                if (gv < 0.0) return fail(msg, "negative ghost volume %d", j + 1);
                // unit normal as update_ghostcell_variables forms it from the
                // first two ghost nodes (update_flow_variables.f90:114-121)
                const double gs1 = g[3] - g[1];
                const double gs2 = -(g[2] - g[0]);
                const double gsav = std::max(std::sqrt(gs1 * gs1 + gs2 * gs2), kSmall);
                m.gnrm[2 * j] = gs1 / gsav;
                m.gnrm[2 * j + 1] = gs2 / gsav;
                ++j;
            }
        }
        if (m.num_ghost[i] + m.neigh[i] != 3) return fail(msg, "error in number of ghost cells for cell %d", i + 1);
    }
    // a cell whose neighbours repeat (two faces, one of them periodic, to the
    // same cell) is refused: the sweeps take one term per neighbour and step
    for (int i = 0; i < ne; ++i)
        for (int a = 0; a < m.neigh[i]; ++a)
            for (int b = a + 1; b < m.neigh[i]; ++b)
                if (m.neighc[3 * i + a] == m.neighc[3 * i + b])
                    return fail(msg, "cell %d has cell %d as a neighbour across two faces", i + 1, m.neighc[3 * i + a] + 1);

    // ---- the faces of each cell's TimeStep sum (TimeStep.f90:16-96): every
    // face adds its alpha to NL and, when NR > 0, to NR, periodic faces too
    // (no FProperty test there); a cell's sum runs in face order
    m.tsx.clear();
    if (!m.periodic.empty()) {
        m.tsx.assign((size_t)kTsFaces * ne, -1);
        std::vector<int> nts(ne, 0);
        for (int f = 0; f < m.nf; ++f) {
            const int sides[2] = {m.fnl[f], m.fnr[f]};
            for (int s = 0; s < 2; ++s) {
                const int c = sides[s];
                if (c < 0) continue;
                if (nts[c] >= kTsFaces) return fail(msg, "cell %d takes more than %d TimeStep faces", c + 1, kTsFaces);
                m.tsx[(size_t)kTsFaces * c + nts[c]++] = f;
            }
        }
    }

    // ---- LSQ matrices (reconstruction_initialization.f90:71-150)
    m.xyh.assign(2 * (size_t)ne, 0.0);
    m.wdis.assign(3 * (size_t)ne, 1.0);
    m.amls.assign(6 * (size_t)ne, 0.0);
    m.stencil.assign(3 * (size_t)ne, 0);
    for (int i = 0; i < ne; ++i) {
        const double *p1 = X + 2 * m.cn[3 * i], *p2 = X + 2 * m.cn[3 * i + 1], *p3 = X + 2 * m.cn[3 * i + 2];
        for (int io = 0; io < 2; ++io) {
            const Integrand fs{3 + io, m.cellxy[2 * i + io], m.reflen[2 * i + io]};
            m.xyh[2 * i + io] = quad_rule(fs, p1, p2, p3, p3) / m.vol[i];
        }
        if (std::fabs(m.xyh[2 * i]) + std::fabs(m.xyh[2 * i + 1]) > 1.e-8) return fail(msg, "error in XYH, cell %d", i + 1);
    }
    for (int i = 0; i < ne; ++i) {
        double a1[3][2], b1[2][3];
        const int nn = m.neigh[i];
        for (int k = 0; k < nn; ++k) {
            const int jc = m.neighc[3 * i + k];
            m.stencil[3 * i + k] = jc;
            // the neighbour's nodes moved by its offset Ofs (:101-109)
            const double o0 = m.nofs[6 * (size_t)i + 2 * k], o1 = m.nofs[6 * (size_t)i + 2 * k + 1];
            double q[3][2];
            for (int a = 0; a < 3; ++a) {
                q[a][0] = X[2 * m.cn[3 * jc + a]] - o0;
                q[a][1] = X[2 * m.cn[3 * jc + a] + 1] - o1;
            }
            const double *q1 = q[0], *q2 = q[1], *q3 = q[2];
            const double dx = (m.cellxy[2 * jc] - o0) - m.cellxy[2 * i];
            const double dy = (m.cellxy[2 * jc + 1] - o1) - m.cellxy[2 * i + 1];
            const double w = powi(std::sqrt(dx * dx + dy * dy), -1);
            m.wdis[3 * i + k] = w;
            for (int io = 0; io < 2; ++io) {
                const Integrand fs{3 + io, m.cellxy[2 * i + io], m.reflen[2 * i + io]};
                a1[k][io] = w * (quad_rule(fs, q1, q2, q3, q3) / m.vol[jc] - m.xyh[2 * i + io]);
            }
        }
        for (int k = 0; k < m.num_ghost[i]; ++k) {
            const int g = m.ghostcell[3 * i + k];
            m.stencil[3 * i + nn + k] = -(g + 1);
            const double *gn = &m.ghost_nodes[8 * (size_t)g];
            const double dx = m.ghost_xy[2 * g] - m.cellxy[2 * i];
            const double dy = m.ghost_xy[2 * g + 1] - m.cellxy[2 * i + 1];
            const double w = powi(std::sqrt(dx * dx + dy * dy), -1);
            m.wdis[3 * i + nn + k] = w;
            for (int io = 0; io < 2; ++io) {
                const Integrand fs{3 + io, m.cellxy[2 * i + io], m.reflen[2 * i + io]};
                a1[nn + k][io] = w * (quad_rule(fs, gn, gn + 2, gn + 4, gn + 6) / m.ghost_vol[g] - m.xyh[2 * i + io]);
            }
        }
        pinv32(a1, b1);
        for (int k = 0; k < 3; ++k) {
            m.amls[6 * (size_t)i + 2 * k] = b1[0][k];
            m.amls[6 * (size_t)i + 2 * k + 1] = b1[1][k];
        }
    }

    // ---- derived kernel operands
    m.lvv.assign(6 * (size_t)ne, 0.0);
    for (int i = 0; i < ne; ++i) {
        for (int kf = 0; kf < 3; ++kf) {
            const int f = m.neighf[3 * i + kf];
            const int n1 = m.fnode[2 * f], n2 = m.fnode[2 * f + 1];
            for (int d = 0; d < 2; ++d) {
                const double xy = 0.5 * (X[2 * n1 + d] + X[2 * n2 + d]);       // limiter.f90:63
                m.lvv[6 * (size_t)i + 2 * kf + d] = (xy - m.cellxy[2 * i + d]) / m.reflen[2 * i + d];
            }
        }
    }
    m.fvl.assign(2 * (size_t)m.nf, 0.0);
    m.fvr.assign(2 * (size_t)m.nf, 0.0);
    for (int f = 0; f < m.nf; ++f) {
        const int n1 = m.fnode[2 * f], n2 = m.fnode[2 * f + 1];
        const int nl = m.fnl[f], nr = m.fnr[f];
        for (int d = 0; d < 2; ++d) {
            const double xy = X[2 * n1 + d] + 0.5 * (X[2 * n2 + d] - X[2 * n1 + d]);  // flux_lax.f90:34
            m.fvl[2 * f + d] = (xy - m.cellxy[2 * nl + d]) / m.reflen[2 * nl + d];
            if (nr >= 0) m.fvr[2 * f + d] = (xy - (m.cellxy[2 * nr + d] - ofs(f, d))) / m.reflen[2 * nr + d];  // :86
        }
    }
    // viscous face operands (flux_viscous.f90:187-209, fbaseGrad_dxdy
    // module.f90:124-137): 1/reflen of the two cells and the centroid
    // difference NL - NR
    m.vgeo.assign(6 * (size_t)m.nf, 0.0);
    for (int f = 0; f < m.nf; ++f) {
        const int nl = m.fnl[f], nr = m.fnr[f];
        double *g = &m.vgeo[6 * (size_t)f];
        g[0] = 1. / m.reflen[2 * nl];
        g[1] = 1. / m.reflen[2 * nl + 1];
        if (nr >= 0) {
            g[2] = 1. / m.reflen[2 * nr];
            g[3] = 1. / m.reflen[2 * nr + 1];
            g[4] = m.cellxy[2 * nl] - (m.cellxy[2 * nr] - ofs(f, 0));        // flux_viscous.f90:212-213
            g[5] = m.cellxy[2 * nl + 1] - (m.cellxy[2 * nr + 1] - ofs(f, 1));
        }
    }

    // ---- LU-SGS schedule of the given order.  The order is cut into runs:
    // a cell joins the run of the cell before it when it shares a face with a
    // cell of that run and the run holds fewer than kRunMax cells.  A run's
    // level is 1 + the highest level of the runs holding its cells' lower
    // neighbours outside it (upper neighbours for the backward sweep are in
    // higher levels by the same rule).  One thread sweeps one run cell by cell
    // in the given order, and the runs of one level are independent, so a
    // level-by-level sweep is the reference's serial sweep.
    std::vector<int> run_of(ne, -1), run_lev, run_first;
    int nlev = 0;
    for (int i = 0; i < ne; ++i) {
        int r = i > 0 ? run_of[i - 1] : -1;
        bool join = false;
        if (r >= 0 && i - run_first[r] < kRunMax)
            for (int k = 0; k < m.neigh[i]; ++k) {
                const int jc = m.neighc[3 * i + k];
                if (jc < i && jc >= run_first[r]) join = true;
            }
        if (!join) {
            r = (int)run_first.size();
            run_first.push_back(i);
            run_lev.push_back(0);
        }
        run_of[i] = r;
        for (int k = 0; k < m.neigh[i]; ++k) {
            const int jc = m.neighc[3 * i + k];
            if (jc < run_first[r]) run_lev[r] = std::max(run_lev[r], run_lev[run_of[jc]] + 1);
        }
        nlev = std::max(nlev, run_lev[r] + 1);
    }
    const int nruns = (int)run_first.size();
    m.lvl_ptr.assign(nlev + 1, 0);
    for (int r = 0; r < nruns; ++r) m.lvl_ptr[run_lev[r] + 1]++;
    for (int l = 0; l < nlev; ++l) m.lvl_ptr[l + 1] += m.lvl_ptr[l];
    m.runs.assign(2 * (size_t)nruns, 0);
    std::vector<int> pos(m.lvl_ptr.begin(), m.lvl_ptr.end() - 1);
    for (int r = 0; r < nruns; ++r) {
        const int t = pos[run_lev[r]]++;
        m.runs[2 * t] = run_first[r];
        m.runs[2 * t + 1] = (r + 1 < nruns ? run_first[r + 1] : ne) - run_first[r];
    }

    // ---- point-implicit sweep schedule (backward Euler): cell levels.  A
    // cell's lower neighbours are in lower levels, its upper neighbours in
    // higher ones, so the forward sweep runs the levels up and the backward
    // sweep down, each level's cells at once.
    std::vector<int> clev(ne, 0);
    int npl = 0;
    for (int i = 0; i < ne; ++i) {
        for (int k = 0; k < m.neigh[i]; ++k) {
            const int jc = m.neighc[3 * i + k];
            if (jc < i) clev[i] = std::max(clev[i], clev[jc] + 1);
        }
        npl = std::max(npl, clev[i] + 1);
    }
    m.pi_lvl_ptr.assign(npl + 1, 0);
    for (int i = 0; i < ne; ++i) m.pi_lvl_ptr[clev[i] + 1]++;
    for (int l = 0; l < npl; ++l) m.pi_lvl_ptr[l + 1] += m.pi_lvl_ptr[l];
    m.pi_cells.assign(ne, 0);
    std::vector<int> cpos(m.pi_lvl_ptr.begin(), m.pi_lvl_ptr.end() - 1);
    for (int i = 0; i < ne; ++i) m.pi_cells[cpos[clev[i]]++] = i;
    return RC_OK;
}

int order_multicolor(int ne, const int *cells0, int block, int *perm, int *ncolors) {
    if (ne <= 0 || block < 1) return RC_ERR_ARGUMENT;
    // face adjacency
    std::unordered_map<uint64_t, int> first;
    first.reserve((size_t)ne * 2);
    std::vector<int> adj(3 * (size_t)ne, -1), deg(ne, 0);
    for (int j = 0; j < ne; ++j) {
        for (int kf = 0; kf < 3; ++kf) {
            const int n1 = cells0[3 * j + kf], n2 = cells0[3 * j + (kf + 1) % 3];
            const uint64_t key = edge_key(n1, n2);
            auto it = first.find(key);
            if (it == first.end()) {
                first.emplace(key, j);
            } else {
                const int o = it->second;
                if (deg[o] >= 3 || deg[j] >= 3) return RC_ERR_MESH;
                adj[3 * o + deg[o]++] = j;
                adj[3 * j + deg[j]++] = o;
            }
        }
    }
    // reverse Cuthill-McKee, component by component, each started from a
    // pseudo-peripheral cell (end of a BFS from the lowest-degree cell)
    std::vector<int> order;
    order.reserve(ne);
    std::vector<char> seen(ne, 0);
    std::vector<int> dist(ne, -1);
    auto bfs_far = [&](int s) {
        std::vector<int> q{s};
        dist[s] = 0;
        int far = s;
        for (size_t h = 0; h < q.size(); ++h) {
            const int u = q[h];
            if (dist[u] > dist[far] || (dist[u] == dist[far] && deg[u] < deg[far])) far = u;
            for (int k = 0; k < deg[u]; ++k)
                if (dist[adj[3 * u + k]] < 0) { dist[adj[3 * u + k]] = dist[u] + 1; q.push_back(adj[3 * u + k]); }
        }
        for (int u : q) dist[u] = -1;
        return far;
    };
    for (int s0 = 0; s0 < ne; ++s0) {
        if (seen[s0]) continue;
        const int start = bfs_far(bfs_far(s0));
        std::vector<int> q{start};
        seen[start] = 1;
        for (size_t h = 0; h < q.size(); ++h) {
            const int u = q[h];
            int nb[3], nn = 0;
            for (int k = 0; k < deg[u]; ++k)
                if (!seen[adj[3 * u + k]]) nb[nn++] = adj[3 * u + k];
            std::sort(nb, nb + nn, [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
            for (int k = 0; k < nn; ++k) { seen[nb[k]] = 1; q.push_back(nb[k]); }
        }
        order.insert(order.end(), q.rbegin(), q.rend());
    }
    // blocks of up to `block` face-connected cells, grown in RCM order (a
    // block takes its seed's unassigned neighbours, lowest RCM rank first),
    // then a greedy colouring of the block adjacency graph.  Cells of one
    // block stay consecutive, in the order they joined the block (each shares
    // a face with a cell before it), so the sweeps keep Gauss-Seidel coupling
    // inside every block and a block is one sweep run; blocks of one colour
    // share no face.
    std::vector<int> rank(ne);
    for (int k = 0; k < ne; ++k) rank[order[k]] = k;
    std::vector<int> blk(ne, -1), bseed, gpos(ne, 0);
    for (int k = 0; k < ne; ++k) {
        const int u = order[k];
        if (blk[u] >= 0) continue;
        const int b = (int)bseed.size();
        bseed.push_back(k);
        blk[u] = b;
        std::vector<int> members{u};
        for (size_t h = 0; h < members.size() && (int)members.size() < block; ++h) {
            const int w = members[h];
            int nb[3], nn = 0;
            for (int t = 0; t < deg[w]; ++t)
                if (blk[adj[3 * w + t]] < 0) nb[nn++] = adj[3 * w + t];
            std::sort(nb, nb + nn, [&](int x, int y) { return rank[x] < rank[y]; });
            for (int t = 0; t < nn && (int)members.size() < block; ++t) {
                blk[nb[t]] = b;
                gpos[nb[t]] = (int)members.size();
                members.push_back(nb[t]);
            }
        }
    }
    const int nblk = (int)bseed.size();
    std::vector<std::vector<int>> badj(nblk);
    for (int u = 0; u < ne; ++u)
        for (int t = 0; t < deg[u]; ++t) {
            const int v = adj[3 * u + t];
            if (blk[u] != blk[v]) badj[blk[u]].push_back(blk[v]);
        }
    for (auto &v : badj) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    // Colouring of the block graph: the number of colours is the number of
    // sweep levels, each a chain of dependent launches, so fewer is faster
    // (a planar block graph needs 4 at least).  Greedy in RCM order first;
    // then DSatur (the most-constrained block next); then iterated greedy
    // passes (Culberson: re-colour greedily taking the colour classes in
    // reverse order, which never needs more colours), keeping the best.
    auto greedy = [&](const std::vector<int> &seq, std::vector<int> &col) {
        col.assign(nblk, -1);
        int n = 0;
        for (int b : seq) {
            uint64_t used = 0;
            for (int v : badj[b])
                if (col[v] >= 0 && col[v] < 64) used |= 1ull << col[v];
            int c = 0;
            while (c < 64 && (used & (1ull << c))) ++c;
            col[b] = c;
            n = std::max(n, c + 1);
        }
        return n;
    };
    std::vector<int> seq(nblk);
    std::iota(seq.begin(), seq.end(), 0);
    std::vector<int> bcol;
    int nc = greedy(seq, bcol);
    {
        // DSatur with a lazy max-heap keyed by (saturation, degree, -index)
        std::vector<int> col(nblk, -1), sat(nblk, 0);
        std::vector<uint64_t> seen_c(nblk, 0);
        std::priority_queue<std::tuple<int, int, int>> pq;
        for (int b = 0; b < nblk; ++b) pq.emplace(0, (int)badj[b].size(), -b);
        int n = 0;
        while (!pq.empty()) {
            const auto [sb, db, nbb] = pq.top();
            pq.pop();
            const int b = -nbb;
            if (col[b] >= 0 || sb != sat[b]) continue;
            int c = 0;
            while (c < 64 && (seen_c[b] & (1ull << c))) ++c;
            col[b] = c;
            n = std::max(n, c + 1);
            for (int v : badj[b])
                if (col[v] < 0 && c < 64 && !(seen_c[v] & (1ull << c))) {
                    seen_c[v] |= 1ull << c;
                    pq.emplace(++sat[v], (int)badj[v].size(), -v);
                }
        }
        if (n < nc) {
            nc = n;
            bcol = col;
        }
    }
    for (int pass = 0; pass < 32 && nc > 4; ++pass) {
        // colour classes in reverse order, blocks of a class in RCM order
        std::vector<int> s2(nblk);
        std::iota(s2.begin(), s2.end(), 0);
        std::stable_sort(s2.begin(), s2.end(), [&](int a, int b) { return bcol[a] > bcol[b]; });
        std::vector<int> col;
        const int n = greedy(s2, col);
        if (n > nc) break;
        nc = n;
        bcol = col;
    }
    if (nc > 4) {
        // four colours by local search (tabu min-conflicts, Hertz and de
        // Werra's TabuCol): blocks of colour >= 4 take the colour with
        // fewest conflicts, then a conflicting block moves to the colour
        // with fewest conflicts that is not tabu, until none conflicts;
        // deterministic (SplitMix64 stream), kept only if it succeeds
        constexpr int K = 4;
        std::vector<int> col(bcol);
        std::vector<int> conf((size_t)nblk * K, 0);
        for (int b = 0; b < nblk; ++b)
            if (col[b] >= K) col[b] = -1;
        auto cnt = [&](int b, int c) -> int & { return conf[(size_t)b * K + c]; };
        for (int b = 0; b < nblk; ++b)
            if (col[b] >= 0)
                for (int v : badj[b]) ++cnt(v, col[b]);
        for (int b = 0; b < nblk; ++b)
            if (col[b] < 0) {
                int best = 0;
                for (int c = 1; c < K; ++c)
                    if (cnt(b, c) < cnt(b, best)) best = c;
                col[b] = best;
                for (int v : badj[b]) ++cnt(v, best);
            }
        std::vector<int> bad, pos(nblk, -1);
        auto refresh = [&](int b) {
            const bool is_bad = cnt(b, col[b]) > 0;
            if (is_bad && pos[b] < 0) {
                pos[b] = (int)bad.size();
                bad.push_back(b);
            } else if (!is_bad && pos[b] >= 0) {
                const int last = bad.back();
                bad[pos[b]] = last;
                pos[last] = pos[b];
                bad.pop_back();
                pos[b] = -1;
            }
        };
        for (int b = 0; b < nblk; ++b) refresh(b);
        std::vector<long> tabu((size_t)nblk * K, -1);
        uint64_t rng = 0x9e3779b97f4a7c15ull;
        auto next = [&]() {
            uint64_t z = (rng += 0x9e3779b97f4a7c15ull);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            return z ^ (z >> 31);
        };
        const long max_iter = 200L * nblk + 100000;
        for (long it = 0; it < max_iter && !bad.empty(); ++it) {
            const int b = bad[next() % bad.size()];
            const int old = col[b];
            int best = -1;
            for (int c = 0; c < K; ++c) {
                if (c == old) continue;
                const bool allowed = tabu[(size_t)b * K + c] < it || cnt(b, c) == 0;
                if (allowed && (best < 0 || cnt(b, c) < cnt(b, best))) best = c;
            }
            if (best < 0) continue;
            col[b] = best;
            tabu[(size_t)b * K + old] = it + 7 + (long)(next() % 10);
            for (int v : badj[b]) {
                --cnt(v, old);
                ++cnt(v, best);
                refresh(v);
            }
            refresh(b);
        }
        if (bad.empty()) {
            bcol = col;
            nc = K;
        }
    }
    {
        // balance the colour classes (cells per sweep level): a level's
        // blocks run as one round of waves while they fit on the device, so
        // the largest class sets the cost; a block moves to the smallest
        // class none of its neighbours is in while that evens the sizes
        std::vector<long> wsz(nblk, 0), csz(nc, 0);
        for (int u = 0; u < ne; ++u) ++wsz[blk[u]];
        for (int b = 0; b < nblk; ++b) csz[bcol[b]] += wsz[b];
        for (int pass = 0; pass < 8; ++pass) {
            bool moved = false;
            for (int b = 0; b < nblk; ++b) {
                const int c0 = bcol[b];
                uint64_t used = 0;
                for (int v : badj[b]) used |= 1ull << bcol[v];
                int best = -1;
                for (int c = 0; c < nc; ++c)
                    if (c != c0 && !(used & (1ull << c)) && (best < 0 || csz[c] < csz[best])) best = c;
                if (best >= 0 && csz[best] + wsz[b] < csz[c0]) {
                    csz[c0] -= wsz[b];
                    csz[best] += wsz[b];
                    bcol[b] = best;
                    moved = true;
                }
            }
            if (!moved) break;
        }
    }
    std::vector<int> idx(ne);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int a, int b) {
        const int ba = blk[a], bb = blk[b];
        if (bcol[ba] != bcol[bb]) return bcol[ba] < bcol[bb];
        if (ba != bb) return bseed[ba] < bseed[bb];
        return gpos[a] < gpos[b];
    });
    for (int k = 0; k < ne; ++k) perm[k] = idx[k];
    if (ncolors) *ncolors = nc;
    return RC_OK;
}

}  // namespace rc
