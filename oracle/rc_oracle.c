/* Test infrastructure only: serial C restatement of the reference's steady
 * residual + LU-SGS path (the oracle).  Not linked into the product; used by
 * tests/ and bench.py's CPU-baseline leg as the checker.
 *
 * Each function cites the reference routine it restates.  Operation order
 * follows the Fortran expressions (left-to-right sums, the same
 * parenthesisation), so results agree with the reference to rounding of
 * the math library.  Built with -ffp-contract=off (no FMA contraction), like
 * the reference on baseline x86-64.
 */
#include "rc_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NFLOW 4
#define SMALL 1.0e-16              /* module.f90:19 */

struct rco_case {
    rco_mech m;
    rco_params p;
    int ne, nnt, nf, nvar, nsp, ng;
    double *coor;                  /* Coor(2,NNT) */
    int *n4;                       /* N(4,NE), 1-based nodes, N(4)=N(3) for triangles */
    int *fnode;                    /* FNode(2,NF) */
    int *neighbor;                 /* Neighbor(2,NF) */
    double *vecx;                  /* VECX(2,NF) */
    double *vol, *cellxy, *reflen; /* vol(NE), CellXY(2,NE), reflen(2,NE) */
    int *neighf;                   /* NeighF(3,NE) */
    int *neighc;                   /* NeighC(0:3,NE) */
    int *neigh;                    /* Neigh(NE) */
    int *hostcell, *hostface, *bctype_g;  /* per ghost */
    int *ghostcell, *num_ghost;    /* Ghostcell(3,NE), Num_ghostcell(NE) */
    double *coor_gn, *vol_g, *cellxy_g;   /* Coor_Ghostnode(2,4,NG), vol_Ghostcell, CellXY_Ghostcell(2,NG) */
    double *wdis, *amls, *xyh;     /* wdis(3,NE), AMLS(2,3,NE), XYH(2,NE) */
    double total_volume;
    double *pa, *wa, *hwa, *gradc, *gpa, *alpaf, *step, *prod, *dd;
    double *fixed;                 /* fixed(Nvar,NE) of the dual-time iteration */
    double *wp, *wp2;              /* WP, WP2 (Nvar,NE): the states at the start of the last two BDF2 steps */
    rco_transport tr;              /* transport / diffusion tables */
    int viscous;                   /* IfViscous */
    int *fprop;                    /* FProperty(NF): 100 on periodic faces */
    double *sideofs;               /* SideOfs(2,NF) */
    double *ofs;                   /* Ofs(2,0:3,NE) */
    char err[256];
};

#define PA(k, i)    c->pa[((i) - 1) * c->nvar + (k) - 1]
#define WA(k, i)    c->wa[((i) - 1) * c->nvar + (k) - 1]
#define HWA(k, i)   c->hwa[((i) - 1) * c->nvar + (k) - 1]
#define GPA(k, j)   c->gpa[((j) - 1) * c->nvar + (k) - 1]
#define GRADC(r, k, i) c->gradc[(((i) - 1) * c->nvar + (k) - 1) * 2 + (r) - 1]
#define COOR(d, n)  c->coor[((n) - 1) * 2 + (d) - 1]
#define CELLXY(d, i) c->cellxy[((i) - 1) * 2 + (d) - 1]
#define REFLEN(d, i) c->reflen[((i) - 1) * 2 + (d) - 1]
#define VECX(d, f)  c->vecx[((f) - 1) * 2 + (d) - 1]
#define NEIGHBOR(s, f) c->neighbor[((f) - 1) * 2 + (s) - 1]
#define FNODE(s, f) c->fnode[((f) - 1) * 2 + (s) - 1]
#define NEIGHF(k, i) c->neighf[((i) - 1) * 3 + (k) - 1]
#define NEIGHC(k, i) c->neighc[((i) - 1) * 4 + (k)]
#define NODE(k, i)  c->n4[((i) - 1) * 4 + (k) - 1]
#define WDIS(k, i)  c->wdis[((i) - 1) * 3 + (k) - 1]
#define AMLS(r, k, i) c->amls[(((i) - 1) * 3 + (k) - 1) * 2 + (r) - 1]
#define GHOSTCELL(k, i) c->ghostcell[((i) - 1) * 3 + (k) - 1]
#define CGN(d, v, j) c->coor_gn[(((j) - 1) * 4 + (v) - 1) * 2 + (d) - 1]
#define CXYG(d, j)  c->cellxy_g[((j) - 1) * 2 + (d) - 1]
#define SIDEOFS(d, f) c->sideofs[((f) - 1) * 2 + (d) - 1]
#define OFS(d, k, i) c->ofs[(((i) - 1) * 4 + (k)) * 2 + (d) - 1]

static int fail(rco_case *c, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof c->err, fmt, ap);
    va_end(ap);
    return 1;
}

/* real**integer as compiler-rt's __powidf2 evaluates it */
static double powi(double a, int b)
{
    const int recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

/* ------------------------------------------------------------------------ */
/* thermochemistry: src/combustion_modeling/thermo_property.f90            */
/* ------------------------------------------------------------------------ */

/* ComputeGasParameter, thermo_property.f90:78-112 */
int rco_gas_param(const rco_mech *m, const double *y, double t, double *out)
{
    double basis[7], cps[RCO_MAXSP];
    const double (*coef)[7];
    int j, i;
    for (j = 1; j <= 7; j++) basis[j - 1] = powi(t, j - 3);
    if (m->t_low <= t && t <= m->t_crit) coef = m->low;
    else if (m->t_crit < t && t <= m->t_high) coef = m->high;
    else return 1;
    for (i = 0; i < m->nsp; i++) {
        double s = 0.0;
        for (j = 0; j < 7; j++) s = s + coef[i][j] * basis[j];
        cps[i] = s * (m->rconstant / m->mw[i]);
    }
    double cp = 0.0, r = 0.0;
    for (i = 0; i < m->nsp; i++) cp = cp + cps[i] * y[i];
    for (i = 0; i < m->nsp; i++) r = r + (m->rconstant / m->mw[i]) * y[i];
    out[0] = cp;
    out[1] = cp / (cp - r);
    out[2] = r;
    return 0;
}

/* compute_enthalpy, thermo_property.f90:114-144 */
int rco_enthalpy(const rco_mech *m, double t, double *h)
{
    double basis[7];
    const double (*coef)[7];
    const double *b1;
    int j, i;
    basis[0] = -powi(t, -2);
    basis[1] = log(t) / t;
    for (j = 3; j <= 7; j++) basis[j - 1] = powi(t, j - 3) / (double)(j - 2);
    if (m->t_low <= t && t <= m->t_crit) { coef = m->low; b1 = m->low_b1; }
    else if (m->t_crit < t && t <= m->t_high) { coef = m->high; b1 = m->high_b1; }
    else return 1;
    for (i = 0; i < m->nsp; i++) {
        double s = 0.0;
        for (j = 0; j < 7; j++) s = s + coef[i][j] * basis[j];
        h[i] = (m->rconstant / m->mw[i]) * (t * s + b1[i]);
    }
    return 0;
}

/* Newton_temperature, thermo_property.f90:146-180 */
int rco_newton(const rco_mech *m, double et, const double *y, double t0, double *out)
{
    double t = t0, h[RCO_MAXSP], g[3] = {0, 0, 0};
    int it = 0, i;
    while (it < 200) {
        it++;
        if (rco_enthalpy(m, t, h)) return 1;
        if (rco_gas_param(m, y, t, g)) return 1;
        double hh = 0.0;
        for (i = 0; i < m->nsp; i++) hh = hh + y[i] * h[i];
        double ft = hh - g[2] * t - et;
        double dft = g[0] - g[2];
        double dtt = -ft / dft;
        t = t + dtt;
        t = fmax(t, m->t_low);
        t = fmin(t, m->t_high);
        if (fabs(dtt) < 1.0e-5) break;
    }
    out[0] = t; out[1] = g[0]; out[2] = g[1]; out[3] = g[2];
    return 0;
}

/* compute_production_rate, chemical_kinetics.f90:109-165 */
void rco_production_rate(const rco_mech *m, double t, double rho, const double *y2, double *w)
{
    double y[RCO_MAXSP], kf;
    int k;
    for (k = 0; k < m->nsp; k++) y[k] = fmin(1.0, fmax(0.0, y2[k]));
    if (m->model == 0) {
        kf = m->a_f * exp(-m->e_act / (m->rconstant * t));
        kf = kf * pow(y[0] * rho / (m->mw[0] * 1000.0), 0.2);
        kf = kf * pow(y[1] * rho / (m->mw[1] * 1000.0), 1.3);
    } else {
        kf = 0.0;
        if (t > 710.0) kf = 1e4;
        kf = kf * pow(y[0] * rho / (m->mw[0] * 1000.0), 1.0);
        kf = kf * pow(y[1] * rho / (m->mw[1] * 1000.0), 0.0);
    }
    for (k = 0; k < m->nsp; k++)
        w[k] = 1000.0 * m->mw[k] * (m->coef_r[k] - m->coef_f[k]) * kf;
}

/* ------------------------------------------------------------------------ */
/* set-up                                                                   */
/* ------------------------------------------------------------------------ */

/* GaussInteg_Tri, gaussInteg4.f90:4-42; fun selects 1, x or y, or the
 * scaled coordinate (x-cx)/rx of compute_reconstruction_matrix:153-167 */
typedef struct { int kind; double cx, rx; } gfun;
static double gval(const gfun *f, const double *v)
{
    switch (f->kind) {
    case 0: return 1.0;
    case 1: return v[0];
    case 2: return v[1];
    case 3: return (v[0] - f->cx) / f->rx;
    default: return (v[1] - f->cx) / f->rx;
    }
}

static double gauss_tri(const gfun *f, const double *xv1, const double *xv2, const double *xv3)
{
    double x1 = xv1[0], y1 = xv1[1], x2 = xv2[0], y2 = xv2[1], x3 = xv3[0], y3 = xv3[1];
    double area = 0.5 * fabs((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1));
    double v[13][2];
    int d;
    for (d = 0; d < 2; d++) {
        double a, b;
        v[0][d] = (xv1[d] + xv2[d] + xv3[d]) / 3.0;
        a = 0.479308067841920; b = 0.260345966079040;
        v[1][d] = a * xv1[d] + b * (xv2[d] + xv3[d]);
        v[2][d] = a * xv2[d] + b * (xv1[d] + xv3[d]);
        v[3][d] = a * xv3[d] + b * (xv1[d] + xv2[d]);
        a = 0.869739794195568; b = 0.065130102902216;
        v[4][d] = a * xv1[d] + b * (xv2[d] + xv3[d]);
        v[5][d] = a * xv2[d] + b * (xv1[d] + xv3[d]);
        v[6][d] = a * xv3[d] + b * (xv1[d] + xv2[d]);
        a = 0.048690315425316; b = 0.312865496004874;
        v[7][d] = a * xv1[d] + b * xv2[d] + (1 - a - b) * xv3[d];
        v[8][d] = a * xv1[d] + b * xv3[d] + (1 - a - b) * xv2[d];
        v[9][d] = a * xv2[d] + b * xv1[d] + (1 - a - b) * xv3[d];
        v[10][d] = a * xv2[d] + b * xv3[d] + (1 - a - b) * xv1[d];
        v[11][d] = a * xv3[d] + b * xv1[d] + (1 - a - b) * xv2[d];
        v[12][d] = a * xv3[d] + b * xv2[d] + (1 - a - b) * xv1[d];
    }
    double r = -0.149570044467682 * gval(f, v[0])
        + 0.175615257433208 * (gval(f, v[1]) + gval(f, v[2]) + gval(f, v[3]))
        + 0.053347235608838 * (gval(f, v[4]) + gval(f, v[5]) + gval(f, v[6]))
        + 0.077113760890257 * (gval(f, v[7]) + gval(f, v[8]) + gval(f, v[9]) + gval(f, v[10])
                               + gval(f, v[11]) + gval(f, v[12]));
    return r * area;
}

/* GaussInteg_Quad, gaussInteg4.f90:44-50 */
static double gauss_quad(const gfun *f, const double *a, const double *b, const double *cc, const double *d)
{
    return gauss_tri(f, a, b, cc) + gauss_tri(f, a, cc, d);
}

/* pseudo_inverse_lapack (la_math.f90:1-62) forms pinv(A) of the 3x2 LSQ
 * matrix with dgesvd.  The restatement uses a thin QR factorisation by
 * Gram-Schmidt with one re-orthogonalisation pass, A = QR, pinv = R^-1 Q^T,
 * which is backward stable like the SVD route (the two agree to
 * O(cond(A) eps)). */
static void pinv32(const double a[3][2], double b[2][3])
{
    double q[3][2], r[2][2];
    double v[3];
    int i;
    /* modified Gram-Schmidt with one re-orthogonalisation pass */
    double n0 = sqrt(a[0][0] * a[0][0] + a[1][0] * a[1][0] + a[2][0] * a[2][0]);
    for (i = 0; i < 3; i++) q[i][0] = a[i][0] / n0;
    r[0][0] = n0;
    for (i = 0; i < 3; i++) v[i] = a[i][1];
    double r01 = q[0][0] * v[0] + q[1][0] * v[1] + q[2][0] * v[2];
    for (i = 0; i < 3; i++) v[i] = v[i] - r01 * q[i][0];
    double s = q[0][0] * v[0] + q[1][0] * v[1] + q[2][0] * v[2];
    for (i = 0; i < 3; i++) v[i] = v[i] - s * q[i][0];
    r01 = r01 + s;
    double n1 = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (i = 0; i < 3; i++) q[i][1] = v[i] / n1;
    r[0][1] = r01; r[1][1] = n1; r[1][0] = 0.0;
    /* B = R^-1 Q^T */
    for (i = 0; i < 3; i++) {
        double y1 = q[i][1] / r[1][1];
        double y0 = (q[i][0] - r[0][1] * y1) / r[0][0];
        b[0][i] = y0;
        b[1][i] = y1;
    }
}

static void *xcalloc(size_t n, size_t s) { void *p = calloc(n ? n : 1, s); if (!p) abort(); return p; }

int rco_create(const rco_mech *mech, const rco_params *par, int nnt, const double *coor,
               int ne, const int *cells, int nb, const int *bedges, const int *btype, rco_case **out)
{
    return rco_create_periodic(mech, par, nnt, coor, ne, cells, nb, bedges, btype, 0, NULL, NULL, out);
}

int rco_create_periodic(const rco_mech *mech, const rco_params *par, int nnt, const double *coor,
                        int ne, const int *cells, int nb, const int *bedges, const int *btype,
                        int ncon, const int *con, const double *dxyz, rco_case **out)
{
    rco_case *c = (rco_case *)xcalloc(1, sizeof *c);
    *out = c;
    c->m = *mech;
    c->p = *par;
    c->ne = ne; c->nnt = nnt; c->nsp = mech->nsp; c->nvar = NFLOW + mech->nsp - 1;
    c->coor = (double *)xcalloc(2 * nnt, sizeof(double));
    memcpy(c->coor, coor, sizeof(double) * 2 * nnt);
    c->n4 = (int *)xcalloc(4 * ne, sizeof(int));
    int i, k;
    for (i = 1; i <= ne; i++) {
        for (k = 1; k <= 3; k++) NODE(k, i) = cells[(i - 1) * 3 + k - 1] + 1;
        NODE(4, i) = NODE(3, i);   /* mesh.f90:33-35 */
    }

    /* change / face, mesh.f90:120-199: faces in order of first appearance */
    int maxf = 3 * ne;
    c->fnode = (int *)xcalloc(2 * maxf, sizeof(int));
    c->neighbor = (int *)xcalloc(2 * maxf, sizeof(int));
    c->vecx = (double *)xcalloc(2 * maxf, sizeof(double));
    int *numnf = (int *)xcalloc(nnt + 1, sizeof(int));
    int *nodeface = (int *)xcalloc((size_t)(nnt + 1) * 50, sizeof(int));
    int nf = 0, j, kf;
    for (j = 1; j <= ne; j++) {
        for (kf = 1; kf <= 3; kf++) {
            int n1 = NODE(kf, j), n2 = NODE(kf % 3 + 1, j);
            int idx = 0, inum;
            for (inum = 1; inum <= numnf[n2]; inum++) {
                int f = nodeface[n2 * 50 + inum - 1];
                if ((FNODE(1, f) == n2 && FNODE(2, f) == n1) || (FNODE(1, f) == n1 && FNODE(2, f) == n2)) { idx = f; break; }
            }
            if (idx == 0) {
                nf++;
                NEIGHBOR(1, nf) = j;
                NEIGHBOR(2, nf) = -100;
                VECX(1, nf) = COOR(2, n2) - COOR(2, n1);
                VECX(2, nf) = -(COOR(1, n2) - COOR(1, n1));
                FNODE(1, nf) = n1; FNODE(2, nf) = n2;
                numnf[n1]++; numnf[n2]++;
                if (numnf[n1] > 50 || numnf[n2] > 50) { free(numnf); free(nodeface); return fail(c, "error in NumNFace"); }
                nodeface[n1 * 50 + numnf[n1] - 1] = nf;
                nodeface[n2 * 50 + numnf[n2] - 1] = nf;
            } else {
                NEIGHBOR(2, idx) = j;
            }
        }
    }
    c->nf = nf;

    /* CreatePeriodConnect, period.f90:15-104: boundary faces by node pair
     * (BoundF, :46-62), then each face of boundary Con(1) takes the owner of
     * the first face of boundary Con(2) whose midpoint (single precision)
     * is its own moved by dxyz (:66-101).  Boundaries are the labels btype
     * >= 100 here; the connections name them (mesh.f90:60-64). */
    c->fprop = (int *)xcalloc(nf + 1, sizeof(int));
    c->sideofs = (double *)xcalloc(2 * (size_t)(nf + 1), sizeof(double));
    int icon, b, b2;
    for (icon = 0; icon < ncon; icon++) {
        for (b = 0; b < nb; b++) {
            if (btype[b] != con[2 * icon]) continue;
            int n1 = bedges[2 * b] + 1, n2 = bedges[2 * b + 1] + 1, iface = 0, ic1;
            for (ic1 = 1; ic1 <= numnf[n1]; ic1++) {
                int f = nodeface[n1 * 50 + ic1 - 1];
                if (FNODE(1, f) == n2 || FNODE(2, f) == n2) { iface = f; break; }
            }
            if (!iface) { free(numnf); free(nodeface); return fail(c, "not find boundary face %d", b + 1); }
            float xfs[2], xft[2], ddd[2];
            xfs[0] = (float)((COOR(1, n1) + COOR(1, n2)) / 2.);
            xfs[1] = (float)((COOR(2, n1) + COOR(2, n2)) / 2.);
            int itface = 0;
            for (b2 = 0; b2 < nb; b2++) {
                if (btype[b2] != con[2 * icon + 1]) continue;
                int m1 = bedges[2 * b2] + 1, m2 = bedges[2 * b2 + 1] + 1;
                xft[0] = (float)((COOR(1, m1) + COOR(1, m2)) / 2.);
                xft[1] = (float)((COOR(2, m1) + COOR(2, m2)) / 2.);
                ddd[0] = xft[0] - xfs[0];
                ddd[1] = xft[1] - xfs[1];
                if (fabs(ddd[0] - dxyz[2 * icon]) + fabs(ddd[1] - dxyz[2 * icon + 1]) < 1.e-7f) {
                    for (ic1 = 1; ic1 <= numnf[m1]; ic1++) {
                        int f = nodeface[m1 * 50 + ic1 - 1];
                        if (FNODE(1, f) == m2 || FNODE(2, f) == m2) { itface = f; break; }
                    }
                    break;
                }
            }
            if (!itface) { free(numnf); free(nodeface); return fail(c, "not find target (boundary face %d)", b + 1); }
            NEIGHBOR(2, iface) = NEIGHBOR(1, itface);
            c->fprop[iface] = 100;
            SIDEOFS(1, iface) = dxyz[2 * icon];
            SIDEOFS(2, iface) = dxyz[2 * icon + 1];
        }
    }
    free(numnf); free(nodeface);

    /* BC tagging, mesh.f90:76-93 (a connection's boundaries have type 100) */
    for (i = 1; i <= nf; i++) {
        if (NEIGHBOR(2, i) < 0) {
            int n1 = FNODE(1, i), n2 = FNODE(2, i), found = 0;
            for (b = 0; b < nb; b++) {
                int a1 = bedges[2 * b] + 1, a2 = bedges[2 * b + 1] + 1;
                if ((n1 == a1 && n2 == a2) || (n2 == a1 && n1 == a2)) {
                    int t = btype[b];
                    for (icon = 0; icon < ncon; icon++)
                        if (t == con[2 * icon] || t == con[2 * icon + 1]) t = 100;
                    NEIGHBOR(2, i) = -t; found = 1; break;
                }
            }
            if (!found) return fail(c, "not find boundary of this type (face %d)", i);
        }
    }

    /* VolumeAndCenter, mesh.f90:206-286 */
    c->vol = (double *)xcalloc(ne, sizeof(double));
    c->cellxy = (double *)xcalloc(2 * ne, sizeof(double));
    c->reflen = (double *)xcalloc(2 * ne, sizeof(double));
    gfun f1 = {0, 0, 0}, fx = {1, 0, 0}, fy = {2, 0, 0};
    for (i = 1; i <= ne; i++) {
        const double *x1 = &COOR(1, NODE(1, i)), *x2 = &COOR(1, NODE(2, i)), *x3 = &COOR(1, NODE(3, i)), *x4 = &COOR(1, NODE(4, i));
        c->vol[i - 1] = gauss_quad(&f1, x1, x2, x3, x4);
        CELLXY(1, i) = gauss_quad(&fx, x1, x2, x3, x4) / c->vol[i - 1];
        CELLXY(2, i) = gauss_quad(&fy, x1, x2, x3, x4) / c->vol[i - 1];
        if (c->vol[i - 1] < 0.0) return fail(c, "negative volume %d", i);
        REFLEN(1, i) = 0.0; REFLEN(2, i) = 0.0;
        int a, b, d;
        for (a = 1; a <= 2; a++)
            for (b = a + 1; b <= 3; b++)
                for (d = 1; d <= 2; d++)
                    REFLEN(d, i) = fmax(0.5 * fabs(COOR(d, NODE(a, i)) - COOR(d, NODE(b, i))), REFLEN(d, i));
    }
    double sv = 0.0;
    for (i = 1; i <= ne; i++) sv = sv + c->vol[i - 1];
    c->total_volume = sv;

    /* compute_reconstruction_matrix, reconstruction_initialization.f90:1-150 */
    c->neighf = (int *)xcalloc(3 * ne, sizeof(int));
    c->neighc = (int *)xcalloc(4 * ne, sizeof(int));
    c->neigh = (int *)xcalloc(ne, sizeof(int));
    int *ntemp = (int *)xcalloc(ne + 1, sizeof(int));
    for (i = 1; i <= nf; i++) {
        int NL = NEIGHBOR(1, i), NR = NEIGHBOR(2, i);
        ntemp[NL]++; NEIGHF(ntemp[NL], NL) = i;
        if (NR > 0 && c->fprop[i] != 100) { ntemp[NR]++; NEIGHF(ntemp[NR], NR) = i; }
    }
    free(ntemp);
    c->ofs = (double *)xcalloc(8 * (size_t)ne, sizeof(double));
    for (i = 1; i <= ne; i++) {
        for (k = 1; k <= 3; k++) {
            int iface = NEIGHF(k, i);
            for (kf = 1; kf <= 2; kf++) {
                int ic = NEIGHBOR(kf, iface);
                if (ic != i && ic > 0) {
                    c->neigh[i - 1]++;
                    NEIGHC(c->neigh[i - 1], i) = ic;
                    OFS(1, c->neigh[i - 1], i) = SIDEOFS(1, iface);   /* :40 */
                    OFS(2, c->neigh[i - 1], i) = SIDEOFS(2, iface);
                }
            }
        }
        NEIGHC(0, i) = i;
    }

    /* construct_ghostcells, ghostcell.f90:19-195 */
    int ng = 0;
    for (i = 1; i <= ne; i++) ng += 3 - c->neigh[i - 1];
    c->ng = ng;
    c->hostcell = (int *)xcalloc(ng, sizeof(int));
    c->hostface = (int *)xcalloc(ng, sizeof(int));
    c->bctype_g = (int *)xcalloc(ng, sizeof(int));
    c->ghostcell = (int *)xcalloc(3 * ne, sizeof(int));
    c->num_ghost = (int *)xcalloc(ne, sizeof(int));
    c->coor_gn = (double *)xcalloc(8 * ng, sizeof(double));
    c->vol_g = (double *)xcalloc(ng, sizeof(double));
    c->cellxy_g = (double *)xcalloc(2 * ng, sizeof(double));
    j = 0;
    for (i = 1; i <= ne; i++) {
        if (c->neigh[i - 1] < 3) {
            for (kf = 1; kf <= 3; kf++) {
                int iface = NEIGHF(kf, i);
                if (NEIGHBOR(2, iface) < 0) {
                    j++;
                    c->hostcell[j - 1] = i; c->hostface[j - 1] = iface;
                    c->bctype_g[j - 1] = -NEIGHBOR(2, iface);
                    c->num_ghost[i - 1]++;
                    GHOSTCELL(c->num_ghost[i - 1], i) = j;
                    int n1 = FNODE(1, iface), n2 = FNODE(2, iface), k1, k2, k3, k4;
                    for (k1 = 1; k1 <= 3; k1++) if (n1 == NODE(k1, i)) break;
                    k2 = k1 % 3 + 1;
                    if (NODE(k2, i) != n2) return fail(c, "error in index setting of nodes when looking for ghostcells");
                    k3 = k2 % 3 + 1; k4 = k3 % 3 + 1;
                    int n3 = NODE(k3, i), n4 = NODE(k4, i);
                    double sav1 = VECX(1, iface), sav2 = VECX(2, iface);
                    double sav = fmax(sqrt(sav1 * sav1 + sav2 * sav2), SMALL);
                    double s1n = sav1 / sav, s2n = sav2 / sav;
                    CGN(1, 1, j) = COOR(1, n2); CGN(2, 1, j) = COOR(2, n2);
                    CGN(1, 2, j) = COOR(1, n1); CGN(2, 2, j) = COOR(2, n1);
                    double d1 = (COOR(1, n1) - COOR(1, n4)) * s1n + (COOR(2, n1) - COOR(2, n4)) * s2n;
                    if (d1 < 0) return fail(c, "error in calculating the distance (n1,n4)");
                    CGN(1, 3, j) = COOR(1, n4) + 2 * d1 * s1n;
                    CGN(2, 3, j) = COOR(2, n4) + 2 * d1 * s2n;
                    double d2 = (COOR(1, n2) - COOR(1, n3)) * s1n + (COOR(2, n2) - COOR(2, n3)) * s2n;
                    if (d2 < 0) return fail(c, "error in calculating the distance (n2,n3)");
                    CGN(1, 4, j) = COOR(1, n3) + 2 * d2 * s1n;
                    CGN(2, 4, j) = COOR(2, n3) + 2 * d2 * s2n;
                    const double *g1 = &CGN(1, 1, j), *g2 = &CGN(1, 2, j), *g3 = &CGN(1, 3, j), *g4 = &CGN(1, 4, j);
                    c->vol_g[j - 1] = gauss_quad(&f1, g1, g2, g3, g4);
                    CXYG(1, j) = gauss_quad(&fx, g1, g2, g3, g4) / c->vol_g[j - 1];
                    CXYG(2, j) = gauss_quad(&fy, g1, g2, g3, g4) / c->vol_g[j - 1];
                    if (c->vol_g[j - 1] < 0.0) return fail(c, "negative ghost volume %d", j);
                }
            }
        }
        if (c->num_ghost[i - 1] + c->neigh[i - 1] != 3) return fail(c, "error in number of ghost cells for cell %d", i);
    }

    c->wdis = (double *)xcalloc(3 * ne, sizeof(double));
    c->amls = (double *)xcalloc(6 * ne, sizeof(double));
    c->xyh = (double *)xcalloc(2 * ne, sizeof(double));
    for (i = 1; i <= ne; i++) {
        const double *x1 = &COOR(1, NODE(1, i)), *x2 = &COOR(1, NODE(2, i)), *x3 = &COOR(1, NODE(3, i)), *x4 = &COOR(1, NODE(4, i));
        int io;
        for (io = 1; io <= 2; io++) {
            gfun fs = {2 + io, CELLXY(io, i), REFLEN(io, i)};
            c->xyh[(i - 1) * 2 + io - 1] = gauss_quad(&fs, x1, x2, x3, x4) / c->vol[i - 1];
        }
        if (fabs(c->xyh[(i - 1) * 2]) + fabs(c->xyh[(i - 1) * 2 + 1]) > 1.e-8) return fail(c, "error in XYH %d", i);
    }
    for (i = 1; i <= ne; i++) {
        double a1[3][2], b1[2][3];
        int nn = c->neigh[i - 1];
        for (k = 1; k <= nn; k++) {
            j = NEIGHC(k, i);
            /* the neighbour's nodes moved by its offset (:101-109) */
            double xv[4][2];
            int v;
            for (v = 0; v < 4; v++) {
                xv[v][0] = COOR(1, NODE(v + 1, j)) - OFS(1, k, i);
                xv[v][1] = COOR(2, NODE(v + 1, j)) - OFS(2, k, i);
            }
            const double *x1 = xv[0], *x2 = xv[1], *x3 = xv[2], *x4 = xv[3];
            double dx = (CELLXY(1, j) - OFS(1, k, i)) - CELLXY(1, i), dy = (CELLXY(2, j) - OFS(2, k, i)) - CELLXY(2, i);
            double dist = sqrt(dx * dx + dy * dy);
            WDIS(k, i) = powi(dist, -1);
            int io;
            for (io = 1; io <= 2; io++) {
                gfun fs = {2 + io, CELLXY(io, i), REFLEN(io, i)};
                a1[k - 1][io - 1] = WDIS(k, i) * (gauss_quad(&fs, x1, x2, x3, x4) / c->vol[j - 1] - c->xyh[(i - 1) * 2 + io - 1]);
            }
        }
        for (k = 1; k <= c->num_ghost[i - 1]; k++) {
            j = GHOSTCELL(k, i);
            const double *g1 = &CGN(1, 1, j), *g2 = &CGN(1, 2, j), *g3 = &CGN(1, 3, j), *g4 = &CGN(1, 4, j);
            double dx = CXYG(1, j) - CELLXY(1, i), dy = CXYG(2, j) - CELLXY(2, i);
            double dist = sqrt(dx * dx + dy * dy);
            WDIS(k + nn, i) = powi(dist, -1);
            int io;
            for (io = 1; io <= 2; io++) {
                gfun fs = {2 + io, CELLXY(io, i), REFLEN(io, i)};
                a1[k + nn - 1][io - 1] = WDIS(k + nn, i) * (gauss_quad(&fs, g1, g2, g3, g4) / c->vol_g[j - 1] - c->xyh[(i - 1) * 2 + io - 1]);
            }
        }
        pinv32((const double(*)[2])a1, b1);
        for (k = 1; k <= 3; k++) { AMLS(1, k, i) = b1[0][k - 1]; AMLS(2, k, i) = b1[1][k - 1]; }
    }

    int nv = c->nvar;
    c->pa = (double *)xcalloc((size_t)nv * ne, sizeof(double));
    c->wa = (double *)xcalloc((size_t)nv * ne, sizeof(double));
    c->hwa = (double *)xcalloc((size_t)nv * ne, sizeof(double));
    c->gradc = (double *)xcalloc((size_t)2 * nv * ne, sizeof(double));
    c->gpa = (double *)xcalloc((size_t)nv * ng, sizeof(double));
    c->alpaf = (double *)xcalloc(nf, sizeof(double));
    c->step = (double *)xcalloc(ne, sizeof(double));
    c->prod = (double *)xcalloc((size_t)c->nsp * ne, sizeof(double));
    c->dd = (double *)xcalloc(ne, sizeof(double));
    c->fixed = (double *)xcalloc((size_t)nv * ne, sizeof(double));
    c->wp = (double *)xcalloc((size_t)nv * ne, sizeof(double));
    c->wp2 = (double *)xcalloc((size_t)nv * ne, sizeof(double));
    return 0;
}

void rco_destroy(rco_case *c)
{
    if (!c) return;
    void *ps[] = {c->fprop, c->sideofs, c->ofs, c->coor, c->n4, c->fnode, c->neighbor, c->vecx, c->vol, c->cellxy, c->reflen,
                  c->neighf, c->neighc, c->neigh, c->hostcell, c->hostface, c->bctype_g,
                  c->ghostcell, c->num_ghost, c->coor_gn, c->vol_g, c->cellxy_g, c->wdis, c->amls,
                  c->xyh, c->pa, c->wa, c->hwa, c->gradc, c->gpa, c->alpaf, c->step, c->prod, c->dd, c->fixed,
                  c->wp, c->wp2};
    size_t i;
    for (i = 0; i < sizeof ps / sizeof ps[0]; i++) free(ps[i]);
    free(c);
}

const char *rco_last_error(const rco_case *c) { return c->err; }

void rco_sizes(const rco_case *c, int *o) { o[0] = c->ne; o[1] = c->nf; o[2] = c->ng; o[3] = c->nvar; }

double *rco_darray(rco_case *c, const char *n)
{
    if (!strcmp(n, "PA")) return c->pa;
    if (!strcmp(n, "WA")) return c->wa;
    if (!strcmp(n, "hWA")) return c->hwa;
    if (!strcmp(n, "gradC")) return c->gradc;
    if (!strcmp(n, "GPA")) return c->gpa;
    if (!strcmp(n, "alpaF")) return c->alpaf;
    if (!strcmp(n, "step")) return c->step;
    if (!strcmp(n, "dd")) return c->dd;
    if (!strcmp(n, "fixed")) return c->fixed;
    if (!strcmp(n, "production_rate")) return c->prod;
    if (!strcmp(n, "VECX")) return c->vecx;
    if (!strcmp(n, "vol")) return c->vol;
    if (!strcmp(n, "cellxy")) return c->cellxy;
    if (!strcmp(n, "reflen")) return c->reflen;
    if (!strcmp(n, "wdis")) return c->wdis;
    if (!strcmp(n, "AMLS")) return c->amls;
    if (!strcmp(n, "vol_Ghostcell")) return c->vol_g;
    if (!strcmp(n, "CellXY_Ghostcell")) return c->cellxy_g;
    if (!strcmp(n, "total_volume")) return &c->total_volume;
    return NULL;
}

int *rco_iarray(rco_case *c, const char *n)
{
    if (!strcmp(n, "Neighbor")) return c->neighbor;
    if (!strcmp(n, "FNode")) return c->fnode;
    if (!strcmp(n, "NeighF")) return c->neighf;
    if (!strcmp(n, "NeighC")) return c->neighc;
    if (!strcmp(n, "Neigh")) return c->neigh;
    if (!strcmp(n, "Num_ghostcell")) return c->num_ghost;
    if (!strcmp(n, "Ghostcell")) return c->ghostcell;
    if (!strcmp(n, "Hostcell")) return c->hostcell;
    if (!strcmp(n, "BCtype_Ghostcell")) return c->bctype_g;
    return NULL;
}

/* ------------------------------------------------------------------------ */
/* state                                                                    */
/* ------------------------------------------------------------------------ */

/* primitve_to_conserved_variables, update_flow_variables.f90:50-84 */
static int prim_to_cons(rco_case *c)
{
    int i, k, nsp = c->nsp;
    for (i = 1; i <= c->ne; i++) {
        double pp = PA(1, i), uu = PA(2, i), vv = PA(3, i), tt = PA(4, i);
        double yy[RCO_MAXSP], g[3], h[RCO_MAXSP], s = 0.0;
        for (k = 0; k < nsp - 1; k++) { yy[k] = PA(NFLOW + 1 + k, i); s = s + yy[k]; }
        yy[nsp - 1] = 1.0 - s;
        if (rco_gas_param(&c->m, yy, tt, g)) return fail(c, "temperature %g out of range (cell %d)", tt, i);
        double rr = pp / (g[2] * tt);
        if (rco_enthalpy(&c->m, tt, h)) return fail(c, "temperature %g out of range (cell %d)", tt, i);
        double hh = 0.0;
        for (k = 0; k < nsp; k++) hh = hh + yy[k] * h[k];
        WA(1, i) = rr;
        WA(2, i) = rr * uu;
        WA(3, i) = rr * vv;
        WA(4, i) = rr * (hh + 0.5 * (uu * uu + vv * vv)) - pp;
        for (k = 0; k < nsp - 1; k++) WA(NFLOW + 1 + k, i) = rr * yy[k];
    }
    return 0;
}

int rco_set_pa(rco_case *c, const double *pa)
{
    memcpy(c->pa, pa, sizeof(double) * c->nvar * c->ne);
    return prim_to_cons(c);
}

/* TimeStep (steady; viscous branch when IfViscous), TimeStep.f90:4-103 */
int rco_timestep(rco_case *c)
{
    int i, k, nsp = c->nsp, nvar = c->nvar;
    double *alpac = (double *)xcalloc(c->ne + 1, sizeof(double));
    for (i = 1; i <= c->nf; i++) {
        int NL = NEIGHBOR(1, i), NR = NEIGHBOR(2, i);
        double dx = VECX(1, i), dy = VECX(2, i), pp, uu, vv, tt, yy[RCO_MAXSP], s = 0.0, g[3];
        if (NR > 0) {
            pp = 0.5 * (PA(1, NL) + PA(1, NR));
            uu = 0.5 * (PA(2, NL) + PA(2, NR));
            vv = 0.5 * (PA(3, NL) + PA(3, NR));
            tt = 0.5 * (PA(4, NL) + PA(4, NR));
            for (k = 0; k < nvar - NFLOW; k++) { yy[k] = 0.5 * (PA(NFLOW + 1 + k, NL) + PA(NFLOW + 1 + k, NR)); s = s + yy[k]; }
        } else {
            pp = PA(1, NL); uu = PA(2, NL); vv = PA(3, NL); tt = PA(4, NL);
            for (k = 0; k < nvar - NFLOW; k++) { yy[k] = PA(NFLOW + 1 + k, NL); s = s + yy[k]; }
        }
        yy[nsp - 1] = 1.0 - s;
        if (rco_gas_param(&c->m, yy, tt, g)) { free(alpac); return fail(c, "timestep: temperature %g out of range", tt); }
        double rr = pp / (g[2] * tt);
        if (rr < 0. || pp < 0. || tt < 0.) { free(alpac); return fail(c, "error in computing timestep, cells %d %d", NL, NR); }
        double un = uu * dx + vv * dy;
        if (c->viscous) {
            /* TimeStep.f90:73-85 */
            double vmul, akmu;
            if (rco_transport_property(&c->m, &c->tr, tt, yy, &vmul, &akmu)) {
                free(alpac);
                return fail(c, "error: temperature %g is not in range of the transport table", tt);
            }
            double prl = vmul * g[0] / akmu;
            double alpa = fabs(un) + sqrt(g[1] * pp / rr) * sqrt(dx * dx + dy * dy);
            double alpavis = 2. * (dx * dx + dy * dy) * fmax(4. / 3., g[1] / prl) * vmul / rr;
            c->alpaf[i - 1] = alpa + alpavis / c->vol[NL - 1];
            alpac[NL] = alpac[NL] + alpa + alpavis / c->vol[NL - 1];
            if (NR > 0) alpac[NR] = alpac[NR] + alpa + alpavis / c->vol[NR - 1];
            continue;
        }
        double alpa = fabs(un) + sqrt(g[1] * pp / rr) * sqrt(dx * dx + dy * dy);
        c->alpaf[i - 1] = alpa;
        alpac[NL] = alpac[NL] + alpa;
        if (NR > 0) alpac[NR] = alpac[NR] + alpa;
    }
    for (i = 1; i <= c->ne; i++) c->step[i - 1] = c->p.cfl * c->vol[i - 1] / alpac[i];
    free(alpac);
    return 0;
}

static void inlet_state(const rco_case *c, int which, double pl, double *pr, double *ur, double *vr,
                        double *tr, double *rr, double *yr, double *g, int *bad)
{
    /* oxidizer_inlet_bc oxidizer.f90:53-68 / fuel_inlet_bc fuel.f90:54-69 */
    const double *yin = which == 3 ? c->p.ox_y : c->p.fu_y;
    int k;
    for (k = 0; k < c->nsp; k++) yr[k] = yin[k];
    *tr = which == 3 ? c->p.ox_t : c->p.fu_t;
    if (rco_gas_param(&c->m, yr, *tr, g)) *bad = 1;
    *pr = pl;
    *rr = *pr / (g[2] * *tr);
    *ur = (which == 3 ? c->p.ox_g : c->p.fu_g) / *rr;
    *vr = 0.0;
}

/* update_ghostcell_variables, update_flow_variables.f90:90-269 */
static int ghost_update(rco_case *c)
{
    int j, k, nsp = c->nsp;
    for (j = 1; j <= c->ng; j++) {
        int i = c->hostcell[j - 1];
        double pl = PA(1, i), ul = PA(2, i), vl = PA(3, i), tl = PA(4, i), yl[RCO_MAXSP], yr[RCO_MAXSP], s = 0.0, g[3];
        for (k = 0; k < nsp - 1; k++) { yl[k] = PA(NFLOW + 1 + k, i); s = s + yl[k]; }
        yl[nsp - 1] = 1.0 - s;
        if (rco_gas_param(&c->m, yl, tl, g)) return fail(c, "ghost: temperature %g out of range", tl);
        double rl = pl / (g[2] * tl);
        const double *xv1 = &CGN(1, 1, j), *xv2 = &CGN(1, 2, j);
        double sav1 = xv2[1] - xv1[1];
        double sav2 = -(xv2[0] - xv1[0]);
        double sav = fmax(sqrt(sav1 * sav1 + sav2 * sav2), SMALL);
        double s1n = sav1 / sav, s2n = sav2 / sav;
        double ac = sqrt(g[1] * pl / rl);
        double pr, ur, vr, tr, rr, vn;
        int bad = 0;
        switch (c->bctype_g[j - 1]) {
        case 1:
            pr = pl; tr = tl; for (k = 0; k < nsp; k++) yr[k] = yl[k];
            ur = -ul; vr = -vl;
            break;
        case 2: case 6:
            pr = pl; tr = tl; for (k = 0; k < nsp; k++) yr[k] = yl[k];
            vn = ul * s1n + vl * s2n;
            ur = ul - 2 * vn * s1n;
            vr = vl - 2 * vn * s2n;
            break;
        case 3: case 4:
            inlet_state(c, c->bctype_g[j - 1], pl, &pr, &ur, &vr, &tr, &rr, yr, g, &bad);
            if (bad) return fail(c, "ghost inlet: temperature out of range");
            break;
        case 5:
            vn = ul * s1n + vl * s2n;
            if (fabs(vn) > ac) { pr = pl; ur = ul; vr = vl; tr = tl; }
            else { pr = c->p.p_out; ur = ul; vr = vl; tr = tl; }
            for (k = 0; k < nsp; k++) yr[k] = yl[k];
            break;
        default:
            return fail(c, "error in BCtype of ghost cells: %d", c->bctype_g[j - 1]);
        }
        GPA(1, j) = pr; GPA(2, j) = ur; GPA(3, j) = vr; GPA(4, j) = tr;
        for (k = 0; k < nsp - 1; k++) GPA(NFLOW + 1 + k, j) = yr[k];
    }
    return 0;
}

/* gradient, gradient.f90:4-45 (the limiter is separate, limiter.f90) */
int rco_gradient(rco_case *c)
{
    int i, k, nei, nvar = c->nvar;
    if (ghost_update(c)) return 1;
    double C[3][RCO_MAXSP + NFLOW];
    for (i = 1; i <= c->ne; i++) {
        int nn = c->neigh[i - 1];
        for (nei = 1; nei <= nn; nei++) {
            int j = NEIGHC(nei, i);
            for (k = 1; k <= nvar; k++) C[nei - 1][k - 1] = WDIS(nei, i) * (PA(k, j) - PA(k, i));
        }
        for (nei = 1; nei <= c->num_ghost[i - 1]; nei++) {
            int j = GHOSTCELL(nei, i);
            for (k = 1; k <= nvar; k++) C[nei + nn - 1][k - 1] = WDIS(nei + nn, i) * (GPA(k, j) - PA(k, i));
        }
        for (k = 1; k <= nvar; k++) {
            int r;
            for (r = 1; r <= 2; r++) {
                double s = 0.0;
                int m;
                for (m = 1; m <= 3; m++) s = s + AMLS(r, m, i) * C[m - 1][k - 1];
                GRADC(r, k, i) = s;
            }
        }
    }
    return 0;
}

/* Limiter_Barth, limiter.f90:5-107 */
int rco_limiter(rco_case *c)
{
    int i, k, nei, kf, nvar = c->nvar;
    double pmin[RCO_MAXSP + NFLOW], pmax[RCO_MAXSP + NFLOW], phi[RCO_MAXSP + NFLOW];
    for (i = 1; i <= c->ne; i++) {
        for (k = 1; k <= nvar; k++) { pmin[k - 1] = PA(k, i); pmax[k - 1] = PA(k, i); }
        for (nei = 1; nei <= c->neigh[i - 1]; nei++) {
            int j = NEIGHC(nei, i);
            for (k = 1; k <= nvar; k++) { pmin[k - 1] = fmin(pmin[k - 1], PA(k, j)); pmax[k - 1] = fmax(pmax[k - 1], PA(k, j)); }
        }
        for (nei = 1; nei <= c->num_ghost[i - 1]; nei++) {
            int j = GHOSTCELL(nei, i);
            for (k = 1; k <= nvar; k++) { pmin[k - 1] = fmin(pmin[k - 1], GPA(k, j)); pmax[k - 1] = fmax(pmax[k - 1], GPA(k, j)); }
        }
        for (k = 1; k <= nvar; k++) phi[k - 1] = 1.0;
        for (k = 1; k <= nvar; k++) {
            for (kf = 1; kf <= 3; kf++) {
                int iface = NEIGHF(kf, i);
                int n1 = FNODE(1, iface), n2 = FNODE(2, iface);
                double xy0 = 0.5 * (COOR(1, n1) + COOR(1, n2));
                double xy1 = 0.5 * (COOR(2, n1) + COOR(2, n2));
                double v0 = (xy0 - CELLXY(1, i)) / REFLEN(1, i);
                double v1 = (xy1 - CELLXY(2, i)) / REFLEN(2, i);
                double fv = PA(k, i) + (GRADC(1, k, i) * v0 + GRADC(2, k, i) * v1);
                double pf;
                if (fv > PA(k, i)) pf = fmin(1.0, (pmax[k - 1] - PA(k, i)) / (fv - PA(k, i)));
                else if (fv < PA(k, i)) pf = fmin(1.0, (pmin[k - 1] - PA(k, i)) / (fv - PA(k, i)));
                else pf = 1.0;
                phi[k - 1] = fmin(phi[k - 1], pf);
            }
        }
        for (k = 1; k <= nvar; k++) { GRADC(1, k, i) = phi[k - 1] * GRADC(1, k, i); GRADC(2, k, i) = phi[k - 1] * GRADC(2, k, i); }
    }
    return 0;
}

/* compute_transport_property, transport_property.f90:78-125: mixture
 * viscosity (Wilke) and conductivity at (T, Y); 1 if T is above the table */
int rco_transport_property(const rco_mech *m, const rco_transport *tr, double tem2, const double *yy,
                           double *vmul, double *akmu)
{
    int i, j, nsp = m->nsp;
    double tem = fmax(tr->t_low, tem2);
    double xx[RCO_MAXSP], mul[RCO_MAXSP], kl[RCO_MAXSP], phil[RCO_MAXSP], s = 0.0;
    for (i = 0; i < nsp; i++) s = s + yy[i] / m->mw[i];
    double ww = 1. / s;
    for (i = 0; i < nsp; i++) xx[i] = ww * yy[i] / m->mw[i];
    const double(*mu)[4], (*kk)[4];
    if (tr->t_low <= tem && tem <= tr->t_crit) {
        mu = tr->mu_low; kk = tr->k_low;
    } else if (tr->t_crit < tem && tem <= tr->t_high) {
        mu = tr->mu_high; kk = tr->k_high;
    } else {
        return 1;
    }
    for (i = 0; i < nsp; i++) {
        mul[i] = exp(mu[i][0] * log(tem) + mu[i][1] / tem + mu[i][2] / (tem * tem) + mu[i][3]);
        kl[i] = exp(kk[i][0] * log(tem) + kk[i][1] / tem + kk[i][2] / (tem * tem) + kk[i][3]);
    }
    for (i = 0; i < nsp; i++) mul[i] = mul[i] * 1.0e-7;
    for (i = 0; i < nsp; i++) kl[i] = kl[i] * 1.0e-4;
    for (i = 0; i < nsp; i++) {
        phil[i] = 0.;
        for (j = 0; j < nsp; j++) {
            double t = 1. + sqrt(mul[i] / mul[j]) * pow(m->mw[j] / m->mw[i], 0.25);
            phil[i] = phil[i] + xx[j] * (t * t) / sqrt(1. + m->mw[i] / m->mw[j]);
        }
        phil[i] = 0.25 * sqrt(2.) * phil[i];
    }
    double v = 0.0, a1 = 0.0, a2 = 0.0;
    for (i = 0; i < nsp; i++) v = v + xx[i] * mul[i] / phil[i];
    for (i = 0; i < nsp; i++) a1 = a1 + xx[i] * kl[i];
    for (i = 0; i < nsp; i++) a2 = a2 + xx[i] / kl[i];
    *vmul = v;
    *akmu = 0.5 * (a1 + 1. / a2);
    return 0;
}

/* compute_diffusion_velocity, diffusion_velocity.f90:55-117: mixture-
 * averaged diffusion coefficients from binary ones (Chapman-Enskog with the
 * Neufeld collision integral) */
void rco_diffusion_velocity(const rco_mech *m, const rco_transport *tr, double pp, double tem, const double *yy,
                            double *dk)
{
    static const double A = 1.06036, B = 0.15610, C = 0.19300, D = 0.47635, E = 1.03587, F = 1.52996,
                        G = 1.76474, H = 3.89411;
    int i, j, nsp = m->nsp;
    double xx[RCO_MAXSP], dij[RCO_MAXSP], s = 0.0;
    for (i = 0; i < nsp; i++) s = s + yy[i] / m->mw[i];
    double ww = 1. / s;
    for (i = 0; i < nsp; i++) xx[i] = ww * yy[i] / m->mw[i];
    for (i = 0; i < nsp; i++) {
        for (j = 0; j < nsp; j++) {
            double mw = 2. / (1. / m->mw[i] + 1. / m->mw[j]) * 1000;
            double cd = 0.5 * (tr->sigma[i] + tr->sigma[j]);
            double tstar = tem / sqrt(tr->eps[i] * tr->eps[j]);
            double om = A * pow(tstar, -B) + C * exp(-D * tstar) + E * exp(-F * tstar) + G * exp(-H * tstar);
            dij[j] = 0.0266 * sqrt(tem * tem * tem) / (pp * sqrt(mw) * (cd * cd) * om);
        }
        double xd = 0.;
        for (j = 0; j < i; j++) xd = xd + xx[j] / dij[j];
        for (j = i + 1; j < nsp; j++) xd = xd + xx[j] / dij[j];
        dk[i] = (1. - xx[i]) / (xd + 1.0e-16);
    }
}

static void removedirection(double *vx, double *vy, double nx, double ny)
{
    /* flux_viscous.f90:354-362 */
    double vn = *vx * nx + *vy * ny;
    *vx = *vx - vn * nx;
    *vy = *vy - vn * ny;
}

/* flux_viscous, flux_viscous.f90:4-350 (one Gauss point, wgg = 1): the
 * viscous, heat-conduction and species-diffusion fluxes at the face
 * midpoint from the reconstructed states and the averaged gradients with
 * the face-normal jump term, accumulated into hWA in face order */
static int flux_viscous(rco_case *c)
{
    int i, k, nsp = c->nsp, nvar = c->nvar, ns1 = nsp - 1;
    const double yita = 4. / 3;
    for (i = 1; i <= c->nf; i++) {
        int NL = NEIGHBOR(1, i), NR = NEIGHBOR(2, i);
        double dx = VECX(1, i), dy = VECX(2, i);
        int n1 = FNODE(1, i), n2 = FNODE(2, i);
        double sav1 = dx, sav2 = dy;
        double sav = fmax(sqrt(dx * dx + dy * dy), SMALL);
        double s1n = sav1 / sav, s2n = sav2 / sav;
        double xy0 = COOR(1, n1) + 0.5 * (COOR(1, n2) - COOR(1, n1));
        double xy1 = COOR(2, n1) + 0.5 * (COOR(2, n2) - COOR(2, n1));
        double v0 = (xy0 - CELLXY(1, NL)) / REFLEN(1, NL);
        double v1 = (xy1 - CELLXY(2, NL)) / REFLEN(2, NL);
        double u_l[RCO_MAXSP + NFLOW], u_r[RCO_MAXSP + NFLOW];
        for (k = 1; k <= nvar; k++) u_l[k - 1] = PA(k, NL) + (GRADC(1, k, NL) * v0 + GRADC(2, k, NL) * v1);
        double pl = u_l[0], ul = u_l[1], vl = u_l[2], tl = u_l[3];
        double yl[RCO_MAXSP], yr[RCO_MAXSP], gl[3], gr[3], s = 0.0;
        for (k = 0; k < ns1; k++) yl[k] = u_l[NFLOW + k];
        for (k = 0; k < ns1; k++) s = s + yl[k];
        yl[nsp - 1] = 1. - s;
        if (rco_gas_param(&c->m, yl, tl, gl)) return fail(c, "flux_viscous: temperature %g out of range", tl);
        double rl = pl / (gl[2] * tl);
        double pr, ur, vr, tr, rr;
        if (NR > 0) {
            double w0 = (xy0 - (CELLXY(1, NR) - SIDEOFS(1, i))) / REFLEN(1, NR);
            double w1 = (xy1 - (CELLXY(2, NR) - SIDEOFS(2, i))) / REFLEN(2, NR);
            for (k = 1; k <= nvar; k++) u_r[k - 1] = PA(k, NR) + (GRADC(1, k, NR) * w0 + GRADC(2, k, NR) * w1);
            pr = u_r[0]; ur = u_r[1]; vr = u_r[2]; tr = u_r[3];
            s = 0.0;
            for (k = 0; k < ns1; k++) yr[k] = u_r[NFLOW + k];
            for (k = 0; k < ns1; k++) s = s + yr[k];
            yr[nsp - 1] = 1. - s;
            if (rco_gas_param(&c->m, yr, tr, gr)) return fail(c, "flux_viscous: temperature %g out of range", tr);
            rr = pr / (gr[2] * tr);
        } else if (NR == -1) {
            rr = rl; pr = pl; tr = tl;
            ul = 0.; vl = 0.;
            ur = 0.; vr = 0.;
            for (k = 0; k < nsp; k++) yr[k] = yl[k];
            gr[0] = gl[0]; gr[1] = gl[1]; gr[2] = gl[2];
        } else if (NR == -2 || NR == -6) {
            removedirection(&ul, &vl, s1n, s2n);
            rr = rl; pr = pl; tr = tl;
            ur = ul; vr = vl;
            for (k = 0; k < nsp; k++) yr[k] = yl[k];
            gr[0] = gl[0]; gr[1] = gl[1]; gr[2] = gl[2];
        } else if (NR == -3 || NR == -4) {
            int bad = 0;
            inlet_state(c, -NR, pl, &pr, &ur, &vr, &tr, &rr, yr, gr, &bad);
            if (bad) return fail(c, "flux_viscous inlet: temperature out of range");
            pl = pr; ul = ur; vl = vr; tl = tr; rl = rr;
            for (k = 0; k < nsp; k++) yl[k] = yr[k];
            gl[0] = gr[0]; gl[1] = gr[1]; gl[2] = gr[2];
        } else if (NR == -5) {
            double ac = sqrt(gl[1] * pl / rl);
            double vn = ul * s1n + vl * s2n;
            if (fabs(vn) > ac) {
                rr = rl; ur = ul; vr = vl; pr = pl; tr = tl;
            } else {
                pr = c->p.p_out;
                ur = ul + s1n * (pl - pr) / (rl * ac);
                vr = vl + s2n * (pl - pr) / (rl * ac);
                rr = rl + (pr - pl) / (ac * ac);
                tr = pr / (gl[2] * rr);
                pl = pr; ul = ur; vl = vr; tl = tr; rl = rr;
            }
            for (k = 0; k < nsp; k++) yr[k] = yl[k];
            gr[0] = gl[0]; gr[1] = gl[1]; gr[2] = gl[2];
        } else {
            rr = rl; pr = pl; tr = tl;
            ur = ul; vr = vl;
            for (k = 0; k < nsp; k++) yr[k] = yl[k];
            gr[0] = gl[0]; gr[1] = gl[1]; gr[2] = gl[2];
        }
        double pp = 0.5 * (pl + pr), uu = 0.5 * (ul + ur), vv = 0.5 * (vl + vr), tem = 0.5 * (tl + tr);
        double yy[RCO_MAXSP], g[3];
        for (k = 0; k < nsp; k++) yy[k] = 0.5 * (yl[k] + yr[k]);
        if (rco_gas_param(&c->m, yy, tem, g)) return fail(c, "flux_viscous: temperature %g out of range", tem);
        double rho = pp / (g[2] * tem);
        /* gradients: fbaseGrad_dxdy (module.f90:124-137) scales by 1/reflen */
        double tgd[2][RCO_MAXSP + NFLOW];
        double ix = 1. / REFLEN(1, NL), iy = 1. / REFLEN(2, NL);
        for (k = 1; k <= nvar; k++) {
            tgd[0][k - 1] = GRADC(1, k, NL) * ix + GRADC(2, k, NL) * 0.;
            tgd[1][k - 1] = GRADC(1, k, NL) * 0. + GRADC(2, k, NL) * iy;
        }
        if (NR > 0) {
            double dlrx = CELLXY(1, NL) - (CELLXY(1, NR) - SIDEOFS(1, i));
            double dlry = CELLXY(2, NL) - (CELLXY(2, NR) - SIDEOFS(2, i));
            double dxk = fabs(dlrx * s1n + dlry * s2n);
            double jx = 1. / REFLEN(1, NR), jy = 1. / REFLEN(2, NR);
            for (k = 1; k <= nvar; k++) {
                tgd[0][k - 1] = 0.5 * (tgd[0][k - 1] + (GRADC(1, k, NR) * jx + GRADC(2, k, NR) * 0.)) +
                                yita * s1n * (u_r[k - 1] - u_l[k - 1]) / dxk;
                tgd[1][k - 1] = 0.5 * (tgd[1][k - 1] + (GRADC(1, k, NR) * 0. + GRADC(2, k, NR) * jy)) +
                                yita * s2n * (u_r[k - 1] - u_l[k - 1]) / dxk;
            }
        }
        double dpdx = tgd[0][0], dpdy = tgd[1][0], dudx = tgd[0][1], dudy = tgd[1][1];
        double dvdx = tgd[0][2], dvdy = tgd[1][2], dtdx = tgd[0][3], dtdy = tgd[1][3];
        double dydx[RCO_MAXSP], dydy[RCO_MAXSP];
        for (k = 0; k < ns1; k++) { dydx[k] = tgd[0][NFLOW + k]; dydy[k] = tgd[1][NFLOW + k]; }
        if (NR == -1 || NR == -6) removedirection(&dtdx, &dtdy, s1n, s2n);
        if (NR == -2) {
            double snx = s1n, sny = s2n, stx = -sny, sty = snx;
            removedirection(&dpdx, &dpdy, snx, sny);
            removedirection(&dtdx, &dtdy, snx, sny);
            for (k = 0; k < ns1; k++) removedirection(&dydx[k], &dydy[k], snx, sny);
            double dutdx = dudx * stx + dvdx * sty, dutdy = dudy * stx + dvdy * sty;
            double dundx = dudx * snx + dvdx * sny, dundy = dudy * snx + dvdy * sny;
            removedirection(&dutdx, &dutdy, snx, sny);
            removedirection(&dundx, &dundy, stx, sty);
            dudx = dutdx * stx + dundx * snx;
            dudy = dutdy * stx + dundy * snx;
            dvdx = dutdx * sty + dundx * sny;
            dvdy = dutdy * sty + dundy * sny;
        }
        (void)dpdx; (void)dpdy;
        s = 0.0;
        for (k = 0; k < ns1; k++) s = s + dydx[k];
        dydx[nsp - 1] = -s;
        s = 0.0;
        for (k = 0; k < ns1; k++) s = s + dydy[k];
        dydy[nsp - 1] = -s;
        double entha[RCO_MAXSP], vmulc, akmu, dk[RCO_MAXSP];
        if (rco_enthalpy(&c->m, tem, entha)) return fail(c, "flux_viscous: temperature %g out of range", tem);
        if (rco_transport_property(&c->m, &c->tr, tem, yy, &vmulc, &akmu))
            return fail(c, "error: temperature %g is not in range of the transport table", tem);
        rco_diffusion_velocity(&c->m, &c->tr, pp, tem, yy, dk);
        double uicx = 0.0, uicy = 0.0;
        for (k = 0; k < nsp; k++) uicx = uicx + dk[k] * dydx[k];
        for (k = 0; k < nsp; k++) uicy = uicy + dk[k] * dydy[k];
        double vlac = -2. / 3. * vmulc;
        double div = dudx + dvdy;
        double txx = 2. * vmulc * dudx + vlac * div;
        double tyy = 2. * vmulc * dvdy + vlac * div;
        double txy = vmulc * (dudy + dvdx);
        double sx = 0.0, sy = 0.0;
        for (k = 0; k < nsp; k++) sx = sx + (dk[k] * dydx[k] - yy[k] * uicx) * entha[k];
        for (k = 0; k < nsp; k++) sy = sy + (dk[k] * dydy[k] - yy[k] * uicy) * entha[k];
        double qx = -akmu * dtdx - rho * sx;
        double qy = -akmu * dtdy - rho * sy;
        double btx = uu * txx + vv * txy - qx;
        double bty = uu * txy + vv * tyy - qy;
        double flu[RCO_MAXSP + NFLOW];
        flu[0] = 0.;
        flu[1] = -(sav1 * txx + sav2 * txy);
        flu[2] = -(sav1 * txy + sav2 * tyy);
        flu[3] = -(sav1 * btx + sav2 * bty);
        for (k = 0; k < ns1; k++) {
            double spx = rho * (dk[k] * dydx[k] - yy[k] * uicx);
            double spy = rho * (dk[k] * dydy[k] - yy[k] * uicy);
            flu[NFLOW + k] = -(sav1 * spx + sav2 * spy);
        }
        for (k = 1; k <= nvar; k++) {
            HWA(k, NL) = HWA(k, NL) - flu[k - 1] * 1.0;
            if (NR > 0 && c->fprop[i] != 100) HWA(k, NR) = HWA(k, NR) + flu[k - 1] * 1.0;
        }
    }
    return 0;
}

void rco_set_transport(rco_case *c, const rco_transport *tr, int viscous)
{
    c->tr = *tr;
    c->viscous = viscous;
}

/* flux_lax, flux_lax.f90:4-299, and sourceterm, sourceterm.f90:2-70,
 * composed as flux_sourceterm, fvm2d.f90:196-224 */
int rco_flux_sourceterm(rco_case *c)
{
    int i, k, nsp = c->nsp, nvar = c->nvar, ns1 = nsp - 1;
    memset(c->hwa, 0, sizeof(double) * nvar * c->ne);
    for (i = 1; i <= c->nf; i++) {
        int NL = NEIGHBOR(1, i), NR = NEIGHBOR(2, i);
        double dx = VECX(1, i), dy = VECX(2, i);
        int n1 = FNODE(1, i), n2 = FNODE(2, i);
        double sav1 = dx, sav2 = dy;
        double sav = fmax(sqrt(dx * dx + dy * dy), SMALL);
        double s1n = sav1 / sav, s2n = sav2 / sav;
        double xy0 = COOR(1, n1) + 0.5 * (COOR(1, n2) - COOR(1, n1));
        double xy1 = COOR(2, n1) + 0.5 * (COOR(2, n2) - COOR(2, n1));
        double v0 = (xy0 - CELLXY(1, NL)) / REFLEN(1, NL);
        double v1 = (xy1 - CELLXY(2, NL)) / REFLEN(2, NL);
        double pl = PA(1, NL) + (GRADC(1, 1, NL) * v0 + GRADC(2, 1, NL) * v1);
        double ul = PA(2, NL) + (GRADC(1, 2, NL) * v0 + GRADC(2, 2, NL) * v1);
        double vl = PA(3, NL) + (GRADC(1, 3, NL) * v0 + GRADC(2, 3, NL) * v1);
        double tl = PA(4, NL) + (GRADC(1, 4, NL) * v0 + GRADC(2, 4, NL) * v1);
        double yl[RCO_MAXSP], yr[RCO_MAXSP], gl[3], gr[3], s = 0.0;
        for (k = 1; k <= ns1; k++) { yl[k - 1] = PA(k + NFLOW, NL) + (GRADC(1, k + NFLOW, NL) * v0 + GRADC(2, k + NFLOW, NL) * v1); }
        for (k = 0; k < ns1; k++) s = s + yl[k];
        yl[nsp - 1] = 1.0 - s;
        if (rco_gas_param(&c->m, yl, tl, gl)) return fail(c, "flux: temperature %g out of range (face %d)", tl, i);
        double rl = pl / (gl[2] * tl);
        double pr, ur, vr, tr, rr, vn;
        int bad = 0;
        if (NR > 0) {
            double w0 = (xy0 - (CELLXY(1, NR) - SIDEOFS(1, i))) / REFLEN(1, NR);
            double w1 = (xy1 - (CELLXY(2, NR) - SIDEOFS(2, i))) / REFLEN(2, NR);
            pr = PA(1, NR) + (GRADC(1, 1, NR) * w0 + GRADC(2, 1, NR) * w1);
            ur = PA(2, NR) + (GRADC(1, 2, NR) * w0 + GRADC(2, 2, NR) * w1);
            vr = PA(3, NR) + (GRADC(1, 3, NR) * w0 + GRADC(2, 3, NR) * w1);
            tr = PA(4, NR) + (GRADC(1, 4, NR) * w0 + GRADC(2, 4, NR) * w1);
            s = 0.0;
            for (k = 1; k <= ns1; k++) yr[k - 1] = PA(k + NFLOW, NR) + (GRADC(1, k + NFLOW, NR) * w0 + GRADC(2, k + NFLOW, NR) * w1);
            for (k = 0; k < ns1; k++) s = s + yr[k];
            yr[nsp - 1] = 1.0 - s;
            if (rco_gas_param(&c->m, yr, tr, gr)) return fail(c, "flux: temperature %g out of range (face %d)", tr, i);
            rr = pr / (gr[2] * tr);
        } else if (NR == -1) {
            rr = rl; pr = pl; tr = tl; ur = -ul; vr = -vl;
            for (k = 0; k < nsp; k++) yr[k] = yl[k];
            gr[0] = gl[0]; gr[1] = gl[1]; gr[2] = gl[2];
        } else if (NR == -2 || NR == -6) {
            rr = rl; pr = pl; tr = tl;
            vn = ul * s1n + vl * s2n;
            ur = ul - 2 * vn * s1n;
            vr = vl - 2 * vn * s2n;
            for (k = 0; k < nsp; k++) yr[k] = yl[k];
            gr[0] = gl[0]; gr[1] = gl[1]; gr[2] = gl[2];
        } else if (NR == -3 || NR == -4) {
            inlet_state(c, -NR, pl, &pr, &ur, &vr, &tr, &rr, yr, gr, &bad);
            if (bad) return fail(c, "flux inlet: temperature out of range");
            pl = pr; ul = ur; vl = vr; tl = tr; rl = rr;
            for (k = 0; k < nsp; k++) yl[k] = yr[k];
            gl[0] = gr[0]; gl[1] = gr[1]; gl[2] = gr[2];
        } else if (NR == -5) {
            double ac = sqrt(gl[1] * pl / rl);
            vn = ul * s1n + vl * s2n;
            if (fabs(vn) > ac) {
                rr = rl; ur = ul; vr = vl; pr = pl; tr = tl;
            } else {
                pr = c->p.p_out;
                ur = ul + s1n * (pl - pr) / (rl * ac);
                vr = vl + s2n * (pl - pr) / (rl * ac);
                rr = rl + (pr - pl) / (ac * ac);
                tr = pr / (gl[2] * rr);
                pl = pr; ul = ur; vl = vr; tl = tr; rl = rr;
            }
            for (k = 0; k < nsp; k++) yr[k] = yl[k];
            gr[0] = gl[0]; gr[1] = gl[1]; gr[2] = gl[2];
        } else {
            return fail(c, "error,no face of this type! %d %d", i, NR);
        }
        if (rl < 0. || pl < 0. || (NR > 0 && (rr < 0. || pr < 0.)))
            return fail(c, "error,if: face %d rl=%g rr=%g pl=%g pr=%g", i, rl, rr, pl, pr);
        double h[RCO_MAXSP], hl = 0.0, hr = 0.0;
        if (rco_enthalpy(&c->m, tl, h)) return fail(c, "flux enthalpy: temperature %g out of range", tl);
        for (k = 0; k < nsp; k++) hl = hl + yl[k] * h[k];
        if (rco_enthalpy(&c->m, tr, h)) return fail(c, "flux enthalpy: temperature %g out of range", tr);
        for (k = 0; k < nsp; k++) hr = hr + yr[k] * h[k];
        double vaml = ul * ul + vl * vl, vamr = ur * ur + vr * vr;
        hl = hl + 0.5 * vaml;
        hr = hr + 0.5 * vamr;
        double el = hl - pl / rl, er = hr - pr / rr;
        double acl = sqrt(gl[1] * pl / rl), acr = sqrt(gr[1] * pr / rr);
        double uln = sav1 * ul + sav2 * vl, urn = sav1 * ur + sav2 * vr;
        double coex = fmax(fabs(uln) + acl * sav, fabs(urn) + acr * sav);
        double ruln = rl * uln, rurn = rr * urn;
        double flu[RCO_MAXSP + NFLOW];
        flu[0] = 0.5 * (ruln + rurn - coex * (rr - rl));
        flu[1] = 0.5 * (ruln * ul + rurn * ur + sav1 * (pl + pr) - coex * (rr * ur - rl * ul));
        flu[2] = 0.5 * (ruln * vl + rurn * vr + sav2 * (pl + pr) - coex * (rr * vr - rl * vl));
        flu[3] = 0.5 * (ruln * hl + rurn * hr - coex * (rr * er - rl * el));
        for (k = 0; k < ns1; k++) flu[NFLOW + k] = 0.5 * (ruln * yl[k] + rurn * yr[k] - coex * (rr * yr[k] - rl * yl[k]));
        for (k = 1; k <= nvar; k++) {
            HWA(k, NL) = HWA(k, NL) - flu[k - 1] * 1.0;
            if (NR > 0 && c->fprop[i] != 100) HWA(k, NR) = HWA(k, NR) + flu[k - 1] * 1.0;
        }
    }
    if (c->viscous && flux_viscous(c)) return 1;
    if (!c->p.combustion) return 0;
    for (i = 1; i <= c->ne; i++) {
        double p = PA(1, i), t = PA(4, i), y[RCO_MAXSP], s = 0.0, g[3], w[RCO_MAXSP];
        for (k = 0; k < ns1; k++) { y[k] = PA(NFLOW + 1 + k, i); s = s + y[k]; }
        y[nsp - 1] = 1.0 - s;
        if (rco_gas_param(&c->m, y, t, g)) return fail(c, "source: temperature %g out of range", t);
        double rho = p / (t * g[2]);
        rco_production_rate(&c->m, t, rho, y, w);
        for (k = 0; k < ns1; k++) c->prod[(i - 1) * nsp + k] = w[k];
        for (k = 1; k <= NFLOW; k++) HWA(k, i) = HWA(k, i) + c->vol[i - 1] * 0.0;
        for (k = 0; k < ns1; k++) HWA(NFLOW + 1 + k, i) = HWA(NFLOW + 1 + k, i) + c->vol[i - 1] * w[k];
    }
    return 0;
}

/* the neighbour term of both LU-SGS sweeps, Time_LUSGS_steady.f90:72-111 */
static int lusgs_neighbor(rco_case *c, int NR, double dx, double dy, double af, double *dflux)
{
    int k, nsp = c->nsp, ns1 = nsp - 1;
    double pp = PA(1, NR), uu = PA(2, NR), vv = PA(3, NR), tt = PA(4, NR), yy[RCO_MAXSP], s = 0.0, g[3], h[RCO_MAXSP];
    for (k = 0; k < ns1; k++) { yy[k] = PA(NFLOW + 1 + k, NR); s = s + yy[k]; }
    yy[nsp - 1] = 1.0 - s;
    if (rco_gas_param(&c->m, yy, tt, g)) return fail(c, "lusgs: temperature %g out of range", tt);
    if (rco_enthalpy(&c->m, tt, h)) return fail(c, "lusgs: temperature %g out of range", tt);
    double hh = 0.0;
    for (k = 0; k < nsp; k++) hh = hh + yy[k] * h[k];
    double gama = g[1], rcp = g[2];
    double rr = pp / (rcp * tt);
    double rou = rr * uu, rov = rr * vv;
    double roe = rr * (hh + 0.5 * (uu * uu + vv * vv)) - pp;
    double ry[RCO_MAXSP], dry[RCO_MAXSP];
    for (k = 0; k < nsp; k++) ry[k] = rr * yy[k];
    double drr = HWA(1, NR), dru = HWA(2, NR), drv = HWA(3, NR), dre = HWA(4, NR);
    s = 0.0;
    for (k = 0; k < ns1; k++) { dry[k] = HWA(NFLOW + 1 + k, NR); s = s + dry[k]; }
    dry[nsp - 1] = drr - s;
    double duu = (dru - uu * drr) / rr;
    double dvv = (drv - vv * drr) / rr;
    double s1 = 0.0, s2 = 0.0;
    for (k = 0; k < nsp; k++) s1 = s1 + h[k] * dry[k];
    for (k = 0; k < nsp; k++) s2 = s2 + (c->m.rconstant / c->m.mw[k]) * dry[k];
    double dpp = (gama - 1) * (dre - s1 + tt * s2 * gama / (gama - 1)
                               - 0.5 * (dru * uu + duu * rou + drv * vv + dvv * rov));
    double un = uu * dx + vv * dy;
    double dun = duu * dx + dvv * dy;
    dflux[0] = dflux[0] + 0.5 * (-af * drr + dru * dx + drv * dy);
    dflux[1] = dflux[1] + 0.5 * (-af * dru + dru * un + rou * dun + dpp * dx);
    dflux[2] = dflux[2] + 0.5 * (-af * drv + drv * un + rov * dun + dpp * dy);
    dflux[3] = dflux[3] + 0.5 * (-af * dre + dun * (roe + pp) + un * (dre + dpp));
    for (k = 0; k < ns1; k++) dflux[NFLOW + k] = dflux[NFLOW + k] + 0.5 * (-af * dry[k] + dry[k] * un + ry[k] * dun);
    return 0;
}

/* update_primitve_variables, update_flow_variables.f90:4-48 */
static int update_primitive(rco_case *c)
{
    int i, k, nsp = c->nsp, ns1 = nsp - 1;
    for (i = 1; i <= c->ne; i++) {
        double rho = WA(1, i), u = WA(2, i) / rho, v = WA(3, i) / rho, y[RCO_MAXSP], s = 0.0, o[4];
        for (k = 0; k < ns1; k++) y[k] = WA(NFLOW + 1 + k, i) / rho;
        for (k = 0; k < ns1; k++) s = s + y[k];
        y[nsp - 1] = 1.0 - s;
        if (rco_newton(&c->m, WA(4, i) / rho - 0.5 * (u * u + v * v), y, PA(4, i), o))
            return fail(c, "update: temperature out of range (cell %d)", i);
        double t = o[0];
        double p = rho * o[3] * t;
        PA(1, i) = p; PA(2, i) = u; PA(3, i) = v; PA(4, i) = t;
        for (k = 0; k < ns1; k++) PA(NFLOW + 1 + k, i) = y[k];
    }
    return 0;
}

/* LU_SGS_steady, Time_LUSGS_steady.f90:32-204 */
int rco_lusgs_steady(rco_case *c)
{
    int i, k, nei, nvar = c->nvar;
    double dflux[RCO_MAXSP + NFLOW];
    for (i = 1; i <= c->ne; i++) {
        double d = 0.0;
        for (nei = 1; nei <= 3; nei++) d = d + 0.5 * c->alpaf[NEIGHF(nei, i) - 1];
        c->dd[i - 1] = d + c->vol[i - 1] / c->step[i - 1];
    }
    for (i = 1; i <= c->ne; i++) {
        for (k = 0; k < nvar; k++) dflux[k] = 0.0;
        for (nei = 1; nei <= 3; nei++) {
            int iface = NEIGHF(nei, i);
            double dx = VECX(1, iface), dy = VECX(2, iface);
            int NR = NEIGHBOR(2, iface);
            if (i == NR) { NR = NEIGHBOR(1, iface); dx = -dx; dy = -dy; }
            if (NR > 0 && NR < i)
                if (lusgs_neighbor(c, NR, dx, dy, c->alpaf[iface - 1], dflux)) return 1;
        }
        for (k = 1; k <= nvar; k++) HWA(k, i) = (HWA(k, i) - dflux[k - 1]) / c->dd[i - 1];
    }
    for (i = c->ne; i >= 1; i--) {
        for (k = 0; k < nvar; k++) dflux[k] = 0.0;
        for (nei = 1; nei <= 3; nei++) {
            int iface = NEIGHF(nei, i);
            double dx = VECX(1, iface), dy = VECX(2, iface);
            int NR = NEIGHBOR(2, iface);
            if (i == NR) { NR = NEIGHBOR(1, iface); dx = -dx; dy = -dy; }
            if (NR > 0 && NR > i)
                if (lusgs_neighbor(c, NR, dx, dy, c->alpaf[iface - 1], dflux)) return 1;
        }
        for (k = 1; k <= nvar; k++) HWA(k, i) = HWA(k, i) - dflux[k - 1] / c->dd[i - 1];
    }
    for (i = 1; i <= c->ne; i++)
        for (k = 1; k <= nvar; k++) WA(k, i) = WA(k, i) + HWA(k, i);
    return update_primitive(c);
}

/* lu_sgs_stage, Time_LUSGS_steady.f90:4-30 (reconstruction = gradient +
 * Limiter_Barth when limit_ID>0, fvm2d.f90:183-191) */
int rco_stage(rco_case *c)
{
    if (rco_timestep(c)) return 1;
    if (rco_gradient(c)) return 1;
    if (c->p.limit_id > 0 && rco_limiter(c)) return 1;
    if (rco_flux_sourceterm(c)) return 1;
    if (rco_lusgs_steady(c)) return 1;
    if (isnan(WA(1, 1))) return fail(c, "there is NAN in WA(1,1)");
    return 0;
}

/* Residual + step_residual normalisation, residual.f90:4-21,61-79 */
void rco_residual(rco_case *c, double *res)
{
    int i, k, nvar = c->nvar;
    double romax = 0.0;
    for (k = 0; k < nvar; k++) res[k] = 0.0;
    for (i = 1; i <= c->ne; i++) {
        for (k = 1; k <= nvar; k++) res[k - 1] = res[k - 1] + c->vol[i - 1] * fabs(HWA(k, i));
        romax = fmax(romax, WA(1, i));
    }
    for (k = 0; k < nvar; k++) res[k] = res[k] / c->total_volume;
    res[nvar] = romax;
}


/* ------------------------------------------------------------------------ */
/* dual-time backward Euler with the point-implicit LU-SGS                  */
/* (implicit_time_method = 3): Time_BackwardEuler.f90                       */
/* ------------------------------------------------------------------------ */

/* Compute_Con_to_Pri_Jaco, Con_to_Pri_Jacobian.f90:1-87: dU/dPA, row-major
 * jaco[r][c] = J(r+1,c+1) */
int rco_con_to_pri_jaco(const rco_mech *m, const double *pri, double *jaco)
{
    int i, j, nsp = m->nsp, ns1 = nsp - 1, nv = NFLOW + ns1;
    double p = pri[0], u = pri[1], v = pri[2], t = pri[3], y[RCO_MAXSP], y2[RCO_MAXSP], s = 0.0, g[3], entha[RCO_MAXSP];
    for (i = 0; i < ns1; i++) { y[i] = pri[NFLOW + i]; y2[i] = y[i]; s = s + y[i]; }
    y2[nsp - 1] = 1.0 - s;
    if (rco_gas_param(m, y2, t, g)) return 1;
    double cp = g[0], rcp = g[2];
    double r = p / (rcp * t);
    if (rco_enthalpy(m, t, entha)) return 1;
    double h = entha[nsp - 1], dhdy[RCO_MAXSP];
    for (i = 0; i < ns1; i++) {
        h = h + y[i] * (entha[i] - entha[nsp - 1]);
        dhdy[i] = entha[i] - entha[nsp - 1];
    }
    double drdp = 1. / (rcp * t), drdt = -p / (rcp * t * t), drdy[RCO_MAXSP];
    const double rn = m->rconstant / m->mw[nsp - 1];
    for (i = 0; i < ns1; i++) drdy[i] = -(m->rconstant / m->mw[i] - rn) * p / (t * rcp * rcp);
    const double e2 = h + 0.5 * (u * u + v * v);
    for (i = 0; i < nv * nv; i++) jaco[i] = 0.0;
#define J(a, b) jaco[(a) * nv + (b)]
    J(0, 0) = drdp; J(0, 3) = drdt;
    J(1, 0) = u * drdp; J(1, 1) = r; J(1, 3) = u * drdt;
    J(2, 0) = v * drdp; J(2, 2) = r; J(2, 3) = v * drdt;
    J(3, 0) = e2 * drdp - 1.; J(3, 1) = r * u; J(3, 2) = r * v; J(3, 3) = r * cp + e2 * drdt;
    for (j = 0; j < ns1; j++) {
        J(0, NFLOW + j) = drdy[j];
        J(1, NFLOW + j) = u * drdy[j];
        J(2, NFLOW + j) = v * drdy[j];
        J(3, NFLOW + j) = e2 * drdy[j] + r * dhdy[j];
    }
    for (i = 0; i < ns1; i++) {
        J(NFLOW + i, 0) = drdp * y[i];
        J(NFLOW + i, 3) = drdt * y[i];
        for (j = 0; j < ns1; j++) J(NFLOW + i, NFLOW + j) = (i == j) ? r + y[i] * drdy[j] : y[i] * drdy[j];
    }
#undef J
    return 0;
}

/* compute_derivative_production_rate, chemical_kinetics.f90:167-265:
 * dwdp[i][k] = d omega_i / d PA_k, i < nsp-1.  In the model-1 branch the
 * reference never assigns dkfdy(3) (:213-236); it is taken as 0 here. */
int rco_dwdp(const rco_mech *m, const double *pri, double *dwdp)
{
    int i, nsp = m->nsp, ns1 = nsp - 1, nv = NFLOW + ns1;
    double p = pri[0], T = pri[3], y[RCO_MAXSP], y2[RCO_MAXSP], s = 0.0, g[3];
    for (i = 0; i < ns1; i++) { y[i] = pri[NFLOW + i]; y2[i] = y[i]; s = s + y[i]; }
    y2[nsp - 1] = 1.0 - s;
    if (rco_gas_param(m, y2, T, g)) return 1;
    double rcp = g[2];
    double rho = p / (rcp * T);
    for (i = 0; i < ns1; i++) y[i] = fmin(1., fmax(0., y[i]));
    double dkf[RCO_MAXSP + NFLOW];
    for (i = 0; i < nv; i++) dkf[i] = 0.0;
    if (!(y[0] < 1.0e-12)) {
        double af = m->a_f * exp(-m->e_act / (m->rconstant * T));
        double r1 = y[0] * rho / (m->mw[0] * 1000), dr1dr = y[0] / (m->mw[0] * 1000), dr1dy1 = rho / (m->mw[0] * 1000);
        double r2 = y[1] * rho / (m->mw[1] * 1000), dr2dr = y[1] / (m->mw[1] * 1000), dr2dy2 = rho / (m->mw[1] * 1000);
        double drdp = rho / p, drdT = -rho / T, drdy[RCO_MAXSP];
        const double rn = m->rconstant / m->mw[nsp - 1];
        for (i = 0; i < ns1; i++) drdy[i] = -(m->rconstant / m->mw[i] - rn) * rho / rcp;
        double dkfdr, dkfdp, dkfdT, dkfdy[RCO_MAXSP];
        for (i = 0; i < ns1; i++) dkfdy[i] = 0.0;
        if (m->model == 0) {
            double dter1_dr = 0.2 * pow(rho, -0.8) * pow(dr1dr, 0.2);
            dkfdr = af * (pow(r1, 0.2) * 1.3 * pow(r2, 0.3) * dr2dr + dter1_dr * pow(r2, 1.3));
            dkfdp = dkfdr * drdp;
            dkfdT = af * pow(r1, 0.2) * pow(r2, 1.3) * m->e_act / (m->rconstant * T * T) + dkfdr * drdT;
            dkfdy[0] = af * 0.2 * pow(r1, -0.8) * dr1dy1 * pow(r2, 1.3) + dkfdr * drdy[0];
            dkfdy[1] = af * pow(r1, 0.2) * 1.3 * pow(r2, 0.3) * dr2dy2 + dkfdr * drdy[1];
            for (i = 2; i < ns1; i++) dkfdy[i] = dkfdr * drdy[i];
        } else {
            af = 0.0;
            if (T > 710.0) af = 1e4;
            double dter1_dr = 1 * powi(rho, 0) * powi(dr1dr, 1);
            dkfdr = af * (powi(r1, 1) * 0 * powi(r2, 0) * dr2dr + dter1_dr * powi(r2, 0));
            dkfdp = dkfdr * drdp;
            dkfdT = 0.;
            dkfdy[0] = af * 1 * powi(r1, 0) * dr1dy1 * powi(r2, 0) + dkfdr * drdy[0];
            dkfdy[1] = af * powi(r1, 1) * 0 * powi(r2, 0) * dr2dy2 + dkfdr * drdy[1];
        }
        dkf[0] = dkfdp; dkf[1] = 0.; dkf[2] = 0.; dkf[3] = dkfdT;
        for (i = 0; i < ns1; i++) dkf[NFLOW + i] = dkfdy[i];
    }
    for (i = 0; i < ns1; i++) {
        int k;
        for (k = 0; k < nv; k++) dwdp[i * nv + k] = 1000 * m->mw[i] * (m->coef_r[i] - m->coef_f[i]) * dkf[k];
    }
    return 0;
}

/* advanced_linear_solve_lapack (la_math.f90:120-213) calls LAPACK dgesvx
 * with FACT='E': equilibrate (dgeequ, applied as dlaqge decides), LU with
 * partial pivoting (dgetrf), solve (dgetrs), iterative refinement (dgerfs),
 * unscale.  Restated on the published algorithm of those routines for an
 * n x n system, a row-major. */
int rco_solve_dgesvx(int n, const double *a_in, const double *b_in, double *x)
{
    double a[64], af[64], b[8], r[8], cs[8], w[8], dx[8];
    int ipiv[8], i, j, k;
    const double eps = 1.1102230246251565e-16, sfmin = 2.2250738585072014e-308;
    const double smlnum = sfmin, bignum = 1.0 / smlnum;
    for (i = 0; i < n * n; i++) a[i] = a_in[i];
    for (i = 0; i < n; i++) b[i] = b_in[i];
    /* dgeequ */
    double rcmin = bignum, rcmax = 0.0;
    for (i = 0; i < n; i++) {
        r[i] = 0.0;
        for (j = 0; j < n; j++) r[i] = fmax(r[i], fabs(a[i * n + j]));
        rcmax = fmax(rcmax, r[i]);
        rcmin = fmin(rcmin, r[i]);
    }
    const double amax = rcmax;
    if (rcmin == 0.0) return 1;
    for (i = 0; i < n; i++) r[i] = 1.0 / fmin(fmax(r[i], smlnum), bignum);
    const double rowcnd = fmax(rcmin, smlnum) / fmin(rcmax, bignum);
    rcmin = bignum; rcmax = 0.0;
    for (j = 0; j < n; j++) {
        cs[j] = 0.0;
        for (i = 0; i < n; i++) cs[j] = fmax(cs[j], fabs(a[i * n + j]) * r[i]);
        rcmin = fmin(rcmin, cs[j]);
        rcmax = fmax(rcmax, cs[j]);
    }
    if (rcmin == 0.0) return 1;
    for (j = 0; j < n; j++) cs[j] = 1.0 / fmin(fmax(cs[j], smlnum), bignum);
    const double colcnd = fmax(rcmin, smlnum) / fmin(rcmax, bignum);
    /* dlaqge */
    const double thresh = 0.1, small = sfmin / (2 * eps), large = 1.0 / small;
    int rows = 0, cols = 0;
    if (rowcnd >= thresh && amax >= small && amax <= large) {
        cols = colcnd < thresh;
    } else {
        rows = 1;
        cols = colcnd < thresh;
    }
    for (j = 0; j < n; j++)
        for (i = 0; i < n; i++) {
            if (rows && cols) a[i * n + j] = cs[j] * r[i] * a[i * n + j];
            else if (rows) a[i * n + j] = r[i] * a[i * n + j];
            else if (cols) a[i * n + j] = cs[j] * a[i * n + j];
        }
    if (rows)
        for (i = 0; i < n; i++) b[i] = r[i] * b[i];
    /* dgetf2 */
    for (i = 0; i < n * n; i++) af[i] = a[i];
    for (j = 0; j < n; j++) {
        int p = j;
        for (i = j + 1; i < n; i++)
            if (fabs(af[i * n + j]) > fabs(af[p * n + j])) p = i;
        ipiv[j] = p;
        if (af[p * n + j] == 0.0) return 1;
        if (p != j)
            for (k = 0; k < n; k++) { double t = af[j * n + k]; af[j * n + k] = af[p * n + k]; af[p * n + k] = t; }
        if (fabs(af[j * n + j]) >= sfmin) {
            const double rp = 1.0 / af[j * n + j];
            for (i = j + 1; i < n; i++) af[i * n + j] = rp * af[i * n + j];
        } else {
            for (i = j + 1; i < n; i++) af[i * n + j] = af[i * n + j] / af[j * n + j];
        }
        for (k = j + 1; k < n; k++) {
            const double t = -af[j * n + k];
            for (i = j + 1; i < n; i++) af[i * n + k] = af[i * n + k] + t * af[i * n + j];
        }
    }
    /* dgetrs on a right-hand side v (in place) */
#define GETRS(v)                                                                   \
    do {                                                                           \
        for (i = 0; i < n; i++)                                                    \
            if (ipiv[i] != i) { double t = v[i]; v[i] = v[ipiv[i]]; v[ipiv[i]] = t; } \
        for (k = 0; k < n; k++)                                                    \
            if (v[k] != 0.0)                                                       \
                for (i = k + 1; i < n; i++) v[i] = v[i] - v[k] * af[i * n + k];    \
        for (k = n - 1; k >= 0; k--)                                               \
            if (v[k] != 0.0) {                                                     \
                v[k] = v[k] / af[k * n + k];                                       \
                for (i = 0; i < k; i++) v[i] = v[i] - v[k] * af[i * n + k];        \
            }                                                                      \
    } while (0)
    for (i = 0; i < n; i++) x[i] = b[i];
    GETRS(x);
    /* dgerfs: at most 5 refinement steps */
    double lstres = 3.0;
    for (int count = 1; count <= 5; count++) {
        for (i = 0; i < n; i++) w[i] = b[i];
        for (j = 0; j < n; j++) {
            const double t = -x[j];
            for (i = 0; i < n; i++) w[i] = w[i] + t * a[i * n + j];
        }
        double berr = 0.0;
        for (i = 0; i < n; i++) {
            double den = fabs(b[i]);
            for (j = 0; j < n; j++) den = den + fabs(a[i * n + j]) * fabs(x[j]);
            const double safe1 = (n + 1) * sfmin, safe2 = safe1 / eps;
            berr = fmax(berr, den > safe2 ? fabs(w[i]) / den : (fabs(w[i]) + safe1) / (den + safe1));
        }
        if (!(berr > eps && 2.0 * berr <= lstres)) break;
        for (i = 0; i < n; i++) dx[i] = w[i];
        GETRS(dx);
        for (i = 0; i < n; i++) x[i] = x[i] + dx[i];
        lstres = berr;
    }
#undef GETRS
    if (cols)
        for (j = 0; j < n; j++) x[j] = cs[j] * x[j];
    return 0;
}

/* the neighbour term of backwardEuler_LU_SGS (Time_BackwardEuler.f90:
 * 131-176): the neighbour's primitive increment hWA(:,NR) -> conserved
 * increment, then the linearised Rusanov flux change */
static int be_neighbor(rco_case *c, int NR, double dx, double dy, double af, double *dflux)
{
    int k, nsp = c->nsp, ns1 = nsp - 1;
    double pp = PA(1, NR), uu = PA(2, NR), vv = PA(3, NR), tt = PA(4, NR), yy[RCO_MAXSP], s = 0.0, g[3],
           entha[RCO_MAXSP];
    for (k = 0; k < ns1; k++) { yy[k] = PA(NFLOW + 1 + k, NR); s = s + yy[k]; }
    yy[nsp - 1] = 1.0 - s;
    if (rco_gas_param(&c->m, yy, tt, g)) return fail(c, "backward Euler: temperature %g out of range", tt);
    if (rco_enthalpy(&c->m, tt, entha)) return fail(c, "backward Euler: temperature %g out of range", tt);
    double cp = g[0], rcp = g[2], hh = 0.0;
    for (k = 0; k < nsp; k++) hh = hh + yy[k] * entha[k];
    double rr = pp / (rcp * tt), rou = rr * uu, rov = rr * vv;
    double roe = rr * (hh + 0.5 * (uu * uu + vv * vv)) - pp;
    double ry[RCO_MAXSP], dyy[RCO_MAXSP], dry[RCO_MAXSP];
    for (k = 0; k < nsp; k++) ry[k] = rr * yy[k];
    double dpp = HWA(1, NR), duu = HWA(2, NR), dvv = HWA(3, NR), dtt = HWA(4, NR);
    s = 0.0;
    for (k = 0; k < ns1; k++) { dyy[k] = HWA(NFLOW + 1 + k, NR); s = s + dyy[k]; }
    dyy[nsp - 1] = -s;
    double s1 = 0.0, s2 = 0.0;
    for (k = 0; k < nsp; k++) s1 = s1 + (c->m.rconstant / c->m.mw[k]) * dyy[k];
    for (k = 0; k < nsp; k++) s2 = s2 + entha[k] * dyy[k];
    double drr = (dpp - rr * rcp * dtt - rr * tt * s1) / (rcp * tt);
    double dru = rr * duu + uu * drr;
    double drv = rr * dvv + vv * drr;
    double dre = rr * cp * dtt + rr * s2 + hh * drr + rr * (uu * duu + vv * dvv) + 0.5 * (uu * uu + vv * vv) * drr - dpp;
    for (k = 0; k < nsp; k++) dry[k] = rr * dyy[k] + yy[k] * drr;
    double un = uu * dx + vv * dy, dun = duu * dx + dvv * dy;
    dflux[0] = dflux[0] + 0.5 * (-af * drr + dru * dx + drv * dy);
    dflux[1] = dflux[1] + 0.5 * (-af * dru + dru * un + rou * dun + dpp * dx);
    dflux[2] = dflux[2] + 0.5 * (-af * drv + drv * un + rov * dun + dpp * dy);
    dflux[3] = dflux[3] + 0.5 * (-af * dre + dun * (roe + pp) + un * (dre + dpp));
    for (k = 0; k < ns1; k++) dflux[NFLOW + k] = dflux[NFLOW + k] + 0.5 * (-af * dry[k] + dry[k] * un + ry[k] * dun);
    return 0;
}

/* the block of cell i: dU/dPA*dd(i) - vol(i)*d(omega)/dPA in the species
 * rows (Time_BackwardEuler.f90:106-116) */
static int be_block(rco_case *c, int i, double *jm)
{
    int nv = c->nvar, j, k;
    double dw[RCO_MAXSP * (RCO_MAXSP + NFLOW)];
    if (rco_con_to_pri_jaco(&c->m, &PA(1, i), jm)) return fail(c, "backward Euler: temperature out of range (cell %d)", i);
    for (k = 0; k < nv * nv; k++) jm[k] = jm[k] * c->dd[i - 1];
    if (rco_dwdp(&c->m, &PA(1, i), dw)) return fail(c, "backward Euler: temperature out of range (cell %d)", i);
    for (j = 0; j < c->nsp - 1; j++)
        for (k = 0; k < nv; k++) jm[(NFLOW + j) * nv + k] = jm[(NFLOW + j) * nv + k] - c->vol[i - 1] * dw[j * nv + k];
    return 0;
}

/* backwardEuler_LU_SGS, Time_BackwardEuler.f90:58-324 */
/* backwardEuler_LU_SGS (Time_BackwardEuler.f90:58-324) and, with bdf2,
 * BDF2_LU_SGS(Inner_rk, phi) (Time_BDF2.f90:71-337): the same sweeps, the
 * BDF2 right-hand side and diagonal weighing the two previous steps' states
 * WP, WP2 with phi */
static int be_lusgs(rco_case *c, double dt, int inner, int bdf2, double phi)
{
    int i, k, nei, nvar = c->nvar;
    double dflux[RCO_MAXSP + NFLOW], jm[(RCO_MAXSP + NFLOW) * (RCO_MAXSP + NFLOW)], rhs[RCO_MAXSP + NFLOW],
           sol[RCO_MAXSP + NFLOW];
    if (inner == 1)
        for (i = 1; i <= c->ne; i++)
            for (k = 1; k <= nvar; k++) {
                const size_t q = (size_t)(i - 1) * nvar + k - 1;
                c->fixed[q] = bdf2 ? c->vol[i - 1] / dt * ((1 + 2 * phi) * c->wp[q] - phi * c->wp2[q])
                                   : c->vol[i - 1] / dt * WA(k, i);
            }
    for (i = 1; i <= c->ne; i++) {
        double d = 0.0;
        for (nei = 1; nei <= 3; nei++) d = d + 0.5 * c->alpaf[NEIGHF(nei, i) - 1];
        c->dd[i - 1] = bdf2 ? d + c->vol[i - 1] * (1. / c->step[i - 1] + (1 + phi) / dt)
                            : d + c->vol[i - 1] * (1. / c->step[i - 1] + 1. / dt);
    }
    for (i = 1; i <= c->ne; i++)
        for (k = 1; k <= nvar; k++)
            HWA(k, i) = bdf2 ? HWA(k, i) - (1 + phi) * c->vol[i - 1] / dt * WA(k, i) + c->fixed[(i - 1) * nvar + k - 1]
                             : HWA(k, i) - c->vol[i - 1] / dt * WA(k, i) + c->fixed[(i - 1) * nvar + k - 1];
    for (int pass = 0; pass < 2; pass++) {
        for (int ii = 1; ii <= c->ne; ii++) {
            i = pass == 0 ? ii : c->ne + 1 - ii;
            if (be_block(c, i, jm)) return 1;
            for (k = 0; k < nvar; k++) dflux[k] = 0.0;
            for (nei = 1; nei <= 3; nei++) {
                int iface = NEIGHF(nei, i);
                double dx = VECX(1, iface), dy = VECX(2, iface);
                int NR = NEIGHBOR(2, iface);
                if (i == NR) { NR = NEIGHBOR(1, iface); dx = -dx; dy = -dy; }
                if (NR > 0 && (pass == 0 ? NR < i : NR > i))
                    if (be_neighbor(c, NR, dx, dy, c->alpaf[iface - 1], dflux)) return 1;
            }
            for (k = 0; k < nvar; k++) rhs[k] = pass == 0 ? HWA(k + 1, i) - dflux[k] : dflux[k];
            if (rco_solve_dgesvx(nvar, jm, rhs, sol)) return fail(c, "error in routine dgesvx (cell %d)", i);
            for (k = 0; k < nvar; k++) HWA(k + 1, i) = pass == 0 ? sol[k] : HWA(k + 1, i) - sol[k];
        }
    }
    for (i = 1; i <= c->ne; i++)
        for (k = 1; k <= nvar; k++) PA(k, i) = PA(k, i) + HWA(k, i);
    return prim_to_cons(c);
}

int rco_be_lusgs(rco_case *c, double dt, int inner) { return be_lusgs(c, dt, inner, 0, 0.0); }
int rco_bdf2_lusgs(rco_case *c, double dt, int inner, double phi) { return be_lusgs(c, dt, inner, 1, phi); }

/* the prologue of BDF2_stage (Time_BDF2.f90:12-20): at the first step WP =
 * WA, WP2 = WP, phi = 0; later WP2 = WP, WP = WA, phi = 0.5; returns phi */
double rco_bdf2_begin_step(rco_case *c, int first)
{
    const size_t n = (size_t)c->nvar * c->ne;
    if (first) {
        memcpy(c->wp, c->wa, n * sizeof(double));
        memcpy(c->wp2, c->wp, n * sizeof(double));
        return 0.0;
    }
    memcpy(c->wp2, c->wp, n * sizeof(double));
    memcpy(c->wp, c->wa, n * sizeof(double));
    return 0.5;
}

/* backwardEuler_stage, Time_BackwardEuler.f90:4-56 (without its closing
 * diagnostic reconstruction/flux_sourceterm): inner iterations until RoRes
 * drops below relTol times its first value or InnerLoop is reached; the
 * number of inner iterations and their RoRes go to *ninner, roress[] */
int rco_be_stage(rco_case *c, double dt, int innerloop, double reltol, int *ninner, double *roress)
{
    double res[RCO_MAXSP + NFLOW + 1], ro1 = 0.0;
    int inner;
    for (inner = 1; inner <= innerloop; inner++) {
        if (rco_timestep(c)) return 1;
        if (rco_gradient(c)) return 1;
        if (c->p.limit_id > 0 && rco_limiter(c)) return 1;
        if (rco_flux_sourceterm(c)) return 1;
        if (rco_be_lusgs(c, dt, inner)) return 1;
        rco_residual(c, res);
        double ro = 0.0;
        for (int k = NFLOW; k < c->nvar; k++) ro = ro + res[k];
        if (roress) roress[inner - 1] = ro;
        if (ninner) *ninner = inner;
        if (inner == 1) ro1 = ro;
        if (ro < ro1 * reltol || inner == innerloop) {
            /* the closing reconstruction + flux_sourceterm of the converged
               inner loop (Time_BackwardEuler.f90:36-38): hWA and GradC then
               hold the residual and gradient of the new state, which
               step_residual reads (fvm2d.f90:137) */
            if (rco_gradient(c)) return 1;
            if (c->p.limit_id > 0 && rco_limiter(c)) return 1;
            if (rco_flux_sourceterm(c)) return 1;
            break;
        }
        if (isnan(WA(1, 1))) return fail(c, "there is NAN in WA(1,1)");
    }
    return 0;
}

/* p_out as compute_p_out sets it for the next evaluations (fvm2d.f90:229-239) */
void rco_set_p_out(rco_case *c, double p_out) { c->p.p_out = p_out; }
