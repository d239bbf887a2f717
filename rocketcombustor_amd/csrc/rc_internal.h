// Internal declarations shared by the host set-up (rc_mesh.cpp), the HIP
// kernels (rc_kernels.hip) and the C-ABI (rc_api.hip).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/rocketcombustor.h"

namespace rc {

constexpr int kNFlow = 4;          // src/module.f90:8
constexpr int kNsp = 4;            // species count the kernels are built for
constexpr int kNvar = kNFlow + kNsp - 1;
constexpr double kSmall = 1.0e-16; // src/module.f90:19
constexpr int kRunMax = 4;         // cells one thread sweeps in order (LU-SGS runs)

// Host-side mesh set-up: topology and geometry in the reference's numbering
// (0-based here), plus the derived per-face / per-cell operands the kernels read.
struct HostMesh {
    int ne = 0, nf = 0, ng = 0, nnt = 0;
    std::vector<double> coor;        // 2*nnt  (x,y)
    std::vector<int> cn;             // 3*ne   0-based nodes, CCW
    std::vector<int> fnode;          // 2*nf   face nodes (n1,n2) as created
    std::vector<int> fnl, fnr;       // nf     owner; neighbour (>=0) or -BCtype
    std::vector<double> vecx;        // 2*nf   (y2-y1, -(x2-x1))
    std::vector<double> vol, cellxy, reflen;   // ne, 2*ne, 2*ne
    std::vector<int> neighf;         // 3*ne   faces of each cell, ascending
    std::vector<int> neigh;          // ne     number of interior neighbours
    std::vector<int> neighc;         // 3*ne   interior neighbours (first neigh[i])
    std::vector<int> num_ghost;      // ne
    std::vector<int> ghostcell;      // 3*ne   ghost ids of a cell
    std::vector<int> ghost_host, ghost_face, ghost_type;  // ng
    std::vector<double> ghost_nodes; // 8*ng   (x,y) of the 4 mirrored nodes
    std::vector<double> ghost_vol, ghost_xy;  // ng, 2*ng
    std::vector<double> xyh, wdis, amls;      // 2*ne, 3*ne, 6*ne ([i][slot][r])
    double total_volume = 0.0;
    // derived operands
    std::vector<int> stencil;        // 3*ne  slot -> cell (>=0) or -(ghost+1)
    std::vector<double> lvv;         // 6*ne  limiter face-midpoint offsets [i][kf][d]
    std::vector<double> fvl, fvr;    // 2*nf  flux-point offsets of owner / neighbour
    std::vector<double> vgeo;        // 6*nf  viscous face operands: 1/reflen of NL (x,y), of NR (x,y), cellxy(NL)-(cellxy(NR)-SideOfs)
    // periodic face pairs (period.f90): a periodic face's neighbour is the
    // owner of its partner face, SideOfs its translation; empty without them
    std::vector<double> sideofs;     // 2*nf  SideOfs (0 on other faces)
    std::vector<char> periodic;      // nf    FProperty == 100
    std::vector<double> nofs;        // 6*ne  stencil offsets Ofs of the interior neighbours [i][k][d]
    std::vector<int> tsx;            // kTsFaces*ne  faces adding to the cell's TimeStep sum, in face order (-1 pad)
    std::vector<double> gnrm;        // 2*ng  unit normal used by the ghost update
    std::vector<int> cface;          // 3*ne  face f if cell is the owner, ~f otherwise
    std::vector<int> lu_nb;          // 3*ne  neighbour across each face (-1: boundary)
    std::vector<double> lu_n;        // 6*ne  face normal oriented out of the cell [i][k][d]
    // LU-SGS schedule: the cell order is cut into runs of up to kRunMax
    // consecutive, face-connected cells; one thread sweeps a run in order.
    std::vector<int> runs;           // 2*nruns (first cell, cell count), sorted by level
    std::vector<int> lvl_ptr;        // nlev+1  runs of level l: runs[lvl_ptr[l] .. lvl_ptr[l+1])
    // point-implicit (backward Euler) sweep schedule: a cell's level is 1 +
    // the highest level of its lower neighbours; one lane group solves one
    // cell, so the cells of a level are independent
    std::vector<int> pi_cells;       // ne      cells sorted by level (ascending within a level)
    std::vector<int> pi_lvl_ptr;     // npl+1   cells of level l: pi_cells[pi_lvl_ptr[l] .. pi_lvl_ptr[l+1])
};

// faces whose alpha a cell's TimeStep sum takes at most: its three, and up to
// three periodic faces of which it is the neighbour
constexpr int kTsFaces = 6;

// Builds HostMesh from 0-based inputs; returns RC_OK or RC_ERR_MESH with msg.
// Boundary labels btype >= 100 name the boundaries of the ncon periodic
// connections con[2c] -> con[2c+1], translation dxyz[2c..2c+1].
int build_mesh(int nnt, const double *coor, int ne, const int *cells0, int nb, const int *bedges0,
               const int *btype, int ncon, const int *con, const double *dxyz, HostMesh &m, std::string &msg);

// Multicolour ordering (no reference counterpart).
int order_multicolor(int ne, const int *cells0, int block, int *perm, int *ncolors);

// Thermo/kinetics tables on the device.
struct DevThermo {
    int nsp;
    int model;
    double rsp[RC_MAXSP];          // R_species
    double mw[RC_MAXSP];
    double t_low, t_crit, t_high;
    double low[RC_MAXSP][7], high[RC_MAXSP][7];
    double low_b1[RC_MAXSP], high_b1[RC_MAXSP];
    double dnu[RC_MAXSP];          // coef_r - coef_f of reaction 1
    double a_f, e_act, rconstant;
};

// Transport and diffusion tables on the device (viscous path), with the
// per-pair constants of the Wilke rule and of the binary diffusion
// coefficients formed once on the host.
struct DevTransport {
    double t_low, t_crit, t_high;
    double mu_low[kNsp][4], mu_high[kNsp][4], k_low[kNsp][4], k_high[kNsp][4];
    double wa[kNsp][kNsp];     // (mw_j/mw_i)^0.25
    double wb[kNsp][kNsp];     // sqrt(1 + mw_i/mw_j)
    double sqmw[kNsp][kNsp];   // sqrt(2/(1/mw_i + 1/mw_j)*1000)
    double cd2[kNsp][kNsp];    // (0.5*(sigma_i + sigma_j))^2
    double sqeps[kNsp][kNsp];  // sqrt(eps_i*eps_j)
};

// Run parameters on the device.
struct DevParams {
    double cfl;
    double p_out;
    double ox_y[RC_MAXSP], ox_t, ox_g;
    double fu_y[RC_MAXSP], fu_t, fu_g;
    int limit_id;
    int combustion;
    int viscous;                   // IfViscous
    int surrogate;                 // source term from the learned-chemistry array S.src
    int periodic;                  // the mesh has periodic faces: TimeStep sums over M.tsx
};

}  // namespace rc
