/* rocketcombustor.h -- C-ABI of the MI355X (gfx950) FP64 steady residual +
 * LU-SGS path of RocketCombustor.
 *
 * The reference has no plugin or FFI layer: its hot operators are
 * argument-less Fortran subroutines sharing `module mainvar` state.  Each
 * entry point below replaces one of them for a host that keeps the
 * reference's call sequence (fvm2d.f90 main loop -> lu_sgs_stage); the
 * Fortran binding is include/rocketcombustor_iface.f90 (ISO_C_BINDING).
 *
 * Conventions
 *  - FP64 throughout, int32 indices.
 *  - Arrays crossing the boundary use the reference's column-major layout:
 *    PA(Nvar,NE), WA(Nvar,NE), hWA(Nvar,NE), GradC(2,Nvar,NE), alpaF(NF),
 *    step(NE); in C terms pa[i*nvar + k] is PA(k+1,i+1).
 *  - Node numbers in cell and boundary lists are 1-based, as in grid<N>.in
 *    and N(4,NE) (src/mesh.f90:20-58).
 *  - The library deep-copies at create/set, keeps all state resident in
 *    device memory between calls and copies back only on rc_get_* calls.
 *  - Calls are synchronous (stream-ordered, the stream is synchronised before
 *    return) unless documented otherwise.  A handle is not re-entrant;
 *    separate handles may be driven from separate host threads.
 *  - Every call returns 0 on success or an RC_ERR_* code; rc_last_error()
 *    gives the message.  Where the reference executes `stop`, the library
 *    returns the code instead (a Fortran host turns it into `stop`).
 */
#ifndef ROCKETCOMBUSTOR_H
#define ROCKETCOMBUSTOR_H
#ifdef __cplusplus
extern "C" {
#endif
/* The library is built with hidden symbols; these entry points are its
 * only exports, so no other library loaded in the process (an MPI
 * runtime, a profiler's tools, a framework) can interpose on its internals
 * or it on theirs. */
#pragma GCC visibility push(default)

#define RC_MAXSP 8

enum {
    RC_OK = 0,
    RC_ERR_ARGUMENT = 1,      /* invalid argument / unsupported input */
    RC_ERR_TEMPERATURE = 2,   /* T outside the polynomial range: thermo_property.f90:97-100,137-141 */
    RC_ERR_NEGATIVE = 3,      /* rho or p < 0: flux_lax.f90:239-242, TimeStep.f90:66-70 */
    RC_ERR_NAN = 4,           /* NaN in WA(1,1): Time_LUSGS_steady.f90:24-27 */
    RC_ERR_MESH = 5,          /* mesh set-up checks of mesh.f90 / ghostcell.f90 / reconstruction_initialization.f90 */
    RC_ERR_DEVICE = 6,        /* HIP runtime error */
    RC_ERR_MEMORY = 7,        /* device memory exhausted / budget exceeded */
    RC_ERR_SINGULAR = 8       /* zero pivot / zero row or column in a point-implicit block: la_math.f90:177-209 */
};

/* Species, thermo and kinetics tables: module species (species.f90:4-12),
 * thermo_property (thermo_property.f90:4-16), chemical_kinetics
 * (chemical_kinetics.f90:4-13), filled by setup_species_info,
 * setup_thermo_property and setup_chemical_kinetics. */
typedef struct rc_mechanism {
    int nsp;                            /* species count (Nvar = 4 + nsp - 1); the kernels are built for nsp = 4 */
    double mole_weight[RC_MAXSP];       /* kg/mol */
    double rconstant;                   /* J/mol/K */
    double t_low, t_crit, t_high;       /* polynomial ranges */
    double low_coef[RC_MAXSP][7];       /* cp/R coefficients of T^(j-3), j=1..7 */
    double high_coef[RC_MAXSP][7];
    double low_b1[RC_MAXSP], high_b1[RC_MAXSP];
    int model;                          /* code_chemical_model: 0 Arrhenius, 1 threshold */
    double coef_f[RC_MAXSP], coef_r[RC_MAXSP];
    double a_f, beta, e_act;
} rc_mechanism;

/* Run parameters that the reference keeps in module globals: CFL
 * (startup.f90:91), limit_ID (:108), If_turn_on_combustion (:61), p_out
 * (initialization.f90:32-33), the inlet states of oxidizer.in / fuel.in with
 * the mass flux after initialization.f90:22-23. */
typedef struct rc_params {
    double cfl;
    int limit_id;                       /* >0: Barth limiter after the gradient */
    int combustion;                     /* 0: no source term (sourceterm.f90:9) */
    double p_out;
    double ox_y[RC_MAXSP], ox_t, ox_mass_flux;
    double fu_y[RC_MAXSP], fu_t, fu_mass_flux;
} rc_params;

typedef struct rc_case rc_case;
typedef struct rc_batch rc_batch;

/* Replaces ReadMesh (mesh.f90:2-115: faces by `change`, BC tags,
 * VolumeAndCenter) and compute_reconstruction_matrix
 * (reconstruction_initialization.f90:1-169: NeighF/NeighC, ghost cells of
 * ghostcell.f90:19-195, inverse-distance weights, pinv(A) of la_math.f90:1-62)
 * for a one-rank triangle mesh.
 *   coor[2*nnt]      node coordinates Coor(2,NNT)
 *   cellnodes[3*ne]  1-based counter-clockwise triangle nodes (N(1:3,NE))
 *   bnodes[2*nb]     1-based node pairs of boundary faces, btype[nb] their
 *                    reference BC type (1 wall, 2 symmetry, 3 oxidizer inlet,
 *                    4 fuel inlet, 5 outlet, 6 slip wall)
 *   device           HIP device ordinal
 * The cell order given is the LU-SGS sweep order, exactly as in the
 * reference (Time_LUSGS_steady.f90:56,127); see rc_mesh_order_multicolor. */
int rc_case_create(const rc_mechanism *mech, const rc_params *par,
                   int nnt, const double *coor, int ne, const int *cellnodes,
                   int nb, const int *bnodes, const int *btype,
                   int device, rc_case **out);
/* rc_case_create for grids with periodic boundaries: replaces, in addition,
 * the periodic part of the grid reader (mesh.f90:60-64: NCon lines
 * Con(1:2), dxyz(1:2)) and CreatePeriodConnect (period.f90:15-104).
 * btype values >= 100 are labels of the boundaries the connections name
 * (their faces in the grid's order); connection c takes every face of
 * boundary con[2c] to the face of boundary con[2c+1] whose midpoint is its
 * own moved by (dxyz[2c], dxyz[2c+1]), matched as period.f90:76-87 matches
 * them.  Such a face gets the owner of its partner as neighbour,
 * FProperty 100 and SideOfs dxyz: flux_lax and flux_viscous update its
 * owner only (flux_lax.f90:289-293), the reconstruction and both fluxes
 * move the partner by SideOfs (reconstruction_initialization.f90:40,
 * flux_lax.f90:86), and TimeStep adds its alpha to both cells.  A face of a
 * labelled boundary left without a partner is refused (flux_lax stops on
 * type 100). */
int rc_case_create_periodic(const rc_mechanism *mech, const rc_params *par,
                            int nnt, const double *coor, int ne, const int *cellnodes,
                            int nb, const int *bnodes, const int *btype,
                            int ncon, const int *con, const double *dxyz,
                            int device, rc_case **out);
void rc_case_destroy(rc_case *h);
const char *rc_last_error(const rc_case *h);           /* thread-local message when h is NULL */

/* sizes[0..5] = NE, NF, number of ghost cells, Nvar, number of LU-SGS
 * wavefront levels, resident device bytes. */
int rc_case_info(const rc_case *h, long long *sizes6);
/* No counterpart in the reference: cells[l] = the number of cells in LU-SGS
 * wavefront level l, for l < n (n at least the level count of
 * rc_case_info); for measurement, whose byte counts follow the levels. */
int rc_case_levels(const rc_case *h, int *cells, int n);

/* Replaces the PA load of Read_Backup (output.f90:276-338) followed by
 * primitve_to_conserved_variables (update_flow_variables.f90:50-84). */
int rc_set_primitive(rc_case *h, const double *pa);

/* Replaces TimeStep, steady inviscid branch (TimeStep.f90:4-103): alpaF,
 * local step(i) = CFL*vol/sum(alpaF). */
int rc_timestep(rc_case *h);
/* Replaces gradient (gradient.f90:4-45) including
 * update_ghostcell_variables (update_flow_variables.f90:90-269). */
int rc_gradient(rc_case *h);
/* Replaces Limiter_Barth (limiter.f90:5-107), applied to the current GradC. */
int rc_limiter_barth(rc_case *h);
/* Replaces reconstruction (fvm2d.f90:183-191): gradient, then
 * Limiter_Barth when limit_ID>0, fused in one kernel. */
int rc_reconstruction(rc_case *h);
/* Replaces flux_sourceterm (fvm2d.f90:196-224): hWA = 0, flux_lax
 * (flux_lax.f90:4-299), sourceterm (sourceterm.f90:2-70). */
int rc_flux_sourceterm(rc_case *h);
/* Replaces LU_SGS_steady (Time_LUSGS_steady.f90:32-204) including
 * update_primitve_variables (update_flow_variables.f90:4-48). */
int rc_lusgs_steady(rc_case *h);
/* Replaces lu_sgs_stage (Time_LUSGS_steady.f90:4-30): the four calls above
 * and the NaN check on WA(1,1). */
int rc_lusgs_stage(rc_case *h);

/* Replaces the steady branch of the main loop (fvm2d.f90:48-150 with
 * IfSteady, .not.IfExplicit): nsteps calls of lu_sgs_stage, and every
 * `every` steps (and at step 1) the step_residual norms
 * (residual.f90:61-79): res_out[(n)*(nvar+1) + k] holds Res(k)/total_volume
 * and, at k = nvar, max(rho), for the n-th reported step.  Runs
 * asynchronously between reports; the device error flags are checked at
 * each report and at the end.  `res_out` may be NULL (then there is no
 * report at step 1, only every `every` steps).  Returns the number of
 * completed steps in *done.  The steps replay captured HIP graphs of up to
 * 8 steps (RC_GRAPH_STEPS sets the largest, RC_NO_GRAPH=1 launches every
 * kernel); with RC_PERSIST=n set when the case is created, up to n steps
 * run as one cooperative launch instead (same results, measured slower). */
int rc_run_steady(rc_case *h, int nsteps, int every, double *res_out, int *done);

/* ---- viscous path (IfViscous, startup.f90:76).
 * The tables of transport.in and diffusion.in as setup_transport_property
 * (transport_property.f90:18-45) and setup_diffusion_velocity
 * (diffusion_velocity.f90:15-38) read them, species in the mechanism's
 * order. */
typedef struct rc_transport {
    double t_low, t_crit, t_high;          /* transport table ranges (K) */
    double mu_low[RC_MAXSP][4];            /* ln mu[micropoise] = A ln T + B/T + C/T^2 + D, T <= t_crit */
    double mu_high[RC_MAXSP][4];           /* the same above t_crit */
    double k_low[RC_MAXSP][4];             /* ln K[microW/cm/K], same form */
    double k_high[RC_MAXSP][4];
    double k_b;                            /* Boltzmann's constant as read (unused by the formulas) */
    double sigma[RC_MAXSP];                /* collision diameter (Angstrom) */
    double eps[RC_MAXSP];                  /* Lennard-Jones energy / k_B (K) */
} rc_transport;

/* Switches the case to the viscous path with these tables: from the next
 * call on, TimeStep adds the viscous spectral radius (TimeStep.f90:73-85)
 * and flux_sourceterm adds flux_viscous (flux_viscous.f90:4-350) after
 * flux_lax, in the steady step, the batch step and the backward-Euler
 * stage alike. */
int rc_set_transport(rc_case *h, const rc_transport *tr);

/* ---- learned chemistry (no reference counterpart): an FP32 MLP for the
 * species source term, omega_1..omega_{Nsp-1} from (T, p, Y_1..Y_Nsp), with
 * `hidden` tanh layers of `width`, trained on the device by the library's own
 * forward, backward and Adam kernels.  Parameters are one vector: per layer
 * W (fan_in x fan_out, row-major) then b. */
typedef struct rc_mlp rc_mlp;
int  rc_mlp_create(int device, int n_in, int width, int hidden, int n_out, rc_mlp **out);
void rc_mlp_destroy(rc_mlp *m);
const char *rc_mlp_last_error(const rc_mlp *m);
long long rc_mlp_num_params(const rc_mlp *m);
int  rc_mlp_set_params(rc_mlp *m, const float *params);      /* also resets the Adam moments */
int  rc_mlp_get_params(rc_mlp *m, float *params);
/* y[n x n_out] = f(x[n x n_in]) */
int  rc_mlp_forward(rc_mlp *m, const float *x, int n, float *y);
/* mean squared error of f(x) against t over the n x n_out outputs, and its
 * gradient with respect to the parameter vector (grad may be NULL) */
int  rc_mlp_gradient(rc_mlp *m, const float *x, const float *t, int n, float *grad, float *loss);
/* the training set, resident on the device (n rows) */
int  rc_mlp_set_data(rc_mlp *m, const float *x, const float *t, long long n);
/* nsteps Adam steps on minibatches of `batch` rows drawn with replacement
 * (SplitMix64 stream `seed`); every `every` steps (and the last) the batch
 * loss goes to loss_out (may be NULL) */
int  rc_mlp_train(rc_mlp *m, int nsteps, int batch, float lr, unsigned long long seed, int every, float *loss_out);

/* The species source term of the case's steady step (and of
 * rc_flux_sourceterm / the backward-Euler stage) from the network m: per
 * cell the inputs (T, log p, Y_1..Y_Nsp) standardised with xform, the
 * outputs mapped back by omega_k = kappa s_k sinh(y_k), and the result
 * accepted when the raw inputs lie in the training range and the rates keep
 * the reaction's mass balance, |sum omega| <= tol (sum |omega| + 1e-3 max
 * s_k); with rtol > 0 the rates must also lie within rtol of the exact FP64
 * rates, |omega_k - exact_k| <= rtol max(|exact_k|, 1e-3 s_k).  A rejected
 * cell takes the exact FP64 rate (sourceterm.f90:17-43).  xform = input mean
 * (2+Nsp), input std (2+Nsp), output scales s_k (Nsp-1), kappa (1), raw input
 * range low and high (2+Nsp each, on T, p, Y).  m = NULL restores the exact
 * source.  The network lives on the case's device, stays owned by the
 * caller, must outlive its use and is not trained while a case uses it
 * (the step reads its parameters on the case's stream); such a case steps
 * alone (not in a batch).  No reference counterpart. */
int rc_set_surrogate(rc_case *h, rc_mlp *m, const double *xform, double tol, double rtol);
/* A calibrated trust region for the learned source's gate, which then needs
 * no exact rate: a grid of dims3[0] x dims3[1] x dims3[2] cells over the
 * features (T, log p, Y_1) between lo3 and hi3; a cell's network rates are
 * accepted only when its features fall in a grid cell with accept[(i0*d1 +
 * i1)*d2 + i2] != 0 (besides the range and mass-balance tests), else it
 * takes the exact FP64 rate.  The caller marks the grid cells where the
 * network's error on held-out states stayed within its bound.  accept =
 * NULL removes the grid.  Needs rc_set_surrogate first; no reference
 * counterpart. */
int rc_set_surrogate_trust(rc_case *h, const int *dims3, const double *lo3, const double *hi3,
                           const unsigned char *accept);
/* accepted and evaluated cell source evaluations since rc_set_surrogate */
int rc_get_surrogate_stats(rc_case *h, long long *accepted_evaluated2);

/* ---- dual-time backward Euler with the point-implicit LU-SGS
 * (implicit_time_method = 3; fvm2d.f90:60-110 unsteady branch). */

/* Replaces compute_p_out (fvm2d.f90:229-239): the exit pressure the ghost
 * update uses from the next call on (the host evaluates
 * p_static*(1 + A_p*sin(2*pi*f_p*ttime)) and passes the value). */
int rc_set_outlet_pressure(rc_case *h, double p_out);

/* Replaces backwardEuler_LU_SGS(Inner_rk) (Time_BackwardEuler.f90:58-324),
 * after rc_timestep, rc_reconstruction and rc_flux_sourceterm: at inner == 1
 * fixed = vol/dt*WA; dd = sum 0.5*alpaF + vol*(1/step + 1/dt); hWA = hWA -
 * vol/dt*WA + fixed; forward and backward point-implicit sweeps solving each
 * cell's block dU/dPA*dd - vol*d(omega)/dPA as dgesvx('E') does; PA += hWA;
 * WA = U(PA).  hWA holds the primitive increment afterwards. */
int rc_backward_euler_lusgs(rc_case *h, double dt, int inner);

/* Replaces backwardEuler_stage(istep) (Time_BackwardEuler.f90:4-56): inner
 * iterations of TimeStep, reconstruction, flux_sourceterm,
 * backwardEuler_LU_SGS and inner_iteration_residual (residual.f90:23-59)
 * until RoRes = sum of the species norms drops below relTol times its value
 * at the first inner iteration, or innerloop iterations.  res_out (may be
 * NULL) receives one row per inner iteration: Res(k)/total_volume, k <
 * nvar, then max(rho); *ninner the number of inner iterations.  After the
 * last inner iteration the closing reconstruction + flux_sourceterm
 * (:36-38) run as in the reference, so hWA and GradC hold the residual and
 * limited gradient of the new state (what step_residual reads, fvm2d.f90:137;
 * rc_residual returns its norms).  The exit pressure is the last one set. */
int rc_backward_euler_stage(rc_case *h, double dt, int innerloop, double reltol, int *ninner, double *res_out);

/* ---- BDF2 with the same point-implicit LU-SGS (implicit_time_method = 5). */

/* The prologue of BDF2_stage(istep) (Time_BDF2.f90:12-20): with first
 * (istep == 1) WP = WA, WP2 = WP and *phi = 0; else WP2 = WP, WP = WA and
 * *phi = 0.5.  WP and WP2 stay resident for rc_bdf2_lusgs. */
int rc_bdf2_begin_step(rc_case *h, int first, double *phi);
/* Replaces BDF2_LU_SGS(Inner_rk, phi) (Time_BDF2.f90:71-337): as
 * rc_backward_euler_lusgs, with fixed = vol/dt*((1+2phi)WP - phi WP2) at
 * inner == 1, dd = sum 0.5*alpaF + vol*(1/step + (1+phi)/dt) and the
 * right-hand side hWA - (1+phi)*vol/dt*WA + fixed. */
int rc_bdf2_lusgs(rc_case *h, double dt, int inner, double phi);
/* Replaces BDF2_stage(istep) (Time_BDF2.f90:4-69): rc_bdf2_begin_step
 * (first = istep == 1), then the inner iterations, the stopping rule, the
 * rows and the closing reconstruction + flux_sourceterm of
 * rc_backward_euler_stage. */
int rc_bdf2_stage(rc_case *h, double dt, int innerloop, double reltol, int first, int *ninner, double *res_out);

/* Replaces Residual + the normalisation of step_residual (residual.f90:4-21,
 * 61-79): res[k] = sum_i vol_i |hWA(k,i)| / total_volume, res[nvar] = max rho. */
int rc_residual(rc_case *h, double *res);

/* Copies of the resident arrays, reference layout (see Conventions). */
int rc_get_primitive(rc_case *h, double *pa);           /* PA(Nvar,NE) */
int rc_get_conserved(rc_case *h, double *wa);           /* WA(Nvar,NE) */
int rc_get_increment(rc_case *h, double *hwa);          /* hWA(Nvar,NE): residual after flux_sourceterm, dW after LU-SGS */
int rc_get_gradient(rc_case *h, double *gradc);         /* GradC(2,Nvar,NE) */
int rc_get_timestep(rc_case *h, double *step, double *alpaf); /* step(NE), alpaF(NF); either may be NULL */
int rc_get_ghost(rc_case *h, double *gpa);              /* GPA(Nvar,Nghost) */
/* The derived cell fields of output_vtk (output.f90:5-227) for the current
 * state: rho(NE), Ma(NE) and Production_rate(Nsp-1,NE) as sourceterm forms
 * it (sourceterm.f90:17-43; zero without combustion).  Any may be NULL. */
int rc_get_derived(rc_case *h, double *rho, double *mach, double *prod);
/* rc_get_derived without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernel stages its output in a static device array of 524288 cells x
 * (Nsp+1) doubles per GPU (held until the process ends); cases with more
 * cells return RC_ERR_ARGUMENT.  Calls of one process share the array and
 * are serialised by a lock; outputs as rc_get_derived. */
int rc_get_derived_noalloc(rc_case *h, double *rho, double *mach, double *prod);

/* Per-kernel device time of the fused step summed over the steps run while
 * timing was on, in ms: t[0] gradient + Barth limiter (ghost states formed
 * in-kernel), t[1] flux faces (+ TimeStep spectral radius), t[2] a separate
 * residual assembly (empty: the fused step forms the residual, source,
 * local step and LU diagonal in its forward sweep), t[3] LU-SGS forward
 * sweep of the wavefronts below the highest, t[4] the highest wavefront
 * (forward and, in one launch, backward sweep; forward only with
 * RC_TOPFUSE=0), t[5] the other backward wavefronts (all of them with
 * RC_TOPFUSE=0), t[6] update; t[7] the number of timed steps.  With the
 * opt-in dataflow sweep (RC_FLOW=1) both sweeps are in t[3].  Collected with
 * HIP events recorded on the case's stream between the kernels, without
 * synchronising; with timing off, rc_run_steady replays the steps as
 * captured HIP graphs (of up to 8 steps) instead. */
int rc_set_timing(rc_case *h, int enable);
int rc_get_timing(rc_case *h, double *t8);

/* No counterpart in the reference: a host-side utility that returns a
 * cell permutation (perm[k] = old 0-based index of new cell k) for block
 * multicolour sweeps: cells are grouped into blocks of up to `block`
 * face-connected cells grown in reverse Cuthill-McKee order, the block graph
 * is coloured greedily, and cells are ordered by (block colour, block,
 * RCM rank).  The LU-SGS sweeps of the reordered mesh then run as at most
 * block*ncolors wavefronts while keeping Gauss-Seidel coupling inside each
 * block (block = 1 is plain per-cell colouring, which stalls the steady
 * nozzle; block >= 2 converges like the serial order).  Writing the same
 * order into grid1.in makes the reference's serial sweeps perform the
 * identical computation. */
int rc_mesh_order_multicolor(int ne, const int *cellnodes, int block, int *perm, int *ncolors);

/* Batches of resident cases on one device, stepped together.  No
 * counterpart in the reference (one case per run there); for hosts that
 * solve many cases at once (parameter sweeps, serving).  Every launch of a
 * batch step covers all members, so many small cases fill the GPU; each
 * member's arithmetic is its own single-case step, so its results are
 * bitwise the same alone and in any batch.  Between batch calls the members
 * are ordinary cases (rc_get_*, rc_set_primitive); a case belongs to at most
 * one batch and is removed from it when destroyed. */
int rc_batch_create(int device, rc_batch **out);
void rc_batch_destroy(rc_batch *b);                    /* members stay valid */
int rc_batch_add(rc_batch *b, rc_case *h);
int rc_batch_remove(rc_batch *b, rc_case *h);
/* info3[0..2] = members, total cells, most LU-SGS levels of a member. */
int rc_batch_info(const rc_batch *b, long long *info3);
/* Launch shape of the batch's step (applied at its next build): threads per
 * block of the cell, face and ghost kernels and of the LU-SGS level kernels
 * (64, 128 or 256; kRunMax = 4 lanes per run), and XCD banding of the work
 * lists (1 on, 0 off, -1 as the members were laid out).  Results do not
 * depend on it: every cell and face is computed by one thread in the same
 * order whatever the shape.  The autotuner's knob; no reference counterpart
 * (the reference steps one case with serial loops). */
int rc_batch_set_launch(rc_batch *b, int cell_threads, int run_threads, int xcd);
/* rc_run_steady for every member at once: nsteps steps; every `every` steps
 * and at step 1 (when res_out is given; reports check the device error
 * flags), res_out[(n*members + c)*(nvar+1) + k] = the step_residual
 * norms of member c (member order = order of rc_batch_add, after removals).
 * On a device error of a member the run stops at that report with the
 * member's code and *err_case = its index. */
int rc_batch_run_steady(rc_batch *b, int nsteps, int every, double *res_out, int *done, int *err_case);

#pragma GCC visibility pop
#ifdef __cplusplus
}
#endif
#endif
