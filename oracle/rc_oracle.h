/* Test infrastructure only (not the product): a plain-C, serial restatement
 * of the reference solver's steady residual + LU-SGS path, used by tests/
 * and by bench.py's CPU-baseline leg as the checker.  Every routine follows
 * the reference loop order and operation order; see rc_oracle.c for the
 * file:line each one restates.
 *
 * Arrays mirror the reference's column-major Fortran arrays with 1-based
 * cell/face/node numbers kept as values (Neighbor, NeighF, NeighC hold
 * 1-based indices, boundary codes are negative) so the restatement reads
 * like the original.  Flat storage is 0-based: PA(k,i) -> pa[(i-1)*nvar+k-1].
 */
#ifndef RC_ORACLE_H
#define RC_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

#define RCO_MAXSP 8

typedef struct rco_mech {
    int nsp;
    double mw[RCO_MAXSP];
    double rconstant;
    double t_low, t_crit, t_high;
    double low[RCO_MAXSP][7], high[RCO_MAXSP][7];
    double low_b1[RCO_MAXSP], high_b1[RCO_MAXSP];
    int model;
    double coef_f[RCO_MAXSP], coef_r[RCO_MAXSP];
    double a_f, e_act;
} rco_mech;

typedef struct rco_params {
    double cfl;
    int limit_id;
    int combustion;
    double p_out;
    double ox_y[RCO_MAXSP], ox_t, ox_g;   /* oxidizer inlet: Y, T, mass flux */
    double fu_y[RCO_MAXSP], fu_t, fu_g;   /* fuel inlet */
} rco_params;

/* transport.in / diffusion.in (transport_property.f90:31-42,
 * diffusion_velocity.f90:30-35); mu, K = exp(A ln T + B/T + C/T^2 + D) */
typedef struct rco_transport {
    double t_low, t_crit, t_high;
    double mu_low[RCO_MAXSP][4], mu_high[RCO_MAXSP][4], k_low[RCO_MAXSP][4], k_high[RCO_MAXSP][4];
    double k_b, sigma[RCO_MAXSP], eps[RCO_MAXSP];
} rco_transport;

typedef struct rco_case rco_case;

int  rco_create(const rco_mech *mech, const rco_params *par, int nnt, const double *coor,
                int ne, const int *cells, int nb, const int *bedges, const int *btype,
                rco_case **out);
/* with periodic connections (mesh.f90:60-64, period.f90:15-104): btype
 * labels >= 100 name the boundaries; connection c joins boundary con[2c]
 * to con[2c+1] with translation dxyz[2c..2c+1] */
int  rco_create_periodic(const rco_mech *mech, const rco_params *par, int nnt, const double *coor,
                         int ne, const int *cells, int nb, const int *bedges, const int *btype,
                         int ncon, const int *con, const double *dxyz, rco_case **out);
void rco_destroy(rco_case *c);
const char *rco_last_error(const rco_case *c);
void rco_sizes(const rco_case *c, int *out4);            /* ne, nf, nghost, nvar */
double *rco_darray(rco_case *c, const char *name);
int *rco_iarray(rco_case *c, const char *name);

int  rco_set_pa(rco_case *c, const double *pa);           /* PA(nvar,ne) -> PA, WA */
int  rco_timestep(rco_case *c);
int  rco_gradient(rco_case *c);
int  rco_limiter(rco_case *c);
int  rco_flux_sourceterm(rco_case *c);
int  rco_lusgs_steady(rco_case *c);
int  rco_stage(rco_case *c);                              /* lu_sgs_stage */
void rco_residual(rco_case *c, double *res);              /* nvar L1 norms / total volume, then romax */

/* leaf thermochemistry, for unit tests */
int  rco_gas_param(const rco_mech *m, const double *y, double t, double *cp_gam_r);
int  rco_enthalpy(const rco_mech *m, double t, double *h);
int  rco_newton(const rco_mech *m, double et, const double *y, double t0, double *t_cp_gam_r);
void rco_production_rate(const rco_mech *m, double t, double rho, const double *y, double *w);
/* dual-time backward Euler (implicit_time_method = 3) */
int  rco_con_to_pri_jaco(const rco_mech *m, const double *pri, double *jaco_rowmajor);
int  rco_dwdp(const rco_mech *m, const double *pri, double *dwdp_rowmajor);
int  rco_solve_dgesvx(int n, const double *a_rowmajor, const double *b, double *x);
int  rco_be_lusgs(rco_case *c, double dt, int inner);
int  rco_bdf2_lusgs(rco_case *c, double dt, int inner, double phi);
double rco_bdf2_begin_step(rco_case *c, int first);
int  rco_be_stage(rco_case *c, double dt, int innerloop, double reltol, int *ninner, double *roress);
void rco_set_p_out(rco_case *c, double p_out);
/* viscous path (IfViscous): transport tables, then TimeStep's viscous
 * branch and flux_viscous inside flux_sourceterm */
void rco_set_transport(rco_case *c, const rco_transport *tr, int viscous);
int  rco_transport_property(const rco_mech *m, const rco_transport *tr, double t, const double *y,
                            double *vmul, double *akmu);
void rco_diffusion_velocity(const rco_mech *m, const rco_transport *tr, double p, double t, const double *y,
                            double *dk);

#ifdef __cplusplus
}
#endif
#endif
