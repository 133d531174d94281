/*
 * hclpp_oracle.h -- CPU restatement of the HERACLES++ hydro step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * (heraclespp_amd/, libhclpp_hip.so) is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  It is
 * never linked into, called by or used as a fallback for the product path.
 *
 * Every function restates the reference algorithm in the reference's own
 * operation order (unfused kernels, same accumulation order, no FMA
 * contraction: build with -ffp-contract=off), citing the reference file:line
 * it follows.  Arrays use the reference layout: Kokkos::LayoutLeft over the
 * ghost-inclusive extents, i fastest, component indices slowest.
 */
#ifndef HCLPP_ORACLE_H
#define HCLPP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_GEOM_CARTESIAN = 0, OR_GEOM_SPHERICAL = 1 };
enum { OR_LIM_CONSTANT = 0, OR_LIM_VANLEER = 1, OR_LIM_MINMOD = 2, OR_LIM_VANALBADA = 3 };
enum { OR_RS_HLL = 0, OR_RS_HLLC = 1, OR_RS_LOWMACH = 2 };
enum { OR_GRAV_UNIFORM = 0, OR_GRAV_POINT_MASS = 1, OR_GRAV_INTERNAL_MASS = 2 };
enum { OR_BC_PERIODIC = 0, OR_BC_NULLGRADIENT = 1, OR_BC_REFLEXIVE = 2 };

typedef struct {
    int ndim;          /* 1..3 */
    int geom;          /* OR_GEOM_* */
    int limiter;       /* OR_LIM_* */
    int riemann;       /* OR_RS_* */
    int gravity;       /* OR_GRAV_* */
    int nfx;           /* passive scalars */
    int ng;            /* ghost layers on active dims (2) */
    int n[3];          /* interior cells per dim (1 on inactive dims) */
    int bc[6];         /* per face (2*idim+iface): OR_BC_* */
    int bc_order[6];   /* order in which border BCs are applied */
    double gamma, mmw;
    double g[3];       /* uniform gravity */
    double central_mass;
    const double* x[3]; /* node coordinates, nwg[d]+1 entries each (host) */
    int pressure_fix;   /* main.cpp:459: "Pressure fix" = On */
    double eps_pf;      /* param.cpp:71 (default 1e-6) */
} or_config;

typedef struct or_ctx or_ctx;

/* scalar physics (ndim given explicitly; arrays of ndim+2 doubles: rho, rhou[ndim], E) */
double or_evol_from_pres(double gamma, double mmw, double rho, double P);
double or_pres_from_evol(double gamma, double mmw, double rho, double evol);
double or_speed_of_sound(double gamma, double rho, double P);
double or_ek_cons(int ndim, const double* cons);
double or_ek_prim(int ndim, const double* prim);
void or_to_prim(int ndim, double gamma, double mmw, const double* cons, double* prim);
void or_to_cons(int ndim, double gamma, double mmw, const double* prim, double* cons);
void or_flux_prim(int ndim, double gamma, double mmw, const double* prim, int locdim, double* flux);
void or_flux_cons(int ndim, double gamma, double mmw, const double* cons, int locdim, double* flux);
void or_riemann(int solver, int ndim, double gamma, double mmw, const double* consL, const double* consR, int locdim, double* flux);
double or_limiter(int limiter, double diffR, double diffL);

/* context: geometry, gravity, and the driver's own state arrays */
or_ctx* or_create(const or_config* cfg);
void or_destroy(or_ctx* ctx);
const char* or_last_error(void);
int64_t or_ncells_wg(const or_ctx* ctx);
void or_extents(const or_ctx* ctx, int* nwg3);
/* geometry arrays: which = 0 dx0,1 dx1,2 dx2,3 xc0,4 xc1,5 xc2 (1D); 6 ds (4D, 3 comps); 7 dv (3D) */
const double* or_geometry(const or_ctx* ctx, int which);

/* array kernels on caller arrays (full ghost-inclusive extents). range: begin[3], end[3] */
void or_conv_prim_to_cons(const or_ctx* c, const int* range, const double* rho, const double* u, const double* P, double* rhou, double* E);
void or_conv_cons_to_prim(const or_ctx* c, const int* range, const double* rho, const double* rhou, const double* E, double* u, double* P);
void or_face_reconstruction(const or_ctx* c, const int* range, const double* var, double* var_rec);
void or_hancock(const or_ctx* c, const int* range, double dt, const double* u_rec, const double* P_rec,
                double* rho_rec, double* rhou_rec, double* E_rec, double* fx_rec);
void or_hydro_reconstruction(const or_ctx* c, const int* range, double dt, const double* rho, const double* u,
                             const double* P, const double* fx, double* rho_rec, double* rhou_rec, double* E_rec,
                             double* fx_rec, double* P_rec, double* u_rec);
void or_godunov(const or_ctx* c, const int* range, double dt, const double* rho, const double* rhou, const double* E,
                const double* fx, const double* rho_rec, const double* rhou_rec, const double* E_rec,
                const double* fx_rec, double* rho_new, double* rhou_new, double* E_new, double* fx_new);
double or_time_step(const or_ctx* c, const int* range, const double* rho, const double* u, const double* P);
double or_min_internal_energy(const or_ctx* c, const int* range, const double* rho, const double* rhou, const double* E);
double or_integrate(const or_ctx* c, const int* range, const double* var);
void or_boundaries(const or_ctx* c, double* rho, double* rhou, double* E, double* fx);
/* hydro/pressure_fix.hpp:19-172 (updates E_new in place) */
void or_pressure_fix(const or_ctx* c, const int* range, double dt, double eps, const double* rho, const double* rhou,
                     const double* E, const double* rho_rec, const double* rhou_rec, const double* E_rec,
                     const double* rho_new, const double* rhou_new, double* E_new);
/* gravity/gravity.cpp:59-128: rebuild the radial gravity table from rho (ghost-filled); g table out (nwg[0]) */
void or_internal_mass_gravity(or_ctx* c, const double* rho);
const double* or_gravity_table(const or_ctx* c);
void or_range(const or_ctx* c, int nghost_eff, int* range6);

/* driver (main.cpp:305-311, 354-531) on the context's own arrays */
int or_init_from_prim(or_ctx* c, const double* rho, const double* u, const double* P, const double* fx);
/* one step; returns 0 ok, 1 invalid internal energy; *exit_flag set when t_end/max_iter reached */
int or_step(or_ctx* c, double cfl, double t_end, int max_iter, double* dt_out, int* exit_flag);
/* or_step with the output schedule of main.cpp:362-385: *output_flag = make_output; returns 2
 * ("the time step is negative") when the next output time is not after t */
int or_step_outputs(or_ctx* c, double cfl, double t_end, int max_iter, double* dt_out, int* exit_flag, int* output_flag);
void or_set_outputs(or_ctx* c, int iter_output_frequency, double time_first_output, double time_output_frequency, int iter_output_id,
                    int time_output_id);
void or_output_ids(const or_ctx* c, int* iter_output_id, int* time_output_id);
/* which: 0 rho,1 rhou,2 E,3 fx,4 u,5 P */
double* or_state(or_ctx* c, int which);
double or_time(const or_ctx* c);
int or_iter(const or_ctx* c);
void or_set_time(or_ctx* c, double t, int iter);

#ifdef __cplusplus
}
#endif
#endif
