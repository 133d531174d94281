/*
 * hclpp_oracle.c -- CPU restatement of the HERACLES++ hydro step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see hclpp_oracle.h).  Not part of the
 * product, never a fallback for it.
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off -fno-fast-math -fPIC -shared
 * (oracle/Makefile).  -ffp-contract=off keeps every a*b+c as two roundings,
 * as in the reference's x86-64 CPU build (no FMA in its baseline ISA).
 *
 * Citations are to the reference tree (src/...).  The reference is C++20 over
 * Kokkos with ndim a compile-time constant; here ndim is a run-time field and
 * every kernel is a plain OpenMP loop over the same cell range, in the same
 * iteration-independent form, so the result does not depend on thread count.
 */
#include "hclpp_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* units/units.hpp:60-63 (SI code units: kg = m = s = Kelvin = 1) */
#define OR_KB 1.380649E-23
#define OR_MP 1.672649E-27
#define OR_G 6.6743015E-11
/* Kokkos::numbers::pi */
#define OR_PI 3.141592653589793238462643383279502884

static char g_err[256];
const char* or_last_error(void) { return g_err; }
static void set_err(const char* m) { snprintf(g_err, sizeof g_err, "%s", m); }

/* ------------------------------------------------------------------ */
/* PerfectGas: thermodynamics/perfect_gas.hpp:46-83, perfect_gas.cpp:13 */
typedef struct { double gamma, gamma_m1, mmw; } eos_t;

static inline double temp_from_pres(const eos_t* e, double rho, double P)
{   /* perfect_gas.hpp:69-72 */
    return P * e->mmw * OR_MP / (rho * OR_KB);
}
static inline double evol_from_temp(const eos_t* e, double rho, double T)
{   /* perfect_gas.hpp:52-56 */
    return rho * OR_KB * T / (e->mmw * OR_MP * e->gamma_m1);
}
static inline double evol_from_pres(const eos_t* e, double rho, double P)
{   /* perfect_gas.hpp:46-50 */
    return evol_from_temp(e, rho, temp_from_pres(e, rho, P));
}
static inline double temp_from_evol(const eos_t* e, double rho, double evol)
{   /* perfect_gas.hpp:74-78 */
    return evol * e->mmw * OR_MP * e->gamma_m1 / (rho * OR_KB);
}
static inline double pres_from_temp(const eos_t* e, double rho, double T)
{   /* perfect_gas.hpp:64-67 */
    return rho * OR_KB * T / (e->mmw * OR_MP);
}
static inline double pres_from_evol(const eos_t* e, double rho, double evol)
{   /* perfect_gas.hpp:58-62 */
    return pres_from_temp(e, rho, temp_from_evol(e, rho, evol));
}
static inline double sound(const eos_t* e, double rho, double P)
{   /* perfect_gas.hpp:80-83 */
    return sqrt(e->gamma * P / rho);
}

/* ------------------------------------------------------------------ */
/* Euler algebra: hydro/euler_equations.hpp:20-134 */
typedef struct { double rho; double m[3]; double E; } cons_t; /* m = rhou */
typedef struct { double rho; double u[3]; double P; } prim_t;
typedef cons_t flux_t;

static inline double ek_cons(int nd, const cons_t* c)
{   /* euler_equations.hpp:41-49 */
    double n = 0;
    for (int d = 0; d < nd; ++d) n += c->m[d] * c->m[d];
    return n / c->rho / 2;
}
static inline double ek_prim(int nd, const prim_t* p)
{   /* euler_equations.hpp:51-59 */
    double n = 0;
    for (int d = 0; d < nd; ++d) n += p->u[d] * p->u[d];
    return p->rho * n / 2;
}
static inline double evol_cons(int nd, const cons_t* c)
{   /* euler_equations.hpp:61-65 */
    return c->E - ek_cons(nd, c);
}
static inline flux_t flux_prim(int nd, const eos_t* e, const prim_t* p, int loc)
{   /* euler_equations.hpp:72-86 */
    flux_t f;
    double const E = ek_prim(nd, p) + evol_from_pres(e, p->rho, p->P);
    f.rho = p->rho * p->u[loc];
    for (int d = 0; d < nd; ++d) f.m[d] = p->rho * p->u[loc] * p->u[d];
    f.m[loc] += p->P;
    f.E = p->u[loc] * (E + p->P);
    return f;
}
static inline flux_t flux_cons(int nd, const eos_t* e, const cons_t* c, int loc)
{   /* euler_equations.hpp:93-109 */
    flux_t f;
    double const evol = evol_cons(nd, c);
    double const P = pres_from_evol(e, c->rho, evol);
    double const u = c->m[loc] / c->rho;
    f.rho = u * c->rho;
    for (int d = 0; d < nd; ++d) f.m[d] = u * c->m[d];
    f.m[loc] += P;
    f.E = u * (c->E + P);
    return f;
}
static inline prim_t to_prim(int nd, const eos_t* e, const cons_t* c)
{   /* euler_equations.hpp:111-122 */
    prim_t p;
    double const evol = evol_cons(nd, c);
    p.rho = c->rho;
    for (int d = 0; d < nd; ++d) p.u[d] = c->m[d] / c->rho;
    for (int d = nd; d < 3; ++d) p.u[d] = 0;
    p.P = pres_from_evol(e, c->rho, evol);
    return p;
}
static inline cons_t to_cons(int nd, const eos_t* e, const prim_t* p)
{   /* euler_equations.hpp:124-134 */
    cons_t c;
    c.rho = p->rho;
    for (int d = 0; d < nd; ++d) c.m[d] = p->rho * p->u[d];
    for (int d = nd; d < 3; ++d) c.m[d] = 0;
    c.E = evol_from_pres(e, p->rho, p->P) + ek_prim(nd, p);
    return c;
}

/* ------------------------------------------------------------------ */
/* Riemann solvers: hydro/riemann_solver.hpp */
static inline double flux_hll1(double UL, double UR, double FL, double FR, double wsL, double wsR)
{   /* riemann_solver.hpp:60-64 */
    return ((wsR * FL) - (wsL * FR) + (wsL * wsR * (UR - UL))) / (wsR - wsL);
}
static flux_t rs_hll(int nd, const eos_t* e, const cons_t* cL, const cons_t* cR, int loc)
{   /* riemann_solver.hpp:30-58 */
    prim_t const pL = to_prim(nd, e, cL);
    prim_t const pR = to_prim(nd, e, cR);
    double const sL = sound(e, pL.rho, pL.P);
    double const sR = sound(e, pR.rho, pR.P);
    double const wsL = fmin(pL.u[loc] - sL, pR.u[loc] - sR);
    double const wsR = fmax(pL.u[loc] + sL, pR.u[loc] + sR);
    double const nwsL = fmin(wsL, 0.);
    double const pwsR = fmax(wsR, 0.);
    flux_t const fL = flux_prim(nd, e, &pL, loc);
    flux_t const fR = flux_prim(nd, e, &pR, loc);
    flux_t f;
    f.rho = flux_hll1(cL->rho, cR->rho, fL.rho, fR.rho, nwsL, pwsR);
    for (int d = 0; d < nd; ++d) f.m[d] = flux_hll1(cL->m[d], cR->m[d], fL.m[d], fR.m[d], nwsL, pwsR);
    f.E = flux_hll1(cL->E, cR->E, fL.E, fR.E, nwsL, pwsR);
    return f;
}
static inline double flux_hllc1(double F, double U, double Us, double ws)
{   /* riemann_solver.hpp:136-140 */
    return F + (ws * (Us - U));
}
static flux_t rs_hllc(int nd, const eos_t* e, const cons_t* cL, const cons_t* cR, int loc)
{   /* riemann_solver.hpp:70-134 */
    prim_t const pL = to_prim(nd, e, cL);
    prim_t const pR = to_prim(nd, e, cR);
    double const sL = sound(e, pL.rho, pL.P);
    double const sR = sound(e, pR.rho, pR.P);
    double const wsL = fmin(pL.u[loc] - sL, pR.u[loc] - sR);
    double const wsR = fmax(pL.u[loc] + sL, pR.u[loc] + sR);
    flux_t const fL = flux_prim(nd, e, &pL, loc);
    flux_t const fR = flux_prim(nd, e, &pR, loc);
    double const rcL = pL.rho * (wsL - pL.u[loc]);
    double const rcR = pR.rho * (wsR - pR.u[loc]);
    double const ustar = (pR.P - pL.P + (rcL * pL.u[loc]) - (rcR * pR.u[loc])) / (rcL - rcR);
    double const pstar = 1. / 2 * (pL.P + pR.P + (rcL * (ustar - pL.u[loc])) + (rcR * (ustar - pR.u[loc])));
    double S;
    const cons_t* cs;
    const prim_t* ps;
    const flux_t* fs;
    if (ustar > 0) {
        S = wsL; cs = cL; ps = &pL; fs = &fL;
    } else {
        S = wsR; cs = cR; ps = &pR; fs = &fR;
    }
    double const un = ps->u[loc];
    flux_t f;
    if (wsL * wsR > 0) {
        f = *fs;
    } else {
        double const rho_star = cs->rho * (S - un) / (S - ustar);
        f.rho = flux_hllc1(fs->rho, cs->rho, rho_star, S);
        for (int d = 0; d < nd; ++d) f.m[d] = rho_star * ustar * ps->u[d];
        f.m[loc] = (rho_star * ustar * ustar) + pstar;
        double const E_star = ((S - un) / (S - ustar) * cs->E) + ((ustar - un) / (S - ustar) * ((ustar * cs->rho * (S - un)) + ps->P));
        f.E = flux_hllc1(fs->E, cs->E, E_star, S);
    }
    return f;
}
static flux_t rs_lowmach(int nd, const eos_t* e, const cons_t* cL, const cons_t* cR, int loc)
{   /* riemann_solver.hpp:146-180 ("Low Mach" Splitting) */
    prim_t const pL = to_prim(nd, e, cL);
    prim_t const pR = to_prim(nd, e, cR);
    double const sL = sound(e, pL.rho, pL.P);
    double const sR = sound(e, pR.rho, pR.P);
    double const a = 1.1 * fmax(pL.rho * sL, pR.rho * sR);
    double const ustar = ((pL.u[loc] + pR.u[loc]) / 2) - (1 / (2 * a) * (pR.P - pL.P));
    double const Ma = fabs(ustar) / fmin(sL, sR);
    double const theta = fmin(1, Ma);
    double const Pstar = ((pL.P + pR.P) / 2) - ((theta * a) / 2 * (pR.u[loc] - pL.u[loc]));
    const cons_t* cs = (ustar > 0) ? cL : cR;
    flux_t f;
    f.rho = ustar * cs->rho;
    for (int d = 0; d < nd; ++d) f.m[d] = ustar * cs->m[d];
    f.m[loc] += Pstar;
    f.E = (ustar * cs->E) + (Pstar * ustar);
    return f;
}
static inline flux_t riemann(int solver, int nd, const eos_t* e, const cons_t* cL, const cons_t* cR, int loc)
{
    if (solver == OR_RS_HLL) return rs_hll(nd, e, cL, cR, loc);
    if (solver == OR_RS_HLLC) return rs_hllc(nd, e, cL, cR, loc);
    return rs_lowmach(nd, e, cL, cR, loc);
}

/* ------------------------------------------------------------------ */
/* slope limiters: hydro/slope_limiters.hpp:18-87 */
static inline double limiter(int kind, double dR, double dL)
{
    if (kind == OR_LIM_CONSTANT) return 0; /* :21-25 */
    if (dL * dR > 0) {
        double const r = dR / dL;
        if (kind == OR_LIM_VANLEER) return (1. / 2) * (dR + dL) * (4 * r) / ((r + 1) * (r + 1)); /* :39-42 */
        if (kind == OR_LIM_MINMOD) {                                                           /* :59-63 */
            double const mm = 2 * fmin(1., r) / (1 + r);
            return mm * (dL + dR) / 2;
        }
        return (1. / 2) * (dR + dL) * (2 * r) / ((r * r) + 1); /* VanAlbada :80-83 */
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* source terms: hydro/source_terms.hpp:18-77 (sin/cos taken from host tables) */
static inline double src_pres(double dtodv, double PL, double PR, double dS, double dSp)
{   /* :18-22 */
    return dtodv * (PL + PR) / 2 * (dSp - dS);
}
static inline double src_u_r(double dtodv, double rL, double rR, double uL, double uR, double dS, double dSp)
{   /* :24-29 */
    return dtodv * ((rL * uL * uL) + (rR * uR * uR)) / 2 * (dSp - dS) / 2;
}
static inline double src_u_idir_r(double dtodv, double x0, double x0p, double rL, double rR, double uLr, double uRr,
                                  double uLo, double uRo, double dS, double dSp)
{   /* :31-48 */
    return dtodv * (x0p - x0) / (x0p + x0) * ((rR * uRo * uRr * dSp) + (rL * uLo * uLr * dS));
}
static inline double src_u_th(double dtodv, double cot_mid, double rL, double rR, double uLp, double uRp, double dS, double dSp)
{   /* :50-56; cot_mid = cos((x1+x1p)/2)/sin((x1+x1p)/2) */
    return dtodv * ((rL * uLp * uLp) + (rR * uRp * uRp)) / 2 * cot_mid * (dSp - dS) / 2;
}
static inline double src_u_phi(double dtodv, double sm, double sp, double rL, double rR, double uLt, double uRt,
                               double uLp, double uRp, double dS, double dSp)
{   /* :58-76; sm = sin(x1), sp = sin(x1p) */
    return dtodv * (sp - sm) / (sp + sm) * ((rR * uRp * uRt * dSp) + (rL * uLp * uLt * dS));
}

/* ------------------------------------------------------------------ */
/* context */
struct or_ctx {
    or_config cfg;
    eos_t eos;
    int nd, nfx, ng;
    int n[3], nwg[3], ngh[3];
    int64_t plane;
    double *x[3], *dx[3], *xc[3];
    double *ds, *dv, *gpm;
    double *cot_mid, *sin_node, *sin_center; /* per x1 index */
    /* driver state */
    double *rho, *rhou, *E, *fx, *u, *P;
    double *rho_rec, *rhou_rec, *E_rec, *fx_rec, *u_rec, *P_rec;
    double *rho_new, *rhou_new, *E_new, *fx_new;
    double t;
    int iter;
    /* output schedule, main.cpp:362-375 (utils/param.cpp:44-47) */
    int iter_output_frequency, iter_output_id, time_output_id;
    double time_first_output, time_output_frequency;
};

#define IDX(c, i, j, k) ((int64_t)(i) + (int64_t)(c)->nwg[0] * ((int64_t)(j) + (int64_t)(c)->nwg[1] * (int64_t)(k)))

static inline double grav(const or_ctx* c, int i, int dir)
{   /* gravity/gravity.hpp:36-40 (uniform), :63-70 (point mass) */
    if (c->cfg.gravity == OR_GRAV_UNIFORM) return c->cfg.g[dir];
    return dir == 0 ? c->gpm[i] : 0;
}

static void* xcalloc(size_t n, size_t s)
{
    void* p = calloc(n == 0 ? 1 : n, s);
    return p;
}

or_ctx* or_create(const or_config* cfg)
{
    if (cfg->ndim < 1 || cfg->ndim > 3) { set_err("ndim must be 1, 2 or 3"); return NULL; }
    /* perfect_gas.cpp:15-21 */
    if (!isfinite(cfg->gamma) || cfg->gamma <= 1) { set_err("Invalid gamma"); return NULL; }
    if (!isfinite(cfg->mmw) || cfg->mmw <= 0) { set_err("Invalid mmw"); return NULL; }
    or_ctx* c = (or_ctx*)xcalloc(1, sizeof(or_ctx));
    if (c) {
        c->iter_output_id = -1; /* main.cpp:209-211 */
        c->time_output_id = -1;
    }
    c->cfg = *cfg;
    c->nd = cfg->ndim;
    c->nfx = cfg->nfx;
    c->ng = cfg->ng;
    c->eos.gamma = cfg->gamma;
    c->eos.gamma_m1 = cfg->gamma - 1;
    c->eos.mmw = cfg->mmw;
    for (int d = 0; d < 3; ++d) {
        c->n[d] = cfg->n[d];
        c->ngh[d] = d < c->nd ? cfg->ng : 0; /* mesh/grid.cpp:56-58 */
        c->nwg[d] = c->n[d] + 2 * c->ngh[d];
    }
    c->plane = (int64_t)c->nwg[0] * c->nwg[1] * c->nwg[2];
    /* mesh/grid.cpp:134-155: copy nodes, compute dx and centers */
    for (int d = 0; d < 3; ++d) {
        int const nn = c->nwg[d];
        c->x[d] = (double*)xcalloc(nn + 1, sizeof(double));
        c->dx[d] = (double*)xcalloc(nn, sizeof(double));
        c->xc[d] = (double*)xcalloc(nn, sizeof(double));
        memcpy(c->x[d], cfg->x[d], sizeof(double) * (nn + 1));
        for (int i = 0; i < nn; ++i) {
            c->dx[d][i] = c->x[d][i + 1] - c->x[d][i];       /* grid.cpp:31-35 */
            c->xc[d][i] = (c->x[d][i] + c->x[d][i + 1]) / 2; /* grid.cpp:37-44 */
        }
    }
    /* host trigonometric tables (sin/cos of x1): used identically by the HIP product */
    int const n1 = c->nwg[1];
    c->cot_mid = (double*)xcalloc(n1, sizeof(double));
    c->sin_node = (double*)xcalloc(n1 + 1, sizeof(double));
    c->sin_center = (double*)xcalloc(n1, sizeof(double));
    for (int j = 0; j <= n1; ++j) c->sin_node[j] = sin(c->x[1][j]);
    for (int j = 0; j < n1; ++j) {
        double const mid = (c->x[1][j] + c->x[1][j + 1]) / 2;
        c->cot_mid[j] = cos(mid) / sin(mid);
        c->sin_center[j] = sin(c->xc[1][j]);
    }
    /* ds / dv over all ghosts: mesh/geometry.cpp:32-126 */
    c->ds = (double*)xcalloc(c->plane * 3, sizeof(double));
    c->dv = (double*)xcalloc(c->plane, sizeof(double));
    for (int k = 0; k < c->nwg[2]; ++k)
        for (int j = 0; j < c->nwg[1]; ++j)
            for (int i = 0; i < c->nwg[0]; ++i) {
                int64_t const id = IDX(c, i, j, k);
                double const *x0 = c->x[0], *x1 = c->x[1];
                double const *dx0 = c->dx[0], *dx1 = c->dx[1], *dx2 = c->dx[2];
                if (cfg->geom == OR_GEOM_CARTESIAN) {
                    /* geometry.cpp:43-57 */
                    for (int idim = 0; idim < 3; ++idim) {
                        double di[3] = {dx0[i], dx1[j], dx2[k]};
                        di[idim] = 1;
                        c->ds[id + c->plane * idim] = di[0] * di[1] * di[2];
                    }
                    c->dv[id] = dx0[i] * dx1[j] * dx2[k];
                } else if (c->nd == 1) {
                    /* geometry.cpp:71-85 */
                    c->ds[id] = 4 * OR_PI * x0[i] * x0[i];
                    c->dv[id] = (4 * OR_PI * ((x0[i + 1] * x0[i + 1] * x0[i + 1]) - (x0[i] * x0[i] * x0[i]))) / 3;
                } else if (c->nd == 2) {
                    /* geometry.cpp:87-105 */
                    double const dcos = cos(x1[j]) - cos(x1[j + 1]);
                    c->ds[id] = 2 * OR_PI * x0[i] * x0[i] * dcos;
                    c->ds[id + c->plane] = (2 * OR_PI / 2) * ((x0[i + 1] * x0[i + 1]) - (x0[i] * x0[i])) * sin(x1[j]);
                    c->dv[id] = (2 * OR_PI / 3) * ((x0[i + 1] * x0[i + 1] * x0[i + 1]) - (x0[i] * x0[i] * x0[i])) * dcos;
                } else {
                    /* geometry.cpp:107-125 */
                    double const dcos = cos(x1[j]) - cos(x1[j + 1]);
                    c->ds[id] = x0[i] * x0[i] * dcos * dx2[k];
                    c->ds[id + c->plane] = (1. / 2) * ((x0[i + 1] * x0[i + 1]) - (x0[i] * x0[i])) * sin(x1[j]) * dx2[k];
                    c->ds[id + 2 * c->plane] = x0[i] * dx0[i] * dx1[j];
                    c->dv[id] = (1. / 3) * ((x0[i + 1] * x0[i + 1] * x0[i + 1]) - (x0[i] * x0[i] * x0[i])) * dcos * dx2[k];
                }
            }
    /* point mass gravity: gravity/gravity.cpp:44-55 */
    c->gpm = (double*)xcalloc(c->nwg[0], sizeof(double));
    for (int i = 0; i < c->nwg[0]; ++i) c->gpm[i] = -OR_G * cfg->central_mass / (c->xc[0][i] * c->xc[0][i]);
    /* driver arrays: main.cpp:187-205 */
    int const nd = c->nd;
    int64_t const P1 = c->plane;
    c->rho = (double*)xcalloc(P1, 8);
    c->rhou = (double*)xcalloc(P1 * nd, 8);
    c->E = (double*)xcalloc(P1, 8);
    c->fx = (double*)xcalloc(P1 * c->nfx, 8);
    c->u = (double*)xcalloc(P1 * nd, 8);
    c->P = (double*)xcalloc(P1, 8);
    c->rho_rec = (double*)xcalloc(P1 * 2 * nd, 8);
    c->rhou_rec = (double*)xcalloc(P1 * 2 * nd * nd, 8);
    c->E_rec = (double*)xcalloc(P1 * 2 * nd, 8);
    c->fx_rec = (double*)xcalloc(P1 * 2 * nd * c->nfx, 8);
    c->u_rec = (double*)xcalloc(P1 * 2 * nd * nd, 8);
    c->P_rec = (double*)xcalloc(P1 * 2 * nd, 8);
    c->rho_new = (double*)xcalloc(P1, 8);
    c->rhou_new = (double*)xcalloc(P1 * nd, 8);
    c->E_new = (double*)xcalloc(P1, 8);
    c->fx_new = (double*)xcalloc(P1 * c->nfx, 8);
    if (!c->rho || !c->rhou_rec || !c->u_rec || !c->fx_new) { set_err("out of memory"); or_destroy(c); return NULL; }
    c->t = 0;
    c->iter = 0;
    return c;
}

void or_destroy(or_ctx* c)
{
    if (!c) return;
    for (int d = 0; d < 3; ++d) { free(c->x[d]); free(c->dx[d]); free(c->xc[d]); }
    free(c->ds); free(c->dv); free(c->gpm); free(c->cot_mid); free(c->sin_node); free(c->sin_center);
    free(c->rho); free(c->rhou); free(c->E); free(c->fx); free(c->u); free(c->P);
    free(c->rho_rec); free(c->rhou_rec); free(c->E_rec); free(c->fx_rec); free(c->u_rec); free(c->P_rec);
    free(c->rho_new); free(c->rhou_new); free(c->E_new); free(c->fx_new);
    free(c);
}

int64_t or_ncells_wg(const or_ctx* c) { return c->plane; }
void or_extents(const or_ctx* c, int* nwg3) { for (int d = 0; d < 3; ++d) nwg3[d] = c->nwg[d]; }
const double* or_geometry(const or_ctx* c, int which)
{
    if (which >= 0 && which < 3) return c->dx[which];
    if (which >= 3 && which < 6) return c->xc[which - 3];
    if (which == 6) return c->ds;
    if (which == 7) return c->dv;
    return NULL;
}

/* mesh/range.cpp:120-135: [Ng-NgEff, Ng+N+NgEff) on active dims, [Ng, Ng+N) inactive */
void or_range(const or_ctx* c, int ngeff, int* r)
{
    for (int d = 0; d < 3; ++d) {
        int const e = d < c->nd ? ngeff : 0;
        r[d] = c->ngh[d] - e;
        r[3 + d] = c->ngh[d] + c->n[d] + e;
    }
}

#define FOR_RANGE(r)                                                   \
    _Pragma("omp parallel for collapse(2) schedule(static)")          \
    for (int k = (r)[2]; k < (r)[5]; ++k)                              \
        for (int j = (r)[1]; j < (r)[4]; ++j)                          \
            for (int i = (r)[0]; i < (r)[3]; ++i)

/* ------------------------------------------------------------------ */
/* hydro/array_conversion.hpp:32-63 (prim -> cons) and :107-138 (cons -> prim) */
void or_conv_prim_to_cons(const or_ctx* c, const int* r, const double* rho, const double* u, const double* P, double* rhou, double* E)
{
    int const nd = c->nd;
    int64_t const pl = c->plane;
    FOR_RANGE(r)
    {
        int64_t const id = IDX(c, i, j, k);
        prim_t p;
        p.rho = rho[id];
        for (int d = 0; d < nd; ++d) p.u[d] = u[id + pl * d];
        p.P = P[id];
        cons_t const q = to_cons(nd, &c->eos, &p);
        for (int d = 0; d < nd; ++d) rhou[id + pl * d] = q.m[d];
        E[id] = q.E;
    }
}

void or_conv_cons_to_prim(const or_ctx* c, const int* r, const double* rho, const double* rhou, const double* E, double* u, double* P)
{
    int const nd = c->nd;
    int64_t const pl = c->plane;
    FOR_RANGE(r)
    {
        int64_t const id = IDX(c, i, j, k);
        cons_t q;
        q.rho = rho[id];
        for (int d = 0; d < nd; ++d) q.m[d] = rhou[id + pl * d];
        q.E = E[id];
        prim_t const p = to_prim(nd, &c->eos, &q);
        for (int d = 0; d < nd; ++d) u[id + pl * d] = p.u[d];
        P[id] = p.P;
    }
}

/* hydro/limited_linear_reconstruction.hpp:34-65; var_rec[wg, side(2), idim(ndim)] */
void or_face_reconstruction(const or_ctx* c, const int* r, const double* var, double* rec)
{
    int const nd = c->nd;
    int64_t const pl = c->plane;
    int const lim = c->cfg.limiter;
    double const *dx0 = c->dx[0], *dx1 = c->dx[1], *dx2 = c->dx[2];
    FOR_RANGE(r)
    {
        int64_t const id = IDX(c, i, j, k);
        for (int idim = 0; idim < nd; ++idim) {
            int const k0 = idim == 0, k1 = idim == 1, k2 = idim == 2; /* kron: utils/kronecker.hpp:15-35 */
            int const im = i - k0, jm = j - k1, km = k - k2;
            int const ip = i + k0, jp = j + k1, kp = k + k2;
            double const dl = (k0 * dx0[i]) + (k1 * dx1[j]) + (k2 * dx2[k]);
            double const dl_m = (k0 * dx0[im]) + (k1 * dx1[jm]) + (k2 * dx2[km]);
            double const dl_p = (k0 * dx0[ip]) + (k1 * dx1[jp]) + (k2 * dx2[kp]);
            double const v = var[id];
            double const slope = limiter(lim, (var[IDX(c, ip, jp, kp)] - v) / ((dl + dl_p) / 2),
                                         (v - var[IDX(c, im, jm, km)]) / ((dl_m + dl) / 2));
            rec[id + pl * (0 + 2 * idim)] = v - ((dl / 2) * slope);
            rec[id + pl * (1 + 2 * idim)] = v + ((dl / 2) * slope);
        }
    }
}

/* hydro/extrapolation_reconstruction.hpp:65-338 */
void or_hancock(const or_ctx* c, const int* r, double dt, const double* u_rec, const double* P_rec,
                double* rho_rec, double* rhou_rec, double* E_rec, double* fx_rec)
{
    int const nd = c->nd, nfx = c->nfx;
    int64_t const pl = c->plane;
    int const sph = c->cfg.geom == OR_GEOM_SPHERICAL;
    double* fx_rec_old = NULL;
    if (nfx > 0) { /* :87-88 */
        fx_rec_old = (double*)malloc(sizeof(double) * pl * 2 * nd * nfx);
        memcpy(fx_rec_old, fx_rec, sizeof(double) * pl * 2 * nd * nfx);
    }
#define RR(s, d) rho_rec[id + pl * ((s) + 2 * (d))]
#define MR(s, d, v) rhou_rec[id + pl * ((s) + 2 * ((d) + nd * (v)))]
#define ER(s, d) E_rec[id + pl * ((s) + 2 * (d))]
#define FR(s, d, f) fx_rec[id + pl * ((s) + 2 * ((d) + nd * (f)))]
    FOR_RANGE(r)
    {
        int64_t const id = IDX(c, i, j, k);
        double rho_old[2][3], rhou_old[2][3][3];
        for (int idim = 0; idim < nd; ++idim) { /* :94-109 */
            for (int s = 0; s < 2; ++s) {
                rho_old[s][idim] = RR(s, idim);
                for (int idr = 0; idr < nd; ++idr) rhou_old[s][idim][idr] = MR(s, idim, idr);
            }
            for (int f = 0; f < nfx; ++f) {
                FR(0, idim, f) *= RR(0, idim);
                FR(1, idim, f) *= RR(1, idim);
            }
        }
        for (int idim = 0; idim < nd; ++idim) {
            int const ip = i + (idim == 0), jp = j + (idim == 1), kp = k + (idim == 2);
            int64_t const idp = IDX(c, ip, jp, kp);
            prim_t pL, pR; /* :114-128 */
            pL.rho = rho_old[0][idim];
            pR.rho = rho_old[1][idim];
            for (int idr = 0; idr < nd; ++idr) {
                pL.u[idr] = u_rec[id + pl * (0 + 2 * (idim + nd * idr))];
                pR.u[idr] = u_rec[id + pl * (1 + 2 * (idim + nd * idr))];
            }
            for (int idr = nd; idr < 3; ++idr) { pL.u[idr] = 0; pR.u[idr] = 0; }
            pL.P = P_rec[id + pl * (0 + 2 * idim)];
            pR.P = P_rec[id + pl * (1 + 2 * idim)];
            flux_t const fL = flux_prim(nd, &c->eos, &pL, idim);
            flux_t const fR = flux_prim(nd, &c->eos, &pR, idim);
            double const dtodv = dt / c->dv[id];
            double const dS = c->ds[id + pl * idim];
            double const dSp = c->ds[idp + pl * idim];
            for (int ipos = 0; ipos < nd; ++ipos) { /* :132-143 */
                for (int s = 0; s < 2; ++s) {
                    RR(s, ipos) += dtodv * ((fL.rho * dS) - (fR.rho * dSp));
                    for (int idr = 0; idr < nd; ++idr) MR(s, ipos, idr) += dtodv * ((fL.m[idr] * dS) - (fR.m[idr] * dSp));
                    ER(s, ipos) += dtodv * ((fL.E * dS) - (fR.E * dSp));
                }
            }
            if (sph) { /* :145-303 */
                for (int ipos = 0; ipos < nd; ++ipos) {
                    for (int s = 0; s < 2; ++s) {
                        if (idim == 0) {
                            MR(s, ipos, 0) += src_pres(dtodv, pL.P, pR.P, dS, dSp);
                            if (nd >= 2) {
                                MR(s, ipos, 0) += src_u_r(dtodv, pL.rho, pR.rho, pL.u[1], pR.u[1], dS, dSp);
                                MR(s, ipos, 1) -= src_u_idir_r(dtodv, c->x[0][i], c->x[0][i + 1], pL.rho, pR.rho, pL.u[0], pR.u[0], pL.u[1], pR.u[1], dS, dSp);
                            }
                            if (nd == 3) {
                                MR(s, ipos, 0) += src_u_r(dtodv, pL.rho, pR.rho, pL.u[2], pR.u[2], dS, dSp);
                                MR(s, ipos, 2) -= src_u_idir_r(dtodv, c->x[0][i], c->x[0][i + 1], pL.rho, pR.rho, pL.u[0], pR.u[0], pL.u[2], pR.u[2], dS, dSp);
                            }
                        }
                        if (idim == 1) {
                            MR(s, ipos, 1) += src_pres(dtodv, pL.P, pR.P, dS, dSp);
                            if (nd == 3) {
                                MR(s, ipos, 1) += src_u_th(dtodv, c->cot_mid[j], pL.rho, pR.rho, pL.u[2], pR.u[2], dS, dSp);
                                MR(s, ipos, 2) -= src_u_phi(dtodv, c->sin_node[j], c->sin_node[j + 1], pL.rho, pR.rho, pL.u[1], pR.u[1], pL.u[2], pR.u[2], dS, dSp);
                            }
                        }
                    }
                }
            }
            double const g = grav(c, i, idim);
            for (int ipos = 0; ipos < nd; ++ipos) { /* :306-312 */
                for (int s = 0; s < 2; ++s) {
                    MR(s, ipos, idim) += dt * g * rho_old[s][ipos];
                    ER(s, ipos) += dt * g * rhou_old[s][ipos][idim];
                }
            }
            for (int f = 0; f < nfx; ++f) { /* :315-323 */
                for (int ipos = 0; ipos < nd; ++ipos) {
                    double const ffL = fx_rec_old[id + pl * (0 + 2 * (idim + nd * f))] * fL.rho;
                    double const ffR = fx_rec_old[id + pl * (1 + 2 * (idim + nd * f))] * fR.rho;
                    FR(0, ipos, f) += dtodv * ((ffL * dS) - (ffR * dSp));
                    FR(1, ipos, f) += dtodv * ((ffL * dS) - (ffR * dSp));
                }
            }
        }
    }
    if (nfx > 0) { /* passive_scalar_extrapolation :327-337 */
        FOR_RANGE(r)
        {
            int64_t const id = IDX(c, i, j, k);
            for (int idim = 0; idim < nd; ++idim)
                for (int f = 0; f < nfx; ++f) {
                    FR(0, idim, f) /= RR(0, idim);
                    FR(1, idim, f) /= RR(1, idim);
                }
        }
        free(fx_rec_old);
    }
#undef RR
#undef MR
#undef ER
#undef FR
}

/* hydro/hydro_reconstruction.hpp:85-129 (MUSCLHancockHydroReconstruction::execute) */
void or_hydro_reconstruction(const or_ctx* c, const int* r, double dt, const double* rho, const double* u,
                             const double* P, const double* fx, double* rho_rec, double* rhou_rec, double* E_rec,
                             double* fx_rec, double* P_rec, double* u_rec)
{
    int const nd = c->nd, nfx = c->nfx;
    int64_t const pl = c->plane;
    or_face_reconstruction(c, r, rho, rho_rec);
    for (int d = 0; d < nd; ++d) or_face_reconstruction(c, r, u + pl * d, u_rec + pl * 2 * nd * d);
    or_face_reconstruction(c, r, P, P_rec);
    for (int f = 0; f < nfx; ++f) or_face_reconstruction(c, r, fx + pl * f, fx_rec + pl * 2 * nd * f);
    for (int idim = 0; idim < nd; ++idim) {
        for (int s = 0; s < 2; ++s) {
            /* face (s, idim) sub-arrays: prim -> cons over the range */
            int64_t const off = pl * (s + 2 * idim);
            FOR_RANGE(r)
            {
                int64_t const id = IDX(c, i, j, k);
                prim_t p;
                p.rho = rho_rec[off + id];
                for (int v = 0; v < nd; ++v) p.u[v] = u_rec[off + pl * 2 * nd * v + id];
                p.P = P_rec[off + id];
                cons_t const q = to_cons(nd, &c->eos, &p);
                for (int v = 0; v < nd; ++v) rhou_rec[off + pl * 2 * nd * v + id] = q.m[v];
                E_rec[off + id] = q.E;
            }
        }
    }
    or_hancock(c, r, dt, u_rec, P_rec, rho_rec, rhou_rec, E_rec, fx_rec);
}

/* hydro/godunov_scheme.hpp:77-302 */
void or_godunov(const or_ctx* c, const int* r, double dt, const double* rho, const double* rhou, const double* E,
                const double* fx, const double* rho_rec, const double* rhou_rec, const double* E_rec,
                const double* fx_rec, double* rho_new, double* rhou_new, double* E_new, double* fx_new)
{
    int const nd = c->nd, nfx = c->nfx;
    int64_t const pl = c->plane;
    int const sph = c->cfg.geom == OR_GEOM_SPHERICAL;
    int const rs = c->cfg.riemann;
    FOR_RANGE(r)
    {
        int64_t const id = IDX(c, i, j, k);
        double rn = rho[id], En = E[id], mn[3];
        for (int d = 0; d < nd; ++d) mn[d] = rhou[id + pl * d];
        for (int f = 0; f < nfx; ++f) fx_new[id + pl * f] = fx[id + pl * f] * rho[id];
        for (int idim = 0; idim < nd; ++idim) {
            int const im = i - (idim == 0), jm = j - (idim == 1), km = k - (idim == 2);
            int const ip = i + (idim == 0), jp = j + (idim == 1), kp = k + (idim == 2);
            int64_t const idm = IDX(c, im, jm, km), idp = IDX(c, ip, jp, kp);
            cons_t mR, vL, vR, pLf;
#define LOADC(q, cell, s)                                                                   \
    do {                                                                                    \
        (q).rho = rho_rec[(cell) + pl * ((s) + 2 * idim)];                                  \
        for (int v = 0; v < nd; ++v) (q).m[v] = rhou_rec[(cell) + pl * ((s) + 2 * (idim + nd * v))]; \
        for (int v = nd; v < 3; ++v) (q).m[v] = 0;                                          \
        (q).E = E_rec[(cell) + pl * ((s) + 2 * idim)];                                      \
    } while (0)
            LOADC(mR, idm, 1);
            LOADC(vL, id, 0);
            flux_t const FL = riemann(rs, nd, &c->eos, &mR, &vL, idim); /* :129-141 */
            LOADC(vR, id, 1);
            LOADC(pLf, idp, 0);
            flux_t const FRx = riemann(rs, nd, &c->eos, &vR, &pLf, idim); /* :143-155 */
#undef LOADC
            double const dtodv = dt / c->dv[id];
            double const dS = c->ds[id + pl * idim];
            double const dSp = c->ds[idp + pl * idim];
            rn += dtodv * ((FL.rho * dS) - (FRx.rho * dSp)); /* :159-163 */
            for (int v = 0; v < nd; ++v) mn[v] += dtodv * ((FL.m[v] * dS) - (FRx.m[v] * dSp));
            En += dtodv * ((FL.E * dS) - (FRx.E * dSp));
            if (sph) { /* :165-256 */
                prim_t const pL = to_prim(nd, &c->eos, &vL);
                prim_t const pR = to_prim(nd, &c->eos, &vR);
                if (idim == 0) {
                    mn[0] += src_pres(dtodv, pL.P, pR.P, dS, dSp);
                    if (nd >= 2) {
                        mn[0] += src_u_r(dtodv, pL.rho, pR.rho, pL.u[1], pR.u[1], dS, dSp);
                        mn[1] -= src_u_idir_r(dtodv, c->x[0][i], c->x[0][i + 1], pL.rho, pR.rho, pL.u[0], pR.u[0], pL.u[1], pR.u[1], dS, dSp);
                    }
                    if (nd == 3) {
                        mn[0] += src_u_r(dtodv, pL.rho, pR.rho, pL.u[2], pR.u[2], dS, dSp);
                        mn[2] -= src_u_idir_r(dtodv, c->x[0][i], c->x[0][i + 1], pL.rho, pR.rho, pL.u[0], pR.u[0], pL.u[2], pR.u[2], dS, dSp);
                    }
                }
                if (idim == 1) {
                    mn[1] += src_pres(dtodv, pL.P, pR.P, dS, dSp);
                    if (nd == 3) {
                        mn[1] += src_u_th(dtodv, c->cot_mid[j], pL.rho, pR.rho, pL.u[2], pR.u[2], dS, dSp);
                        mn[2] -= src_u_phi(dtodv, c->sin_node[j], c->sin_node[j + 1], pL.rho, pR.rho, pL.u[1], pR.u[1], pL.u[2], pR.u[2], dS, dSp);
                    }
                }
            }
            double const g = grav(c, i, idim); /* :258-260 */
            mn[idim] += dt * g * rho[id];
            En += dt * g * rhou[id + pl * idim];
            for (int f = 0; f < nfx; ++f) { /* :262-290 */
                int64_t cl = idm, cr = id;
                int fl = 1, fr = 1;
                if (FL.rho < 0) { cl = id; fl = 0; }
                if (FRx.rho < 0) { cr = idp; fr = 0; }
                double const ffL = fx_rec[cl + pl * (fl + 2 * (idim + nd * f))] * FL.rho;
                double const ffR = fx_rec[cr + pl * (fr + 2 * (idim + nd * f))] * FRx.rho;
                fx_new[id + pl * f] += dtodv * ((ffL * dS) - (ffR * dSp));
            }
        }
        rho_new[id] = rn;
        for (int d = 0; d < nd; ++d) rhou_new[id + pl * d] = mn[d];
        E_new[id] = En;
    }
    if (nfx > 0) { /* passive_scalar_Godunov :294-301 */
        FOR_RANGE(r)
        {
            int64_t const id = IDX(c, i, j, k);
            for (int f = 0; f < nfx; ++f) {
                double const y = fx_new[id + pl * f] / rho_new[id];
                fx_new[id + pl * f] = y < 0. ? 0. : (y > 1. ? 1. : y);
            }
        }
    }
}

/* hydro/time_step.hpp:35-82; returns 1 / max(...) as the reference does */
double or_time_step(const or_ctx* c, const int* r, const double* rho, const double* u, const double* P)
{
    int const nd = c->nd;
    int64_t const pl = c->plane;
    int const sph = c->cfg.geom == OR_GEOM_SPHERICAL;
    double inv = 0;
#pragma omp parallel for collapse(2) reduction(max : inv) schedule(static)
    for (int k = r[2]; k < r[5]; ++k)
        for (int j = r[1]; j < r[4]; ++j)
            for (int i = r[0]; i < r[3]; ++i) {
                int64_t const id = IDX(c, i, j, k);
                double const s = sound(&c->eos, rho[id], P[id]);
                double dxg[3] = {c->dx[0][i], c->dx[1][j], c->dx[2][k]};
                if (sph) {
                    if (nd == 2) dxg[1] *= c->xc[0][i];
                    if (nd == 3) {
                        dxg[1] *= c->xc[0][i];
                        dxg[2] *= c->xc[0][i] * c->sin_center[j];
                    }
                }
                double ci = 0;
                for (int d = 0; d < nd; ++d) ci += (fabs(u[id + pl * d]) + s) / dxg[d];
                inv = inv > ci ? inv : ci; /* Kokkos::max */
            }
    return 1 / inv;
}

/* analysis/internal_energy.cpp:20-45 */
double or_min_internal_energy(const or_ctx* c, const int* r, const double* rho, const double* rhou, const double* E)
{
    int const nd = c->nd;
    int64_t const pl = c->plane;
    double mn = 0; /* the reference initialises its result to 0 and reduces with Kokkos::Min */
    int first = 1;
    for (int k = r[2]; k < r[5]; ++k)
        for (int j = r[1]; j < r[4]; ++j)
            for (int i = r[0]; i < r[3]; ++i) {
                int64_t const id = IDX(c, i, j, k);
                cons_t q;
                q.rho = rho[id];
                for (int d = 0; d < nd; ++d) q.m[d] = rhou[id + pl * d];
                q.E = E[id];
                double const ev = evol_cons(nd, &q);
                if (isnan(ev)) return ev;
                if (first || ev < mn) { mn = ev; first = 0; }
            }
    return mn;
}

/* analysis/integration.cpp:22-37 (sum in i-fastest order) */
double or_integrate(const or_ctx* c, const int* r, const double* var)
{
    double s = 0;
    for (int k = r[2]; k < r[5]; ++k)
        for (int j = r[1]; j < r[4]; ++j)
            for (int i = r[0]; i < r[3]; ++i) {
                int64_t const id = IDX(c, i, j, k);
                s += var[id] * c->dv[id];
            }
    return s;
}

/* ------------------------------------------------------------------ */
/* boundaries: boundary_distribute.hpp:42-76, boundary_distribute.cpp:75-120, boundary.cpp:33-108 */
static void copy_slab(const or_ctx* c, double* a, int idim, int src0, int dst0, int ng)
{
    /* copy ng layers [src0, src0+ng) -> [dst0, dst0+ng) along idim over full transverse extents */
    int lo[3] = {0, 0, 0}, hi[3] = {c->nwg[0], c->nwg[1], c->nwg[2]};
    lo[idim] = 0;
    hi[idim] = ng;
    for (int k = lo[2]; k < hi[2]; ++k)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int i = lo[0]; i < hi[0]; ++i) {
                int s[3] = {i, j, k}, d[3] = {i, j, k};
                s[idim] += src0;
                d[idim] += dst0;
                a[IDX(c, d[0], d[1], d[2])] = a[IDX(c, s[0], s[1], s[2])];
            }
}

static void border_bc(const or_ctx* c, int kind, int idim, int iface, double** views, int nviews, int imom)
{
    int const ng = c->ngh[idim];
    int const ext = c->nwg[idim];
    int lo[3] = {0, 0, 0}, hi[3] = {c->nwg[0], c->nwg[1], c->nwg[2]};
    if (iface == 1) lo[idim] = ext - ng;
    hi[idim] = lo[idim] + ng;
    int const offset = iface == 0 ? hi[idim] : lo[idim] - 1;            /* boundary.cpp:53 */
    int const mirror = iface == 0 ? ((2 * ng) - 1) : ((2 * (ext - ng)) - 1); /* boundary.cpp:91 */
    for (int k = lo[2]; k < hi[2]; ++k)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int i = lo[0]; i < hi[0]; ++i) {
                int o[3] = {i, j, k};
                if (kind == OR_BC_NULLGRADIENT) o[idim] = offset;
                else o[idim] = mirror - o[idim];
                int64_t const dst = IDX(c, i, j, k), src = IDX(c, o[0], o[1], o[2]);
                for (int v = 0; v < nviews; ++v) views[v][dst] = views[v][src];
                if (kind == OR_BC_REFLEXIVE) views[imom + idim][dst] = -views[imom + idim][dst];
            }
}

void or_boundaries(const or_ctx* c, double* rho, double* rhou, double* E, double* fx)
{
    int const nd = c->nd;
    int64_t const pl = c->plane;
    double* views[3 + 2 + 16];
    int nv = 0;
    views[nv++] = rho;
    for (int d = 0; d < nd; ++d) views[nv++] = rhou + pl * d;
    views[nv++] = E;
    for (int f = 0; f < c->nfx && nv < 21; ++f) views[nv++] = fx + pl * f;
    /* single-rank periodic self-exchange in dim order, face 0 then 1 */
    for (int idim = 0; idim < nd; ++idim) {
        int const ng = c->ngh[idim], ext = c->nwg[idim];
        for (int v = 0; v < nv; ++v) {
            copy_slab(c, views[v], idim, ng, ext - ng, ng);          /* face 0: first interior -> right ghosts */
        }
        for (int v = 0; v < nv; ++v) {
            copy_slab(c, views[v], idim, ext - 2 * ng, 0, ng);       /* face 1: last interior -> left ghosts */
        }
    }
    for (int b = 0; b < 2 * nd; ++b) {
        int const id = c->cfg.bc_order[b];
        int const kind = c->cfg.bc[id];
        if (kind == OR_BC_PERIODIC) continue; /* boundary.hpp:124-136: no-op */
        border_bc(c, kind, id / 2, id % 2, views, nv, 1);
    }
}

/* hydro/pressure_fix.hpp:19-172, over `r` (no_ghosts in main.cpp:459-475) */
void or_pressure_fix(const or_ctx* c, const int* r, double dt, double eps, const double* rho, const double* rhou,
                     const double* E, const double* rho_rec, const double* rhou_rec, const double* E_rec,
                     const double* rho_new, const double* rhou_new, double* E_new)
{
    int const nd = c->nd;
    int64_t const pl = c->plane;
    double const eps2 = eps / 10;
    FOR_RANGE(r)
    {
        int64_t const id = IDX(c, i, j, k);
        cons_t cons, cons_new;
        cons.rho = rho[id];
        cons_new.rho = rho_new[id];
        for (int d = 0; d < 3; ++d) {
            cons.m[d] = d < nd ? rhou[id + pl * d] : 0;
            cons_new.m[d] = d < nd ? rhou_new[id + pl * d] : 0;
        }
        cons.E = E[id];
        cons_new.E = E_new[id];
        prim_t const prim = to_prim(nd, &c->eos, &cons);
        double const ekin = ek_cons(nd, &cons);
        double const evol = E[id] - ekin;
        double const evol_new = E_new[id] - ek_cons(nd, &cons_new);
        double const de_god = evol_new - evol;
        if (evol < eps * ekin) {
            E_new[id] = E[id];
            double divU = 0, divUE = 0, einL = 0, einR = 0;
            double alpha = (evol * evol) / (ekin * ekin * eps * eps);
            if (evol < eps2 * ekin) alpha = 0;
            for (int idim = 0; idim < nd; ++idim) {
                int const im = i - (idim == 0), jm = j - (idim == 1), km = k - (idim == 2);
                int const ip = i + (idim == 0), jp = j + (idim == 1), kp = k + (idim == 2);
                int64_t const idm = IDX(c, im, jm, km), idp = IDX(c, ip, jp, kp);
                for (int ipos = 0; ipos < nd; ++ipos) {
                    cons_t vL, vR, mR, pL;
#define LOADF(q, cell, s)                                                                                \
    do {                                                                                                 \
        (q).rho = rho_rec[(cell) + pl * ((s) + 2 * idim)];                                               \
        for (int v = 0; v < 3; ++v) (q).m[v] = v < nd ? rhou_rec[(cell) + pl * ((s) + 2 * (idim + nd * v))] : 0; \
        (q).E = E_rec[(cell) + pl * ((s) + 2 * idim)];                                                   \
    } while (0)
                    LOADF(vL, id, 0);
                    LOADF(vR, id, 1);
                    LOADF(mR, idm, 1);
                    LOADF(pL, idp, 0);
#undef LOADF
                    prim_t const primL = to_prim(nd, &c->eos, &vL);
                    prim_t const primR = to_prim(nd, &c->eos, &vR);
                    prim_t const prim_mR = to_prim(nd, &c->eos, &mR);
                    prim_t const prim_pL = to_prim(nd, &c->eos, &pL);
                    cons_t cons_p, cons_m;
                    cons_p.rho = rho[idp];
                    cons_m.rho = rho[idm];
                    for (int d = 0; d < 3; ++d) {
                        cons_p.m[d] = d < nd ? rhou[idp + pl * d] : 0;
                        cons_m.m[d] = d < nd ? rhou[idm + pl * d] : 0;
                    }
                    cons_p.E = E[idp];
                    cons_m.E = E[idm];
                    double const us = (prim_mR.u[ipos] + primL.u[ipos]) / 2;   /* :129 */
                    double const us_p = (primR.u[ipos] + prim_pL.u[ipos]) / 2; /* :130 */
                    double const dsp = c->ds[idp + pl * ipos], ds0 = c->ds[id + pl * ipos], dv = c->dv[id];
                    divU += ((us_p * dsp) - (us * ds0)) / dv;                  /* :132 */
                    if (us_p > 0) einR = E[id] - ek_cons(nd, &cons);
                    else einR = E[idp] - ek_cons(nd, &cons_p);
                    if (us < 0) einL = E[id] - ek_cons(nd, &cons);
                    else einL = E[idm] - ek_cons(nd, &cons_m);
                    divUE += ((us_p * einR * dsp) - (us * einL * ds0)) / dv;   /* :146 */
                    double const de_pf = dt * ((-prim.P * divU) - divUE);
                    double sum = 0;
                    for (int idr = 0; idr < nd; ++idr) sum += cons.m[idr] * cons.m[idr] / cons.rho;
                    double const de_tot = (alpha * de_god) + ((1 - alpha) * de_pf);
                    E_new[id] = evol + de_tot + (1. / 2 * sum);              /* :160 */
                }
            }
        }
    }
}

/* gravity/gravity.cpp:59-128 (one rank along dim 0): shell-averaged density of each radial
 * layer (interior j, k), enclosed mass at each cell centre, g = -G M(r) / r^2 */
void or_internal_mass_gravity(or_ctx* c, const double* rho)
{
    int const ng0 = c->ngh[0], n0 = c->n[0];
    int const nsh = n0 + ng0;
    double* mass = (double*)xcalloc(nsh, sizeof(double));
    double* vol = (double*)xcalloc(nsh, sizeof(double));
    double* Mr = (double*)xcalloc(c->nwg[0], sizeof(double));
    for (int s = 0; s < nsh; ++s) { /* :75-87, sums in k-major, j-minor order */
        int const i = s + ng0;
        double sm = 0, sv = 0;
        for (int k = c->ngh[2]; k < c->ngh[2] + c->n[2]; ++k)
            for (int j = c->ngh[1]; j < c->ngh[1] + c->n[1]; ++j) {
                int64_t const id = IDX(c, i, j, k);
                sm += rho[id] * c->dv[id];
                sv += c->dv[id];
            }
        mass[s] = sm;
        vol[s] = sv;
    }
    for (int s = 0; s < nsh; ++s) mass[s] /= vol[s];                       /* :96-99 rho_mean */
    for (int i = 0; i < ng0; ++i) Mr[i] = c->cfg.central_mass;             /* :101 */
    double partial = 0;                                                     /* :103-118 */
    double const vol_s3 = 4 * OR_PI / 3;
    for (int i = ng0; i < n0 + 2 * ng0; ++i) {
        int const off = i - ng0;
        double const *x0 = c->x[0], *xc = c->xc[0];
        double const x3 = x0[i] * x0[i] * x0[i];
        if (i == ng0) {
            partial += c->cfg.central_mass + (vol_s3 * ((xc[i] * xc[i] * xc[i]) - x3) * mass[off]);
        } else {
            partial += (vol_s3 * (x3 - (xc[i - 1] * xc[i - 1] * xc[i - 1])) * mass[off - 1])
                       + (vol_s3 * ((xc[i] * xc[i] * xc[i]) - x3) * mass[off]);
        }
        Mr[i] = partial;
    }
    for (int i = 0; i < c->nwg[0]; ++i) c->gpm[i] = -OR_G * Mr[i] / (c->xc[0][i] * c->xc[0][i]); /* :120-123 */
    free(mass);
    free(vol);
    free(Mr);
}

const double* or_gravity_table(const or_ctx* c) { return c->gpm; }

/* ------------------------------------------------------------------ */
/* driver: main.cpp:305-311 (init) and :354-531 (loop body) */
int or_init_from_prim(or_ctx* c, const double* rho, const double* u, const double* P, const double* fx)
{
    int64_t const pl = c->plane;
    int const nd = c->nd;
    memcpy(c->rho, rho, sizeof(double) * pl);
    memcpy(c->u, u, sizeof(double) * pl * nd);
    memcpy(c->P, P, sizeof(double) * pl);
    if (c->nfx > 0 && fx) memcpy(c->fx, fx, sizeof(double) * pl * c->nfx);
    int r0[6], rall[6];
    or_range(c, 0, r0);
    or_range(c, c->ng, rall);
    or_conv_prim_to_cons(c, r0, c->rho, c->u, c->P, c->rhou, c->E);
    or_boundaries(c, c->rho, c->rhou, c->E, c->fx);
    if (c->cfg.gravity == OR_GRAV_INTERNAL_MASS) or_internal_mass_gravity(c, c->rho); /* :309 */
    or_conv_cons_to_prim(c, rall, c->rho, c->rhou, c->E, c->u, c->P);
    c->t = 0;
    c->iter = 0;
    return 0;
}

void or_set_outputs(or_ctx* c, int iter_output_frequency, double time_first_output, double time_output_frequency, int iter_output_id,
                    int time_output_id)
{
    c->iter_output_frequency = iter_output_frequency;
    c->time_first_output = time_first_output;
    c->time_output_frequency = time_output_frequency;
    c->iter_output_id = iter_output_id;
    c->time_output_id = time_output_id;
}

void or_output_ids(const or_ctx* c, int* iter_output_id, int* time_output_id)
{
    *iter_output_id = c->iter_output_id;
    *time_output_id = c->time_output_id;
}

int or_step(or_ctx* c, double cfl, double t_end, int max_iter, double* dt_out, int* exit_flag)
{
    return or_step_outputs(c, cfl, t_end, max_iter, dt_out, exit_flag, NULL);
}

int or_step_outputs(or_ctx* c, double cfl, double t_end, int max_iter, double* dt_out, int* exit_flag, int* output_flag)
{
    int64_t const pl = c->plane;
    int const nd = c->nd;
    int r0[6], r1[6], rall[6];
    or_range(c, 0, r0);
    or_range(c, 1, r1);
    or_range(c, c->ng, rall);
    double dt = cfl * or_time_step(c, rall, c->rho, c->u, c->P); /* :355 */
    int ex = 0, out = 0;
    if (c->iter_output_frequency > 0) {                             /* :362-368 (iter_ini = 0) */
        int const next_output = (c->iter_output_id + 1) * c->iter_output_frequency;
        if ((c->iter + 1) >= next_output) { out = 1; ++c->iter_output_id; }
    }
    if (c->time_output_frequency > 0) {                             /* :369-375 */
        double const next_output = c->time_first_output + ((c->time_output_id + 1) * c->time_output_frequency);
        if ((c->t + dt) >= next_output) {
            dt = next_output - c->t;
            if (dt <= 0) { set_err("Error: the time step is negative"); return 2; }
            out = 1;
            ++c->time_output_id;
        }
    }
    if ((c->t + dt) >= t_end) { dt = t_end - c->t; ex = 1; out = 1; } /* :377-381 */
    if ((c->iter + 1) >= max_iter) { ex = 1; out = 1; }             /* :382-385 */
    double const me = or_min_internal_energy(c, r0, c->rho, c->rhou, c->E); /* :397-398 */
    if (isnan(me) || me < 0) { set_err("detected invalid volumic internal energy"); return 1; }
    or_hydro_reconstruction(c, r1, dt / 2, c->rho, c->u, c->P, c->fx, c->rho_rec, c->rhou_rec, c->E_rec, c->fx_rec,
                            c->P_rec, c->u_rec);                   /* :427-439 */
    or_godunov(c, r0, dt, c->rho, c->rhou, c->E, c->fx, c->rho_rec, c->rhou_rec, c->E_rec, c->fx_rec, c->rho_new,
               c->rhou_new, c->E_new, c->fx_new);                   /* :441-457 */
    if (c->cfg.pressure_fix)                                            /* :459-475 */
        or_pressure_fix(c, r0, dt, c->cfg.eps_pf, c->rho, c->rhou, c->E, c->rho_rec, c->rhou_rec, c->E_rec, c->rho_new,
                        c->rhou_new, c->E_new);
    or_boundaries(c, c->rho_new, c->rhou_new, c->E_new, c->fx_new); /* :479 */
    memcpy(c->rho, c->rho_new, sizeof(double) * pl);                /* :481-484 */
    memcpy(c->rhou, c->rhou_new, sizeof(double) * pl * nd);
    memcpy(c->E, c->E_new, sizeof(double) * pl);
    if (c->nfx > 0) memcpy(c->fx, c->fx_new, sizeof(double) * pl * c->nfx);
    or_conv_cons_to_prim(c, rall, c->rho, c->rhou, c->E, c->u, c->P); /* :526 */
    if (c->cfg.gravity == OR_GRAV_INTERNAL_MASS) or_internal_mass_gravity(c, c->rho); /* :528 */
    c->t += dt;                                                        /* :530-531 */
    c->iter += 1;
    if (dt_out) *dt_out = dt;
    if (exit_flag) *exit_flag = ex;
    if (output_flag) *output_flag = out;
    return 0;
}

double* or_state(or_ctx* c, int which)
{
    switch (which) {
    case 0: return c->rho;
    case 1: return c->rhou;
    case 2: return c->E;
    case 3: return c->fx;
    case 4: return c->u;
    case 5: return c->P;
    default: return NULL;
    }
}
double or_time(const or_ctx* c) { return c->t; }
int or_iter(const or_ctx* c) { return c->iter; }
void or_set_time(or_ctx* c, double t, int iter) { c->t = t; c->iter = iter; }

/* ------------------------------------------------------------------ */
/* scalar exports for the known-answer tests */
double or_evol_from_pres(double gamma, double mmw, double rho, double P)
{
    eos_t e = {gamma, gamma - 1, mmw};
    return evol_from_pres(&e, rho, P);
}
double or_pres_from_evol(double gamma, double mmw, double rho, double evol)
{
    eos_t e = {gamma, gamma - 1, mmw};
    return pres_from_evol(&e, rho, evol);
}
double or_speed_of_sound(double gamma, double rho, double P)
{
    eos_t e = {gamma, gamma - 1, 1};
    return sound(&e, rho, P);
}
static cons_t unpack_cons(int nd, const double* a)
{
    cons_t q;
    q.rho = a[0];
    for (int d = 0; d < 3; ++d) q.m[d] = d < nd ? a[1 + d] : 0;
    q.E = a[1 + nd];
    return q;
}
static prim_t unpack_prim(int nd, const double* a)
{
    prim_t p;
    p.rho = a[0];
    for (int d = 0; d < 3; ++d) p.u[d] = d < nd ? a[1 + d] : 0;
    p.P = a[1 + nd];
    return p;
}
static void pack(int nd, const cons_t* q, double* a)
{
    a[0] = q->rho;
    for (int d = 0; d < nd; ++d) a[1 + d] = q->m[d];
    a[1 + nd] = q->E;
}
double or_ek_cons(int nd, const double* a) { cons_t q = unpack_cons(nd, a); return ek_cons(nd, &q); }
double or_ek_prim(int nd, const double* a) { prim_t p = unpack_prim(nd, a); return ek_prim(nd, &p); }
void or_to_prim(int nd, double gamma, double mmw, const double* a, double* out)
{
    eos_t e = {gamma, gamma - 1, mmw};
    cons_t q = unpack_cons(nd, a);
    prim_t p = to_prim(nd, &e, &q);
    out[0] = p.rho;
    for (int d = 0; d < nd; ++d) out[1 + d] = p.u[d];
    out[1 + nd] = p.P;
}
void or_to_cons(int nd, double gamma, double mmw, const double* a, double* out)
{
    eos_t e = {gamma, gamma - 1, mmw};
    prim_t p = unpack_prim(nd, a);
    cons_t q = to_cons(nd, &e, &p);
    pack(nd, &q, out);
}
void or_flux_prim(int nd, double gamma, double mmw, const double* a, int loc, double* out)
{
    eos_t e = {gamma, gamma - 1, mmw};
    prim_t p = unpack_prim(nd, a);
    flux_t f = flux_prim(nd, &e, &p, loc);
    pack(nd, &f, out);
}
void or_flux_cons(int nd, double gamma, double mmw, const double* a, int loc, double* out)
{
    eos_t e = {gamma, gamma - 1, mmw};
    cons_t q = unpack_cons(nd, a);
    flux_t f = flux_cons(nd, &e, &q, loc);
    pack(nd, &f, out);
}
void or_riemann(int solver, int nd, double gamma, double mmw, const double* a, const double* b, int loc, double* out)
{
    eos_t e = {gamma, gamma - 1, mmw};
    cons_t qL = unpack_cons(nd, a), qR = unpack_cons(nd, b);
    flux_t f = riemann(solver, nd, &e, &qL, &qR, loc);
    pack(nd, &f, out);
}
double or_limiter(int kind, double dR, double dL) { return limiter(kind, dR, dL); }
