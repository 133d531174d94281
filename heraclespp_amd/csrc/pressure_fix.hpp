// pressure_fix.hpp -- the per-cell pressure fix of hydro/pressure_fix.hpp:41-171, shared by the
// operator kernel (ref_kernels.hip, face states read from the rec arrays) and the fused step
// (pressure_fix_fused.hip, face states recomputed by the Hancock predictor for the few cells
// the fix touches).
#pragma once

#include "grid.hpp"
#include "physics.hpp"

namespace hclpp {

// Whether the pressure fix rewrites a cell's energy: hydro/pressure_fix.hpp:61-68, on the
// state at time n only (E - e_k < eps e_k).
template <int ND>
__host__ __device__ __forceinline__ bool pressure_fix_applies(const State<ND>& cons, double eps)
{
    double const ekin = kinetic_cons(cons);
    double const evol = cons[ND + 1] - ekin;
    return evol < eps * ekin;
}

// hydro/pressure_fix.hpp:43-169 for cell (i, j, k), where pressure_fix_applies holds.
//   cell(c)         conservative state at time n of the cell with flat index c;
//   face(c, s, d)   its Hancock face state (conservative) on side s of dim d;
//   cons_new        the Godunov update of the cell (before the fix).
// Returns the new total energy.  The double loop over (idim, ipos) -- neighbours along idim,
// velocities and face areas of direction ipos, divU and divUE accumulated over both -- and
// the operation order are the reference's.
template <int ND, int GEOM, class Cell, class Face>
__device__ __forceinline__ double pressure_fix_energy(const GridDev& G, const PerfectGas& eos, double dt, double eps, int i, int j, int k,
                                                      const State<ND>& cons, const State<ND>& cons_new, Cell&& cell, Face&& face)
{
    int64_t const id = flat(G, i, j, k);
    State<ND> const prim = cons_to_prim(cons, eos);
    double const ekin = kinetic_cons(cons);
    double const evol = cons[ND + 1] - ekin;
    double const evol_new = cons_new[ND + 1] - kinetic_cons(cons_new);
    double const de_god = evol_new - evol;
    double const eps2 = eps / 10;
    double divU = 0, divUE = 0, einL = 0, einR = 0;
    double alpha = (evol * evol) / (ekin * ekin * eps * eps);
    if (evol < eps2 * ekin) alpha = 0;
    double out = cons[ND + 1];
#pragma unroll
    for (int idim = 0; idim < ND; ++idim) {
        int const di = idim == 0, dj = idim == 1, dk = idim == 2;
        int64_t const idm = flat(G, i - di, j - dj, k - dk);
        int64_t const idp = flat(G, i + di, j + dj, k + dk);
        State<ND> const primL = cons_to_prim(face(id, 0, idim), eos);
        State<ND> const primR = cons_to_prim(face(id, 1, idim), eos);
        State<ND> const prim_mR = cons_to_prim(face(idm, 1, idim), eos);
        State<ND> const prim_pL = cons_to_prim(face(idp, 0, idim), eos);
        State<ND> const cons_p = cell(idp);
        State<ND> const cons_m = cell(idm);
        double const dv = cell_volume<ND, GEOM>(G, i, j, k);
#pragma unroll
        for (int ipos = 0; ipos < ND; ++ipos) {
            double const us = (prim_mR[1 + ipos] + primL[1 + ipos]) / 2;
            double const us_p = (primR[1 + ipos] + prim_pL[1 + ipos]) / 2;
            double const dsp = face_area<ND, GEOM>(G, ipos, i + di, j + dj, k + dk);
            double const ds0 = face_area<ND, GEOM>(G, ipos, i, j, k);
            divU += ((us_p * dsp) - (us * ds0)) / dv;
            einR = (us_p > 0) ? cons[ND + 1] - kinetic_cons(cons) : cons_p[ND + 1] - kinetic_cons(cons_p);
            einL = (us < 0) ? cons[ND + 1] - kinetic_cons(cons) : cons_m[ND + 1] - kinetic_cons(cons_m);
            divUE += ((us_p * einR * dsp) - (us * einL * ds0)) / dv;
            double const de_pf = dt * ((-prim[ND + 1] * divU) - divUE);
            double sum = 0;
#pragma unroll
            for (int idr = 0; idr < ND; ++idr) sum += cons[1 + idr] * cons[1 + idr] / cons[0];
            double const de_tot = (alpha * de_god) + ((1 - alpha) * de_pf);
            out = evol + de_tot + (0.5 * sum);
        }
    }
    return out;
}

}  // namespace hclpp
