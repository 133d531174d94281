// adjoint_host.cpp -- host build of the reverse-mode step bodies (adjoint_step.hpp), for
// CPU tests: the same per-cell functions the GPU kernels run, looped over the cells in
// the pass order of launch_step_adjoint, on host arrays.  Cartesian or spherical geometry,
// uniform gravity.  Test infrastructure only.
#include <cmath>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "adjoint_step.hpp"

using namespace hclpp;

namespace {

template <int ND, int GEOM, int RS, int LIM>
double run(const GridDev& G, const PerfectGas& eos, const BcSpec& bc, const adj::AdjArrays& A)
{
    int lo[3], hi[3];
    double lg = 0;
    auto loop = [&](auto&& body) {
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) body(i, j, k);
    };
    for (int d = 0; d < 3; ++d) { lo[d] = 0; hi[d] = G.nwg[d]; }
    loop([&](int i, int j, int k) { adj::cell_prim<ND>(G, eos, A, i, j, k); });
    for (int d = 0; d < 3; ++d) {
        lo[d] = d < ND ? G.ngh[d] - 1 : 0;
        hi[d] = d < ND ? G.ngh[d] + G.n[d] + 1 : 1;
    }
    loop([&](int i, int j, int k) { adj::cell_faces<ND, LIM, GEOM>(G, eos, A, i, j, k); });
    auto flux = [&](auto dir) {
        constexpr int f = decltype(dir)::value;
        for (int d = 0; d < 3; ++d) {
            lo[d] = G.ngh[d];
            hi[d] = G.ngh[d] + G.n[d] + (d == f ? 1 : 0);
        }
        loop([&](int i, int j, int k) { lg += adj::face_flux<ND, RS, f, GEOM>(G, eos, A, i, j, k); });
    };
    flux(std::integral_constant<int, 0>{});
    if constexpr (ND >= 2) flux(std::integral_constant<int, 1>{});
    if constexpr (ND == 3) flux(std::integral_constant<int, 2>{});
    for (int d = 0; d < 3; ++d) {
        lo[d] = d < ND ? G.ngh[d] - 1 : 0;
        hi[d] = d < ND ? G.ngh[d] + G.n[d] + 1 : 1;
    }
    loop([&](int i, int j, int k) { lg += adj::cell_hancock<ND, LIM, GEOM>(G, eos, A, i, j, k); });
    for (int d = 0; d < 3; ++d) { lo[d] = G.ngh[d]; hi[d] = G.ngh[d] + G.n[d]; }
    loop([&](int i, int j, int k) { lg += adj::cell_gather<ND, LIM>(G, eos, bc, A, i, j, k); });
    return lg;
}

template <int ND, int GEOM, int RS>
double run_lim(int lim, const GridDev& G, const PerfectGas& eos, const BcSpec& bc, const adj::AdjArrays& A)
{
    switch (lim) {
    case kConstant: return run<ND, GEOM, RS, kConstant>(G, eos, bc, A);
    case kVanLeer: return run<ND, GEOM, RS, kVanLeer>(G, eos, bc, A);
    case kMinmod: return run<ND, GEOM, RS, kMinmod>(G, eos, bc, A);
    default: return run<ND, GEOM, RS, kVanAlbada>(G, eos, bc, A);
    }
}

template <int ND, int GEOM>
double run_rs(int rs, int lim, const GridDev& G, const PerfectGas& eos, const BcSpec& bc, const adj::AdjArrays& A)
{
    return rs == kHLL ? run_lim<ND, GEOM, kHLL>(lim, G, eos, bc, A) : run_lim<ND, GEOM, kHLLC>(lim, G, eos, bc, A);
}

template <int ND>
double run_geom(int geom, int rs, int lim, const GridDev& G, const PerfectGas& eos, const BcSpec& bc, const adj::AdjArrays& A)
{
    return geom == kSpherical ? run_rs<ND, kSpherical>(rs, lim, G, eos, bc, A) : run_rs<ND, kCartesian>(rs, lim, G, eos, bc, A);
}

}  // namespace

extern "C" {

// One reverse step.  geom 0 Cartesian, 1 spherical; n[3] interior extents (1 on inactive dims),
// ng ghost layers, x[d] node coordinates (n[d] + 2 ng + 1 entries on active dims, 2 on inactive
// ones), bc[6] kinds, g[3] uniform gravity.  State and adjoint arrays are ghost-inclusive
// LayoutLeft planes; rho/rhou/E must have their ghosts filled.  Returns dL/dgamma of this step in
// *dgamma.
int adjh_step(int nd, int geom, int rs, int lim, const int* n, int ng, const double* const* x, const int* bcs, double gamma, const double* g,
              double dt, const double* rho, const double* rhou, const double* E, const double* lrho_new, const double* lrhou_new,
              const double* lE_new, double* lrho, double* lrhou, double* lE, double* dgamma)
{
    if (nd < 1 || nd > 3 || ng < 2 || ng > adj::kMaxGhosts || (rs != kHLL && rs != kHLLC) || (geom != kCartesian && geom != kSpherical))
        return 1;
    GridDev G{};
    std::vector<double> tab;
    size_t off[3][3];
    for (int d = 0; d < 3; ++d) {
        G.n[d] = n[d];
        G.ngh[d] = d < nd ? ng : 0;
        G.nwg[d] = G.n[d] + 2 * G.ngh[d];
    }
    for (int d = 0; d < 3; ++d) {
        int const m = G.nwg[d];
        off[d][0] = tab.size();
        tab.insert(tab.end(), x[d], x[d] + m + 1);
        off[d][1] = tab.size();
        for (int i = 0; i < m; ++i) tab.push_back(x[d][i + 1] - x[d][i]);
        off[d][2] = tab.size();
        for (int i = 0; i < m; ++i) tab.push_back((x[d][i] + x[d][i + 1]) / 2);
    }
    // the spherical geometry's trigonometric tables (mesh/geometry.cpp), as the library builds them
    int const n1 = G.nwg[1];
    size_t const off_dcos = tab.size();
    for (int j = 0; j < n1; ++j) tab.push_back(std::cos(x[1][j]) - std::cos(x[1][j + 1]));
    size_t const off_sin = tab.size();
    for (int j = 0; j <= n1; ++j) tab.push_back(std::sin(x[1][j]));
    size_t const off_cot = tab.size();
    for (int j = 0; j < n1; ++j) {
        double const mid = (x[1][j] + x[1][j + 1]) / 2;
        tab.push_back(std::cos(mid) / std::sin(mid));
    }
    size_t const off_sinc = tab.size();
    for (int j = 0; j < n1; ++j) tab.push_back(std::sin((x[1][j] + x[1][j + 1]) / 2));
    for (int d = 0; d < 3; ++d) {
        G.x[d] = tab.data() + off[d][0];
        G.dx[d] = tab.data() + off[d][1];
        G.xc[d] = tab.data() + off[d][2];
        G.g[d] = g[d];
    }
    G.dcos = tab.data() + off_dcos;
    G.sin_node = tab.data() + off_sin;
    G.cot_mid = tab.data() + off_cot;
    G.sin_center = tab.data() + off_sinc;
    G.plane = (int64_t)G.nwg[0] * G.nwg[1] * G.nwg[2];
    G.gravity = kUniformGravity;
    PerfectGas eos{gamma, gamma - 1, 1.0};
    BcSpec bc{};
    for (int f = 0; f < 6; ++f) bc.kind[f] = f < 2 * nd ? bcs[f] : 0;
    std::vector<double> prim((size_t)G.plane * (nd + 2)), face((size_t)G.plane * adj::face_planes(nd));
    adj::AdjArrays A{rho, rhou, E, prim.data(), face.data(), lrho_new, lrhou_new, lE_new, lrho, lrhou, lE, dt};
    double lg;
    switch (nd) {
    case 1: lg = run_geom<1>(geom, rs, lim, G, eos, bc, A); break;
    case 2: lg = run_geom<2>(geom, rs, lim, G, eos, bc, A); break;
    default: lg = run_geom<3>(geom, rs, lim, G, eos, bc, A); break;
    }
    *dgamma = lg;
    return 0;
}

}  // extern "C"
