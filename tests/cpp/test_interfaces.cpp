// test_interfaces.cpp -- the C++ host side (include/hclpp/hydro.hpp) driven the way the
// reference's tests and main loop drive the operator interfaces, checked against the oracle.
// Built by __graft_entry__.build(); run by tests/test_gpu_cpp_interfaces.py on a GPU.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "hclpp/hydro.hpp"
#include "hclpp_oracle.h"

using namespace hclpp::hip;

static int g_fail = 0;
static int g_pass = 0;
#define EXPECT(cond, ...)                                                    \
    do {                                                                     \
        if (cond) {                                                          \
            ++g_pass;                                                        \
        } else {                                                             \
            ++g_fail;                                                        \
            std::printf("FAIL %s:%d: %s ", __FILE__, __LINE__, #cond);      \
            std::printf(__VA_ARGS__);                                        \
            std::printf("\n");                                               \
        }                                                                    \
    } while (0)

template <class F>
static bool throws_runtime(F f, char const* needle)
{
    try {
        f();
    } catch (std::runtime_error const& e) {
        return std::strstr(e.what(), needle) != nullptr;
    }
    return false;
}

// tests/conversions.cpp:17-103 on the device through conv_prim_to_cons / conv_cons_to_prim
static void test_conversions()
{
    int const n = 10;
    PerfectGas const eos(2, 1);
    for (int nd = 1; nd <= 3; ++nd) {
        Grid const grid(nd, {n, n, n}, 0, Geometry::Geom_cartesian, {0, 0, 0}, {1, 1, 1});
        auto const& w = grid.Nx_local_wg;
        KV_double_3d rho("rho", w[0], w[1], w[2]), P("P", w[0], w[1], w[2]), E("E", w[0], w[1], w[2]);
        KV_double_4d u("u", w[0], w[1], w[2], nd), rhou("rhou", w[0], w[1], w[2], nd);
        size_t const cells = rho.size();
        rho.copy_from_host(std::vector<double>(cells, 2.));
        u.copy_from_host(std::vector<double>(cells * nd, -1.));
        P.copy_from_host(std::vector<double>(cells, 10.));
        conv_prim_to_cons(grid.range.all_ghosts(), eos, grid, rho, u, P, rhou, E);
        double prim[5] = {2., -1., -1., -1., 10.}, cons[5];
        prim[1 + nd] = 10.;
        or_to_cons(nd, 2., 1., prim, cons);
        auto const hr = rhou.to_host();
        auto const he = E.to_host();
        bool ok = true;
        for (size_t c = 0; c < cells; ++c) {
            ok = ok && he[c] == cons[1 + nd];
            for (int d = 0; d < nd; ++d) ok = ok && hr[c + cells * d] == cons[1 + d];
        }
        EXPECT(ok, "prim->cons nd=%d", nd);
        rhou.copy_from_host(std::vector<double>(cells * nd, -2.));
        E.copy_from_host(std::vector<double>(cells, 10.));
        conv_cons_to_prim(grid.range.all_ghosts(), eos, grid, rho, rhou, E, u, P);
        double c2[5] = {2., -2., -2., -2., 10.}, p2[5];
        c2[1 + nd] = 10.;
        or_to_prim(nd, 2., 1., c2, p2);
        auto const hu = u.to_host();
        auto const hp = P.to_host();
        ok = true;
        for (size_t c = 0; c < cells; ++c) {
            ok = ok && hp[c] == p2[1 + nd];
            for (int d = 0; d < nd; ++d) ok = ok && hu[c + cells * d] == p2[1 + d];
        }
        EXPECT(ok, "cons->prim nd=%d", nd);
    }
}

static void test_errors()
{
    EXPECT(throws_runtime([] { factory_face_reconstruction("Superbee"); }, "Unknown face reconstruction algorithm: Superbee."), "factory");
    EXPECT(throws_runtime([] { factory_godunov_scheme<UniformGravity>("Roe", PerfectGas(1.4, 1)); }, "Invalid riemann solver: Roe."), "factory");
    bool dom = false;
    try {
        PerfectGas(1.0, 1.0);
    } catch (std::domain_error const&) {
        dom = true;
    }
    EXPECT(dom, "PerfectGas(1, 1) must throw std::domain_error");
    Range const r({0, 0, 0}, {4, 4, 4}, 2, 3);
    EXPECT(throws_runtime([&] { (void)r.with_ghosts(3); }, "NgEff > Ng"), "with_ghosts");
}

// KH-like random state: the reference main loop (main.cpp:354-531) written with the interface
// classes, against the oracle's driver; every conservative array must be bit-identical.
static void test_operator_loop()
{
    int const nd = 2, n = 40;
    double const cfl = 0.4, gamma = 1.4;
    Grid const grid(nd, {n, n, 1}, 2, Geometry::Geom_cartesian, {0, 0, 0}, {1, 1, 1});
    PerfectGas const eos(gamma, 1.);
    UniformGravity const g({0.0, -0.1, 0.0});
    auto const& w = grid.Nx_local_wg;
    size_t const cells = static_cast<size_t>(w[0]) * w[1] * w[2];
    std::vector<double> hrho(cells, 0.), hu(cells * nd, 0.), hP(cells, 0.);
    for (int j = 2; j < n + 2; ++j)
        for (int i = 2; i < n + 2; ++i) {
            size_t const c = i + w[0] * j;
            double const x = (i - 1.5) / n, y = (j - 1.5) / n;
            hrho[c] = std::abs(y - 0.5) < 0.25 ? 2.0 : 1.0;
            hu[c] = std::abs(y - 0.5) < 0.25 ? 0.5 : -0.5;
            hu[c + cells] = 0.01 * std::sin(4 * M_PI * x) * std::sin(7.0 * x + 3.0 * y);
            hP[c] = 2.5;
        }
    // oracle driver
    or_config oc {};
    oc.ndim = nd;
    oc.geom = OR_GEOM_CARTESIAN;
    oc.limiter = OR_LIM_VANLEER;
    oc.riemann = OR_RS_HLLC;
    oc.gravity = OR_GRAV_UNIFORM;
    oc.ng = 2;
    oc.n[0] = n;
    oc.n[1] = n;
    oc.n[2] = 1;
    for (int f = 0; f < 6; ++f) {
        oc.bc[f] = f < 2 ? OR_BC_PERIODIC : OR_BC_REFLEXIVE;
        oc.bc_order[f] = f;
    }
    oc.gamma = gamma;
    oc.mmw = 1.;
    oc.g[1] = -0.1;
    for (int d = 0; d < 3; ++d) oc.x[d] = grid.x[d].data();
    or_ctx* o = or_create(&oc);
    or_init_from_prim(o, hrho.data(), hu.data(), hP.data(), nullptr);
    int const nsteps = 8;
    for (int s = 0; s < nsteps; ++s) {
        double dt;
        int ex;
        or_step(o, cfl, 1e30, nsteps, &dt, &ex);
    }
    // device: the reference sequence through the interfaces
    KV_double_3d rho("rho", w[0], w[1], w[2]), P("P", w[0], w[1], w[2]), E("E", w[0], w[1], w[2]);
    KV_double_4d u("u", w[0], w[1], w[2], nd), rhou("rhou", w[0], w[1], w[2], nd), fx("fx", w[0], w[1], w[2], 0);
    KV_double_3d rho_new("rho_new", w[0], w[1], w[2]), E_new("E_new", w[0], w[1], w[2]);
    KV_double_4d rhou_new("rhou_new", w[0], w[1], w[2], nd), fx_new("fx_new", w[0], w[1], w[2], 0);
    KV_double_5d rho_rec("rho_rec", w[0], w[1], w[2], 2, nd), E_rec("E_rec", w[0], w[1], w[2], 2, nd), P_rec("P_rec", w[0], w[1], w[2], 2, nd);
    KV_double_6d rhou_rec("rhou_rec", w[0], w[1], w[2], 2, nd, nd), u_rec("u_rec", w[0], w[1], w[2], 2, nd, nd),
            fx_rec("fx_rec", w[0], w[1], w[2], 2, nd, 0);
    rho.copy_from_host(hrho);
    u.copy_from_host(hu);
    P.copy_from_host(hP);
    DistributedBoundaryCondition const bcs(grid, {"Periodic", "Periodic", "Reflexive", "Reflexive"});
    auto reconstruction = std::make_unique<MUSCLHancockHydroReconstruction<UniformGravity>>(
            factory_face_reconstruction("VanLeer"), std::make_unique<HancockExtrapolationReconstruction<UniformGravity>>(eos), eos, P_rec, u_rec);
    auto godunov = factory_godunov_scheme<UniformGravity>("HLLC", eos);
    conv_prim_to_cons(grid.range.no_ghosts(), eos, grid, rho, u, P, rhou, E);  // main.cpp:305
    bcs(grid, rho, rhou, E, fx);                                                 // :307
    conv_cons_to_prim(grid.range.all_ghosts(), eos, grid, rho, rhou, E, u, P);  // :311
    for (int s = 0; s < nsteps; ++s) {
        double const dt = cfl * time_step(grid.range.all_ghosts(), eos, grid, rho, u, P);  // :355
        double const emin = minimum_internal_energy(grid.range.no_ghosts(), grid, rho, rhou, E);
        EXPECT(emin > 0, "internal energy");
        reconstruction->execute(grid.range.with_ghosts(1), grid, g, dt / 2, rho, u, P, fx, rho_rec, rhou_rec, E_rec, fx_rec);
        godunov->execute(grid.range.no_ghosts(), grid, g, dt, rho, rhou, E, fx, rho_rec, rhou_rec, E_rec, fx_rec, rho_new, rhou_new, E_new, fx_new);
        bcs(grid, rho_new, rhou_new, E_new, fx_new);
        std::swap(rho, rho_new);
        std::swap(rhou, rhou_new);
        std::swap(E, E_new);
        conv_cons_to_prim(grid.range.all_ghosts(), eos, grid, rho, rhou, E, u, P);
    }
    auto const r = rho.to_host();
    auto const m = rhou.to_host();
    auto const e = E.to_host();
    EXPECT(std::memcmp(r.data(), or_state(o, 0), r.size() * 8) == 0, "rho differs from the oracle");
    EXPECT(std::memcmp(m.data(), or_state(o, 1), m.size() * 8) == 0, "rhou differs from the oracle");
    EXPECT(std::memcmp(e.data(), or_state(o, 2), e.size() * 8) == 0, "E differs from the oracle");

    // the fused driver on the same problem
    HydroDriver<UniformGravity> drv(grid, eos, g, "VanLeer", "HLLC", bcs, cfl);
    rho.copy_from_host(hrho);
    u.copy_from_host(hu);
    P.copy_from_host(hP);
    drv.init_from_prim(rho, u, P);
    int const steps = drv.run(1e30, nsteps);
    EXPECT(steps == nsteps, "steps %d", steps);
    EXPECT(std::memcmp(drv.rho().to_host().data(), or_state(o, 0), r.size() * 8) == 0, "fused rho differs from the oracle");
    EXPECT(std::memcmp(drv.rhou().to_host().data(), or_state(o, 1), m.size() * 8) == 0, "fused rhou differs from the oracle");
    EXPECT(std::memcmp(drv.E().to_host().data(), or_state(o, 2), e.size() * 8) == 0, "fused E differs from the oracle");
    EXPECT(drv.time() == or_time(o), "time %.17g vs %.17g", drv.time(), or_time(o));
    or_destroy(o);
}

// main.cpp:305-531 with internal-mass gravity (rebuilt from rho after every step, :309, :528),
// 1D spherical, against the oracle driver (one cell per shell: the shell sums are exact).
static void test_internal_mass_loop()
{
    int const nd = 1, n = 64, nsteps = 20;
    double const cfl = 0.4, gamma = 1.4, M = 2.0e10;
    Grid const grid(nd, {n, 1, 1}, 2, Geometry::Geom_spherical, {0.5, 0, 0}, {1.5, 1, 1});
    PerfectGas const eos(gamma, 1.);
    auto const& w = grid.Nx_local_wg;
    size_t const cells = static_cast<size_t>(w[0]);
    std::vector<double> hrho(cells, 0.), hu(cells, 0.), hP(cells, 0.);
    for (int i = 2; i < n + 2; ++i) {
        double const r = 0.5 + (i - 1.5) / n;
        hrho[i] = 1.0 + 0.5 * std::exp(-(r - 1.0) * (r - 1.0) / 0.01);
        hu[i] = 0.1 * std::sin(9.0 * r);
        hP[i] = 1.0 + 0.1 * std::cos(5.0 * r);
    }
    or_config oc {};
    oc.ndim = nd;
    oc.geom = OR_GEOM_SPHERICAL;
    oc.limiter = OR_LIM_VANLEER;
    oc.riemann = OR_RS_HLL;
    oc.gravity = OR_GRAV_INTERNAL_MASS;
    oc.central_mass = M;
    oc.ng = 2;
    oc.n[0] = n;
    oc.n[1] = 1;
    oc.n[2] = 1;
    oc.bc[0] = OR_BC_REFLEXIVE;
    oc.bc[1] = OR_BC_NULLGRADIENT;
    oc.bc_order[0] = 0;
    oc.bc_order[1] = 1;
    oc.gamma = gamma;
    oc.mmw = 1.;
    for (int d = 0; d < 3; ++d) oc.x[d] = grid.x[d].data();
    or_ctx* o = or_create(&oc);
    or_init_from_prim(o, hrho.data(), hu.data(), hP.data(), nullptr);
    for (int s = 0; s < nsteps; ++s) {
        double dt;
        int ex;
        or_step(o, cfl, 1e30, nsteps, &dt, &ex);
    }
    KV_double_3d rho("rho", w[0], 1, 1), P("P", w[0], 1, 1), E("E", w[0], 1, 1);
    KV_double_4d u("u", w[0], 1, 1, nd), rhou("rhou", w[0], 1, 1, nd), fx("fx", w[0], 1, 1, 0);
    KV_double_3d rho_new("rho_new", w[0], 1, 1), E_new("E_new", w[0], 1, 1);
    KV_double_4d rhou_new("rhou_new", w[0], 1, 1, nd), fx_new("fx_new", w[0], 1, 1, 0);
    KV_double_5d rho_rec("rho_rec", w[0], 1, 1, 2, nd), E_rec("E_rec", w[0], 1, 1, 2, nd), P_rec("P_rec", w[0], 1, 1, 2, nd);
    KV_double_6d rhou_rec("rhou_rec", w[0], 1, 1, 2, nd, nd), u_rec("u_rec", w[0], 1, 1, 2, nd, nd), fx_rec("fx_rec", w[0], 1, 1, 2, nd, 0);
    rho.copy_from_host(hrho);
    u.copy_from_host(hu);
    P.copy_from_host(hP);
    DistributedBoundaryCondition const bcs(grid, {"Reflexive", "NullGradient"});
    auto reconstruction = std::make_unique<MUSCLHancockHydroReconstruction<InternalMassGravity>>(
            factory_face_reconstruction("VanLeer"), std::make_unique<HancockExtrapolationReconstruction<InternalMassGravity>>(eos), eos, P_rec,
            u_rec);
    auto godunov = factory_godunov_scheme<InternalMassGravity>("HLL", eos);
    conv_prim_to_cons(grid.range.no_ghosts(), eos, grid, rho, u, P, rhou, E);  // main.cpp:305
    bcs(grid, rho, rhou, E, fx);                                                 // :307
    auto g = make_internal_mass_gravity(M, grid, rho);                          // :309
    conv_cons_to_prim(grid.range.all_ghosts(), eos, grid, rho, rhou, E, u, P);  // :311
    for (int s = 0; s < nsteps; ++s) {
        double const dt = cfl * time_step(grid.range.all_ghosts(), eos, grid, rho, u, P);
        reconstruction->execute(grid.range.with_ghosts(1), grid, g, dt / 2, rho, u, P, fx, rho_rec, rhou_rec, E_rec, fx_rec);
        godunov->execute(grid.range.no_ghosts(), grid, g, dt, rho, rhou, E, fx, rho_rec, rhou_rec, E_rec, fx_rec, rho_new, rhou_new, E_new, fx_new);
        bcs(grid, rho_new, rhou_new, E_new, fx_new);
        std::swap(rho, rho_new);
        std::swap(rhou, rhou_new);
        std::swap(E, E_new);
        conv_cons_to_prim(grid.range.all_ghosts(), eos, grid, rho, rhou, E, u, P);
        g = make_internal_mass_gravity(M, grid, rho);  // :528
    }
    auto const r = rho.to_host();
    auto const m = rhou.to_host();
    auto const e = E.to_host();
    EXPECT(std::memcmp(r.data(), or_state(o, 0), r.size() * 8) == 0, "rho differs from the oracle (internal mass)");
    EXPECT(std::memcmp(m.data(), or_state(o, 1), m.size() * 8) == 0, "rhou differs from the oracle (internal mass)");
    EXPECT(std::memcmp(e.data(), or_state(o, 2), e.size() * 8) == 0, "E differs from the oracle (internal mass)");
    or_destroy(o);
}

// pressure_fix after the Godunov update (main.cpp:459-475) on a cold uniform 2D stream, against the oracle.
static void test_pressure_fix_loop()
{
    int const nd = 2, n = 32, nsteps = 10;
    double const cfl = 0.4, gamma = 1.4, eps_pf = 1e-5;
    Grid const grid(nd, {n, n, 1}, 2, Geometry::Geom_cartesian, {0, 0, 0}, {1, 1, 1});
    PerfectGas const eos(gamma, 1.);
    auto const& w = grid.Nx_local_wg;
    size_t const cells = static_cast<size_t>(w[0]) * w[1];
    std::vector<double> hrho(cells, 1.), hu(cells * nd, 3.), hP(cells, 0.);
    for (size_t c = 0; c < cells; ++c) hP[c] = 1e-5 * (1.0 + 0.5 * std::sin(0.37 * c) * std::sin(0.11 * c));
    or_config oc {};
    oc.ndim = nd;
    oc.geom = OR_GEOM_CARTESIAN;
    oc.limiter = OR_LIM_VANLEER;
    oc.riemann = OR_RS_HLL;
    oc.gravity = OR_GRAV_UNIFORM;
    oc.ng = 2;
    oc.n[0] = n;
    oc.n[1] = n;
    oc.n[2] = 1;
    for (int f = 0; f < 6; ++f) {
        oc.bc[f] = OR_BC_PERIODIC;
        oc.bc_order[f] = f;
    }
    oc.gamma = gamma;
    oc.mmw = 1.;
    oc.pressure_fix = 1;
    oc.eps_pf = eps_pf;
    for (int d = 0; d < 3; ++d) oc.x[d] = grid.x[d].data();
    or_ctx* o = or_create(&oc);
    or_init_from_prim(o, hrho.data(), hu.data(), hP.data(), nullptr);
    for (int s = 0; s < nsteps; ++s) {
        double dt;
        int ex;
        or_step(o, cfl, 1e30, nsteps, &dt, &ex);
    }
    UniformGravity const g({0.0, 0.0, 0.0});
    KV_double_3d rho("rho", w[0], w[1], 1), P("P", w[0], w[1], 1), E("E", w[0], w[1], 1);
    KV_double_4d u("u", w[0], w[1], 1, nd), rhou("rhou", w[0], w[1], 1, nd), fx("fx", w[0], w[1], 1, 0);
    KV_double_3d rho_new("rho_new", w[0], w[1], 1), E_new("E_new", w[0], w[1], 1);
    KV_double_4d rhou_new("rhou_new", w[0], w[1], 1, nd), fx_new("fx_new", w[0], w[1], 1, 0);
    KV_double_5d rho_rec("rho_rec", w[0], w[1], 1, 2, nd), E_rec("E_rec", w[0], w[1], 1, 2, nd), P_rec("P_rec", w[0], w[1], 1, 2, nd);
    KV_double_6d rhou_rec("rhou_rec", w[0], w[1], 1, 2, nd, nd), u_rec("u_rec", w[0], w[1], 1, 2, nd, nd),
            fx_rec("fx_rec", w[0], w[1], 1, 2, nd, 0);
    rho.copy_from_host(hrho);
    u.copy_from_host(hu);
    P.copy_from_host(hP);
    DistributedBoundaryCondition const bcs(grid, {"Periodic", "Periodic", "Periodic", "Periodic"});
    auto reconstruction = std::make_unique<MUSCLHancockHydroReconstruction<UniformGravity>>(
            factory_face_reconstruction("VanLeer"), std::make_unique<HancockExtrapolationReconstruction<UniformGravity>>(eos), eos, P_rec, u_rec);
    auto godunov = factory_godunov_scheme<UniformGravity>("HLL", eos);
    conv_prim_to_cons(grid.range.no_ghosts(), eos, grid, rho, u, P, rhou, E);
    bcs(grid, rho, rhou, E, fx);
    conv_cons_to_prim(grid.range.all_ghosts(), eos, grid, rho, rhou, E, u, P);
    for (int s = 0; s < nsteps; ++s) {
        double const dt = cfl * time_step(grid.range.all_ghosts(), eos, grid, rho, u, P);
        reconstruction->execute(grid.range.with_ghosts(1), grid, g, dt / 2, rho, u, P, fx, rho_rec, rhou_rec, E_rec, fx_rec);
        godunov->execute(grid.range.no_ghosts(), grid, g, dt, rho, rhou, E, fx, rho_rec, rhou_rec, E_rec, fx_rec, rho_new, rhou_new, E_new, fx_new);
        pressure_fix(grid.range.no_ghosts(), eos, grid, dt, eps_pf, rho, rhou, E, rho_rec, rhou_rec, E_rec, rho_new, rhou_new, E_new);
        bcs(grid, rho_new, rhou_new, E_new, fx_new);
        std::swap(rho, rho_new);
        std::swap(rhou, rhou_new);
        std::swap(E, E_new);
        conv_cons_to_prim(grid.range.all_ghosts(), eos, grid, rho, rhou, E, u, P);
    }
    auto const e = E.to_host();
    auto const r = rho.to_host();
    EXPECT(std::memcmp(r.data(), or_state(o, 0), r.size() * 8) == 0, "rho differs from the oracle (pressure fix)");
    EXPECT(std::memcmp(e.data(), or_state(o, 2), e.size() * 8) == 0, "E differs from the oracle (pressure fix)");
    or_destroy(o);
}

// main.cpp:305-531 with one passive scalar (nfx = 1, a tracer of the dense band): the reference
// operator sequence through the interface classes and the fused driver (hclpp_step_fx), both
// against the oracle driver; every array, the scalar included, must be bit-identical.
static void test_passive_scalar_loop(int const nfx)
{
    int const nd = 2, n = 32, nsteps = 10;
    double const cfl = 0.4, gamma = 1.4;
    Grid const grid(nd, {n, n, 1}, 2, Geometry::Geom_cartesian, {0, 0, 0}, {1, 1, 1});
    PerfectGas const eos(gamma, 1.);
    UniformGravity const g({0.0, 0.0, 0.0});
    auto const& w = grid.Nx_local_wg;
    size_t const cells = static_cast<size_t>(w[0]) * w[1] * w[2];
    std::vector<double> hrho(cells, 0.), hu(cells * nd, 0.), hP(cells, 0.), hfx(cells * nfx, 0.);
    for (int j = 2; j < n + 2; ++j)
        for (int i = 2; i < n + 2; ++i) {
            size_t const c = i + w[0] * j;
            double const x = (i - 1.5) / n, y = (j - 1.5) / n;
            bool const band = std::abs(y - 0.5) < 0.25;
            hrho[c] = band ? 2.0 : 1.0;
            hu[c] = band ? 0.5 : -0.5;
            hu[c + cells] = 0.01 * std::sin(4 * M_PI * x) * std::sin(5.0 * x + 2.0 * y);
            hP[c] = 2.5;
            // scalar 0 traces the band, 1 its complement, 2 a smooth field (nfx > 1: one fused
            // launch per scalar)
            for (int f = 0; f < nfx; ++f)
                hfx[c + cells * f] = f == 0 ? (band ? 1.0 : 0.0) : f == 1 ? (band ? 0.0 : 1.0) : 0.5 + 0.4 * std::sin(2 * M_PI * (x + f * y));
        }
    or_config oc {};
    oc.ndim = nd;
    oc.geom = OR_GEOM_CARTESIAN;
    oc.limiter = OR_LIM_VANLEER;
    oc.riemann = OR_RS_HLLC;
    oc.gravity = OR_GRAV_UNIFORM;
    oc.nfx = nfx;
    oc.ng = 2;
    oc.n[0] = n;
    oc.n[1] = n;
    oc.n[2] = 1;
    for (int f = 0; f < 6; ++f) {
        oc.bc[f] = OR_BC_PERIODIC;
        oc.bc_order[f] = f;
    }
    oc.gamma = gamma;
    oc.mmw = 1.;
    for (int d = 0; d < 3; ++d) oc.x[d] = grid.x[d].data();
    or_ctx* o = or_create(&oc);
    or_init_from_prim(o, hrho.data(), hu.data(), hP.data(), hfx.data());
    for (int s = 0; s < nsteps; ++s) {
        double dt;
        int ex;
        or_step(o, cfl, 1e30, nsteps, &dt, &ex);
    }
    size_t const bytes = cells * 8;
    // the reference sequence through the interfaces
    KV_double_3d rho("rho", w[0], w[1], w[2]), P("P", w[0], w[1], w[2]), E("E", w[0], w[1], w[2]);
    KV_double_4d u("u", w[0], w[1], w[2], nd), rhou("rhou", w[0], w[1], w[2], nd), fx("fx", w[0], w[1], w[2], nfx);
    KV_double_3d rho_new("rho_new", w[0], w[1], w[2]), E_new("E_new", w[0], w[1], w[2]);
    KV_double_4d rhou_new("rhou_new", w[0], w[1], w[2], nd), fx_new("fx_new", w[0], w[1], w[2], nfx);
    KV_double_5d rho_rec("rho_rec", w[0], w[1], w[2], 2, nd), E_rec("E_rec", w[0], w[1], w[2], 2, nd), P_rec("P_rec", w[0], w[1], w[2], 2, nd);
    KV_double_6d rhou_rec("rhou_rec", w[0], w[1], w[2], 2, nd, nd), u_rec("u_rec", w[0], w[1], w[2], 2, nd, nd),
            fx_rec("fx_rec", w[0], w[1], w[2], 2, nd, nfx);
    rho.copy_from_host(hrho);
    u.copy_from_host(hu);
    P.copy_from_host(hP);
    fx.copy_from_host(hfx);
    DistributedBoundaryCondition const bcs(grid, {"Periodic", "Periodic", "Periodic", "Periodic"});
    auto reconstruction = std::make_unique<MUSCLHancockHydroReconstruction<UniformGravity>>(
            factory_face_reconstruction("VanLeer"), std::make_unique<HancockExtrapolationReconstruction<UniformGravity>>(eos), eos, P_rec, u_rec);
    auto godunov = factory_godunov_scheme<UniformGravity>("HLLC", eos);
    conv_prim_to_cons(grid.range.no_ghosts(), eos, grid, rho, u, P, rhou, E);
    bcs(grid, rho, rhou, E, fx);
    conv_cons_to_prim(grid.range.all_ghosts(), eos, grid, rho, rhou, E, u, P);
    for (int s = 0; s < nsteps; ++s) {
        double const dt = cfl * time_step(grid.range.all_ghosts(), eos, grid, rho, u, P);
        reconstruction->execute(grid.range.with_ghosts(1), grid, g, dt / 2, rho, u, P, fx, rho_rec, rhou_rec, E_rec, fx_rec);
        godunov->execute(grid.range.no_ghosts(), grid, g, dt, rho, rhou, E, fx, rho_rec, rhou_rec, E_rec, fx_rec, rho_new, rhou_new, E_new, fx_new);
        bcs(grid, rho_new, rhou_new, E_new, fx_new);
        std::swap(rho, rho_new);
        std::swap(rhou, rhou_new);
        std::swap(E, E_new);
        std::swap(fx, fx_new);
        conv_cons_to_prim(grid.range.all_ghosts(), eos, grid, rho, rhou, E, u, P);
    }
    EXPECT(std::memcmp(rho.to_host().data(), or_state(o, 0), bytes) == 0, "scalar loop: rho differs from the oracle");
    EXPECT(std::memcmp(rhou.to_host().data(), or_state(o, 1), bytes * nd) == 0, "scalar loop: rhou differs from the oracle");
    EXPECT(std::memcmp(E.to_host().data(), or_state(o, 2), bytes) == 0, "scalar loop: E differs from the oracle");
    EXPECT(std::memcmp(fx.to_host().data(), or_state(o, 3), bytes * nfx) == 0, "scalar loop: fx differs from the oracle");
    // the fused driver
    HydroDriver<UniformGravity> drv(grid, eos, g, "VanLeer", "HLLC", bcs, cfl, nfx);
    KV_double_3d rho0("rho0", w[0], w[1], w[2]), P0("P0", w[0], w[1], w[2]);
    KV_double_4d u0("u0", w[0], w[1], w[2], nd), fx0("fx0", w[0], w[1], w[2], nfx);
    rho0.copy_from_host(hrho);
    u0.copy_from_host(hu);
    P0.copy_from_host(hP);
    fx0.copy_from_host(hfx);
    drv.init_from_prim(rho0, u0, P0, fx0);
    int const steps = drv.run(1e30, nsteps);
    EXPECT(steps == nsteps, "steps %d", steps);
    EXPECT(std::memcmp(drv.rho().to_host().data(), or_state(o, 0), bytes) == 0, "fused scalar run: rho differs from the oracle");
    EXPECT(std::memcmp(drv.rhou().to_host().data(), or_state(o, 1), bytes * nd) == 0, "fused scalar run: rhou differs from the oracle");
    EXPECT(std::memcmp(drv.E().to_host().data(), or_state(o, 2), bytes) == 0, "fused scalar run: E differs from the oracle");
    EXPECT(std::memcmp(drv.fx().to_host().data(), or_state(o, 3), bytes * nfx) == 0, "fused scalar run: fx differs from the oracle");
    EXPECT(drv.time() == or_time(o), "time %.17g vs %.17g", drv.time(), or_time(o));
    or_destroy(o);
}

// A user's own face reconstruction: the reference's limited linear reconstruction
// (limited_linear_reconstruction.hpp:34-65) with the Minmod limiter, computed on the host.  Being
// a user type, it makes MUSCLHancockHydroReconstruction compose the steps through the interfaces.
class HostMinmodReconstruction : public IFaceReconstruction
{
public:
    mutable int calls = 0;
    void execute(Range const& range, Grid const& grid, KV_cdouble_3d const& var, KV_double_5d const& var_rec) const override
    {
        ++calls;
        std::vector<double> const v = var.to_host();
        std::vector<double> rec = var_rec.to_host();
        auto const& w = grid.Nx_local_wg;
        size_t const plane = static_cast<size_t>(w[0]) * w[1] * w[2];
        hclpp_range const r = range.cells();
        for (int k = r.begin[2]; k < r.end[2]; ++k)
            for (int j = r.begin[1]; j < r.end[1]; ++j)
                for (int i = r.begin[0]; i < r.end[0]; ++i) {
                    int const idx[3] = {i, j, k};
                    size_t const c = i + static_cast<size_t>(w[0]) * (j + static_cast<size_t>(w[1]) * k);
                    for (int d = 0; d < grid.ndim; ++d) {
                        size_t const st = d == 0 ? 1 : (d == 1 ? static_cast<size_t>(w[0]) : static_cast<size_t>(w[0]) * w[1]);
                        auto const& x = grid.x[d];
                        int const a = idx[d];
                        double const dl = x[a + 1] - x[a], dlm = x[a] - x[a - 1], dlp = x[a + 2] - x[a + 1];
                        double const slope = or_limiter(OR_LIM_MINMOD, (v[c + st] - v[c]) / ((dl + dlp) / 2), (v[c] - v[c - st]) / ((dlm + dl) / 2));
                        rec[c + plane * (0 + 2 * d)] = v[c] - (dl / 2) * slope;
                        rec[c + plane * (1 + 2 * d)] = v[c] + (dl / 2) * slope;
                    }
                }
        var_rec.copy_from_host(rec);
    }
};

// Another user type: forwards to the library's reconstruction and counts the calls.
class CountingReconstruction : public IFaceReconstruction
{
    std::unique_ptr<IFaceReconstruction> m_inner;

public:
    mutable int calls = 0;
    explicit CountingReconstruction(std::unique_ptr<IFaceReconstruction> inner) : m_inner(std::move(inner)) {}
    void execute(Range const& range, Grid const& grid, KV_cdouble_3d const& var, KV_double_5d const& var_rec) const override
    {
        ++calls;
        m_inner->execute(range, grid, var, var_rec);
    }
};

// hydro_reconstruction.hpp:85-129 with user strategies: MUSCLHancockHydroReconstruction must call
// the injected face reconstruction for rho, each velocity component, P and each passive scalar, and
// the composed result must be bit-identical to the oracle's reconstruction (2D, Minmod, uniform
// gravity, one passive scalar; the library's own strategies take the one-call path and must agree).
static void test_user_face_reconstruction()
{
    int const nd = 2, n = 24, nfx = 1;
    double const gamma = 1.4, dt = 3e-3;
    Grid const grid(nd, {n, n - 5, 1}, 2, Geometry::Geom_cartesian, {0, 0, 0}, {1, 0.8, 1});
    PerfectGas const eos(gamma, 1.);
    UniformGravity const g({0.02, -0.1, 0.0});
    auto const& w = grid.Nx_local_wg;
    size_t const cells = static_cast<size_t>(w[0]) * w[1] * w[2];
    std::vector<double> hrho(cells), hu(cells * nd), hP(cells), hfx(cells * nfx);
    for (size_t c = 0; c < cells; ++c) {
        hrho[c] = 1.0 + 0.4 * std::sin(0.37 * c) * std::cos(0.05 * c);
        hu[c] = 0.3 * std::sin(0.11 * c + 1.0);
        hu[c + cells] = -0.2 * std::cos(0.07 * c);
        hP[c] = 1.0 + 0.3 * std::cos(0.23 * c);
        hfx[c] = 0.5 + 0.5 * std::sin(0.19 * c);
    }
    or_config oc {};
    oc.ndim = nd;
    oc.geom = OR_GEOM_CARTESIAN;
    oc.limiter = OR_LIM_MINMOD;
    oc.riemann = OR_RS_HLLC;
    oc.gravity = OR_GRAV_UNIFORM;
    oc.nfx = nfx;
    oc.ng = 2;
    oc.n[0] = n;
    oc.n[1] = n - 5;
    oc.n[2] = 1;
    for (int f = 0; f < 6; ++f) {
        oc.bc[f] = OR_BC_PERIODIC;
        oc.bc_order[f] = f;
    }
    oc.gamma = gamma;
    oc.mmw = 1.;
    oc.g[0] = 0.02;
    oc.g[1] = -0.1;
    for (int d = 0; d < 3; ++d) oc.x[d] = grid.x[d].data();
    or_ctx* o = or_create(&oc);
    int r6[6];
    or_range(o, 1, r6);
    size_t const fcells = cells * 2 * nd;
    std::vector<double> orho(fcells, 0.), orhou(fcells * nd, 0.), oE(fcells, 0.), ofx(fcells * nfx, 0.), oP(fcells, 0.), ou(fcells * nd, 0.);
    or_hydro_reconstruction(o, r6, dt, hrho.data(), hu.data(), hP.data(), hfx.data(), orho.data(), orhou.data(), oE.data(), ofx.data(),
                            oP.data(), ou.data());
    KV_double_3d rho("rho", w[0], w[1], w[2]), P("P", w[0], w[1], w[2]);
    KV_double_4d u("u", w[0], w[1], w[2], nd), fx("fx", w[0], w[1], w[2], nfx);
    rho.copy_from_host(hrho);
    u.copy_from_host(hu);
    P.copy_from_host(hP);
    fx.copy_from_host(hfx);
    for (int variant = 0; variant < 3; ++variant) {
        KV_double_5d rho_rec("rho_rec", w[0], w[1], w[2], 2, nd), E_rec("E_rec", w[0], w[1], w[2], 2, nd), P_rec("P_rec", w[0], w[1], w[2], 2, nd);
        KV_double_6d rhou_rec("rhou_rec", w[0], w[1], w[2], 2, nd, nd), u_rec("u_rec", w[0], w[1], w[2], 2, nd, nd),
                fx_rec("fx_rec", w[0], w[1], w[2], 2, nd, nfx);
        std::unique_ptr<IFaceReconstruction> face;
        HostMinmodReconstruction* host = nullptr;
        CountingReconstruction* counting = nullptr;
        if (variant == 0) {
            face = factory_face_reconstruction("Minmod");  // the library's own: one C-ABI call
        } else if (variant == 1) {
            auto c = std::make_unique<CountingReconstruction>(factory_face_reconstruction("Minmod"));
            counting = c.get();
            face = std::move(c);
        } else {
            auto h = std::make_unique<HostMinmodReconstruction>();
            host = h.get();
            face = std::move(h);
        }
        MUSCLHancockHydroReconstruction<UniformGravity> rec(std::move(face), std::make_unique<HancockExtrapolationReconstruction<UniformGravity>>(eos),
                                                            eos, P_rec, u_rec);
        rec.execute(grid.range.with_ghosts(1), grid, g, dt, rho, u, P, fx, rho_rec, rhou_rec, E_rec, fx_rec);
        int const expected_calls = 1 + nd + 1 + nfx;
        if (counting) EXPECT(counting->calls == expected_calls, "counting strategy called %d times", counting->calls);
        if (host) EXPECT(host->calls == expected_calls, "host strategy called %d times", host->calls);
        EXPECT(std::memcmp(rho_rec.to_host().data(), orho.data(), fcells * 8) == 0, "variant %d: rho_rec differs from the oracle", variant);
        EXPECT(std::memcmp(rhou_rec.to_host().data(), orhou.data(), fcells * nd * 8) == 0, "variant %d: rhou_rec differs from the oracle", variant);
        EXPECT(std::memcmp(E_rec.to_host().data(), oE.data(), fcells * 8) == 0, "variant %d: E_rec differs from the oracle", variant);
        EXPECT(std::memcmp(fx_rec.to_host().data(), ofx.data(), fcells * nfx * 8) == 0, "variant %d: fx_rec differs from the oracle", variant);
        EXPECT(std::memcmp(P_rec.to_host().data(), oP.data(), fcells * 8) == 0, "variant %d: P_rec differs from the oracle", variant);
    }
    or_destroy(o);
}

int main()
{
    try {
        test_errors();
        test_conversions();
        test_operator_loop();
        test_internal_mass_loop();
        test_pressure_fix_loop();
        test_passive_scalar_loop(1);
        test_passive_scalar_loop(3);
        test_user_face_reconstruction();
    } catch (std::exception const& e) {
        std::printf("FAIL: exception %s\n", e.what());
        return 2;
    }
    std::printf("%d passed, %d failed\n", g_pass, g_fail);
    return g_fail == 0 ? 0 : 1;
}
