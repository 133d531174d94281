// hclpp/hydro.hpp -- C++ host side of the MI355X hydro step.
//
// Header-only adapter that re-exposes the reference operator interfaces
// (IFaceReconstruction, IExtrapolationReconstruction, IHydroReconstruction,
// IGodunovScheme, their factories, time_step, conv_*, integrate,
// minimum_internal_energy, DistributedBoundaryCondition) over the C-ABI of
// libhclpp_hip.so, with the reference's names, argument meaning and error
// behaviour: factories throw std::runtime_error on unknown names, PerfectGas
// throws std::domain_error on invalid parameters, Range::with_ghosts throws on
// NgEff > Ng, and any non-zero C-ABI status becomes std::runtime_error.
//
// Views: DeviceView below is the owning stand-in for Kokkos::View<..., LayoutLeft>
// (extent 0 fastest), with the reference's aliases KV_double_Nd and the read-only
// KV_cdouble_Nd (utils/kokkos_shortcut.hpp:22-48); a writable view converts to a
// read-only one, and slice(i) is Kokkos::subview(v, ALL, ..., ALL, i) of the slowest
// extent.  The free functions (conv_*) take any type with data() and extent_int(i).
// All views hold device memory.  Citations are to the reference tree (src/...).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "hclpp_hip.h"

namespace hclpp::hip {

inline void check(int rc)
{
    if (rc != HCLPP_OK) throw std::runtime_error(hclpp_last_error());
}

inline void check_hip(hipError_t e, char const* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------ views
// Owning LayoutLeft device array: extent(0) fastest.  Copies share ownership; T = const double
// gives the read-only view a writable one converts to.
template <int Rank, class T = double>
class DeviceView
{
    template <int, class>
    friend class DeviceView;
    std::shared_ptr<T> m_data;
    std::array<int, Rank> m_ext {};
    std::string m_label;

public:
    DeviceView() = default;

    template <class... I>
    explicit DeviceView(std::string label, I... ext) : m_ext {static_cast<int>(ext)...}, m_label(std::move(label))
    {
        static_assert(sizeof...(I) == Rank, "one extent per rank");
        static_assert(!std::is_const_v<T>, "allocate a writable view, then convert it");
        double* p = nullptr;
        size_t const n = size();
        if (n > 0) {
            check_hip(hipMalloc(&p, n * sizeof(double)), "hipMalloc");
            check_hip(hipMemset(p, 0, n * sizeof(double)), "hipMemset");
        }
        m_data = std::shared_ptr<double>(p, [](double* q) {
            if (q) (void)hipFree(q);
        });
    }

    // read-only view of a writable one (KV_double_Nd -> KV_cdouble_Nd)
    template <class U, class = std::enable_if_t<std::is_same_v<T, U const> && !std::is_same_v<T, U>>>
    DeviceView(DeviceView<Rank, U> const& o) : m_data(o.m_data), m_ext(o.m_ext), m_label(o.m_label)
    {
    }

    // Kokkos::subview(v, ALL, ..., ALL, i): index i of the slowest extent, sharing ownership
    DeviceView<Rank - 1, T> slice(int i) const
    {
        static_assert(Rank >= 2, "slice of a rank >= 2 view");
        if (i < 0 || i >= m_ext[Rank - 1]) throw std::out_of_range("slice index out of range in " + m_label);
        size_t const stride = m_ext[Rank - 1] > 0 ? size() / static_cast<size_t>(m_ext[Rank - 1]) : 0;
        DeviceView<Rank - 1, T> s;
        s.m_data = std::shared_ptr<T>(m_data, data() + stride * static_cast<size_t>(i));
        for (int r = 0; r < Rank - 1; ++r) s.m_ext[r] = m_ext[r];
        s.m_label = m_label + "[" + std::to_string(i) + "]";
        return s;
    }

    T* data() const { return m_data.get(); }
    int extent_int(int i) const { return i < Rank ? m_ext[i] : 1; }
    size_t extent(int i) const { return static_cast<size_t>(extent_int(i)); }
    size_t size() const
    {
        size_t n = 1;
        for (int e : m_ext) n *= static_cast<size_t>(e);
        return n;
    }
    std::string const& label() const { return m_label; }

    void copy_from_host(std::vector<double> const& h) const
    {
        static_assert(!std::is_const_v<T>, "read-only view");
        if (h.size() != size()) throw std::runtime_error("copy_from_host: size mismatch for " + m_label);
        check_hip(hipMemcpy(data(), h.data(), size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    }
    std::vector<double> to_host() const
    {
        std::vector<double> h(size());
        check_hip(hipMemcpy(h.data(), data(), size() * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
        return h;
    }
};

using KV_double_3d = DeviceView<3>;
using KV_double_4d = DeviceView<4>;
using KV_double_5d = DeviceView<5>;
using KV_double_6d = DeviceView<6>;
using KV_cdouble_3d = DeviceView<3, const double>;
using KV_cdouble_4d = DeviceView<4, const double>;
using KV_cdouble_5d = DeviceView<5, const double>;
using KV_cdouble_6d = DeviceView<6, const double>;

// ------------------------------------------------------------ thermodynamics/perfect_gas.hpp:17-90
class PerfectGas
{
    double m_gamma;
    double m_mmw;

public:
    PerfectGas(double gamma, double mmw) : m_gamma(gamma), m_mmw(mmw)
    {
        if (!std::isfinite(gamma) || gamma <= 1) throw std::domain_error("Invalid gamma");  // perfect_gas.cpp:15-17
        if (!std::isfinite(mmw) || mmw <= 0) throw std::domain_error("Invalid mmw");        // :19-21
    }
    double adiabatic_index() const noexcept { return m_gamma; }
    double mean_molecular_weight() const noexcept { return m_mmw; }
    static bool is_valid(double rho, double P) noexcept { return std::isfinite(rho) && rho > 0 && std::isfinite(P) && P > 0; }
};

// ------------------------------------------------------------ mesh/range.hpp:19-55, range.cpp:23-141
class Range
{
public:
    std::array<int, 3> Corner_min {};
    std::array<int, 3> Corner_max {};
    int Ng = 0;
    int NgEff = 0;
    int ndim = 1;
    std::array<int, 3> Nghost {};

    Range() = default;
    Range(std::array<int, 3> const& cmin, std::array<int, 3> const& cmax, int ng, int nd) : Corner_min(cmin), Corner_max(cmax), Ng(ng), ndim(nd)
    {
        for (int d = 0; d < 3; ++d) {
            if (cmin[d] < 0 || cmax[d] < 0) throw std::runtime_error("Cmin < 0 || Cmax < 0");
            if (cmin[d] > cmax[d]) throw std::runtime_error("Cmin > Cmax");
            Nghost[d] = d < nd ? ng : 0;
        }
    }
    Range with_ghosts(int ngeff) const
    {
        if (ngeff > Ng) throw std::runtime_error("NgEff > Ng");
        Range r(*this);
        r.NgEff = ngeff;
        return r;
    }
    Range no_ghosts() const { return with_ghosts(0); }
    Range all_ghosts() const { return with_ghosts(Ng); }
    // cell_range: range.cpp:120-135
    hclpp_range cells() const
    {
        hclpp_range r;
        for (int d = 0; d < 3; ++d) {
            int const e = d < ndim ? NgEff : 0;
            r.begin[d] = Nghost[d] - e;
            r.end[d] = Nghost[d] + Corner_max[d] - Corner_min[d] + e;
        }
        return r;
    }
};

enum class Geometry { Geom_cartesian, Geom_spherical };

// ------------------------------------------------------------ gravity/gravity.hpp:18-71
class UniformGravity
{
public:
    std::array<double, 3> g;
    explicit UniformGravity(std::array<double, 3> const& gv) : g(gv) {}
    double operator()(int, int, int, int dir) const noexcept { return g[dir]; }
};

inline UniformGravity make_uniform_gravity(double gx0, double gx1, double gx2) { return UniformGravity({gx0, gx1, gx2}); }

class PointMassGravity
{
public:
    double central_mass;
    explicit PointMassGravity(double m) : central_mass(m) {}
};

class Grid;
inline PointMassGravity make_point_mass_gravity(double central_mass, Grid const&) { return PointMassGravity(central_mass); }

// gravity/gravity.hpp:75-103, gravity.cpp:59-128: radial gravity of the central mass plus the
// shell-averaged density inside each cell centre.  The table is built on the device from `rho`
// (ghost-filled), on the first operator call that uses this object on a given context; as in
// main.cpp:309/:528 a new object is made from the current density after every step.
class InternalMassGravity
{
public:
    double central_mass;
    double const* rho;
    long long generation;
    InternalMassGravity(double m, double const* r) : central_mass(m), rho(r), generation(next_generation()) {}

private:
    static long long next_generation()
    {
        static long long g = 0;
        return ++g;
    }
};

template <class View>
InternalMassGravity make_internal_mass_gravity(double central_mass, Grid const&, View const& rho)
{
    return InternalMassGravity(central_mass, rho.data());
}

// ------------------------------------------------------------ mesh/grid.hpp:23-75 (one rank)
class Grid
{
public:
    int ndim;
    Geometry geom;
    int Ng;
    std::array<int, 3> Nghost {};
    std::array<int, 3> Nx_glob_ng {};
    std::array<int, 3> Nx_local_wg {};
    std::array<std::vector<double>, 3> x;  // node coordinates, Nx_local_wg + 1 each
    Range range;

    // Regular grid: mesh/grid_type.cpp:23-58
    Grid(int nd, std::array<int, 3> nx, int ng, Geometry g, std::array<double, 3> xmin, std::array<double, 3> xmax)
        : ndim(nd), geom(g), Ng(ng)
    {
        for (int d = 0; d < 3; ++d) {
            Nx_glob_ng[d] = d < nd ? nx[d] : 1;
            Nghost[d] = d < nd ? ng : 0;
            Nx_local_wg[d] = Nx_glob_ng[d] + 2 * Nghost[d];
            double const dx = (xmax[d] - xmin[d]) / Nx_glob_ng[d];
            x[d].resize(Nx_local_wg[d] + 1);
            for (int i = 0; i <= Nx_local_wg[d]; ++i) x[d][i] = xmin[d] + ((i - Nghost[d]) * dx);
        }
        range = Range({0, 0, 0}, Nx_glob_ng, ng, nd);
    }

    // One device context per (EoS, gravity, boundaries, nfx, scheme) seen on this grid.
    struct CtxKey {
        double gamma, mmw;
        int gravity;
        double g0, g1, g2, mass;
        std::array<int, 6> bc, order;
        int nfx, limiter, riemann;
        bool operator<(CtxKey const& o) const
        {
            return std::tie(gamma, mmw, gravity, g0, g1, g2, mass, bc, order, nfx, limiter, riemann)
                   < std::tie(o.gamma, o.mmw, o.gravity, o.g0, o.g1, o.g2, o.mass, o.bc, o.order, o.nfx, o.limiter, o.riemann);
        }
    };

    hclpp_ctx* context(PerfectGas const& eos, int gravity, std::array<double, 3> g, double mass, std::array<int, 6> bc,
                       std::array<int, 6> order, int nfx, int limiter = HCLPP_LIM_VANLEER, int riemann = HCLPP_RS_HLL) const
    {
        CtxKey const key {eos.adiabatic_index(), eos.mean_molecular_weight(), gravity, g[0], g[1], g[2], mass, bc, order, nfx, limiter, riemann};
        auto it = m_ctx->find(key);
        if (it != m_ctx->end()) return it->second.get();
        hclpp_config c {};
        c.ndim = ndim;
        c.geom = geom == Geometry::Geom_spherical ? HCLPP_GEOM_SPHERICAL : HCLPP_GEOM_CARTESIAN;
        c.limiter = limiter;
        c.riemann = riemann;
        c.gravity = gravity;
        c.nfx = nfx;
        c.ng = Ng;
        for (int d = 0; d < 3; ++d) {
            c.n[d] = Nx_glob_ng[d];
            c.g[d] = g[d];
            c.x[d] = x[d].data();
        }
        for (int f = 0; f < 6; ++f) {
            c.bc[f] = bc[f];
            c.bc_order[f] = order[f];
        }
        c.gamma = eos.adiabatic_index();
        c.mmw = eos.mean_molecular_weight();
        c.central_mass = mass;
        c.precision = HCLPP_FP64;
        hclpp_ctx* h = nullptr;
        check(hclpp_ctx_create(&c, &h));
        auto& slot = (*m_ctx)[key];
        slot = std::shared_ptr<hclpp_ctx>(h, [](hclpp_ctx* p) { hclpp_ctx_destroy(p); });
        return h;
    }

    // The device copy of this grid's tables for the operators that read only the grid and the
    // arrays they are given (face reconstruction, integrate, minimum_internal_energy, the ghost
    // fill): a context whose equation of state, gravity and scheme fields those operators never
    // read (they are set to valid neutral values, as hclpp_ctx_create checks them).
    hclpp_ctx* geometry_context(std::array<int, 6> bc = {0, 0, 0, 0, 0, 0}, std::array<int, 6> const* order = nullptr, int nfx = 0) const
    {
        return context(PerfectGas(kNeutralGamma, 1.0), HCLPP_GRAV_UNIFORM, {0.0, 0.0, 0.0}, 0.0, bc, order ? *order : default_order(), nfx);
    }

    std::array<int, 6> default_order() const
    {
        std::array<int, 6> o {0, 0, 0, 0, 0, 0};
        for (int f = 0; f < 2 * ndim; ++f) o[f] = f;
        return o;
    }

    // generation of the internal-mass gravity table last built on a context
    long long& gravity_generation(hclpp_ctx* c) const { return (*m_gen)[c]; }

private:
    static constexpr double kNeutralGamma = 1.4;
    std::shared_ptr<std::map<CtxKey, std::shared_ptr<hclpp_ctx>>> m_ctx = std::make_shared<std::map<CtxKey, std::shared_ptr<hclpp_ctx>>>();
    std::shared_ptr<std::map<hclpp_ctx*, long long>> m_gen = std::make_shared<std::map<hclpp_ctx*, long long>>();
};

namespace detail {
inline std::tuple<int, std::array<double, 3>, double> gravity_params(UniformGravity const& g) { return {HCLPP_GRAV_UNIFORM, g.g, 0.0}; }
inline std::tuple<int, std::array<double, 3>, double> gravity_params(PointMassGravity const& g)
{
    return {HCLPP_GRAV_POINT_MASS, {0.0, 0.0, 0.0}, g.central_mass};
}
inline std::tuple<int, std::array<double, 3>, double> gravity_params(InternalMassGravity const& g)
{
    return {HCLPP_GRAV_INTERNAL_MASS, {0.0, 0.0, 0.0}, g.central_mass};
}
template <class Gravity>
hclpp_ctx* ctx_for(Grid const& grid, PerfectGas const& eos, Gravity const& gravity, int nfx)
{
    auto const [kind, g, m] = gravity_params(gravity);
    hclpp_ctx* c = grid.context(eos, kind, g, m, {0, 0, 0, 0, 0, 0}, grid.default_order(), nfx);
    if constexpr (std::is_same_v<Gravity, InternalMassGravity>) {
        long long& gen = grid.gravity_generation(c);
        if (gen != gravity.generation) {
            check(hclpp_internal_mass_gravity(c, gravity.rho));
            gen = gravity.generation;
        }
    }
    return c;
}
inline hclpp_ctx* ctx_for(Grid const& grid, PerfectGas const& eos, int nfx = 0)
{
    return ctx_for(grid, eos, UniformGravity({0.0, 0.0, 0.0}), nfx);
}
template <class V>
int nfx_of(V const& fx, int rank_index)
{
    return fx.data() ? fx.extent_int(rank_index) : 0;
}
}  // namespace detail

// ------------------------------------------------------------ hydro/array_conversion.hpp:32-172
template <class V3, class V4>
void conv_prim_to_cons(Range const& range, PerfectGas const& eos, Grid const& grid, V3 const& rho, V4 const& u, V3 const& P, V4 const& rhou,
                       V3 const& E)
{
    hclpp_range const r = range.cells();
    check(hclpp_conv_prim_to_cons(detail::ctx_for(grid, eos), &r, rho.data(), u.data(), P.data(), rhou.data(), E.data()));
}

template <class V3, class V4>
void conv_cons_to_prim(Range const& range, PerfectGas const& eos, Grid const& grid, V3 const& rho, V4 const& rhou, V3 const& E, V4 const& u,
                       V3 const& P)
{
    hclpp_range const r = range.cells();
    check(hclpp_conv_cons_to_prim(detail::ctx_for(grid, eos), &r, rho.data(), rhou.data(), E.data(), u.data(), P.data()));
}

// ------------------------------------------------------------ hydro/face_reconstruction.hpp:21-43
class IFaceReconstruction
{
public:
    virtual ~IFaceReconstruction() noexcept = default;
    //! @param[in] range output iteration range
    //! @param[in] grid provides grid information
    //! @param[in] var cell values
    //! @param[out] var_rec reconstructed values at interfaces
    virtual void execute(Range const& range, Grid const& grid, KV_cdouble_3d const& var, KV_double_5d const& var_rec) const = 0;
};

// hydro/limited_linear_reconstruction.hpp:25-66; the slope limiter is one of the HCLPP_LIM_* kinds
class LimitedLinearReconstruction : public IFaceReconstruction
{
    int m_lim;

public:
    explicit LimitedLinearReconstruction(int slope_limiter) : m_lim(slope_limiter) {}
    void execute(Range const& range, Grid const& grid, KV_cdouble_3d const& var, KV_double_5d const& var_rec) const final
    {
        hclpp_range const r = range.cells();
        check(hclpp_face_reconstruction(grid.geometry_context(), m_lim, &r, var.data(), var_rec.data()));
    }
    int slope_limiter() const noexcept { return m_lim; }
};

// hydro/face_reconstruction.cpp:31-50
inline std::unique_ptr<IFaceReconstruction> factory_face_reconstruction(std::string const& slope)
{
    int const lim = hclpp_limiter_from_name(slope.c_str());
    if (lim < 0) throw std::runtime_error("Unknown face reconstruction algorithm: " + slope + ".");
    return std::make_unique<LimitedLinearReconstruction>(lim);
}

// ------------------------------------------------------------ hydro/extrapolation_reconstruction.hpp:27-54
template <class Gravity>
class IExtrapolationReconstruction
{
public:
    virtual ~IExtrapolationReconstruction() noexcept = default;
    virtual void execute(Range const& range, Grid const& grid, Gravity const& gravity, double dt_reconstruction, KV_cdouble_6d const& u_rec,
                         KV_cdouble_5d const& P_rec, KV_double_5d const& rho_rec, KV_double_6d const& rhou_rec, KV_double_5d const& E_rec,
                         KV_double_6d const& fx_rec) const = 0;
};

// :56-339
template <class Gravity>
class HancockExtrapolationReconstruction : public IExtrapolationReconstruction<Gravity>
{
    PerfectGas m_eos;

public:
    explicit HancockExtrapolationReconstruction(PerfectGas const& eos) : m_eos(eos) {}
    void execute(Range const& range, Grid const& grid, Gravity const& gravity, double dt, KV_cdouble_6d const& u_rec, KV_cdouble_5d const& P_rec,
                 KV_double_5d const& rho_rec, KV_double_6d const& rhou_rec, KV_double_5d const& E_rec, KV_double_6d const& fx_rec) const final
    {
        hclpp_range const r = range.cells();
        check(hclpp_hancock_extrapolation(detail::ctx_for(grid, m_eos, gravity, detail::nfx_of(fx_rec, 5)), &r, dt, u_rec.data(), P_rec.data(),
                                          rho_rec.data(), rhou_rec.data(), E_rec.data(), fx_rec.data()));
    }
    PerfectGas const& eos() const noexcept { return m_eos; }
};

// ------------------------------------------------------------ hydro/hydro_reconstruction.hpp:30-59
template <class Gravity>
class IHydroReconstruction
{
public:
    virtual ~IHydroReconstruction() noexcept = default;
    virtual void execute(Range const& range, Grid const& grid, Gravity const& gravity, double dt, KV_cdouble_3d const& rho, KV_cdouble_4d const& u,
                         KV_cdouble_3d const& P, KV_cdouble_4d const& fx, KV_double_5d const& rho_rec, KV_double_6d const& rhou_rec,
                         KV_double_5d const& E_rec, KV_double_6d const& fx_rec) const = 0;
};

// :61-130 -- P_rec and u_rec are the operator's own scratch views.  execute composes the injected
// strategies exactly as the reference does: the face reconstruction of rho, each velocity
// component, P and each passive scalar, the prim -> cons conversion of every face state, then the
// extrapolation step.  When both strategies are the library's own (LimitedLinearReconstruction and
// HancockExtrapolationReconstruction with this operator's EoS), the same sequence runs as one
// C-ABI call (hclpp_hydro_reconstruction), which launches the same kernels.
template <class Gravity>
class MUSCLHancockHydroReconstruction : public IHydroReconstruction<Gravity>
{
    std::unique_ptr<IFaceReconstruction> m_face_reconstruction;
    std::unique_ptr<IExtrapolationReconstruction<Gravity>> m_hancock_reconstruction;
    PerfectGas m_eos;
    KV_double_5d m_p_rec;
    KV_double_6d m_u_rec;

public:
    MUSCLHancockHydroReconstruction(std::unique_ptr<IFaceReconstruction> face_reconstruction,
                                    std::unique_ptr<IExtrapolationReconstruction<Gravity>> hancock_reconstruction, PerfectGas const& eos,
                                    KV_double_5d P_rec, KV_double_6d u_rec)
        : m_face_reconstruction(std::move(face_reconstruction))
        , m_hancock_reconstruction(std::move(hancock_reconstruction))
        , m_eos(eos)
        , m_p_rec(std::move(P_rec))
        , m_u_rec(std::move(u_rec))
    {
    }
    void execute(Range const& range, Grid const& grid, Gravity const& gravity, double dt, KV_cdouble_3d const& rho, KV_cdouble_4d const& u,
                 KV_cdouble_3d const& P, KV_cdouble_4d const& fx, KV_double_5d const& rho_rec, KV_double_6d const& rhou_rec, KV_double_5d const& E_rec,
                 KV_double_6d const& fx_rec) const final
    {
        hclpp_range const r = range.cells();
        int const nfx = detail::nfx_of(fx, 3);
        auto const* face = dynamic_cast<LimitedLinearReconstruction const*>(m_face_reconstruction.get());
        auto const* hancock = dynamic_cast<HancockExtrapolationReconstruction<Gravity> const*>(m_hancock_reconstruction.get());
        if (face && hancock && hancock->eos().adiabatic_index() == m_eos.adiabatic_index()
            && hancock->eos().mean_molecular_weight() == m_eos.mean_molecular_weight()) {
            check(hclpp_hydro_reconstruction(detail::ctx_for(grid, m_eos, gravity, nfx), face->slope_limiter(), &r, dt, rho.data(), u.data(),
                                             P.data(), fx.data(), rho_rec.data(), rhou_rec.data(), E_rec.data(), fx_rec.data(), m_p_rec.data(),
                                             m_u_rec.data()));
            return;
        }
        // hydro_reconstruction.hpp:99-109
        m_face_reconstruction->execute(range, grid, rho, rho_rec);
        for (int idim = 0; idim < grid.ndim; ++idim) m_face_reconstruction->execute(range, grid, u.slice(idim), m_u_rec.slice(idim));
        m_face_reconstruction->execute(range, grid, P, m_p_rec);
        for (int ifx = 0; ifx < nfx; ++ifx) m_face_reconstruction->execute(range, grid, fx.slice(ifx), fx_rec.slice(ifx));
        // :111-126 (conv_prim_to_cons of every face state)
        check(hclpp_face_prim_to_cons(detail::ctx_for(grid, m_eos), &r, rho_rec.data(), m_u_rec.data(), m_p_rec.data(), rhou_rec.data(),
                                      E_rec.data()));
        // :128
        m_hancock_reconstruction->execute(range, grid, gravity, dt, m_u_rec, m_p_rec, rho_rec, rhou_rec, E_rec, fx_rec);
    }
};

// ------------------------------------------------------------ hydro/godunov_scheme.hpp:31-64
template <class Gravity>
class IGodunovScheme
{
public:
    virtual ~IGodunovScheme() noexcept = default;
    virtual void execute(Range const& range, Grid const& grid, Gravity const& gravity, double dt, KV_cdouble_3d const& rho,
                         KV_cdouble_4d const& rhou, KV_cdouble_3d const& E, KV_cdouble_4d const& fx, KV_cdouble_5d const& rho_rec,
                         KV_cdouble_6d const& rhou_rec, KV_cdouble_5d const& E_rec, KV_cdouble_6d const& fx_rec, KV_double_3d const& rho_new,
                         KV_double_4d const& rhou_new, KV_double_3d const& E_new, KV_double_4d const& fx_new) const = 0;
};

// :66-303
template <class Gravity>
class RiemannBasedGodunovScheme : public IGodunovScheme<Gravity>
{
    int m_solver;
    PerfectGas m_eos;

public:
    RiemannBasedGodunovScheme(int solver, PerfectGas const& eos) : m_solver(solver), m_eos(eos) {}
    void execute(Range const& range, Grid const& grid, Gravity const& gravity, double dt, KV_cdouble_3d const& rho, KV_cdouble_4d const& rhou,
                 KV_cdouble_3d const& E, KV_cdouble_4d const& fx, KV_cdouble_5d const& rho_rec, KV_cdouble_6d const& rhou_rec,
                 KV_cdouble_5d const& E_rec, KV_cdouble_6d const& fx_rec, KV_double_3d const& rho_new, KV_double_4d const& rhou_new,
                 KV_double_3d const& E_new, KV_double_4d const& fx_new) const final
    {
        hclpp_range const r = range.cells();
        check(hclpp_godunov(detail::ctx_for(grid, m_eos, gravity, detail::nfx_of(fx, 3)), m_solver, &r, dt, rho.data(), rhou.data(), E.data(),
                            fx.data(), rho_rec.data(), rhou_rec.data(), E_rec.data(), fx_rec.data(), rho_new.data(), rhou_new.data(), E_new.data(),
                            fx_new.data()));
    }
};

// :305-321
template <class Gravity>
std::unique_ptr<IGodunovScheme<Gravity>> factory_godunov_scheme(std::string const& riemann_solver, PerfectGas const& eos)
{
    int const rs = hclpp_riemann_from_name(riemann_solver.c_str());
    if (rs < 0) throw std::runtime_error("Invalid riemann solver: " + riemann_solver + ".");
    return std::make_unique<RiemannBasedGodunovScheme<Gravity>>(rs, eos);
}

// ------------------------------------------------------------ hydro/time_step.hpp:35-82
[[nodiscard]] inline double time_step(Range const& range, PerfectGas const& eos, Grid const& grid, KV_cdouble_3d const& rho,
                                      KV_cdouble_4d const& u, KV_cdouble_3d const& P)
{
    hclpp_range const r = range.cells();
    double dt = 0;
    check(hclpp_time_step(detail::ctx_for(grid, eos), &r, rho.data(), u.data(), P.data(), &dt));
    return dt;
}

// hydro/pressure_fix.hpp:19-172 (E_new updated in place)
inline void pressure_fix(Range const& range, PerfectGas const& eos, Grid const& grid, double dt, double eps, KV_cdouble_3d const& rho,
                         KV_cdouble_4d const& rhou, KV_cdouble_3d const& E, KV_cdouble_5d const& rho_rec, KV_cdouble_6d const& rhou_rec,
                         KV_cdouble_5d const& E_rec, KV_cdouble_3d const& rho_new, KV_cdouble_4d const& rhou_new, KV_double_3d const& E_new)
{
    hclpp_range const r = range.cells();
    check(hclpp_pressure_fix(detail::ctx_for(grid, eos), &r, dt, eps, rho.data(), rhou.data(), E.data(), rho_rec.data(), rhou_rec.data(),
                             E_rec.data(), rho_new.data(), rhou_new.data(), E_new.data()));
}

// analysis/integration.cpp:22-37
[[nodiscard]] inline double integrate(Range const& range, Grid const& grid, KV_cdouble_3d const& var)
{
    hclpp_range const r = range.cells();
    double s = 0;
    check(hclpp_integrate(grid.geometry_context(), &r, var.data(), &s));
    return s;
}

// analysis/internal_energy.cpp:20-45
[[nodiscard]] inline double minimum_internal_energy(Range const& range, Grid const& grid, KV_cdouble_3d const& rho, KV_cdouble_4d const& rhou,
                                                    KV_cdouble_3d const& E)
{
    hclpp_range const r = range.cells();
    double m = 0;
    check(hclpp_min_internal_energy(grid.geometry_context(), &r, rho.data(), rhou.data(), E.data(), &m));
    return m;
}

// ------------------------------------------------------------ boundary_distribute.hpp:28-77 (one rank)
class DistributedBoundaryCondition
{
    std::array<int, 6> m_bc {};
    std::array<int, 6> m_order {};

public:
    // boundary kinds by face (2*dim + side) as the reference factory names them; order as m_bc_order
    DistributedBoundaryCondition(Grid const& grid, std::vector<std::string> const& bc_names, std::vector<int> order = {})
    {
        if (static_cast<int>(bc_names.size()) != 2 * grid.ndim) throw std::runtime_error("boundary condition not fully defined");
        for (int f = 0; f < 2 * grid.ndim; ++f) {
            int const k = hclpp_bc_from_name(bc_names[f].c_str());
            if (k < 0) throw std::runtime_error("Unknown boundary condition : " + bc_names[f] + ".");
            m_bc[f] = k;
            m_order[f] = order.empty() ? f : order[f];
        }
    }
    void operator()(Grid const& grid, KV_double_3d const& rho, KV_double_4d const& rhou, KV_double_3d const& E, KV_double_4d const& fx) const
    {
        hclpp_ctx* c = grid.geometry_context(m_bc, &m_order, detail::nfx_of(fx, 3));
        check(hclpp_fill_ghosts(c, rho.data(), rhou.data(), E.data(), fx.data()));
    }
    std::array<int, 6> const& kinds() const { return m_bc; }
    std::array<int, 6> const& order() const { return m_order; }
};

// ------------------------------------------------------------ main.cpp:305-311, 354-531 (fused path)
// Owns a double-buffered conservative state (and passive scalars) and
// runs fused steps with the device clock: no host synchronisation inside the loop.
template <class Gravity>
class HydroDriver
{
    Grid const& m_grid;
    hclpp_ctx* m_ctx;
    double m_cfl;
    std::array<KV_double_3d, 2> m_rho;
    std::array<KV_double_4d, 2> m_rhou;
    std::array<KV_double_3d, 2> m_E;
    std::array<KV_double_4d, 2> m_fx;  // mass fractions (nfx components), allocated when nfx > 0
    int m_nfx;
    int m_cur = 0;
    double m_t = 0;
    int m_iter = 0;

public:
    HydroDriver(Grid const& grid, PerfectGas const& eos, Gravity const& gravity, std::string const& reconstruction, std::string const& riemann,
                DistributedBoundaryCondition const& bcs, double cfl, int nfx = 0)
        : m_grid(grid), m_cfl(cfl), m_nfx(nfx)
    {
        if (nfx < 0) throw std::runtime_error("nfx must be >= 0");
        int const lim = hclpp_limiter_from_name(reconstruction.c_str());
        if (lim < 0) throw std::runtime_error("Unknown face reconstruction algorithm: " + reconstruction + ".");
        int const rs = hclpp_riemann_from_name(riemann.c_str());
        if (rs < 0) throw std::runtime_error("Invalid riemann solver: " + riemann + ".");
        auto const [kind, g, m] = detail::gravity_params(gravity);
        m_ctx = grid.context(eos, kind, g, m, bcs.kinds(), bcs.order(), nfx, lim, rs);
        auto const& n = grid.Nx_local_wg;
        for (int b = 0; b < 2; ++b) {
            m_rho[b] = KV_double_3d("rho", n[0], n[1], n[2]);
            m_rhou[b] = KV_double_4d("rhou", n[0], n[1], n[2], grid.ndim);
            m_E[b] = KV_double_3d("E", n[0], n[1], n[2]);
            if (nfx > 0) m_fx[b] = KV_double_4d("fx", n[0], n[1], n[2], nfx);
        }
    }

    // main.cpp:305-311: prim -> cons on the interior, boundaries on the conservative state
    // fx: the passive scalars (mass fractions) when the driver has nfx > 0
    void init_from_prim(KV_double_3d const& rho, KV_double_4d const& u, KV_double_3d const& P, KV_double_4d const& fx = {})
    {
        if (m_nfx > 0 && fx.size() != m_fx[0].size()) throw std::runtime_error("fx: expected nfx mass fractions per cell");
        m_cur = 0;
        check_hip(hipMemcpy(m_rho[0].data(), rho.data(), rho.size() * sizeof(double), hipMemcpyDeviceToDevice), "hipMemcpy");
        if (m_nfx > 0) check_hip(hipMemcpy(m_fx[0].data(), fx.data(), fx.size() * sizeof(double), hipMemcpyDeviceToDevice), "hipMemcpy");
        hclpp_range const r = m_grid.range.no_ghosts().cells();
        check(hclpp_conv_prim_to_cons(m_ctx, &r, m_rho[0].data(), u.data(), P.data(), m_rhou[0].data(), m_E[0].data()));
        check(hclpp_fill_ghosts(m_ctx, m_rho[0].data(), m_rhou[0].data(), m_E[0].data(), m_nfx > 0 ? m_fx[0].data() : nullptr));
        m_t = 0;
        m_iter = 0;
    }

    // Advance until t_end or max_iter; returns the number of steps.
    int run(double t_end, int max_iter, int check_every = 64)
    {
        int const it0 = m_iter, cur0 = m_cur;
        check(hclpp_clock_start(m_ctx, m_t, m_iter, t_end, m_cfl, max_iter, m_rho[m_cur].data(), m_rhou[m_cur].data(), m_E[m_cur].data()));
        hclpp_clock clk {};
        int launched = 0;
        for (;;) {
            for (int k = 0; k < check_every; ++k, ++launched) {
                int const a = (cur0 + launched) % 2, b = 1 - a;
                if (m_nfx > 0)
                    check(hclpp_step_fx(m_ctx, m_rho[a].data(), m_rhou[a].data(), m_E[a].data(), m_fx[a].data(), m_rho[b].data(),
                                        m_rhou[b].data(), m_E[b].data(), m_fx[b].data()));
                else
                    check(hclpp_step(m_ctx, m_rho[a].data(), m_rhou[a].data(), m_E[a].data(), m_rho[b].data(), m_rhou[b].data(),
                                     m_E[b].data()));
            }
            check(hclpp_clock_read(m_ctx, &clk));
            if (clk.done || clk.status) break;
        }
        int const steps = clk.iter - it0;
        m_cur = (cur0 + steps) % 2;
        m_t = clk.t;
        m_iter = clk.iter;
        if (clk.status) {
            throw std::runtime_error("Time = " + std::to_string(m_t) + ", iteration = " + std::to_string(m_iter)
                                     + ": detected invalid volumic internal energy");
        }
        return steps;
    }

    KV_double_3d const& rho() const { return m_rho[m_cur]; }
    KV_double_4d const& rhou() const { return m_rhou[m_cur]; }
    KV_double_3d const& E() const { return m_E[m_cur]; }
    KV_double_4d const& fx() const { return m_fx[m_cur]; }
    double time() const { return m_t; }
    int iter() const { return m_iter; }
};

}  // namespace hclpp::hip
