// abi.cpp -- C-ABI of libhclpp_hip.so (declared in include/hclpp_hip.h).
//
// Owns the context: grid tables in device memory (node coordinates, cell
// sizes and centres, trigonometric tables computed once on the host), the
// equation of state, the stream, reduction scratch and the device clock of
// the fused step.  No exception crosses the C boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <string>
#include <vector>

#include "adjoint_step.hpp"
#include "fused.hpp"
#include "hclpp_hip.h"
#include "launch.hpp"

namespace hclpp {
// reference-shaped kernels reading conservative arrays (ref_kernels.hip)
void launch_time_step_cons(int nd, int geom, const GridDev& G, const Range3& r, const PerfectGas& eos, const double* rho, const double* rhou,
                           const double* E, double* partials, double* out, hipStream_t s);
int64_t reduce_blocks(const Range3& r);
}  // namespace hclpp

using namespace hclpp;

struct hclpp_ctx {
    hclpp_config cfg;
    GridDev G;
    PerfectGas eos;
    PerfectGasMixed eos32;
    PerfectGasFast eos_fast;
    int fast_on = 0;  // hclpp_set_fast_arithmetic: the FP64 fused step uses eos_fast
    BcSpec bc;
    hipStream_t stream = nullptr;
    int device = 0;
    double* tables = nullptr;       // all grid tables, one allocation
    double* partials = nullptr;     // reduction partials
    int64_t partials_cap = 0;       // doubles
    double* ghost_partials = nullptr;
    int64_t ghost_cap = 0;
    double* scalars = nullptr;      // device scalars (reduction outputs)
    double* host_scalars = nullptr; // pinned
    DevClock* clock = nullptr;
    DevClock* host_clock = nullptr; // pinned
    double* fx_old = nullptr;       // Hancock scratch for passive scalars
    int64_t fx_old_cap = 0;
    GhostShell shell{};
    MarchPlan march {};  // launch plan of the fused kernel
    int march_occ = 0;    // its resident workgroups per CU
    int64_t fused_nb = 0;
    double* gtab = nullptr;         // radial gravity table (inside tables; rebuilt for internal-mass gravity)
    double* shells = nullptr;       // internal-mass gravity: per-layer sums of rho dV and dV
    double* adj_prim = nullptr;     // reverse step: primitives (NV planes)
    double* adj_face = nullptr;     // reverse step: face slots and slope-adjoint shares
    double* adj_partials = nullptr; // reverse step: dgamma partials
    // pressure fix of the fused step (hclpp_set_pressure_fix)
    int pf_on = 0;
    double pf_eps = 1e-6;
    double* fix_partials = nullptr;  // two doubles per workgroup of the fused (march) kernel
    int64_t fix_nb = 0;
    // output schedule applied by the next clock start (hclpp_clock_set_outputs)
    int out_iter_freq = 0;
    double out_time_first = 0, out_time_freq = 0;
    int out_iter_id = -1, out_time_id = -1;
    std::vector<hipEvent_t> prof_events;  // pairs (start, stop) per timed fused launch
    int prof_max = 0;
    int prof_count = 0;
};

namespace {

thread_local std::string g_error;

// ---- scratch of hclpp_time_step_noalloc: the workgroup partials of its reduction live in this
// static device array, so that entry point calls no allocating routine.  Every GPU the library
// runs on holds its own copy, at its own address, until the process ends.  In host code the
// array's name is not a device address: kernels and copies get the address hipGetSymbolAddress
// returns for the current device, looked up once per device when a context is created there
// (noalloc_scratch_setup) and kept in g_scratch.
constexpr int64_t kNoallocScratchDoubles = int64_t(1) << 20;  // 8 MiB: one partial per 64 x 4 workgroup
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoallocScratchDoubles];
std::mutex g_scratch_mutex;
double* g_scratch[kMaxDevices] = {};

int fail(int code, const std::string& msg)
{
    g_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what)
{
    return fail(HCLPP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HCLPP_HIP(call)                                  \
    do {                                                 \
        hipError_t const e_ = (call);                    \
        if (e_ != hipSuccess) return hip_fail(e_, #call); \
    } while (0)

#define HCLPP_LAUNCHED(what)                                \
    do {                                                    \
        hipError_t const e_ = hipGetLastError();            \
        if (e_ != hipSuccess) return hip_fail(e_, what);    \
    } while (0)

// Look up this device's copy of `scratch` once (the device must be current).
int noalloc_scratch_setup(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(HCLPP_ERR_UNSUPPORTED, "device index beyond the no-allocation scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[dev]) return HCLPP_OK;
    void* ptr = nullptr;
    hipError_t const e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch[dev] = static_cast<double*>(ptr);
    return HCLPP_OK;
}

double* noalloc_scratch(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch[dev];
}

// the interior cells of the grid (Range::no_ghosts)
Range3 interior_range(const GridDev& G)
{
    Range3 r;
    for (int d = 0; d < 3; ++d) {
        r.b[d] = G.ngh[d];
        r.e[d] = G.ngh[d] + G.n[d];
    }
    return r;
}

Range3 to_range(const hclpp_range* r)
{
    Range3 o;
    for (int d = 0; d < 3; ++d) {
        o.b[d] = r->begin[d];
        o.e[d] = r->end[d];
    }
    return o;
}

int check_range(const hclpp_ctx* c, const hclpp_range* r, int halo)
{
    if (!r) return fail(HCLPP_ERR_ARG, "null range");
    for (int d = 0; d < 3; ++d) {
        int const h = d < c->cfg.ndim ? halo : 0;
        if (r->begin[d] - h < 0 || r->end[d] + h > c->G.nwg[d] || r->begin[d] > r->end[d])
            return fail(HCLPP_ERR_ARG, "range outside the grid (with the operator's stencil)");
    }
    return HCLPP_OK;
}

int ensure_partials(hclpp_ctx* c, int64_t n)
{
    if (n <= c->partials_cap) return HCLPP_OK;
    if (c->partials) HCLPP_HIP(hipFree(c->partials));
    c->partials = nullptr;
    HCLPP_HIP(hipMalloc(&c->partials, sizeof(double) * n));
    c->partials_cap = n;
    return HCLPP_OK;
}

// The clock's done/status pair as a kernel gate (the two ints are adjacent in hclpp_clock).
const DevFlag* clock_gate(const hclpp_ctx* c)
{
    static_assert(offsetof(hclpp_clock, status) == offsetof(hclpp_clock, done) + sizeof(int), "done, status adjacent");
    return reinterpret_cast<const DevFlag*>(&c->clock->pub.done);
}

int read_scalar(hclpp_ctx* c, double* out)
{
    HCLPP_HIP(hipMemcpyAsync(c->host_scalars, c->scalars, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HCLPP_HIP(hipStreamSynchronize(c->stream));
    *out = c->host_scalars[0];
    return HCLPP_OK;
}

}  // namespace

extern "C" {

const char* hclpp_last_error(void) { return g_error.c_str(); }

int hclpp_limiter_from_name(const char* name)
{
    if (!name) return -1;
    if (!std::strcmp(name, "Constant")) return HCLPP_LIM_CONSTANT;
    if (!std::strcmp(name, "VanLeer")) return HCLPP_LIM_VANLEER;
    if (!std::strcmp(name, "Minmod")) return HCLPP_LIM_MINMOD;
    if (!std::strcmp(name, "VanAlbada")) return HCLPP_LIM_VANALBADA;
    return -1;
}

int hclpp_riemann_from_name(const char* name)
{
    if (!name) return -1;
    if (!std::strcmp(name, "HLL")) return HCLPP_RS_HLL;
    if (!std::strcmp(name, "HLLC")) return HCLPP_RS_HLLC;
    if (!std::strcmp(name, "Low Mach")) return HCLPP_RS_LOWMACH;
    return -1;
}

int hclpp_bc_from_name(const char* name)
{
    if (!name) return -1;
    if (!std::strcmp(name, "Periodic")) return HCLPP_BC_PERIODIC;
    if (!std::strcmp(name, "NullGradient")) return HCLPP_BC_NULLGRADIENT;
    if (!std::strcmp(name, "Reflexive")) return HCLPP_BC_REFLEXIVE;
    return -1;
}

int hclpp_ctx_create(const hclpp_config* cfg, hclpp_ctx** out)
{
    if (!cfg || !out) return fail(HCLPP_ERR_ARG, "null argument");
    *out = nullptr;
    int const nd = cfg->ndim;
    if (nd < 1 || nd > 3) return fail(HCLPP_ERR_ARG, "ndim must be 1, 2 or 3");
    // thermodynamics/perfect_gas.cpp:15-21
    if (!std::isfinite(cfg->gamma) || cfg->gamma <= 1) return fail(HCLPP_ERR_ARG, "Invalid gamma");
    if (!std::isfinite(cfg->mmw) || cfg->mmw <= 0) return fail(HCLPP_ERR_ARG, "Invalid mmw");
    if (cfg->geom != HCLPP_GEOM_CARTESIAN && cfg->geom != HCLPP_GEOM_SPHERICAL) return fail(HCLPP_ERR_ARG, "unknown geometry");
    if (cfg->limiter < 0 || cfg->limiter > 3) return fail(HCLPP_ERR_ARG, "unknown limiter");
    if (cfg->riemann < 0 || cfg->riemann > 2) return fail(HCLPP_ERR_ARG, "unknown Riemann solver");
    if (cfg->gravity != HCLPP_GRAV_UNIFORM && cfg->gravity != HCLPP_GRAV_POINT_MASS && cfg->gravity != HCLPP_GRAV_INTERNAL_MASS)
        return fail(HCLPP_ERR_ARG, "unknown gravity");
    if (cfg->nfx < 0) return fail(HCLPP_ERR_ARG, "nfx < 0");
    if (cfg->ng < 0) return fail(HCLPP_ERR_ARG, "ng < 0");
    if (cfg->precision != HCLPP_FP64 && cfg->precision != HCLPP_FP32_STORAGE) return fail(HCLPP_ERR_ARG, "unknown precision");
    int seen[6] = {0, 0, 0, 0, 0, 0};
    for (int f = 0; f < 2 * nd; ++f) {
        if (cfg->bc[f] < 0 || cfg->bc[f] > 2) return fail(HCLPP_ERR_ARG, "unknown boundary condition");
        int const o = cfg->bc_order[f];
        if (o < 0 || o >= 2 * nd || seen[o]) return fail(HCLPP_ERR_ARG, "boundary priority not fully defined");
        seen[o] = 1;
    }
    for (int d = 0; d < 3; ++d) {
        if (!cfg->x[d]) return fail(HCLPP_ERR_ARG, "null node coordinates");
        if (d < nd && (cfg->n[d] < cfg->ng || cfg->n[d] < 1)) return fail(HCLPP_ERR_ARG, "interior extent smaller than the ghost width");
        if (d >= nd && cfg->n[d] != 1) return fail(HCLPP_ERR_ARG, "inactive dims must have extent 1");
    }

    auto* c = new hclpp_ctx();
    c->cfg = *cfg;
    for (int d = 0; d < 3; ++d) c->cfg.x[d] = nullptr;
    GridDev& G = c->G;
    for (int d = 0; d < 3; ++d) {
        G.n[d] = cfg->n[d];
        G.ngh[d] = d < nd ? cfg->ng : 0;  // mesh/grid.cpp:56-58
        G.nwg[d] = G.n[d] + 2 * G.ngh[d];
    }
    G.plane = (int64_t)G.nwg[0] * G.nwg[1] * G.nwg[2];
    for (int d = 0; d < 3; ++d) G.g[d] = cfg->g[d];
    // internal-mass gravity is a radial table like the point-mass one, rebuilt from rho every step
    G.gravity = cfg->gravity == HCLPP_GRAV_UNIFORM ? kUniformGravity : kPointMassGravity;
    c->eos.gamma = cfg->gamma;
    c->eos.gamma_m1 = cfg->gamma - 1;  // perfect_gas.cpp:13
    c->eos.mmw = cfg->mmw;
    c->eos32.gamma = (float)cfg->gamma;
    c->eos32.inv_gamma_m1 = (float)(1.0 / (cfg->gamma - 1));
    c->eos32.gamma_m1 = cfg->gamma - 1;
    c->eos32.gamma_d = cfg->gamma;
    c->eos_fast.gamma = cfg->gamma;
    c->eos_fast.gamma_m1 = cfg->gamma - 1;
    c->eos_fast.inv_gamma_m1 = 1.0 / (cfg->gamma - 1);
    for (int f = 0; f < 6; ++f) c->bc.kind[f] = f < 2 * nd ? cfg->bc[f] : 0;

    // host tables (mesh/grid.cpp:31-44, :134-155; geometry.cpp trigonometry; gravity.cpp:44-55)
    std::vector<double> h;
    size_t off_x[3], off_dx[3], off_xc[3];
    for (int d = 0; d < 3; ++d) {
        int const nn = G.nwg[d];
        off_x[d] = h.size();
        h.insert(h.end(), cfg->x[d], cfg->x[d] + nn + 1);
        off_dx[d] = h.size();
        for (int i = 0; i < nn; ++i) h.push_back(cfg->x[d][i + 1] - cfg->x[d][i]);
        off_xc[d] = h.size();
        for (int i = 0; i < nn; ++i) h.push_back((cfg->x[d][i] + cfg->x[d][i + 1]) / 2);
    }
    int const n1 = G.nwg[1];
    double const* x1 = cfg->x[1];
    size_t const off_dcos = h.size();
    for (int j = 0; j < n1; ++j) h.push_back(std::cos(x1[j]) - std::cos(x1[j + 1]));
    size_t const off_sin = h.size();
    for (int j = 0; j <= n1; ++j) h.push_back(std::sin(x1[j]));
    size_t const off_cot = h.size();
    for (int j = 0; j < n1; ++j) {
        double const mid = (x1[j] + x1[j + 1]) / 2;
        h.push_back(std::cos(mid) / std::sin(mid));
    }
    size_t const off_sinc = h.size();
    for (int j = 0; j < n1; ++j) h.push_back(std::sin(h[off_xc[1] + j]));
    size_t const off_gpm = h.size();
    for (int i = 0; i < G.nwg[0]; ++i) {
        double const xc = h[off_xc[0] + i];
        h.push_back(-kGrav * cfg->central_mass / (xc * xc));
    }

    auto cleanup = [&](int code) {
        hclpp_ctx_destroy(c);
        return code;
    };
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipGetDevice"));
    {
        int const rc = noalloc_scratch_setup(c->device);
        if (rc != HCLPP_OK) return cleanup(rc);
    }
    e = hipMalloc(&c->tables, sizeof(double) * h.size());
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipMalloc(tables)"));
    e = hipMemcpy(c->tables, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipMemcpy(tables)"));
    for (int d = 0; d < 3; ++d) {
        G.x[d] = c->tables + off_x[d];
        G.dx[d] = c->tables + off_dx[d];
        G.xc[d] = c->tables + off_xc[d];
    }
    G.dcos = c->tables + off_dcos;
    G.sin_node = c->tables + off_sin;
    G.cot_mid = c->tables + off_cot;
    G.sin_center = c->tables + off_sinc;
    G.gpm = c->tables + off_gpm;
    c->gtab = c->tables + off_gpm;
    if (cfg->gravity == HCLPP_GRAV_INTERNAL_MASS) {
        e = hipMalloc(&c->shells, sizeof(double) * 2 * (G.n[0] + G.ngh[0]));
        if (e != hipSuccess) return cleanup(hip_fail(e, "hipMalloc(shells)"));
    }

    e = hipMalloc(&c->scalars, sizeof(double) * 16);
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipMalloc(scalars)"));
    e = hipHostMalloc(&c->host_scalars, sizeof(double) * 16, hipHostMallocDefault);
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipHostMalloc(scalars)"));
    e = hipMalloc(&c->clock, sizeof(DevClock));
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipMalloc(clock)"));
    e = hipHostMalloc(&c->host_clock, sizeof(DevClock), hipHostMallocDefault);
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipHostMalloc(clock)"));
    std::memset(c->host_clock, 0, sizeof(DevClock));
    c->host_clock->pub.done = 1;  // no step runs before hclpp_clock_start
    e = hipMemcpy(c->clock, c->host_clock, sizeof(DevClock), hipMemcpyHostToDevice);
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipMemcpy(clock)"));

    // fused-step scratch
    c->shell = make_ghost_shell(G, nd);
    c->ghost_cap = ghost_blocks(c->shell);
    if (c->ghost_cap > 0) {
        e = hipMalloc(&c->ghost_partials, sizeof(double) * c->ghost_cap);
        if (e != hipSuccess) return cleanup(hip_fail(e, "hipMalloc(ghost partials)"));
    }
    {  // the fused kernel this context launches, and how many of its workgroups the chip holds
        int dev = 0, ncu = 0;
        e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return cleanup(hip_fail(e, "hipDeviceGetAttribute(multiProcessorCount)"));
        int const occ = march_occupancy(nd, cfg->geom, cfg->riemann, cfg->limiter, cfg->precision == HCLPP_FP32_STORAGE, cfg->nfx >= 1);
        c->march = plan_march(G, nd, occ, ncu);
        c->march_occ = occ;
    }
    c->fused_nb = c->march.nb;
    Range3 all;
    for (int d = 0; d < 3; ++d) {
        all.b[d] = 0;
        all.e[d] = G.nwg[d];
    }
    int64_t const need = std::max<int64_t>(c->fused_nb * kPartialStride, reduce_blocks(all));
    int const rc = ensure_partials(c, need);
    if (rc != HCLPP_OK) return cleanup(rc);
    *out = c;
    return HCLPP_OK;
}

int hclpp_ctx_destroy(hclpp_ctx* c)
{
    if (!c) return HCLPP_OK;
    if (c->tables) (void)hipFree(c->tables);
    if (c->partials) (void)hipFree(c->partials);
    if (c->ghost_partials) (void)hipFree(c->ghost_partials);
    if (c->fix_partials) (void)hipFree(c->fix_partials);
    if (c->scalars) (void)hipFree(c->scalars);
    if (c->host_scalars) (void)hipHostFree(c->host_scalars);
    if (c->clock) (void)hipFree(c->clock);
    if (c->host_clock) (void)hipHostFree(c->host_clock);
    if (c->fx_old) (void)hipFree(c->fx_old);
    if (c->shells) (void)hipFree(c->shells);
    if (c->adj_prim) (void)hipFree(c->adj_prim);
    if (c->adj_face) (void)hipFree(c->adj_face);
    if (c->adj_partials) (void)hipFree(c->adj_partials);
    for (hipEvent_t e : c->prof_events) (void)hipEventDestroy(e);
    delete c;
    return HCLPP_OK;
}

int hclpp_set_stream(hclpp_ctx* c, void* stream)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    c->stream = (hipStream_t)stream;
    return HCLPP_OK;
}

int hclpp_range_with_ghosts(const hclpp_ctx* c, int ngeff, hclpp_range* out)
{
    if (!c || !out) return fail(HCLPP_ERR_ARG, "null argument");
    if (ngeff < 0 || ngeff > c->cfg.ng) return fail(HCLPP_ERR_ARG, "NgEff > Ng");  // mesh/range.cpp:99-107
    for (int d = 0; d < 3; ++d) {
        int const e = d < c->cfg.ndim ? ngeff : 0;
        out->begin[d] = c->G.ngh[d] - e;
        out->end[d] = c->G.ngh[d] + c->G.n[d] + e;
    }
    return HCLPP_OK;
}

int hclpp_conv_prim_to_cons(hclpp_ctx* c, const hclpp_range* r, const double* rho, const double* u, const double* P, double* rhou, double* E)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    int rc = check_range(c, r, 0);
    if (rc) return rc;
    launch_prim_to_cons(c->cfg.ndim, c->G, to_range(r), c->eos, rho, u, P, rhou, E, c->stream);
    HCLPP_LAUNCHED("conv_prim_to_cons");
    return HCLPP_OK;
}

int hclpp_conv_cons_to_prim(hclpp_ctx* c, const hclpp_range* r, const double* rho, const double* rhou, const double* E, double* u, double* P)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    int rc = check_range(c, r, 0);
    if (rc) return rc;
    launch_cons_to_prim(c->cfg.ndim, c->G, to_range(r), c->eos, rho, rhou, E, u, P, c->stream);
    HCLPP_LAUNCHED("conv_cons_to_prim");
    return HCLPP_OK;
}

int hclpp_face_reconstruction(hclpp_ctx* c, int limiter, const hclpp_range* r, const double* var, double* var_rec)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    if (limiter < 0 || limiter > 3) return fail(HCLPP_ERR_ARG, "Unknown face reconstruction algorithm");
    int rc = check_range(c, r, 1);
    if (rc) return rc;
    launch_face_reconstruction(c->cfg.ndim, limiter, c->G, to_range(r), var, var_rec, c->stream);
    HCLPP_LAUNCHED("face_reconstruction");
    return HCLPP_OK;
}

static int hancock_impl(hclpp_ctx* c, const hclpp_range* r, double dt, const double* u_rec, const double* P_rec, double* rho_rec,
                        double* rhou_rec, double* E_rec, double* fx_rec)
{
    int const nd = c->cfg.ndim, nfx = c->cfg.nfx;
    if (nfx > 0) {  // extrapolation_reconstruction.hpp:87-88: the mass fractions before the update
        int64_t const n = c->G.plane * 2 * nd * nfx;
        if (n > c->fx_old_cap) {
            if (c->fx_old) HCLPP_HIP(hipFree(c->fx_old));
            c->fx_old = nullptr;
            HCLPP_HIP(hipMalloc(&c->fx_old, sizeof(double) * n));
            c->fx_old_cap = n;
        }
        HCLPP_HIP(hipMemcpyAsync(c->fx_old, fx_rec, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    }
    launch_hancock(nd, c->cfg.geom, c->G, to_range(r), c->eos, dt, u_rec, P_rec, rho_rec, rhou_rec, E_rec, fx_rec, c->fx_old, nfx, c->stream);
    HCLPP_LAUNCHED("Hancock_extrapolation");
    return HCLPP_OK;
}

int hclpp_hancock_extrapolation(hclpp_ctx* c, const hclpp_range* r, double dt, const double* u_rec, const double* P_rec, double* rho_rec,
                                double* rhou_rec, double* E_rec, double* fx_rec)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    int rc = check_range(c, r, 1);
    if (rc) return rc;
    return hancock_impl(c, r, dt, u_rec, P_rec, rho_rec, rhou_rec, E_rec, fx_rec);
}

int hclpp_face_prim_to_cons(hclpp_ctx* c, const hclpp_range* r, const double* rho_rec, const double* u_rec, const double* P_rec, double* rhou_rec,
                            double* E_rec)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    int const rc = check_range(c, r, 0);
    if (rc) return rc;
    if (!rho_rec || !u_rec || !P_rec || !rhou_rec || !E_rec) return fail(HCLPP_ERR_ARG, "null array");
    launch_face_prim_to_cons(c->cfg.ndim, c->G, to_range(r), c->eos, rho_rec, u_rec, P_rec, rhou_rec, E_rec, c->stream);
    HCLPP_LAUNCHED("conv_prim_to_cons (face states)");
    return HCLPP_OK;
}

int hclpp_hydro_reconstruction(hclpp_ctx* c, int limiter, const hclpp_range* r, double dt, const double* rho, const double* u, const double* P,
                               const double* fx, double* rho_rec, double* rhou_rec, double* E_rec, double* fx_rec, double* P_rec,
                               double* u_rec)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    if (limiter < 0 || limiter > 3) return fail(HCLPP_ERR_ARG, "Unknown face reconstruction algorithm");
    int rc = check_range(c, r, 1);
    if (rc) return rc;
    int const nd = c->cfg.ndim, nfx = c->cfg.nfx;
    int64_t const pl = c->G.plane;
    Range3 const R = to_range(r);
    // hydro_reconstruction.hpp:99-109
    launch_face_reconstruction(nd, limiter, c->G, R, rho, rho_rec, c->stream);
    for (int d = 0; d < nd; ++d) launch_face_reconstruction(nd, limiter, c->G, R, u + pl * d, u_rec + pl * 2 * nd * d, c->stream);
    launch_face_reconstruction(nd, limiter, c->G, R, P, P_rec, c->stream);
    for (int f = 0; f < nfx; ++f) launch_face_reconstruction(nd, limiter, c->G, R, fx + pl * f, fx_rec + pl * 2 * nd * f, c->stream);
    // :111-126
    launch_face_prim_to_cons(nd, c->G, R, c->eos, rho_rec, u_rec, P_rec, rhou_rec, E_rec, c->stream);
    HCLPP_LAUNCHED("face_reconstruction / conv_prim_to_cons");
    // :128
    return hancock_impl(c, r, dt, u_rec, P_rec, rho_rec, rhou_rec, E_rec, fx_rec);
}

int hclpp_godunov(hclpp_ctx* c, int riemann, const hclpp_range* r, double dt, const double* rho, const double* rhou, const double* E,
                  const double* fx, const double* rho_rec, const double* rhou_rec, const double* E_rec, const double* fx_rec, double* rho_new,
                  double* rhou_new, double* E_new, double* fx_new)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    if (riemann < 0 || riemann > 2) return fail(HCLPP_ERR_ARG, "Invalid riemann solver");
    int rc = check_range(c, r, 1);
    if (rc) return rc;
    launch_godunov(c->cfg.ndim, c->cfg.geom, riemann, c->G, to_range(r), c->eos, dt, rho, rhou, E, fx, rho_rec, rhou_rec, E_rec, fx_rec, rho_new,
                   rhou_new, E_new, fx_new, c->cfg.nfx, c->stream);
    HCLPP_LAUNCHED("Riemann_based_Godunov_scheme");
    return HCLPP_OK;
}

int hclpp_time_step(hclpp_ctx* c, const hclpp_range* r, const double* rho, const double* u, const double* P, double* dt_out)
{
    if (!c || !dt_out) return fail(HCLPP_ERR_ARG, "null argument");
    int rc = check_range(c, r, 0);
    if (rc) return rc;
    Range3 const R = to_range(r);
    rc = ensure_partials(c, reduce_blocks(R));
    if (rc) return rc;
    launch_time_step_rate(c->cfg.ndim, c->cfg.geom, c->G, R, c->eos, rho, u, P, c->partials, c->scalars, c->stream);
    HCLPP_LAUNCHED("time_step");
    double inv;
    rc = read_scalar(c, &inv);
    if (rc) return rc;
    *dt_out = 1 / inv;  // time_step.hpp:81
    return HCLPP_OK;
}

int hclpp_time_step_noalloc(hclpp_ctx* c, const hclpp_range* r, const double* rho, const double* u, const double* P, double* dt_dev)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    int rc = check_range(c, r, 0);
    if (rc) return rc;
    Range3 const R = to_range(r);
    int64_t const nb = reduce_blocks(R);
    if (nb > kNoallocScratchDoubles)
        return fail(HCLPP_ERR_ARG, "hclpp_time_step_noalloc: the range needs " + std::to_string(nb) + " workgroup partials; the "
                                       "no-allocation scratch holds " + std::to_string(kNoallocScratchDoubles) +
                                       " (64 x 4 cells each): use hclpp_time_step");
    if (nb == 0) return fail(HCLPP_ERR_ARG, "empty range");
    if (!rho || !u || !P || !dt_dev) return fail(HCLPP_ERR_ARG, "null argument");
    double* const ptr = noalloc_scratch(c->device);
    if (!ptr) return fail(HCLPP_ERR_STATE, "no-allocation scratch not set up on the context's device");
    // every partial the fold reads is written by this launch: nothing needs clearing
    launch_time_step_inverse(c->cfg.ndim, c->cfg.geom, c->G, R, c->eos, rho, u, P, ptr, dt_dev, c->stream);
    HCLPP_LAUNCHED("time_step (no allocation)");
    return HCLPP_OK;
}

int hclpp_temperature(hclpp_ctx* c, const hclpp_range* r, const double* rho, const double* P, double* T)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    int const rc = check_range(c, r, 0);
    if (rc) return rc;
    if (!rho || !P || !T) return fail(HCLPP_ERR_ARG, "null array");
    launch_temperature(c->G, to_range(r), c->eos, rho, P, T, c->stream);
    HCLPP_LAUNCHED("fill_temperature_array");
    return HCLPP_OK;
}

int hclpp_min_internal_energy(hclpp_ctx* c, const hclpp_range* r, const double* rho, const double* rhou, const double* E, double* out)
{
    if (!c || !out) return fail(HCLPP_ERR_ARG, "null argument");
    int rc = check_range(c, r, 0);
    if (rc) return rc;
    Range3 const R = to_range(r);
    rc = ensure_partials(c, reduce_blocks(R));
    if (rc) return rc;
    launch_min_internal_energy(c->cfg.ndim, c->G, R, rho, rhou, E, c->partials, c->scalars, c->stream);
    HCLPP_LAUNCHED("min_internal_energy");
    return read_scalar(c, out);
}

int hclpp_integrate(hclpp_ctx* c, const hclpp_range* r, const double* var, double* out)
{
    if (!c || !out) return fail(HCLPP_ERR_ARG, "null argument");
    int rc = check_range(c, r, 0);
    if (rc) return rc;
    Range3 const R = to_range(r);
    rc = ensure_partials(c, reduce_blocks(R));
    if (rc) return rc;
    launch_integrate(c->G, c->cfg.ndim, c->cfg.geom, R, var, c->partials, c->scalars, c->stream);
    HCLPP_LAUNCHED("integration");
    return read_scalar(c, out);
}

int hclpp_pressure_fix(hclpp_ctx* c, const hclpp_range* r, double dt, double eps, const double* rho, const double* rhou, const double* E,
                       const double* rho_rec, const double* rhou_rec, const double* E_rec, const double* rho_new, const double* rhou_new,
                       double* E_new)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    int rc = check_range(c, r, 1);
    if (rc) return rc;
    if (!rho || !rhou || !E || !rho_rec || !rhou_rec || !E_rec || !rho_new || !rhou_new || !E_new) return fail(HCLPP_ERR_ARG, "null array");
    launch_pressure_fix(c->cfg.ndim, c->cfg.geom, c->G, to_range(r), c->eos, dt, eps, rho, rhou, E, rho_rec, rhou_rec, E_rec, rho_new,
                        rhou_new, E_new, c->stream);
    HCLPP_LAUNCHED("pressure_fix");
    return HCLPP_OK;
}

int hclpp_internal_mass_gravity(hclpp_ctx* c, const double* rho)
{
    if (!c || !rho) return fail(HCLPP_ERR_ARG, "null argument");
    if (c->cfg.gravity != HCLPP_GRAV_INTERNAL_MASS) return fail(HCLPP_ERR_ARG, "the context was not created with internal-mass gravity");
    launch_internal_mass_gravity(c->cfg.ndim, c->cfg.geom, c->G, c->cfg.central_mass, rho, c->shells, c->gtab, nullptr, c->stream);
    HCLPP_LAUNCHED("internal_mass_gravity");
    return HCLPP_OK;
}

int hclpp_gravity_table(hclpp_ctx* c, double* host_out)
{
    if (!c || !host_out) return fail(HCLPP_ERR_ARG, "null argument");
    HCLPP_HIP(hipMemcpyAsync(host_out, c->gtab, sizeof(double) * c->G.nwg[0], hipMemcpyDeviceToHost, c->stream));
    HCLPP_HIP(hipStreamSynchronize(c->stream));
    return HCLPP_OK;
}

int hclpp_fill_ghosts(hclpp_ctx* c, double* rho, double* rhou, double* E, double* fx)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    launch_fill_ghosts(c->cfg.ndim, c->G, c->bc, rho, rhou, E, fx, c->cfg.nfx, c->stream);
    HCLPP_LAUNCHED("boundary conditions");
    return HCLPP_OK;
}

// One fused step (hclpp_step, hclpp_step_f32 and their passive-scalar forms): S is the storage
// type.  fx = NULL for the scalar-free forms, which refuse contexts with nfx > 0.
extern "C++" {
template <class S>
static int fused_step(hclpp_ctx* c, const S* rho, const S* rhou, const S* E, const S* fx, S* rho_new, S* rhou_new, S* E_new, S* fx_new)
{
    constexpr bool f32 = std::is_same_v<S, float>;
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    int const nfx = c->cfg.nfx;
    bool const with_fx = fx || fx_new;
    if (!with_fx && nfx > 0)
        return fail(HCLPP_ERR_UNSUPPORTED, f32 ? "nfx > 0: use hclpp_step_fx_f32" : "nfx > 0: use hclpp_step_fx (or the operator entry points)");
    if (f32 && c->cfg.precision != HCLPP_FP32_STORAGE) return fail(HCLPP_ERR_ARG, "the context was not created for FP32 storage");
    if (!f32 && c->cfg.precision != HCLPP_FP64) return fail(HCLPP_ERR_UNSUPPORTED, "hclpp_step takes FP64 storage");
    if (with_fx && c->fast_on) return fail(HCLPP_ERR_UNSUPPORTED, "the fast FP64 arithmetic carries no passive scalars");
    if (c->cfg.ng < 2) return fail(HCLPP_ERR_ARG, "the fused step needs ng >= 2 (two-cell reconstruction stencil)");
    if (!rho || !rhou || !E || !rho_new || !rhou_new || !E_new) return fail(HCLPP_ERR_ARG, "null array");
    if (nfx > 0 && (!fx || !fx_new)) return fail(HCLPP_ERR_ARG, "null passive scalar array");
    if (rho == rho_new || E == E_new || rhou == rhou_new || (nfx > 0 && fx == fx_new))
        return fail(HCLPP_ERR_ARG, "input and output arrays must differ");
    StepArraysT<S> A;
    A.rho = rho;
    A.rhou = rhou;
    A.E = E;
    A.rho_new = rho_new;
    A.rhou_new = rhou_new;
    A.E_new = E_new;
    A.fx = nfx > 0 ? fx : nullptr;
    A.fx_new = nfx > 0 ? fx_new : nullptr;
    A.nfx = nfx;
    A.scalars_only = 0;
    A.pf = f32 ? 0 : c->pf_on;
    A.eps_pf = c->pf_eps;
    A.partials = c->partials;
    A.ghost_partials = c->ghost_partials;
    A.clock = c->clock;
    int const nd = c->cfg.ndim;
    auto const& eos = [&]() -> const auto& {
        if constexpr (f32) return c->eos32;
        else return c->eos;
    }();
    if (c->cfg.gravity == HCLPP_GRAV_INTERNAL_MASS) {  // main.cpp:309, :528: gravity from the state the step reads
        launch_internal_mass_gravity(nd, c->cfg.geom, c->G, c->cfg.central_mass, rho, c->shells, c->gtab, clock_gate(c), c->stream);
        HCLPP_LAUNCHED("internal_mass_gravity");
    }
    bool const timed = c->prof_count < c->prof_max;
    if (timed) HCLPP_HIP(hipEventRecord(c->prof_events[2 * c->prof_count], c->stream));
    if (nfx > 0) {
        // one launch of the one-scalar kernel per passive scalar (the scalars do not couple: each
        // is upwinded with the same mass flux, godunov_scheme.hpp:263-301); the first also writes
        // the hydro update and the partials
        for (int f = 0; f < nfx; ++f) {
            StepArraysT<S> Af = A;
            Af.fx = fx + c->G.plane * f;
            Af.fx_new = fx_new + c->G.plane * f;
            Af.nfx = 1;
            Af.scalars_only = f > 0;
            launch_fused_step_fx(nd, c->cfg.geom, c->cfg.riemann, c->cfg.limiter, c->G, c->march, eos, Af, c->stream);
        }
    } else {
        bool fast = false;
        if constexpr (!f32) fast = c->fast_on;
        if constexpr (!f32) {
            if (fast) launch_fused_step(nd, c->cfg.geom, c->cfg.riemann, c->cfg.limiter, c->G, c->march, c->eos_fast, A, c->stream);
        }
        if (!fast) launch_fused_step(nd, c->cfg.geom, c->cfg.riemann, c->cfg.limiter, c->G, c->march, eos, A, c->stream);
    }
    HCLPP_LAUNCHED("fused step");
    if (timed) {
        HCLPP_HIP(hipEventRecord(c->prof_events[2 * c->prof_count + 1], c->stream));
        c->prof_count += 1;
    }
    if constexpr (!f32) {
        if (c->pf_on) {  // main.cpp:459-475: after the update, before the boundaries
            launch_pressure_fix_fused(nd, c->cfg.geom, c->cfg.limiter, c->G, eos, A, c->march, c->fix_partials, c->stream);
            HCLPP_LAUNCHED("pressure fix (fused step)");
        }
    }
    launch_ghosts_rate(nd, c->cfg.geom, c->G, eos, c->bc, c->shell, A, c->stream);
    HCLPP_LAUNCHED("ghost fill");
    launch_clock_advance(nd, c->clock, c->partials, c->fused_nb, c->ghost_partials, c->ghost_cap, c->pf_on && !f32 ? c->fix_partials : nullptr,
                         c->pf_on && !f32 ? c->fix_nb : 0, c->stream);
    HCLPP_LAUNCHED("clock advance");
    return HCLPP_OK;
}
}  // extern "C++"

static int clock_reset(hclpp_ctx* c, double t0, int iter0, double t_end, double cfl, int max_iter)
{
    if (!(cfl > 0) || max_iter < 1) return fail(HCLPP_ERR_ARG, "cfl must be > 0 and max_iter >= 1");
    if (iter0 < 0) return fail(HCLPP_ERR_ARG, "iter0 must be >= 0");
    if (c->cfg.ng < 2) return fail(HCLPP_ERR_ARG, "the fused step needs ng >= 2 (two-cell reconstruction stencil)");
    DevClock* h = c->host_clock;
    HCLPP_HIP(hipStreamSynchronize(c->stream));  // the pinned staging buffer is reused
    std::memset(h, 0, sizeof(DevClock));
    h->pub.t = t0;
    h->pub.iter = iter0;
    h->pub.t_end = t_end;
    h->pub.cfl = cfl;
    h->pub.max_iter = max_iter;
    h->pub.min_eint = NAN;
    h->pub.iter_output_id = c->out_iter_id;
    h->pub.time_output_id = c->out_time_id;
    h->iter_output_frequency = c->out_iter_freq;
    h->time_first_output = c->out_time_first;
    h->time_output_frequency = c->out_time_freq;
    HCLPP_HIP(hipMemcpyAsync(c->clock, h, sizeof(DevClock), hipMemcpyHostToDevice, c->stream));
    Range3 all;
    for (int d = 0; d < 3; ++d) {
        all.b[d] = 0;
        all.e[d] = c->G.nwg[d];
    }
    return ensure_partials(c, reduce_blocks(all));
}

int hclpp_set_pressure_fix(hclpp_ctx* c, int on, double eps)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    if (on && c->cfg.precision != HCLPP_FP64) return fail(HCLPP_ERR_UNSUPPORTED, "the pressure fix runs with FP64 storage only");
    if (on && c->fast_on) return fail(HCLPP_ERR_UNSUPPORTED, "the pressure fix runs with the reference arithmetic only");
    if (on && !std::isfinite(eps)) return fail(HCLPP_ERR_ARG, "eps must be finite");
    if (on && (!c->fix_partials || c->fix_nb != c->fused_nb)) {  // one pair per workgroup of the march kernel's plan
        if (c->fix_partials) {
            HCLPP_HIP(hipStreamSynchronize(c->stream));
            HCLPP_HIP(hipFree(c->fix_partials));
            c->fix_partials = nullptr;
        }
        int64_t const nb = c->fused_nb;
        HCLPP_HIP(hipMalloc(&c->fix_partials, sizeof(double) * 2 * (nb > 0 ? nb : 1)));
        c->fix_nb = nb;
    }
    c->pf_on = on ? 1 : 0;
    c->pf_eps = eps;
    return HCLPP_OK;
}

// Launch plan of the fused kernel the context runs (the plan depends on its occupancy) and the
// partials buffer it fills.  Waits for the stream: the buffer may be reallocated.
static int replan_march(hclpp_ctx* c, int occ)
{
    int dev = 0, ncu = 0;
    HCLPP_HIP(hipGetDevice(&dev));
    HCLPP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    HCLPP_HIP(hipStreamSynchronize(c->stream));
    c->march = plan_march(c->G, c->cfg.ndim, occ, ncu);
    c->march_occ = occ;
    c->fused_nb = c->march.nb;
    return ensure_partials(c, c->fused_nb * kPartialStride);
}

int hclpp_set_fast_arithmetic(hclpp_ctx* c, int on)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    if (on && c->cfg.precision != HCLPP_FP64) return fail(HCLPP_ERR_UNSUPPORTED, "the fast arithmetic is an FP64-storage mode");
    if (on && c->cfg.nfx > 0) return fail(HCLPP_ERR_UNSUPPORTED, "the fast FP64 arithmetic carries no passive scalars");
    if (on && c->pf_on) return fail(HCLPP_ERR_UNSUPPORTED, "the pressure fix runs with the reference arithmetic only");
    if ((on != 0) == (c->fast_on != 0)) return HCLPP_OK;
    int const nd = c->cfg.ndim;
    int const occ = on ? march_occupancy_fast(nd, c->cfg.geom, c->cfg.riemann, c->cfg.limiter)
                       : march_occupancy(nd, c->cfg.geom, c->cfg.riemann, c->cfg.limiter, false, false);
    int const rc = replan_march(c, occ);
    if (rc != HCLPP_OK) return rc;
    c->fast_on = on ? 1 : 0;
    return HCLPP_OK;
}

int hclpp_clock_set_outputs(hclpp_ctx* c, int iter_output_frequency, double time_first_output, double time_output_frequency, int iter_output_id,
                            int time_output_id)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    if (iter_output_frequency < 0 || !(time_output_frequency >= 0) || !std::isfinite(time_first_output) || iter_output_id < -1
        || time_output_id < -1)
        return fail(HCLPP_ERR_ARG, "invalid output schedule");
    c->out_iter_freq = iter_output_frequency;
    c->out_time_first = time_first_output;
    c->out_time_freq = time_output_frequency;
    c->out_iter_id = iter_output_id;
    c->out_time_id = time_output_id;
    return HCLPP_OK;
}

// Continue a paused clock (output_due) on the state it stopped at: t, iter, the output counters and
// the schedule carry over, and dt is computed afresh from the state, as the reference's next loop
// iteration does (main.cpp:355).
static int clock_resume_reset(hclpp_ctx* c)
{
    DevClock* h = c->host_clock;
    HCLPP_HIP(hipMemcpyAsync(h, c->clock, sizeof(DevClock), hipMemcpyDeviceToHost, c->stream));
    HCLPP_HIP(hipStreamSynchronize(c->stream));
    if (h->pub.status) return fail(HCLPP_ERR_STATE, "the clock stopped on an error");
    if (h->pub.exiting) return fail(HCLPP_ERR_ARG, "the run has ended (t_end or max_iter reached)");
    h->pub.done = 0;
    h->pub.output_due = 0;
    h->pending_exit = 0;
    h->pending_output = 0;
    HCLPP_HIP(hipMemcpyAsync(c->clock, h, sizeof(DevClock), hipMemcpyHostToDevice, c->stream));
    return HCLPP_OK;
}

int hclpp_clock_resume(hclpp_ctx* c, const double* rho, const double* rhou, const double* E)
{
    if (!c || !rho || !rhou || !E) return fail(HCLPP_ERR_ARG, "null argument");
    if (c->cfg.precision != HCLPP_FP64) return fail(HCLPP_ERR_ARG, "the context was created for FP32 storage");
    int rc = clock_resume_reset(c);
    if (rc) return rc;
    Range3 all;
    for (int d = 0; d < 3; ++d) {
        all.b[d] = 0;
        all.e[d] = c->G.nwg[d];
    }
    launch_time_step_cons(c->cfg.ndim, c->cfg.geom, c->G, all, c->eos, rho, rhou, E, c->partials, c->scalars, c->stream);
    launch_clock_start(c->clock, c->scalars, c->stream);
    HCLPP_LAUNCHED("clock resume");
    HCLPP_HIP(hipStreamSynchronize(c->stream));
    return HCLPP_OK;
}

int hclpp_clock_resume_f32(hclpp_ctx* c, const float* rho, const float* rhou, const float* E)
{
    if (!c || !rho || !rhou || !E) return fail(HCLPP_ERR_ARG, "null argument");
    if (c->cfg.precision != HCLPP_FP32_STORAGE) return fail(HCLPP_ERR_ARG, "the context was not created for FP32 storage");
    int rc = clock_resume_reset(c);
    if (rc) return rc;
    launch_state_rate_f32(c->cfg.ndim, c->cfg.geom, c->G, c->eos32, rho, rhou, E, c->partials, c->scalars, c->stream);
    launch_clock_start(c->clock, c->scalars, c->stream);
    HCLPP_LAUNCHED("clock resume (fp32)");
    HCLPP_HIP(hipStreamSynchronize(c->stream));
    return HCLPP_OK;
}

int hclpp_clock_start_f32(hclpp_ctx* c, double t0, int iter0, double t_end, double cfl, int max_iter, const float* rho,
                          const float* rhou, const float* E)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    if (c->cfg.precision != HCLPP_FP32_STORAGE) return fail(HCLPP_ERR_ARG, "the context was not created for FP32 storage");
    int rc = clock_reset(c, t0, iter0, t_end, cfl, max_iter);
    if (rc) return rc;
    launch_state_rate_f32(c->cfg.ndim, c->cfg.geom, c->G, c->eos32, rho, rhou, E, c->partials, c->scalars, c->stream);
    launch_clock_start(c->clock, c->scalars, c->stream);
    HCLPP_LAUNCHED("clock start (fp32)");
    HCLPP_HIP(hipStreamSynchronize(c->stream));
    return HCLPP_OK;
}

int hclpp_step_f32(hclpp_ctx* c, const float* rho, const float* rhou, const float* E, float* rho_new, float* rhou_new, float* E_new)
{
    return fused_step<float>(c, rho, rhou, E, nullptr, rho_new, rhou_new, E_new, nullptr);
}

int hclpp_step_fx_f32(hclpp_ctx* c, const float* rho, const float* rhou, const float* E, const float* fx, float* rho_new, float* rhou_new,
                      float* E_new, float* fx_new)
{
    return fused_step<float>(c, rho, rhou, E, fx, rho_new, rhou_new, E_new, fx_new);
}

int hclpp_cast_f64_to_f32(hclpp_ctx* c, const double* in, float* out, long long n)
{
    if (!c || (n > 0 && (!in || !out))) return fail(HCLPP_ERR_ARG, "null argument");
    launch_cast(in, out, (int64_t)n, c->stream);
    HCLPP_LAUNCHED("cast f64->f32");
    return HCLPP_OK;
}

int hclpp_cast_f32_to_f64(hclpp_ctx* c, const float* in, double* out, long long n)
{
    if (!c || (n > 0 && (!in || !out))) return fail(HCLPP_ERR_ARG, "null argument");
    launch_cast(in, out, (int64_t)n, c->stream);
    HCLPP_LAUNCHED("cast f32->f64");
    return HCLPP_OK;
}

int hclpp_clock_start(hclpp_ctx* c, double t0, int iter0, double t_end, double cfl, int max_iter, const double* rho, const double* rhou,
                      const double* E)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    if (c->cfg.precision != HCLPP_FP64) return fail(HCLPP_ERR_ARG, "the context was created for FP32 storage");
    int rc = clock_reset(c, t0, iter0, t_end, cfl, max_iter);
    if (rc) return rc;
    Range3 all;
    for (int d = 0; d < 3; ++d) {
        all.b[d] = 0;
        all.e[d] = c->G.nwg[d];
    }
    // main.cpp:355: dt from the ghost-inclusive state
    launch_time_step_cons(c->cfg.ndim, c->cfg.geom, c->G, all, c->eos, rho, rhou, E, c->partials, c->scalars, c->stream);
    launch_clock_start(c->clock, c->scalars, c->stream);
    HCLPP_LAUNCHED("clock start");
    HCLPP_HIP(hipStreamSynchronize(c->stream));
    return HCLPP_OK;
}

int hclpp_clock_read(hclpp_ctx* c, hclpp_clock* out)
{
    if (!c || !out) return fail(HCLPP_ERR_ARG, "null argument");
    HCLPP_HIP(hipMemcpyAsync(c->host_clock, c->clock, sizeof(DevClock), hipMemcpyDeviceToHost, c->stream));
    HCLPP_HIP(hipStreamSynchronize(c->stream));
    *out = c->host_clock->pub;
    return HCLPP_OK;
}

int hclpp_step(hclpp_ctx* c, const double* rho, const double* rhou, const double* E, double* rho_new, double* rhou_new, double* E_new)
{
    return fused_step<double>(c, rho, rhou, E, nullptr, rho_new, rhou_new, E_new, nullptr);
}

int hclpp_step_fx(hclpp_ctx* c, const double* rho, const double* rhou, const double* E, const double* fx, double* rho_new, double* rhou_new,
                  double* E_new, double* fx_new)
{
    return fused_step<double>(c, rho, rhou, E, fx, rho_new, rhou_new, E_new, fx_new);
}

int hclpp_step_dt(hclpp_ctx* c, double dt, const double* rho, const double* rhou, const double* E, double* rho_new, double* rhou_new,
                  double* E_new)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    if (!std::isfinite(dt) || !(dt > 0)) return fail(HCLPP_ERR_ARG, "dt must be finite and > 0");
    if (c->cfg.precision != HCLPP_FP64) return fail(HCLPP_ERR_UNSUPPORTED, "hclpp_step_dt takes FP64 storage");
    launch_clock_set_dt(c->clock, dt, c->stream);
    HCLPP_LAUNCHED("clock dt");
    return hclpp_step(c, rho, rhou, E, rho_new, rhou_new, E_new);
}

int hclpp_clock_record_dt(hclpp_ctx* c, double* log, int cap)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    if (log && cap < 0) return fail(HCLPP_ERR_ARG, "cap < 0");
    launch_clock_set_log(c->clock, log, cap, c->stream);
    HCLPP_LAUNCHED("clock dt log");
    return HCLPP_OK;
}

namespace {

// One reverse step (see hclpp_step_adjoint).  prims_ready: the primitives of (rho, rhou, E) are
// already in the scratch buffer (the previous call formed them); next_*: the state of the next
// reverse step, whose primitives this step's gather pass forms, or null.  *pending: in, the
// dgamma partials the previous step left unfolded (0 if none); out, this step's (0 if fold_now).
int step_adjoint(hclpp_ctx* c, double dt, const double* rho, const double* rhou, const double* E, const double* lrho_new,
                 const double* lrhou_new, const double* lE_new, double* lrho, double* lrhou, double* lE, double* dgamma, bool prims_ready,
                 const double* next_rho, const double* next_rhou, const double* next_E, int64_t* pending, bool fold_now)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    if (c->cfg.precision != HCLPP_FP64) return fail(HCLPP_ERR_UNSUPPORTED, "the reverse step takes FP64 storage");
    if (c->cfg.riemann != HCLPP_RS_HLL && c->cfg.riemann != HCLPP_RS_HLLC)
        return fail(HCLPP_ERR_UNSUPPORTED, "the reverse step supports the HLL and HLLC solvers only");
    if (c->cfg.nfx > 0) return fail(HCLPP_ERR_UNSUPPORTED, "the reverse step has no passive scalars");
    if (c->pf_on) return fail(HCLPP_ERR_UNSUPPORTED, "the reverse step does not differentiate the pressure fix");
    if (c->cfg.gravity == HCLPP_GRAV_INTERNAL_MASS)
        return fail(HCLPP_ERR_UNSUPPORTED, "the reverse step does not differentiate the state-dependent internal-mass gravity");
    if (c->cfg.ng < 2 || c->cfg.ng > adj::kMaxGhosts) return fail(HCLPP_ERR_UNSUPPORTED, "the reverse step needs 2 <= ng <= 4");
    if (!std::isfinite(dt) || !(dt > 0)) return fail(HCLPP_ERR_ARG, "dt must be finite and > 0");
    if (!rho || !rhou || !E || !lrho_new || !lrhou_new || !lE_new || !lrho || !lrhou || !lE || !dgamma)
        return fail(HCLPP_ERR_ARG, "null array");
    if (lrho == lrho_new || lrhou == lrhou_new || lE == lE_new) return fail(HCLPP_ERR_ARG, "input and output adjoints must differ");
    int const nd = c->cfg.ndim;
    int const nv = nd + 2;
    if (!c->adj_prim) {
        HCLPP_HIP(hipMalloc(&c->adj_prim, sizeof(double) * c->G.plane * nv));
        HCLPP_HIP(hipMalloc(&c->adj_face, sizeof(double) * c->G.plane * adj::face_planes(nd)));
        HCLPP_HIP(hipMalloc(&c->adj_partials, sizeof(double) * adjoint_partials(c->G, nd)));
    }
    adj::AdjArrays A{rho, rhou, E, c->adj_prim, c->adj_face, lrho_new, lrhou_new, lE_new, lrho, lrhou, lE, dt};
    *pending = launch_step_adjoint(nd, c->cfg.geom, c->cfg.riemann, c->cfg.limiter, c->G, c->eos, c->bc, A, prims_ready, next_rho, next_rhou, next_E,
                                   *pending, fold_now, c->adj_partials, dgamma, c->stream);
    HCLPP_LAUNCHED("reverse step");
    return HCLPP_OK;
}

}  // namespace

int hclpp_step_adjoint(hclpp_ctx* c, double dt, const double* rho, const double* rhou, const double* E, const double* lrho_new,
                       const double* lrhou_new, const double* lE_new, double* lrho, double* lrhou, double* lE, double* dgamma)
{
    int64_t pending = 0;
    return step_adjoint(c, dt, rho, rhou, E, lrho_new, lrhou_new, lE_new, lrho, lrhou, lE, dgamma, false, nullptr, nullptr, nullptr,
                        &pending, true);
}

int hclpp_steps_dt(hclpp_ctx* c, int n, const double* dts, double* const* rho, double* const* rhou, double* const* E)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    if (n < 0 || (n > 0 && (!dts || !rho || !rhou || !E))) return fail(HCLPP_ERR_ARG, "null argument");
    for (int k = 0; k < n; ++k) {
        int const rc = hclpp_step_dt(c, dts[k], rho[k], rhou[k], E[k], rho[k + 1], rhou[k + 1], E[k + 1]);
        if (rc) return rc;
    }
    return HCLPP_OK;
}

int hclpp_steps_adjoint(hclpp_ctx* c, int n, const double* dts, const double* const* rho, const double* const* rhou, const double* const* E,
                        double* const* lrho, double* const* lrhou, double* const* lE, double* dgamma)
{
    if (!c) return fail(HCLPP_ERR_ARG, "null context");
    if (n < 0 || (n > 0 && (!dts || !rho || !rhou || !E || !lrho || !lrhou || !lE))) return fail(HCLPP_ERR_ARG, "null argument");
    int64_t pending = 0;  // dgamma partials of the step just run, folded by the next one (or below)
    for (int k = n - 1; k >= 0; --k) {
        bool const more = k > 0;
        int rc = (more && (!rho[k - 1] || !rhou[k - 1] || !E[k - 1])) ? fail(HCLPP_ERR_ARG, "null array") : HCLPP_OK;
        if (!rc)
            rc = step_adjoint(c, dts[k], rho[k], rhou[k], E[k], lrho[k + 1], lrhou[k + 1], lE[k + 1], lrho[k], lrhou[k], lE[k], dgamma,
                              k < n - 1, more ? rho[k - 1] : nullptr, more ? rhou[k - 1] : nullptr, more ? E[k - 1] : nullptr, &pending, !more);
        if (rc) {
            if (pending > 0) launch_adjoint_fold(c->adj_partials, pending, dgamma, c->stream);  // the steps already run stay counted
            return rc;
        }
    }
    return HCLPP_OK;
}

int hclpp_profile_begin(hclpp_ctx* c, int max_steps)
{
    if (!c || max_steps < 0) return fail(HCLPP_ERR_ARG, "invalid argument");
    HCLPP_HIP(hipStreamSynchronize(c->stream));
    while ((int)c->prof_events.size() < 2 * max_steps) {
        hipEvent_t e;
        HCLPP_HIP(hipEventCreate(&e));
        c->prof_events.push_back(e);
    }
    c->prof_max = max_steps;
    c->prof_count = 0;
    return HCLPP_OK;
}

int hclpp_profile_end(hclpp_ctx* c, double* kernel_ms, int* launches)
{
    if (!c || !kernel_ms || !launches) return fail(HCLPP_ERR_ARG, "null argument");
    HCLPP_HIP(hipStreamSynchronize(c->stream));
    double total = 0;
    for (int i = 0; i < c->prof_count; ++i) {
        float ms = 0;
        HCLPP_HIP(hipEventElapsedTime(&ms, c->prof_events[2 * i], c->prof_events[2 * i + 1]));
        total += ms;
    }
    *kernel_ms = total;
    *launches = c->prof_count;
    c->prof_max = 0;
    c->prof_count = 0;
    return HCLPP_OK;
}

int hclpp_fused_plan(hclpp_ctx* c, int* seglen, long long* workgroups, int* blocks_per_cu)
{
    if (!c || !seglen || !workgroups || !blocks_per_cu) return fail(HCLPP_ERR_ARG, "null argument");
    *seglen = c->march.L.seglen;
    *workgroups = c->march.nb;
    *blocks_per_cu = c->march_occ;
    return HCLPP_OK;
}

}  // extern "C"
