// adjoint.hip -- reverse-mode hydro step on gfx950 (Cartesian or spherical geometry, HLL/HLLC,
// every limiter, uniform or point-mass gravity, every boundary kind), dt held fixed.
//
// Five passes over 64 x 4 workgroups (x fastest, coalesced SoA planes); each output element
// is written by exactly one thread, and the dL/dgamma contributions are reduced per
// workgroup and folded in a fixed order, so the gradient is bit-reproducible run to run.
// In a run of reverse steps, each step's gather pass also forms the primitives the next one
// needs, so the primitive pass runs once per run, and each step's faces pass folds the
// previous step's dL/dgamma partials, so the fold runs on its own once per run.
#include "adjoint_step.hpp"
#include "reduce.hpp"

namespace hclpp {

using adj::AdjArrays;

namespace {

constexpr int kT = kBX * kBY;

__device__ __forceinline__ int64_t block_lin() { return blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z); }

__device__ __forceinline__ void put_partial(double v, double* partials)
{
    __shared__ double lds[kT / 64];
    int const tid = threadIdx.x + kBX * threadIdx.y;
    v = block_reduce<kSum>(v, lds, tid, kT);
    if (tid == 0) partials[block_lin()] = v;
}

// dgamma[0] += the sum of n partials in a fixed order, by one workgroup of kT threads: thread t
// adds partials t, t + kT, ... in turn (the loads of 16 rounds issued before their adds), then the
// workgroup tree.  The same order wherever it runs, so dgamma does not depend on where.
__device__ __forceinline__ void fold_partials(const double* __restrict__ partials, int64_t n, double* __restrict__ dgamma)
{
    constexpr int kU = 16;
    __shared__ double lds[kT / 64];
    int const tid = threadIdx.x + kBX * threadIdx.y;
    double v = 0;
    for (int64_t t0 = tid; t0 < n; t0 += (int64_t)kU * kT) {
        double x[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            int64_t const t = t0 + (int64_t)u * kT;
            x[u] = t < n ? partials[t] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (t0 + (int64_t)u * kT < n) v += x[u];
    }
    v = block_reduce<kSum>(v, lds, tid, kT);
    if (tid == 0) dgamma[0] += v;
}

template <int ND>
__global__ __launch_bounds__(256) void k_adj_prim(GridDev G, Range3 r, PerfectGas eos, AdjArrays A)
{
    int i, j, k;
    if (range_cell(r, i, j, k)) adj::cell_prim<ND>(G, eos, A, i, j, k);
}

// Face states; the first workgroup also folds the `pending` partials the previous reverse step
// left (the kernel boundary has published them, and this step writes its own only later).
template <int ND, int LIM, int GEOM>
__global__ __launch_bounds__(256) void k_adj_faces(GridDev G, Range3 r, PerfectGas eos, AdjArrays A, const double* __restrict__ partials,
                                                   int64_t pending, double* __restrict__ dgamma)
{
    int i, j, k;
    if (range_cell(r, i, j, k)) adj::cell_faces<ND, LIM, GEOM>(G, eos, A, i, j, k);
    if (pending > 0 && block_lin() == 0) fold_partials(partials, pending, dgamma);
}

// Cell c of the face range along d: the face between c - e_d and c touches an interior cell.
__device__ __forceinline__ bool on_face_range(const GridDev& G, int d, int nd, const int* c)
{
    for (int e = 0; e < nd; ++e) {
        bool const ok = (e == d) ? (c[e] >= G.ngh[e] && c[e] <= G.ngh[e] + G.n[e]) : adj::inside(G, e, c[e]);
        if (!ok) return false;
    }
    return true;
}

// Every face direction in one pass: cell c handles the face on its low side along each dim
// whose face range holds c.  Each face reads and rewrites its own two slots, so no two faces
// touch the same memory.
template <int ND, int RS, int GEOM>
__global__ __launch_bounds__(256) void k_adj_flux(GridDev G, Range3 r, PerfectGas eos, AdjArrays A, double* __restrict__ partials)
{
    int i, j, k;
    double lg = 0;
    if (range_cell(r, i, j, k)) {
        int const c[3] = {i, j, k};
        if (on_face_range(G, 0, ND, c)) lg += adj::face_flux<ND, RS, 0, GEOM>(G, eos, A, i, j, k);
        if constexpr (ND >= 2)
            if (on_face_range(G, 1, ND, c)) lg += adj::face_flux<ND, RS, 1, GEOM>(G, eos, A, i, j, k);
        if constexpr (ND == 3)
            if (on_face_range(G, 2, ND, c)) lg += adj::face_flux<ND, RS, 2, GEOM>(G, eos, A, i, j, k);
    }
    put_partial(lg, partials);
}

template <int ND, int LIM, int GEOM>
__global__ __launch_bounds__(256) void k_adj_hancock(GridDev G, Range3 r, PerfectGas eos, AdjArrays A, double* __restrict__ partials)
{
    int i, j, k;
    double lg = 0;
    if (range_cell(r, i, j, k)) lg = adj::cell_hancock<ND, LIM, GEOM>(G, eos, A, i, j, k);
    put_partial(lg, partials);
}

// The state the next reverse step differentiates (the step before in time), whose primitives
// the gather pass computes; null when there is none.
struct NextState {
    const double* rho;
    const double* rhou;
    const double* E;
};

__device__ __forceinline__ bool in_range(const Range3& r, int i, int j, int k)
{
    return i >= r.b[0] && i < r.e[0] && j >= r.b[1] && j < r.e[1] && k >= r.b[2] && k < r.e[2];
}

// Gather pass over `rg` (the cells within 2 of the interior): interior cells run the gather;
// with `nx`, every cell also stores the next state's primitives for the next reverse step (the
// gather reads no primitive, so the buffer is free).
template <int ND, int LIM>
__global__ __launch_bounds__(256) void k_adj_gather(GridDev G, Range3 rg, Range3 interior, PerfectGas eos, BcSpec bc, AdjArrays A, NextState nx,
                                                    double* __restrict__ partials)
{
    // a workgroup whose cells all lie where no ghost copies from (adj::no_ghost_copies) takes the
    // gather without the preimage search; the test is uniform over the workgroup
    int const i0 = rg.b[0] + (int)blockIdx.x * kBX, j0 = rg.b[1] + (int)blockIdx.y * kBY;
    bool deep = adj::no_ghost_copies(G, 0, i0, i0 + kBX);
    if (ND >= 2) deep = deep && adj::no_ghost_copies(G, 1, j0, j0 + kBY);
    if (ND == 3) deep = deep && adj::no_ghost_copies(G, 2, rg.b[2] + (int)blockIdx.z, rg.b[2] + (int)blockIdx.z + 1);
    int i, j, k;
    double lg = 0;
    if (range_cell(rg, i, j, k)) {
        if (deep) lg = adj::cell_gather<ND, LIM, true>(G, eos, bc, A, i, j, k);
        else if (in_range(interior, i, j, k)) lg = adj::cell_gather<ND, LIM>(G, eos, bc, A, i, j, k);
        if (nx.rho) {
            AdjArrays B = A;
            B.rho = nx.rho;
            B.rhou = nx.rhou;
            B.E = nx.E;
            adj::cell_prim<ND>(G, eos, B, i, j, k);
        }
    }
    put_partial(lg, partials);
}

// The fold on its own, for the last step of a run.  (Folds run after a kernel boundary, which
// publishes every workgroup's partial at once; a fold in the last workgroup of the gather pass
// would need a device-scope fence, an L2 write-back on each XCD, in every workgroup.)
__global__ __launch_bounds__(256) void k_adj_fold(const double* __restrict__ partials, int64_t n, double* __restrict__ dgamma)
{
    fold_partials(partials, n, dgamma);
}

struct AdjRanges {
    Range3 ring2, ring1, interior, flux;  // flux: the union of the face ranges (interior + 1 cell on the high side of every dim)
};

AdjRanges adj_ranges(const GridDev& G, int nd)
{
    AdjRanges R;
    for (int d = 0; d < 3; ++d) {
        bool const act = d < nd;
        R.ring2.b[d] = act ? G.ngh[d] - 2 : 0;
        R.ring2.e[d] = act ? G.ngh[d] + G.n[d] + 2 : 1;
        R.ring1.b[d] = act ? G.ngh[d] - 1 : 0;
        R.ring1.e[d] = act ? G.ngh[d] + G.n[d] + 1 : 1;
        R.interior.b[d] = G.ngh[d];
        R.interior.e[d] = G.ngh[d] + G.n[d];
    }
    R.flux = R.interior;
    for (int d = 0; d < nd; ++d) R.flux.e[d] += 1;
    return R;
}

int64_t blocks_of(const Range3& r)
{
    dim3 const g = range_grid(r);
    return (int64_t)g.x * g.y * g.z;
}

template <int ND, int GEOM>
void launch_flux(int rs, const GridDev& G, const AdjRanges& R, const PerfectGas& eos, const AdjArrays& A, double* partials, int64_t& off,
                 hipStream_t s)
{
    Range3 const& r = R.flux;
    if (rs == kHLL)
        hipLaunchKernelGGL((k_adj_flux<ND, kHLL, GEOM>), range_grid(r), dim3(kBX, kBY), 0, s, G, r, eos, A, partials + off);
    else
        hipLaunchKernelGGL((k_adj_flux<ND, kHLLC, GEOM>), range_grid(r), dim3(kBX, kBY), 0, s, G, r, eos, A, partials + off);
    off += blocks_of(r);
}

}  // namespace

int64_t adjoint_partials(const GridDev& G, int nd)
{
    AdjRanges const R = adj_ranges(G, nd);
    return blocks_of(R.ring1) + blocks_of(R.ring2) + blocks_of(R.flux);
}

#define HCLPP_ADJ_LIM_SWITCH(lim, BODY)                      \
    switch (lim) {                                           \
    case kConstant: { constexpr int LIM = kConstant; BODY; } break;   \
    case kVanLeer: { constexpr int LIM = kVanLeer; BODY; } break;     \
    case kMinmod: { constexpr int LIM = kMinmod; BODY; } break;       \
    default: { constexpr int LIM = kVanAlbada; BODY; } break;         \
    }
#define HCLPP_ADJ_GEOM_SWITCH(geom, BODY)                                       \
    if (geom == kSpherical) { constexpr int GEOM = kSpherical; BODY; }          \
    else { constexpr int GEOM = kCartesian; BODY; }
#define HCLPP_ADJ_ND_SWITCH(nd, BODY)                  \
    switch (nd) {                                      \
    case 1: { constexpr int ND = 1; BODY; } break;     \
    case 2: { constexpr int ND = 2; BODY; } break;     \
    default: { constexpr int ND = 3; BODY; } break;    \
    }

int64_t launch_step_adjoint(int nd, int geom, int rs, int lim, const GridDev& G, const PerfectGas& eos, const BcSpec& bc, const AdjArrays& A,
                            bool prims_ready, const double* next_rho, const double* next_rhou, const double* next_E, int64_t pending,
                            bool fold_now, double* partials, double* dgamma, hipStream_t s)
{
    AdjRanges const R = adj_ranges(G, nd);
    dim3 const blk(kBX, kBY);
    NextState const nx{next_rho, next_rhou, next_E};
    // The gather grid covers the cells within 2 of the interior whether or not it forms the next
    // primitives, so its workgroups, and with them the dgamma partials, are the same either way:
    // dgamma does not depend on how a run of reverse steps is split into calls.
    Range3 const& rg = R.ring2;
    int64_t off = 0;
    HCLPP_ADJ_ND_SWITCH(nd, HCLPP_ADJ_GEOM_SWITCH(geom, {
        if (!prims_ready) hipLaunchKernelGGL(k_adj_prim<ND>, range_grid(R.ring2), blk, 0, s, G, R.ring2, eos, A);
        HCLPP_ADJ_LIM_SWITCH(lim, hipLaunchKernelGGL((k_adj_faces<ND, LIM, GEOM>), range_grid(R.ring1), blk, 0, s, G, R.ring1, eos, A,
                                                     partials, pending, dgamma));
        (launch_flux<ND, GEOM>)(rs, G, R, eos, A, partials, off, s);
        HCLPP_ADJ_LIM_SWITCH(lim, hipLaunchKernelGGL((k_adj_hancock<ND, LIM, GEOM>), range_grid(R.ring1), blk, 0, s, G, R.ring1, eos, A,
                                                     partials + off));
        off += blocks_of(R.ring1);
        HCLPP_ADJ_LIM_SWITCH(lim, hipLaunchKernelGGL((k_adj_gather<ND, LIM>), range_grid(rg), blk, 0, s, G, rg, R.interior, eos, bc, A, nx,
                                                     partials + off));
        off += blocks_of(rg);
    }));
    if (!fold_now) return off;
    launch_adjoint_fold(partials, off, dgamma, s);
    return 0;
}

void launch_adjoint_fold(const double* partials, int64_t n, double* dgamma, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_adj_fold, dim3(1), dim3(kBX, kBY), 0, s, partials, n, dgamma);
}

}  // namespace hclpp
