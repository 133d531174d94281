// GPU check of hclpp::fdiv (physics.hpp), the FP64 division of the hydro kernels: on operands
// inside its stated range it must return the IEEE quotient bit for bit.  Operand pairs are
// random and adversarial: divisor significands near all ones or near a power of two, quotients
// near rounding midpoints, exact quotients, signed zeros, zero, infinite and NaN operands,
// numerators down to 2^-969, quotients down to the normal minimum, exponent differences up to
// 767.  Pairs outside the range (smaller numerators, subnormal divisors or quotients,
// overflowing quotients) are counted and reported with their distance to the IEEE quotient;
// pairs with a subnormal divisor are counted apart, with the NaNs fdiv returns where a / b is
// finite (its stated behaviour there).
//
// Second check: the four slope limiters (limited_slope_t, slope_limiters.hpp:18-87), whose
// divisions are IEEE divisions because their operands range over every magnitude, on 2^22
// (diffR, diffL) pairs -- random, hydro-sized, with subnormal diffL and products dL * dR at the
// edge of underflow, with quotients r that overflow or whose square does -- must equal the
// host's evaluation of the same expressions bit for bit (NaN where the host gives NaN).
//
// Prints "in-range: N pairs, M mismatches" and "limiters: N pairs, M mismatches" and exits 0
// only when both M are 0.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "physics.hpp"

namespace {

uint64_t next(uint64_t& s)
{
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
}

double make(uint64_t sig, int e, bool neg)
{
    uint64_t bits = ((uint64_t)(e + 1023) << 52) | (sig & ((1ull << 52) - 1));
    if (neg) bits |= 1ull << 63;
    double d;
    std::memcpy(&d, &bits, 8);
    return d;
}

uint64_t bits_of(double d)
{
    uint64_t b;
    std::memcpy(&b, &d, 8);
    return b;
}

int biased_exponent(double d) { return (int)((bits_of(d) >> 52) & 0x7FF); }

// The operands on which fdiv is stated to equal the IEEE quotient: a zero, infinite or NaN
// operand (v_div_fixup_f64 decides those), or the normal range of v_div_scale_f64's identity.
bool in_range(double a, double b)
{
    if (a == 0 || b == 0 || !std::isfinite(a) || !std::isfinite(b)) return true;
    if (!std::isnormal(a) || !std::isnormal(b)) return false;
    if (std::fabs(a) < std::ldexp(1.0, -969) || std::fabs(b) > std::ldexp(1.0, 1022)) return false;
    double const q = a / b;
    if (!std::isnormal(q)) return false;
    return biased_exponent(a) - biased_exponent(b) < 768;
}

// distance in units of the last place between two finite doubles of the same sign
uint64_t ulp_distance(double x, double y)
{
    if (x == y) return 0;
    if (std::signbit(x) != std::signbit(y)) return ~0ull;
    uint64_t const bx = bits_of(x) & ~(1ull << 63), by = bits_of(y) & ~(1ull << 63);
    return bx > by ? bx - by : by - bx;
}

__global__ void k_fdiv(const double* a, const double* b, double* q, long long n)
{
    long long const i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = hclpp::fdiv(a[i], b[i]);
}

template <int LIM>
__host__ __device__ double slope(double dR, double dL)
{
    return hclpp::limited_slope_t<LIM, hclpp::PerfectGas, double>(dR, dL);
}

// out[4 i + l] = limiter l of (dR[i], dL[i])
__global__ void k_limiters(const double* dR, const double* dL, double* out, long long n)
{
    long long const i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[4 * i + 0] = slope<hclpp::kConstant>(dR[i], dL[i]);
    out[4 * i + 1] = slope<hclpp::kVanLeer>(dR[i], dL[i]);
    out[4 * i + 2] = slope<hclpp::kMinmod>(dR[i], dL[i]);
    out[4 * i + 3] = slope<hclpp::kVanAlbada>(dR[i], dL[i]);
}

}  // namespace

#define HIP_OK(x)                                                      \
    do {                                                               \
        hipError_t e_ = (x);                                           \
        if (e_ != hipSuccess) {                                        \
            std::printf("FAIL: %s: %s\n", #x, hipGetErrorString(e_)); \
            return 2;                                                  \
        }                                                              \
    } while (0)

int main()
{
    long long const n = 1ll << 24;
    std::vector<double> a(n), b(n), q(n);
    uint64_t s = 0x9E3779B97F4A7C15ull;
    auto rnd = [&](int lo, int hi) { return lo + (int)(next(s) % (uint64_t)(hi - lo + 1)); };
    for (long long t = 0; t < n; ++t) {
        int const mode = (int)(t % 12);
        uint64_t sa = next(s), sb = next(s);
        bool const na = next(s) & 1, nb = next(s) & 1;
        double aa = make(sa, rnd(-380, 380), na), bb = make(sb, rnd(-380, 380), nb);
        switch (mode) {
            case 1: bb = make(sb | (((1ull << 52) - 1) ^ (next(s) & 0xFF)), rnd(-380, 380), nb); break;  // divisor near all ones
            case 2: bb = make(sb & (next(s) & 0xFF), rnd(-380, 380), nb); break;                         // divisor near 2^k
            case 3: {  // quotient near a rounding midpoint
                double const qq = make(next(s), rnd(-100, 100), false);
                aa = bb * qq + bb * (std::nextafter(qq, INFINITY) - qq) / 2;
                break;
            }
            case 4: aa = na ? -0.0 : 0.0; break;                                          // signed zeros
            case 5: aa = make(sa, rnd(-969, -900), na); bb = make(sb, rnd(-50, 50), nb); break;  // smallest numerators
            case 6: aa = make(sa, rnd(380, 383), na); bb = make(sb, rnd(-384, -380), nb); break;  // exponent gap near 767
            case 7: aa = make(sa, rnd(-600, -590), na); bb = make(sb, rnd(405, 420), nb); break;  // quotients near 2^-1010
            case 8: aa = make(sa, rnd(-40, 40), na); bb = make(sb, rnd(-40, 40), nb); break;      // hydro-sized operands
            case 9: aa = bb * (double)(1 + (next(s) % 4096)); break;                               // exact quotients
            case 10: {  // outside the range: tiny numerators, subnormal quotients, huge gaps, subnormal
                        // numerators, subnormal divisors (with a / b finite)
                int const k = (int)(next(s) % 5);
                if (k == 0) aa = make(sa, rnd(-1022, -970), na);
                if (k == 1) { aa = make(sa, rnd(-700, -600), na); bb = make(sb, rnd(420, 500), nb); }
                if (k == 2) { aa = make(sa, rnd(500, 1000), na); bb = make(sb, rnd(-300, -1), nb); }
                if (k == 3) aa = std::ldexp((double)(next(s) % 1000 + 1), -1074);
                if (k == 4) {
                    bb = std::ldexp((double)(next(s) % (1ull << 52) + 1), -1074) * (nb ? -1 : 1);
                    aa = make(sa, rnd(-1060, -960), na);
                }
                break;
            }
            case 11: {  // zero, infinite and NaN operands on either side
                double const sp[6] = {INFINITY, -INFINITY, NAN, 0.0, -0.0, 1.5};
                aa = (next(s) & 1) ? sp[next(s) % 6] : aa;
                bb = sp[next(s) % 6];
                break;
            }
            default: break;  // mode 0: random operands within 2^+-380
        }
        a[t] = aa;
        b[t] = bb;
    }
    double *da, *db, *dq;
    HIP_OK(hipMalloc(&da, n * 8));
    HIP_OK(hipMalloc(&db, n * 8));
    HIP_OK(hipMalloc(&dq, n * 8));
    HIP_OK(hipMemcpy(da, a.data(), n * 8, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(db, b.data(), n * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_fdiv, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, da, db, dq, n);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(q.data(), dq, n * 8, hipMemcpyDeviceToHost));
    HIP_OK(hipFree(da));
    HIP_OK(hipFree(db));
    HIP_OK(hipFree(dq));

    long long in_n = 0, in_bad = 0, out_n = 0, out_diff = 0, out_far = 0, out_nonfinite = 0;
    long long sub_n = 0, sub_nan = 0;
    uint64_t out_max_ulp = 0;
    for (long long t = 0; t < n; ++t) {
        double const ref = a[t] / b[t];
        bool const same = bits_of(ref) == bits_of(q[t]) || (std::isnan(ref) && std::isnan(q[t]));
        if (b[t] != 0 && std::fpclassify(b[t]) == FP_SUBNORMAL) {
            ++sub_n;
            if (std::isfinite(ref) && std::isnan(q[t])) ++sub_nan;
        }
        if (in_range(a[t], b[t])) {
            ++in_n;
            if (!same) {
                if (in_bad < 5) std::printf("mismatch: %a / %a = %a, fdiv %a\n", a[t], b[t], ref, q[t]);
                ++in_bad;
            }
        } else {
            ++out_n;
            if (same) continue;
            ++out_diff;
            if (!std::isfinite(ref) || !std::isfinite(q[t])) {
                ++out_nonfinite;
            } else {
                uint64_t const u = ulp_distance(ref, q[t]);
                if (u > 1) ++out_far;
                if (u > out_max_ulp) out_max_ulp = u;
            }
        }
    }
    std::printf("in-range: %lld pairs, %lld mismatches\n", in_n, in_bad);
    std::printf("out-of-range: %lld pairs, %lld differ (%lld non-finite on one side, %lld finite by more than one ulp, "
                "largest finite distance %llu ulp)\n",
                out_n, out_diff, out_nonfinite, out_far, (unsigned long long)out_max_ulp);
    std::printf("subnormal divisors: %lld pairs, %lld NaN where a / b is finite\n", sub_n, sub_nan);

    // ---- the slope limiters
    long long const m = 1ll << 22;
    std::vector<double> dR(m), dL(m), sl(4 * m);
    long long n_sub = 0;
    for (long long t = 0; t < m; ++t) {
        int const mode = (int)(t % 8);
        bool const nr = next(s) & 1, nl = (mode >= 4) ? nr : (next(s) & 1);  // modes 4..7: same sign
        double r = make(next(s), rnd(-40, 40), nr), l = make(next(s), rnd(-40, 40), nl);
        switch (mode) {
            case 1: l = r * (1 + (double)(next(s) % 1000) * 1e-3); break;                        // r near 1
            case 2: r = (next(s) & 1) ? 0.0 : -0.0; break;                                        // flat
            case 4: l = std::ldexp((double)(next(s) % (1ull << 52) + 1), -1074) * (nl ? -1 : 1);  // subnormal dL
                    r = make(next(s), rnd(-60, 4), nr);
                    break;
            case 5: {  // dL * dR at the edge of underflow, r between 2^960 and overflow
                l = std::ldexp((double)(next(s) % (1ull << 52) + 1), -1074) * (nl ? -1 : 1);
                r = std::ldexp(1.0, rnd(-60, 2)) * (nr ? -1 : 1) * (1 + (double)(next(s) % 1000) * 1e-3);
                break;
            }
            case 6: l = make(next(s), rnd(-1022, -900), nl); r = make(next(s), rnd(-10, 10), nr); break;  // r^2 overflows
            case 7: l = make(next(s), rnd(-520, -500), nl); r = make(next(s), rnd(-10, 10), nr); break;   // (r + 1)^2 near 2^1024
            default: break;
        }
        dR[t] = r;
        dL[t] = l;
        if (std::fpclassify(l) == FP_SUBNORMAL) ++n_sub;
    }
    double *dr, *dl, *dsl;
    HIP_OK(hipMalloc(&dr, m * 8));
    HIP_OK(hipMalloc(&dl, m * 8));
    HIP_OK(hipMalloc(&dsl, 4 * m * 8));
    HIP_OK(hipMemcpy(dr, dR.data(), m * 8, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dl, dL.data(), m * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_limiters, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, 0, dr, dl, dsl, m);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(sl.data(), dsl, 4 * m * 8, hipMemcpyDeviceToHost));
    HIP_OK(hipFree(dr));
    HIP_OK(hipFree(dl));
    HIP_OK(hipFree(dsl));
    long long lim_bad = 0, lim_nan = 0;
    for (long long t = 0; t < m; ++t) {
        double const host[4] = {slope<hclpp::kConstant>(dR[t], dL[t]), slope<hclpp::kVanLeer>(dR[t], dL[t]),
                                slope<hclpp::kMinmod>(dR[t], dL[t]), slope<hclpp::kVanAlbada>(dR[t], dL[t])};
        for (int l = 0; l < 4; ++l) {
            double const g = sl[4 * t + l];
            if (std::isnan(host[l])) ++lim_nan;
            bool const same = bits_of(host[l]) == bits_of(g) || (std::isnan(host[l]) && std::isnan(g));
            if (!same) {
                if (lim_bad < 5) std::printf("limiter %d mismatch: dR %a dL %a host %a gpu %a\n", l, dR[t], dL[t], host[l], g);
                ++lim_bad;
            }
        }
    }
    std::printf("limiters: %lld pairs, %lld mismatches (%lld with a subnormal diffL, %lld NaN slopes on the host)\n", m, lim_bad,
                n_sub, lim_nan);
    return (in_bad == 0 && lim_bad == 0) ? 0 : 1;
}
