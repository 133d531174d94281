// GPU check of hclpp::fdiv (physics.hpp), the FP64 division of the hydro kernels: on operands
// inside its stated range it must return the IEEE quotient bit for bit.  Operand pairs are
// random and adversarial: divisor significands near all ones or near a power of two, quotients
// near rounding midpoints, exact quotients, signed zeros, zero, infinite and NaN operands,
// numerators down to 2^-969, quotients down to the normal minimum, exponent differences up to
// 767.  Pairs outside the range (smaller numerators, subnormal divisors or quotients,
// overflowing quotients) are counted and reported with their distance to the IEEE quotient.
//
// Prints "in-range: N pairs, M mismatches" and exits 0 only when M == 0.
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
            case 10: {  // outside the range: tiny numerators, subnormal quotients, huge gaps, subnormal divisors
                int const k = (int)(next(s) % 4);
                if (k == 0) aa = make(sa, rnd(-1022, -970), na);
                if (k == 1) { aa = make(sa, rnd(-700, -600), na); bb = make(sb, rnd(420, 500), nb); }
                if (k == 2) { aa = make(sa, rnd(500, 1000), na); bb = make(sb, rnd(-300, -1), nb); }
                if (k == 3) aa = std::ldexp((double)(next(s) % 1000 + 1), -1074);
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
    uint64_t out_max_ulp = 0;
    for (long long t = 0; t < n; ++t) {
        double const ref = a[t] / b[t];
        bool const same = bits_of(ref) == bits_of(q[t]) || (std::isnan(ref) && std::isnan(q[t]));
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
    return in_bad == 0 ? 0 : 1;
}
