// Coefficient-brick precompute: the hybrid thermal + power-law synchrotron model evaluated for every
// cell of a GRMHD snapshot, one lane per cell, FP64.
//
// Behaviour follows python/blackhole_sim/coefficient_bricks.py:93-152 (cell frames with the pitch
// angle to the line of sight at 90 deg, the cast to the brick dtype and nan_to_num) and
// python/blackhole_sim/synchrotron.py:122-325 (kernel table, distributions, pitch-angle integral,
// Kirchhoff absorption, Faraday terms, EVPA rotation).
//
// Design (gfx950):
//  * One lane per cell.  Each block stages in LDS the kernel table (log x, log F, log G and the
//    segment slopes at 256 knots), the Gauss-Legendre nodes and the power-law Lorentz-factor grid
//    with its normalised distribution.
//  * Gauss-Legendre nodes come in +-mu pairs with equal weights, and the integrand depends on
//    sin(alpha) only, so a pair costs one evaluation; the node sum keeps the reference's order.
//  * x = nu / nu_c enters only through log x = log nu - log(1.5 nu_B sin(alpha)) - 2 log gamma,
//    so an evaluation is an index, two linear interpolations and two exponentials.
//  * The Lorentz-factor loop is the outer loop: each gamma's distribution value is formed once and
//    the trapezoid sums of all nodes advance together (the thermal part takes two passes: its
//    normalisation integral, then the emissivities).
//  * K_nu by the trapezoidal rule on its cosh integral gives the x K_{2/3}(x) table on the host and
//    the K_2 normalisation of the Maxwell-Juttner distribution per cell.

#include "bh_internal.h"

#include <cmath>
#include <cstring>
#include <limits>

namespace bh {

namespace syn {

constexpr double kPi = 3.14159265358979323846;
// calibration.py:18-26 and synchrotron.py:200 (cgs)
constexpr double kC = 2.99792458e10;
constexpr double kMp = 1.67262192369e-24;
constexpr double kMe = 9.1093837015e-28;
constexpr double kKb = 1.380649e-16;
constexpr double kE = 4.80320471257e-10;
constexpr double kH = 6.62607015e-27;

constexpr int kTable = 256;         // synchrotron.py:158 (n = 256 knots over [1e-8, 1e3])
constexpr int kMaxNodes = 16;       // Gauss-Legendre pitch nodes per model
constexpr int kMaxUnique = (kMaxNodes + 1) / 2;
constexpr int kMaxGamma = 1024;     // Lorentz-factor grid points per model

// K_nu(x) for real nu >= 0 and x > 0 from the integral representation (DLMF 10.32.9)
//     K_nu(x) = int_0^inf exp(-x cosh t) cosh(nu t) dt,
// summed by the trapezoidal rule.  The integrand is even in t and analytic in the strip
// |Im t| < pi/2, so the rule's error falls like exp(-2 pi d / h) for any d inside the strip
// (Trefethen & Weideman, SIAM Review 56 (2014) 385, section 5); for large x the integrand is close
// to a Gaussian of width 1/sqrt(x), whose trapezoid error is exp(-2 pi^2 / (x h^2)).  The step
// h = min(1/8, 1/(2 sqrt x)) puts both exponents beyond 60.  The factor exp(-x) is taken out of the
// sum and cosh t - 1 is formed as 2 sinh(t/2)^2 with expm1, so no term loses digits to cancellation.
// The integrand rises to one peak and then falls (its log-derivative -x sinh t + nu tanh(nu t) changes
// sign once), so the sum stops on the falling side once a term is below 1e-18 of it.  exp(-x)
// underflows for x > 745 and K_nu is then 0, as scipy.special.kv returns.
__host__ __device__ inline double bessel_k(double nu, double x) {
    const double scale = exp(-x);
    if (scale == 0.0) return 0.0;
    const double h = fmin(0.125, 0.5 / sqrt(x));
    double sum = 0.5;   // the t = 0 node, half weight: exp(0) cosh(0)
    double carry = 0.0; // compensated summation
    double last = 1.0;
    for (int k = 1; k < 16384; ++k) {
        const double t = double(k) * h;
        const double em = expm1(0.5 * t);              // e^{t/2} - 1
        const double sh = 0.5 * (em + em / (1.0 + em)); // sinh(t/2)
        const double en = exp(nu * t);
        const double term = exp(-2.0 * x * sh * sh) * 0.5 * (en + 1.0 / en);
        const double y = term - carry;
        const double next = sum + y;
        carry = (next - sum) - y;
        sum = next;
        if (term <= last && term < 1.0e-18 * sum) break;
        last = term;
    }
    return scale * h * sum;
}

// Legendre polynomial P_n(x) and its derivative P_n'(x) in x87 extended precision, from the
// three-term recurrence (k+1) P_{k+1} = (2k+1) x P_k - k P_{k-1} and the derivative identity
// P'_{k+1} = P'_{k-1} + (2k+1) P_k (DLMF 18.9.1 and 18.9.19), with no division by 1 - x^2.
static void legendre_eval(int n, long double x, long double &pn, long double &dpn) {
    long double p_prev = 1.0L, p = x;     // P_0, P_1
    long double d_prev = 0.0L, d = 1.0L;  // P_0', P_1'
    for (int k = 1; k < n; ++k) {
        const long double p_next = ((2 * k + 1) * x * p - k * p_prev) / (k + 1);
        const long double d_next = d_prev + (2 * k + 1) * p;
        p_prev = p;
        p = p_next;
        d_prev = d;
        d = d_next;
    }
    pn = p;
    dpn = d;
}

// Gauss-Legendre nodes (ascending) and weights on [-1, 1].  The positive roots of P_n start from
// Tricomi's asymptotic estimate x_k = (1 - (n - 1) / (8 n^3)) cos(pi (4k - 1) / (4n + 2))
// (F. G. Tricomi, 1950; see N. Hale & A. Townsend, SIAM J. Sci. Comput. 35 (2013) A652) and are
// polished by Newton steps on P_n / P_n' until the update stops shrinking; the weight is
// w = 2 / ((1 - x^2) P_n'(x)^2) (Abramowitz & Stegun 25.4.29).  Both are formed in extended precision
// (64-bit significand), so the doubles they round to are within an ulp of the exact values: in double
// precision the weights of the outer nodes would inherit the node's rounding error amplified by
// 2 x / (1 - x^2).  Nodes are mirrored so that x[i] = -x[n-1-i] and w[i] = w[n-1-i] exactly.
static void gauss_legendre(int n, double *x, double *w) {
    const long double pi = 3.141592653589793238462643383279502884L;
    const long double shrink = 1.0L - (long double)(n - 1) / (8.0L * n * n * n);
    for (int k = 1; k <= (n + 1) / 2; ++k) {   // k-th largest root (the middle one of odd n is 0)
        long double z = (2 * k - 1 == n) ? 0.0L : shrink * cosl(pi * (4 * k - 1) / (4 * n + 2));
        long double pn, dpn, prev_step = 1.0L;
        for (int it = 0; it < 64 && z != 0.0L; ++it) {
            legendre_eval(n, z, pn, dpn);
            const long double step = pn / dpn;
            z -= step;
            const long double mag = fabsl(step);
            if (mag >= prev_step || mag < 1.0e-21L) break;   // rounding-bound, or converged
            prev_step = mag;
        }
        legendre_eval(n, z, pn, dpn);
        const long double wk = 2.0L / (((1.0L - z) * (1.0L + z)) * dpn * dpn);
        x[k - 1] = -double(z);
        x[n - k] = double(z);   // after the mirror, so the middle node of odd n is +0
        w[n - k] = w[k - 1] = double(wk);
    }
}

// ---- device tables -------------------------------------------------------------------------
// One buffer of doubles, staged into LDS by every block:
//   [0, 1530)     255 knot segments of 6: log x_i, log x_{i+1}, log F_i, slope F_i, log G_i, slope G_i
//   [1536, 1552)  thermal sin(alpha) of the distinct pitch nodes,  [1552, 1568) thermal 0.5 w
//   [1568, 1584)  power-law sin(alpha),                            [1584, 1600) power-law 0.5 w
//   [1600, 1600 + 2 G)  power-law log gamma, then w_k f_k (trapezoid weight times the normalised
//                       distribution; G = grid size)
constexpr int kSeg = 6;
constexpr int kOffThSin = 1536, kOffThHw = 1552, kOffNtSin = 1568, kOffNtHw = 1584, kOffNtGamma = 1600;

struct SynArgs {
    const double *rho, *theta_e, *b_con;
    int64_t n;
    void *out;
    int out_dtype;
    const double *tables;
    int table_words;
    double nu, log_nu, nu2, nu3;
    double rho_scale, b_scale, mpe;                 // mpe = max(mu_e m_p, 1e-300)
    int th_ng, th_nodes, nt_ng, nt_nodes;
    double th_frs, th_fcs;                          // Faraday rotation / conversion scales
    double nt_frac, nt_gmin, nt_abs, nt_pol, nt_fs, nt_log_ratio, nt_gmin2;
    double k_pref, k_mc2, k_nub, k_v, k_q, four_pi;
    double lx_first, lx_last, inv_dlx;
    double lF_top, lG_top;                          // log F and log G at the last knot
};

template <typename T>
__device__ __forceinline__ T pymax(T a, T b) { return (b > a) ? b : a; }   // Python's max(a, b)
template <typename T>
__device__ __forceinline__ T pymin(T a, T b) { return (b < a) ? b : a; }   // Python's min(a, b)

// exp(np.interp(lx, log x, log F)) and the same for G, for lx already clipped to the knot range
// (synchrotron.py:166-174).  The knots are evenly spaced in log x up to rounding, so the scaled
// offset names the segment except within rounding of a knot, where a short search fixes it.
__device__ __forceinline__ void kernel_interp(const double *t, const SynArgs &A, double lx, double &F, double &G) {
#pragma clang fp contract(off)
    const bool top = lx >= A.lx_last;   // x == xp[-1]: fp[-1]
    int j = int((lx - A.lx_first) * A.inv_dlx);
    j = j < 0 ? 0 : (j > kTable - 2 ? kTable - 2 : j);
    const double *e = t + kSeg * j;
    if (!top && (lx < e[0] || lx >= e[1])) {
        while (j > 0 && lx < t[kSeg * j]) --j;
        while (j < kTable - 2 && lx >= t[kSeg * j + 1]) ++j;
        e = t + kSeg * j;
    }
    const double dx = lx - e[0];
    F = exp(top ? A.lF_top : e[3] * dx + e[2]);
    G = exp(top ? A.lG_top : e[5] * dx + e[4]);
}

// The pitch-angle integral of synchrotron.py:210-228 for one cell and one electron population:
// sum over nodes of 0.5 w * trapz(f * pref * F(x)), and the same with G.  The trapezoid sums are
// written as sum_k w_k y_k (w_k the trapezoid weight of gamma_k); DIST(k, log gamma_k, w_k f_k)
// is called for k = 0, 1, ... in order.
template <typename Dist>
__device__ __forceinline__ void emissivities(const double *t, const SynArgs &A, int sin_off, int hw_off, int nodes,
                                             int ng, double ne, double b, Dist &dist, double &j_i, double &j_q) {
    const int nu_n = (nodes + 1) / 2;
    const double nu_b = (kE * b) / A.k_nub;
    double cu[kMaxUnique], pref[kMaxUnique], ti[kMaxUnique], tq[kMaxUnique];
#pragma unroll
    for (int u = 0; u < kMaxUnique; ++u) {
        const double s = t[sin_off + u];
        cu[u] = A.log_nu - log((1.5 * nu_b) * s);
        pref[u] = (A.k_pref * (b * s)) / A.k_mc2;
        ti[u] = tq[u] = 0.0;
    }
    for (int k = 0; k < ng; ++k) {
        double lg, wf;
        dist(k, lg, wf);
#pragma unroll
        for (int u = 0; u < kMaxUnique; ++u) {
            if (u < nu_n) {
                double lx = cu[u] - 2.0 * lg;
                lx = lx < A.lx_first ? A.lx_first : (lx > A.lx_last ? A.lx_last : lx);
                double F, G;
                kernel_interp(t, A, lx, F, G);
                ti[u] += wf * (pref[u] * F);
                tq[u] += wf * (pref[u] * G);
            }
        }
    }
    // node order of the reference: the negative-mu half ascending, then its mirror
    double tot_i = 0.0, tot_q = 0.0;
#pragma unroll
    for (int u = 0; u < kMaxUnique; ++u) {
        if (u < nu_n) {
            tot_i += t[hw_off + u] * ti[u];
            tot_q += t[hw_off + u] * tq[u];
        }
    }
    const int mirror = nodes / 2;   // nodes in the positive half (the middle node of an odd count is not repeated)
#pragma unroll
    for (int u = kMaxUnique - 1; u >= 0; --u) {
        if (u < mirror) {
            tot_i += t[hw_off + u] * ti[u];
            tot_q += t[hw_off + u] * tq[u];
        }
    }
    j_i = ne * tot_i / A.four_pi;
    j_q = ne * tot_q / A.four_pi;
}

// PolarizedCoefficients.rotated (synchrotron.py:68-79) added into acc.
__device__ __forceinline__ void add_rotated(double acc[kCoeffs], const double c[kCoeffs], double evpa) {
    double s2, c2;
    sincos(2.0 * evpa, &s2, &c2);
    acc[0] += c[0];
    acc[3] += c[3];
    acc[4] += c[4];
    acc[7] += c[7];
    acc[8] += c[8];
    const int qs[3] = {1, 5, 9};
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const int q = qs[m], u = q + 1;
        acc[q] += c[q] * c2 - c[u] * s2;
        acc[u] += c[q] * s2 + c[u] * c2;
    }
}

__device__ __forceinline__ double planck_nu(const SynArgs &A, double th) {
    const double T = th * kMe * kC * kC / kKb;
    const double x = kH * A.nu / pymax(kKb * T, 1.0e-300);
    if (x < 1.0e-3) return 2.0 * A.nu * A.nu * kKb * T / (kC * kC);
    return 2.0 * kH * A.nu3 / (kC * kC) / pymax(expm1(pymin(x, 700.0)), 1.0e-300);
}

// f64 -> f16 rounded once to nearest even: round toward zero to f32 and set the last bit when
// inexact (round to odd keeps the second rounding exact), then the hardware f32 -> f16 conversion.
__device__ __forceinline__ _Float16 to_half(double v) {
    float f = __double2float_rz(v);
    if (__builtin_isfinite(v) && double(f) != v) f = __builtin_bit_cast(float, __builtin_bit_cast(unsigned, f) | 1u);
    return _Float16(f);
}

// three waves per SIMD (<= 168 VGPRs; a few values spill outside the evaluation loop) hide the LDS
// and exp latencies better than two
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void synchro_kernel(SynArgs A) {
    extern __shared__ double lds[];
    for (int i = threadIdx.x; i < A.table_words; i += blockDim.x) lds[i] = A.tables[i];
    __syncthreads();
    const int64_t cell = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (cell >= A.n) return;
    const double *t = lds;

    // the cell's plasma frame (coefficient_bricks.py:110-116, calibration.py:113-119)
    const double b1 = A.b_con[4 * cell + 1], b2 = A.b_con[4 * cell + 2], b3 = A.b_con[4 * cell + 3];
    const double b_code = sqrt(b1 * b1 + b2 * b2 + b3 * b3);
    const double b = pymax(b_code * A.b_scale, 0.0);
    const double ne = A.rho[cell] * A.rho_scale / A.mpe;
    const double evpa = b_code > 0.0 ? atan2(b3, b2) : 0.0;
    const double th = pymax(A.theta_e[cell], 1.0e-8);

    double acc[kCoeffs];
#pragma unroll
    for (int k = 0; k < kCoeffs; ++k) acc[k] = 0.0;

    if (ne > 0.0 && b > 0.0 && A.nu > 0.0) {
        // thermal part (synchrotron.py:240-266)
        const double gmax = pymax(10.0, 1.0 + 45.0 * th);
        const double l0 = log10(1.0e-5), l1 = log10(gmax - 1.0);
        const int ng = A.th_ng;
        const double step = (l1 - l0) / double(ng - 1);
        const double tq = pymax(th, 1.0e-12);
        const double norm = th * bessel_k(2.0, 1.0 / tq);
        const double norm_den = pymax(norm, 1.0e-300);
        auto gamma_at = [&](int k) {   // 1 + geomspace(1e-5, gmax - 1, ng)[k], as 10^y = exp(y ln 10)
#pragma clang fp contract(off)
            const double y = double(k) * step + l0;
            const double gs = k == 0 ? 1.0e-5 : (k == ng - 1 ? gmax - 1.0 : exp(y * 2.302585092994045684));
            return 1.0 + gs;
        };
        const double neg_inv_tq = -1.0 / tq, inv_norm = 1.0 / norm_den;
        auto mj = [&](double g) {   // unnormalised Maxwell-Juttner density (synchrotron.py:192-196)
            const double gm = g > 1.0 ? g : 1.0;
            const double beta = sqrt(fmax(1.0 - 1.0 / (gm * gm), 0.0));
            return g * g * beta * exp(g * neg_inv_tq) * inv_norm;
        };
        // trapezoid weight of gamma_k: (gamma_{k+1} - gamma_{k-1}) / 2, one-sided at the ends
        double integ = 0.0;
        {
            double gp = 0.0, g = gamma_at(0);
            for (int k = 0; k < ng; ++k) {
                const double gn = k + 1 < ng ? gamma_at(k + 1) : g;
                const double w = ((k > 0 ? g - gp : 0.0) + (k + 1 < ng ? gn - g : 0.0)) / 2.0;
                integ += w * mj(g);
                gp = g;
                g = gn;
            }
        }
        const double inv_integ = 1.0 / pymax(integ, 1.0e-300);
        double tg_prev = 0.0, tg = gamma_at(0);
        auto thermal_dist = [&](int k, double &lg, double &wf) {
            const double gn = k + 1 < ng ? gamma_at(k + 1) : tg;
            const double w = ((k > 0 ? tg - tg_prev : 0.0) + (k + 1 < ng ? gn - tg : 0.0)) / 2.0;
            lg = log(tg);
            wf = w * (mj(tg) * inv_integ);
            tg_prev = tg;
            tg = gn;
        };
        double j_i, j_q;
        emissivities(t, A, kOffThSin, kOffThHw, A.th_nodes, ng, ne, b, thermal_dist, j_i, j_q);
        const double bnu = planck_nu(A, th);
        const double a_i = j_i / pymax(bnu, 1.0e-300), a_q = j_q / pymax(bnu, 1.0e-300);
        const double rel = pymax(1.0 / (th * th + 1.0), 1.0e-8);
        const double b_par = b * 0.0, b_perp = fabs(b) * 1.0;
        double c[kCoeffs] = {};
        c[0] = pymax(j_i, 0.0);
        c[1] = pymax(j_q, 0.0);
        c[4] = pymax(a_i, 0.0);
        c[5] = pymax(a_q, 0.0);
        c[8] = (A.th_frs * A.k_v) * (ne * b_par / pymax(A.nu2, 1.0e-300) * rel);
        c[9] = (A.th_fcs * A.k_q) * (ne * (b_perp * b_perp) / pymax(A.nu3, 1.0e-300) * pymax(th, 1.0e-8));
        add_rotated(acc, c, evpa);

        // power-law part (synchrotron.py:284-312)
        if (A.nt_frac > 0.0) {
            const double sne = ne * A.nt_frac;
            const int go = kOffNtGamma, gn = A.nt_ng;
            double p_i = 0.0, p_q = 0.0;
            auto powerlaw_dist = [&](int k, double &lg, double &wf) {
                lg = t[go + k];
                wf = t[go + gn + k];
            };
            if (sne > 0.0) emissivities(t, A, kOffNtSin, kOffNtHw, A.nt_nodes, gn, sne, b, powerlaw_dist, p_i, p_q);
            const double cg = pymax(A.nt_gmin, sqrt(pymax(A.nu / pymax(4.2e6 * b_perp, 1.0e-300), 1.0)));
            const double energy = cg * kMe * kC * kC;
            const double pa_i = A.nt_abs * p_i / pymax(energy * A.nu * A.nu / (kC * kC), 1.0e-300);
            const double pq = pymin(p_q, A.nt_pol * p_i);
            const double pa_q = pymin(pa_i * A.nt_pol, pa_i);
            double d[kCoeffs] = {};
            d[0] = pymax(p_i, 0.0);
            d[1] = pymax(pq, 0.0);
            d[4] = pymax(pa_i, 0.0);
            d[5] = pymax(pa_q, 0.0);
            d[8] = (A.nt_fs * A.k_v) * (sne * b_par / pymax(A.nu2, 1.0e-300) / A.nt_gmin2);
            d[9] = (A.nt_fs * A.k_q) * (sne * (b_perp * b_perp) / pymax(A.nu3, 1.0e-300) * A.nt_log_ratio);
            add_rotated(acc, d, evpa);
        }
    }

    // cast to the brick dtype, then nan_to_num (coefficient_bricks.py:139-150)
    if (A.out_dtype == BH_STORE_F64) {
        double *o = static_cast<double *>(A.out) + cell * kCoeffs;
#pragma unroll
        for (int k = 0; k < kCoeffs; ++k) {
            const double v = acc[k];
            o[k] = __builtin_isnan(v) ? 0.0 : (__builtin_isinf(v) ? (v > 0 ? std::numeric_limits<double>::max() : 0.0) : v);
        }
    } else if (A.out_dtype == BH_STORE_F32) {
        float *o = static_cast<float *>(A.out) + cell * kCoeffs;
#pragma unroll
        for (int k = 0; k < kCoeffs; ++k) {
            const float v = float(acc[k]);
            o[k] = __builtin_isnan(v) ? 0.f : (__builtin_isinf(v) ? (v > 0 ? std::numeric_limits<float>::max() : 0.f) : v);
        }
    } else {
        _Float16 *o = static_cast<_Float16 *>(A.out) + cell * kCoeffs;
#pragma unroll
        for (int k = 0; k < kCoeffs; ++k) {
            const _Float16 v = to_half(acc[k]);
            const float vf = float(v);
            o[k] = __builtin_isnan(vf) ? _Float16(0.f) : (__builtin_isinf(vf) ? (vf > 0 ? _Float16(65504.f) : _Float16(0.f)) : v);
        }
    }
}

// Tables and scalar arguments for one desc (everything but the pointers and the cell count).
static int prepare(const bh_synchrotron_desc *d, int out_dtype, std::vector<double> &tab, SynArgs &A) {
    if (!d) return set_error(BH_ERR_VALUE, "desc is NULL");
    if (out_dtype != BH_STORE_F32 && out_dtype != BH_STORE_F16 && out_dtype != BH_STORE_F64)
        return set_error(BH_ERR_VALUE, "dtype must be float32, float16 or float64");
    const double scal[] = {d->nu_hz, d->rho_cgs_per_code, d->b_gauss_per_code, d->mean_molecular_weight_per_electron,
                           d->th_faraday_rotation_scale, d->th_faraday_conversion_scale, d->nt_p, d->nt_gamma_min,
                           d->nt_gamma_max, d->nt_fraction, d->nt_absorption_scale, d->nt_faraday_scale};
    for (double v : scal)
        if (!std::isfinite(v)) return set_error(BH_ERR_VALUE, "synchrotron parameters must be finite");
    if (d->th_gamma_grid_size < 2 || d->th_gamma_grid_size > kMaxGamma || d->nt_gamma_grid_size < 2 ||
        d->nt_gamma_grid_size > kMaxGamma)
        return set_error(BH_ERR_VALUE, "gamma_grid_size must be in [2, %d]", kMaxGamma);
    if (d->th_pitch_nodes < 1 || d->th_pitch_nodes > kMaxNodes || d->nt_pitch_nodes < 1 || d->nt_pitch_nodes > kMaxNodes)
        return set_error(BH_ERR_VALUE, "pitch_nodes must be in [1, %d]", kMaxNodes);
    if (!(d->nt_gamma_min > 0.0) || !(d->nt_gamma_max > 0.0))
        return set_error(BH_ERR_VALUE, "gamma_min and gamma_max must be positive");

    const int ng = d->nt_gamma_grid_size;
    tab.assign(size_t(kOffNtGamma + 2 * ng), 0.0);
    // kernel table (synchrotron.py:158-163): F by the closed form, G = x K_{2/3}(x)
    const double la = std::log10(1.0e-8), lb = std::log10(1.0e3);
    const double lstep = (lb - la) / double(kTable - 1);
    double lxs[kTable], lfs[kTable], lgs[kTable];
    for (int i = 0; i < kTable; ++i) {
        double x = i == 0 ? 1.0e-8 : (i == kTable - 1 ? 1.0e3 : std::pow(10.0, double(i) * lstep + la));
        const double xx = std::fmin(std::fmax(x, 1.0e-12), 1.0e3);
        double F = 1.25 * std::pow(xx, 1.0 / 3.0) * std::exp(-xx) * std::pow(648.0 + xx * xx, 1.0 / 12.0);
        double G = xx * bessel_k(2.0 / 3.0, xx);
        if (!std::isfinite(F)) F = 0.0;
        if (!std::isfinite(G)) G = 0.0;
        lxs[i] = std::log(x);
        lfs[i] = std::log(std::fmax(F, 1.0e-300));
        lgs[i] = std::log(std::fmax(G, 1.0e-300));
    }
    for (int i = 0; i + 1 < kTable; ++i) {
        const double dx = lxs[i + 1] - lxs[i];
        double *e = tab.data() + kSeg * i;
        e[0] = lxs[i];
        e[1] = lxs[i + 1];
        e[2] = lfs[i];
        e[3] = (lfs[i + 1] - lfs[i]) / dx;
        e[4] = lgs[i];
        e[5] = (lgs[i + 1] - lgs[i]) / dx;
    }
    // Gauss-Legendre pitch nodes (synchrotron.py:219): sin(alpha) and 0.5 w of the negative-mu half
    double xs[kMaxNodes], ws[kMaxNodes];
    gauss_legendre(d->th_pitch_nodes, xs, ws);
    for (int u = 0; u < (d->th_pitch_nodes + 1) / 2; ++u) {
        tab[kOffThSin + u] = std::sqrt(std::fmax(0.0, 1.0 - xs[u] * xs[u]));
        tab[kOffThHw + u] = 0.5 * ws[u];
    }
    gauss_legendre(d->nt_pitch_nodes, xs, ws);
    for (int u = 0; u < (d->nt_pitch_nodes + 1) / 2; ++u) {
        tab[kOffNtSin + u] = std::sqrt(std::fmax(0.0, 1.0 - xs[u] * xs[u]));
        tab[kOffNtHw + u] = 0.5 * ws[u];
    }
    // power-law grid and its normalised distribution (synchrotron.py:202-207, 290-291), stored as
    // log gamma_k and w_k f_k with w_k the trapezoid weight of gamma_k
    std::vector<double> gam(ng), fpl(ng);
    const double ga = std::log10(d->nt_gamma_min), gb = std::log10(d->nt_gamma_max);
    const double gstep = (gb - ga) / double(ng - 1);
    for (int k = 0; k < ng; ++k) {
        gam[k] = k == 0 ? d->nt_gamma_min : (k == ng - 1 ? d->nt_gamma_max : std::pow(10.0, double(k) * gstep + ga));
        fpl[k] = (gam[k] >= d->nt_gamma_min && gam[k] <= d->nt_gamma_max) ? std::pow(gam[k], -d->nt_p) : 0.0;
    }
    double integ = 0.0;
    for (int k = 1; k < ng; ++k) integ += (gam[k] - gam[k - 1]) * (fpl[k] + fpl[k - 1]) / 2.0;
    double *lgam = tab.data() + kOffNtGamma, *wf = lgam + ng;
    for (int k = 0; k < ng; ++k) {
        const double w = ((k > 0 ? gam[k] - gam[k - 1] : 0.0) + (k + 1 < ng ? gam[k + 1] - gam[k] : 0.0)) / 2.0;
        lgam[k] = std::log(gam[k]);
        wf[k] = w * (fpl[k] / std::fmax(integ, 1.0e-300));
    }

    std::memset(&A, 0, sizeof(A));
    A.out_dtype = out_dtype;
    A.table_words = int(tab.size());
    A.nu = d->nu_hz;
    A.log_nu = d->nu_hz > 0.0 ? std::log(d->nu_hz) : 0.0;
    A.nu2 = d->nu_hz * d->nu_hz;
    A.nu3 = std::pow(d->nu_hz, 3.0);
    A.rho_scale = d->rho_cgs_per_code;
    A.b_scale = d->b_gauss_per_code;
    A.mpe = std::fmax(d->mean_molecular_weight_per_electron * kMp, 1.0e-300);
    A.th_ng = d->th_gamma_grid_size;
    A.th_nodes = d->th_pitch_nodes;
    A.nt_ng = ng;
    A.nt_nodes = d->nt_pitch_nodes;
    A.th_frs = d->th_faraday_rotation_scale;
    A.th_fcs = d->th_faraday_conversion_scale;
    A.nt_frac = d->nt_fraction;
    A.nt_gmin = d->nt_gamma_min;
    A.nt_abs = d->nt_absorption_scale;
    A.nt_pol = std::fmin((d->nt_p + 1.0) / (d->nt_p + 7.0 / 3.0), 0.85);
    A.nt_fs = d->nt_faraday_scale;
    A.nt_log_ratio = std::log(std::fmax(d->nt_gamma_max / d->nt_gamma_min, 1.0));
    A.nt_gmin2 = std::fmax(std::pow(d->nt_gamma_min, 2.0), 1.0);
    A.k_pref = std::sqrt(3.0) * std::pow(kE, 3.0);
    A.k_mc2 = kMe * std::pow(kC, 2.0);
    A.k_nub = 2.0 * kPi * kMe * kC;
    A.k_v = std::pow(kE, 3.0) / (kPi * std::pow(kMe, 2.0) * std::pow(kC, 2.0));
    A.k_q = std::pow(kE, 4.0) / (4.0 * std::pow(kPi, 2.0) * std::pow(kMe, 3.0) * std::pow(kC, 3.0));
    A.four_pi = 4.0 * kPi;
    A.lx_first = lxs[0];
    A.lx_last = lxs[kTable - 1];
    A.lF_top = lfs[kTable - 1];
    A.lG_top = lgs[kTable - 1];
    A.inv_dlx = double(kTable - 1) / (A.lx_last - A.lx_first);
    return BH_OK;
}

static size_t out_elem(int dtype) { return dtype == BH_STORE_F64 ? 8 : (dtype == BH_STORE_F32 ? 4 : 2); }

static int launch(bh_ctx *ctx, SynArgs &A, const std::vector<double> &tab) {
    const size_t bytes = tab.size() * sizeof(double);
    if (ctx->synchro_cap < bytes) {
        if (ctx->d_synchro) BH_HIP(hipFree(ctx->d_synchro));
        ctx->d_synchro = nullptr;
        ctx->synchro_cap = 0;
        BH_HIP(hipMalloc(&ctx->d_synchro, bytes));
        ctx->synchro_cap = bytes;
    }
    ctx->h_synchro = tab;   // the copy source outlives the asynchronous copy
    BH_HIP(hipMemcpyAsync(ctx->d_synchro, ctx->h_synchro.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    A.tables = static_cast<const double *>(ctx->d_synchro);
    const int threads = 256;
    const int64_t blocks = (A.n + threads - 1) / threads;
    begin_timing(ctx);
    hipLaunchKernelGGL(synchro_kernel, dim3(unsigned(blocks)), dim3(threads), bytes, ctx->stream, A);
    BH_HIP(hipGetLastError());
    end_timing(ctx, 1);
    return BH_OK;
}

}  // namespace syn
}  // namespace bh

using namespace bh;

extern "C" {

int bh_bessel_k(double nu, double x, double *out) {
    clear_error();
    if (!out) return set_error(BH_ERR_VALUE, "out is NULL");
    if (!(nu >= 0.0) || !std::isfinite(nu)) return set_error(BH_ERR_VALUE, "nu must be finite and >= 0");
    if (!(x > 0.0) || !std::isfinite(x)) return set_error(BH_ERR_VALUE, "x must be finite and > 0");
    *out = syn::bessel_k(nu, x);
    return BH_OK;
}

int bh_gauss_legendre(int32_t n, double *x, double *w) {
    clear_error();
    if (n < 1 || n > 64) return set_error(BH_ERR_VALUE, "n must be in [1, 64]");
    if (!x || !w) return set_error(BH_ERR_VALUE, "x and w must not be NULL");
    syn::gauss_legendre(n, x, w);
    return BH_OK;
}

int bh_precompute_coefficient_bricks_device(bh_ctx *ctx, const double *d_rho, const double *d_theta_e,
                                            const double *d_b_con, int64_t n_cells, const bh_synchrotron_desc *desc,
                                            int32_t out_dtype, void *d_out) {
    clear_error();
    int st = ctx_check(ctx);
    if (st) return st;
    if (n_cells < 0) return set_error(BH_ERR_VALUE, "n_cells must be non-negative");
    if (n_cells > (int64_t(1) << 31) * 256) return set_error(BH_ERR_VALUE, "too many cells");
    std::vector<double> tab;
    syn::SynArgs A;
    if ((st = syn::prepare(desc, out_dtype, tab, A))) return st;
    if (n_cells == 0) return BH_OK;
    if (!d_rho || !d_theta_e || !d_b_con || !d_out) return set_error(BH_ERR_VALUE, "buffers must not be NULL");
    A.rho = d_rho;
    A.theta_e = d_theta_e;
    A.b_con = d_b_con;
    A.n = n_cells;
    A.out = d_out;
    return syn::launch(ctx, A, tab);
}

int bh_precompute_coefficient_bricks(const double *rho, const double *theta_e, const double *b_con, int64_t n_cells,
                                     const bh_synchrotron_desc *desc, int32_t out_dtype, void *out) {
    clear_error();
    if (n_cells < 0) return set_error(BH_ERR_VALUE, "n_cells must be non-negative");
    std::vector<double> tab;
    syn::SynArgs A;
    int st;
    if ((st = syn::prepare(desc, out_dtype, tab, A))) return st;
    if (n_cells == 0) return BH_OK;
    if (!rho || !theta_e || !b_con || !out) return set_error(BH_ERR_VALUE, "buffers must not be NULL");
    std::lock_guard<std::mutex> lock(default_ctx_mutex());
    bh_ctx *ctx = default_ctx(&st);
    if (!ctx) return st;
    const size_t n = size_t(n_cells);
    auto rnd = [](size_t v) { return (v + 255) & ~size_t(255); };
    const size_t out_bytes = n * kCoeffs * syn::out_elem(out_dtype);
    void *buf = nullptr;
    if ((st = staging(rnd(n * 8) * 2 + rnd(n * 32) + rnd(out_bytes), &buf))) return st;
    char *base = static_cast<char *>(buf);
    double *drho = reinterpret_cast<double *>(base);
    double *dth = reinterpret_cast<double *>(base + rnd(n * 8));
    double *db = reinterpret_cast<double *>(base + 2 * rnd(n * 8));
    void *dout = base + 2 * rnd(n * 8) + rnd(n * 32);
    BH_HIP(hipMemcpyAsync(drho, rho, n * 8, hipMemcpyHostToDevice, ctx->stream));
    BH_HIP(hipMemcpyAsync(dth, theta_e, n * 8, hipMemcpyHostToDevice, ctx->stream));
    BH_HIP(hipMemcpyAsync(db, b_con, n * 32, hipMemcpyHostToDevice, ctx->stream));
    A.rho = drho;
    A.theta_e = dth;
    A.b_con = db;
    A.n = n_cells;
    A.out = dout;
    if ((st = syn::launch(ctx, A, tab))) return st;
    BH_HIP(hipMemcpyAsync(out, dout, out_bytes, hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    return BH_OK;
}

}  // extern "C"
