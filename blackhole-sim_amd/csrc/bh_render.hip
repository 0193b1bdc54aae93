// Fused per-ray Kerr Stokes render: RK2 midpoint geodesic + SoA brick sampler +
// RK2 Stokes IQUV transfer, one lane per ray, persistent ray queues.
//
// Behaviour follows the reference's CPU emulation of the GPU hot loop
// (python/blackhole_sim/accelerated_renderer.py:100-158) with the RK2 midpoint
// scheme and termination tests of native/rocm/kerr_stokes_kernel.hip:13, 20
// and python/blackhole_sim/kerr.py:357-377; the camera launch is kerr.py:188-233.
//
// Design (gfx950):
//  * Hamiltonian in separated Carter form: with E = -p_t, L = p_phi,
//    P = (r^2 + a^2) E - a L and T = L^2/sin^2 - 2 a E L + a^2 E^2 sin^2,
//    2 Sigma H = Delta p_r^2 + p_theta^2 + T - P^2/Delta.  Its gradients are
//    algebraically identical to the reference's -1/2 p (d g^{mu nu}) p
//    (kerr.py:236-303) at a fraction of the operations: one sincos and three
//    reciprocals per right-hand side.
//  * The brick is SoA [11][nr][ntheta][nphi] (f32, bf16 or f16).  Loads are raw
//    buffer loads: the eight corner offsets are computed once per sample and the
//    channel plane offset rides in the scalar soffset operand, so a channel costs
//    one load instruction and no address arithmetic.  Out-of-range offsets read 0.
//  * The r/theta/phi grids are staged in LDS.  Brackets take an O(1) index guess
//    (log2 for geometric r grids, affine otherwise) and correct it against the
//    grid with compares, which reproduces searchsorted(side='right') - 1 exactly.
//  * Ray scheduling: persistent waves pull rays from a few ray queues (block b
//    serves queue b % queues and steals from the others when it runs dry).  Ray
//    ids are ordered in 8x8 pixel tiles so a wave's lanes trace neighbouring
//    rays and share cache lines.  Every steps_per_sync steps a wave ballots its
//    idle lanes, one lane reserves that many ids with a single atomicAdd, and the
//    idle lanes take consecutive ids by mbcnt rank (wave-level compaction).

#include "bh_internal.h"
#include "bh_stokes_exact.h"

#include <cmath>
#include <algorithm>
#include <cstring>

namespace bh {

constexpr int kMaxGrid = 2048;   // r + theta + phi grid values staged in LDS
constexpr double kIslandTheta = 0.1;    // rad from a pole: FP64 geodesic steps in the mixed path
constexpr double kIslandPtheta = 0.25;  // |p_theta| below this: near a theta turning point

struct CamConst {
    double tan_x, tan_y, cos_roll, sin_roll;
    int roll;
    double e_t0, e_t3, e_r1, e_th2, e_ph3;   // ZAMO tetrad non-zeros (kerr.py:144-148)
    double g00, g03, g11, g22, g33;          // g_{mu nu} at the camera (kerr.py:67-83)
    double r, th, ph;
};

struct RenderArgs;
template <typename R>
struct HotParams {
    R step, r_plus, r_cap, r_esc, L, tau_max;
    int max_steps, adaptive;
};

struct RenderArgs {
    CamConst cam;
    double a, r_plus, r_cap, step, r_esc, L, tau_max;
    double r_island, th_island, pth_island;   // FP64 islands of the mixed-precision path
    float step_f, r_plus_f, r_cap_f, r_esc_f, L_f, tau_max_f;   // FP32 copies for the hot loop
    int W, H, max_steps, adaptive;
    const void *coeffs;
    int nr, nt, np;
    unsigned plane_bytes, total_bytes;
    unsigned brick_base;        // byte offset of this brick inside the coefficient buffer (batches merge bricks), low 32 bits
    int grid_off;               // offset of this brick's r grid in the LDS grid table
    long long out_off;          // first pixel of this render in the batch output arrays
    const float *grids;
    GridHint hr, ht, hp;
    void *iquv;
    int *steps;
    uint8_t *status;
    unsigned *counters;
    unsigned long long *work;   // [0] RK2 steps, [1] in-brick samples, [2] lane slots, [3] split lone rays, [4] time-outs
    const int *tile_order;      // queue slot -> tile, or null for row-major
    const int *ray_list;        // list mode: queue slot -> pixel index (tile_order unused)
    const unsigned *n_list;     // list mode: number of listed pixels (device, read at kernel start)
    int tiles_x, n_ids, queues, q_len;
    int steps_per_sync, refill_min;
    int coop_steps;             // lone steps before a lone ray asks an idle wave for help (0: never)
    int lerp32;                 // FP64 render: sample as the reference renderer does on a float32 brick
    unsigned brick_base_hi;     // high 32 bits of brick_base (merged corner bricks of 4 GiB and more)
};

template <typename R>
__device__ __forceinline__ HotParams<R> hot_params(const RenderArgs &A) {
    if constexpr (sizeof(R) == 4)
        return {A.step_f, A.r_plus_f, A.r_cap_f, A.r_esc_f, A.L_f, A.tau_max_f, A.max_steps, A.adaptive};
    else return {A.step, A.r_plus, A.r_cap, A.r_esc, A.L, A.tau_max, A.max_steps, A.adaptive};
}

// A copy of a wave-uniform 32-bit value that lives in a VGPR.  The step loop's constants would
// otherwise compete for SGPRs with the launch arguments and the branch masks, be spilled to VGPR
// lanes and come back through v_readlane on every step.
__device__ __forceinline__ float in_vgpr(float x) {
    float y;
    asm("v_mov_b32 %0, %1" : "=v"(y) : "s"(x));
    return y;
}
__device__ __forceinline__ int in_vgpr(int x) {
    int y;
    asm("v_mov_b32 %0, %1" : "=v"(y) : "s"(x));
    return y;
}

// The step loop's constants; FP32 ones are moved to VGPRs (FP64 ones stay where the compiler puts them).
template <typename R>
__device__ __forceinline__ HotParams<R> loop_params(const RenderArgs &A) {
    HotParams<R> P = hot_params<R>(A);
    if constexpr (sizeof(R) == 4) {
        P.step = in_vgpr(P.step);
        P.r_plus = in_vgpr(P.r_plus);
        P.r_cap = in_vgpr(P.r_cap);
        P.r_esc = in_vgpr(P.r_esc);
        P.L = in_vgpr(P.L);
        P.tau_max = in_vgpr(P.tau_max);
        P.max_steps = in_vgpr(P.max_steps);
    }
    return P;
}

// ---- precision traits ------------------------------------------------------

template <typename R>
struct Mth;

template <>
struct Mth<float> {
    static __device__ __forceinline__ void sincos(float x, float &s, float &c) {
        s = __sinf(x);
        c = __cosf(x);
    }
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float abs(float x) { return __builtin_fabsf(x); }
    static __device__ __forceinline__ float two_pi() { return 6.28318530717958648f; }
    static __device__ __forceinline__ float pi_minus_eps() { return float(3.14159265358979323846 - kEpsTheta); }
    static __device__ __forceinline__ float wrap(float phi, float base) {
        float p = phi - base;
        p = __builtin_fmaf(-two_pi(), floorf(p * 0.159154943091895336f), p);
        if (p < 0.f) p += two_pi();
        if (p >= two_pi()) p -= two_pi();
        return p + base;
    }
};

template <>
struct Mth<double> {
    static __device__ __forceinline__ void sincos(double x, double &s, double &c) { ::sincos(x, &s, &c); }
    static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
    static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
    static __device__ __forceinline__ double abs(double x) { return __builtin_fabs(x); }
    static __device__ __forceinline__ double two_pi() { return kTwoPi; }
    static __device__ __forceinline__ double pi_minus_eps() { return 3.14159265358979323846 - kEpsTheta; }
    // Python float % (native_kernels.py:133-134): exact fmod, shifted when negative.
    static __device__ __forceinline__ double wrap(double phi, double base) {
        double m = fmod(phi - base, kTwoPi);
        if (m < 0.0) m += kTwoPi;
        return m + base;
    }
};

// Python max(a, b): b if b > a else a (NaN in a survives).
template <typename R>
__device__ __forceinline__ R pymax(R a, R b) { return (b > a) ? b : a; }

// Midpoint of a segment, rounded on its own: contracted, the halving could fuse into whichever
// subtraction consumes it (the phi wrap), and whether it does depends on the code around it, so two
// paths sampling the same segment could round it differently.
template <typename R>
__device__ __forceinline__ R midpoint(R a, R b) {
#pragma clang fp contract(off)
    return R(0.5) * (a + b);
}

// np.clip to [eps, pi - eps], NaN-propagating (kerr.py:35-36).
template <typename R>
__device__ __forceinline__ R clamp_theta(R th) {
    const R lo = R(kEpsTheta), hi = Mth<R>::pi_minus_eps();
    return th < lo ? lo : (th > hi ? hi : th);
}

template <typename R>
struct RayK {
    R a, a2, E, L, L2, a2E2, twoaEL, aE, aL, fourE;
};

// FP64 math for the islands of the mixed kernels, cheaper than the library's and still within a few ulp
// (the FP64 reference kernels keep the library's).
// sin and cos of x in [0, pi]: x = k pi/2 + t with k = rint(2x/pi) in {0, 1, 2}, pi/2 taken as a 33-bit head
// (so k * head is exact) plus the double nearest the rest, |t| <= pi/4 + 2^-50; then the Taylor series of
// sin t to t^17 and of cos t to t^18 (truncation below 1e-19 for |t| <= pi/4), coefficients +-1/n!.
__device__ __forceinline__ void island_sincos(double x, double &s, double &c) {
    const double k = __builtin_rint(x * 0.63661977236758134);
    double t = __builtin_fma(-k, 0x1.921fb544p+0, x);
    t = __builtin_fma(-k, 0x1.0b4611a626331p-34, t);
    const double t2 = t * t;
    double p = 2.8114572543455206e-15;           // 1/17!
    p = __builtin_fma(p, t2, -7.647163731819816e-13);   // -1/15!
    p = __builtin_fma(p, t2, 1.6059043836821613e-10);   // 1/13!
    p = __builtin_fma(p, t2, -2.505210838544172e-08);   // -1/11!
    p = __builtin_fma(p, t2, 2.7557319223985893e-06);   // 1/9!
    p = __builtin_fma(p, t2, -0.0001984126984126984);   // -1/7!
    p = __builtin_fma(p, t2, 0.008333333333333333);     // 1/5!
    p = __builtin_fma(p, t2, -0.16666666666666666);     // -1/3!
    const double st = __builtin_fma(t * t2, p, t);
    double q = -1.5619206968586225e-16;          // -1/18!
    q = __builtin_fma(q, t2, 4.779477332387385e-14);    // 1/16!
    q = __builtin_fma(q, t2, -1.1470745597729725e-11);  // -1/14!
    q = __builtin_fma(q, t2, 2.08767569878681e-09);     // 1/12!
    q = __builtin_fma(q, t2, -2.755731922398589e-07);   // -1/10!
    q = __builtin_fma(q, t2, 2.48015873015873e-05);     // 1/8!
    q = __builtin_fma(q, t2, -0.001388888888888889);    // -1/6!
    q = __builtin_fma(q, t2, 0.041666666666666664);     // 1/4!
    q = __builtin_fma(q, t2, -0.5);                     // -1/2!
    const double ct = __builtin_fma(t2, q, 1.0);
    // quadrant k: (sin, cos) = (st, ct), (ct, -st), (-st, -ct)
    s = k == 0.0 ? st : (k == 1.0 ? ct : -st);
    c = k == 0.0 ? ct : (k == 1.0 ? -st : -ct);
}

// 1/x from the hardware estimate and two Newton steps (each squares the relative error).
__device__ __forceinline__ double island_rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}

// d/dlambda (r, theta, phi, p_r, p_theta) of the Kerr null Hamiltonian, as the FP64 reference kernels take it:
// every step in the ray's own lane, library math, the compiler's multiply-add fusion (the form their parity
// with the oracle is measured on).
template <typename R>
__device__ __forceinline__ void kerr_rhs_free(const RayK<R> &k, R r, R th_in, R pr, R pth, R &dr, R &dth, R &dph, R &dpr,
                                              R &dpth) {
    static_assert(sizeof(R) == 8, "the FP32 kernels and the FP64 islands take kerr_rhs_pinned");
    const R th = clamp_theta(th_in);
    R s, c;
    Mth<R>::sincos(th, s, c);
    R s2 = s * s;
    s2 = (s2 < R(kEpsTheta * kEpsTheta)) ? R(kEpsTheta * kEpsTheta) : s2;   // kerr.py:93 floor
    const R r2 = r * r;
    const R r2a2 = r2 + k.a2;
    const R sig = r2 + k.a2 * (c * c);
    const R dlt = r2a2 - R(2) * r;
    const R P = r2a2 * k.E - k.aL;
    const R inv_s2 = Mth<R>::rcp(s2), inv_sig = Mth<R>::rcp(sig), inv_dlt = Mth<R>::rcp(dlt);
    const R Wd = (P * P) * inv_dlt;
    const R T = k.L2 * inv_s2 - k.twoaEL + k.a2E2 * s2;
    const R pr2 = pr * pr;
    const R N = dlt * pr2 + pth * pth + T - Wd;   // 2 Sigma H
    dr = dlt * pr * inv_sig;
    dth = pth * inv_sig;
    dph = (k.L * inv_s2 - k.aE + k.a * P * inv_dlt) * inv_sig;
    const R dlt_p = R(2) * r - R(2);
    const R dNr = dlt_p * pr2 - k.fourE * r * P * inv_dlt + Wd * dlt_p * inv_dlt;
    dpr = R(-0.5) * inv_sig * (dNr - R(2) * r * N * inv_sig);
    const R sc = s * c;
    const R dNth = R(2) * sc * (k.a2E2 - k.L2 * inv_s2 * inv_s2);
    dpth = R(-0.5) * inv_sig * (dNth + R(2) * k.a2 * sc * N * inv_sig);
}

// The same right-hand side in the FP32 kernels and, ISLAND, in the FP64 islands of a mixed kernel (with the
// island math above).  Its arithmetic, and that of the steps built on it, is written with explicit fused
// multiply-adds and compiled with contraction off, so that a step rounds the same in every code path that
// takes it (the ray's own lane, a lone ray's batch, the helper of a split lone ray): left to the compiler,
// the fusion of a * b + c depends on the code the step is inlined into.
template <typename R, bool ISLAND = false>
__device__ __forceinline__ void kerr_rhs_pinned(const RayK<R> &k, R r, R th_in, R pr, R pth, R &dr, R &dth, R &dph, R &dpr,
                                                R &dpth) {
#pragma clang fp contract(off)
    static_assert(sizeof(R) == 4 || ISLAND, "the FP64 reference kernels take kerr_rhs_free");
    const R th = clamp_theta(th_in);
    R s, c;
    if constexpr (ISLAND) island_sincos(th, s, c);
    else Mth<R>::sincos(th, s, c);
    R s2 = s * s;
    s2 = (s2 < R(kEpsTheta * kEpsTheta)) ? R(kEpsTheta * kEpsTheta) : s2;   // kerr.py:93 floor
    const R r2 = r * r;
    const R r2a2 = r2 + k.a2;
    const R sig = fma_hd(k.a2, c * c, r2);
    const R dlt = fma_hd(R(-2), r, r2a2);
    const R P = fma_hd(r2a2, k.E, -k.aL);
    // one reciprocal of Sigma * Delta * sin^2 serves all three (|product| stays well inside FP32 range)
    const R sd = sig * dlt;
    R inv;
    if constexpr (ISLAND) inv = island_rcp(sd * s2);
    else inv = Mth<R>::rcp(sd * s2);
    const R inv_s2 = sd * inv, inv_sig = (dlt * s2) * inv, inv_dlt = (sig * s2) * inv;
    const R Wd = (P * P) * inv_dlt;
    const R T = fma_hd(k.a2E2, s2, fma_hd(k.L2, inv_s2, -k.twoaEL));
    const R pr2 = pr * pr;
    const R N = fma_hd(dlt, pr2, fma_hd(pth, pth, T)) - Wd;   // 2 Sigma H
    dr = (dlt * pr) * inv_sig;
    dth = pth * inv_sig;
    dph = fma_hd(k.a * P, inv_dlt, fma_hd(k.L, inv_s2, -k.aE)) * inv_sig;
    const R dlt_p = fma_hd(R(2), r, R(-2));
    const R dNr = fma_hd(Wd * dlt_p, inv_dlt, fma_hd(dlt_p, pr2, -((k.fourE * r) * P) * inv_dlt));
    dpr = (R(-0.5) * inv_sig) * fma_hd(-((R(2) * r) * N), inv_sig, dNr);
    const R sc = s * c;
    const R dNth = (R(2) * sc) * fma_hd(-(k.L2 * inv_s2), inv_s2, k.a2E2);
    dpth = (R(-0.5) * inv_sig) * fma_hd(((R(2) * k.a2) * sc) * N, inv_sig, dNth);
}


// ---- brick sampler -----------------------------------------------------------

// Storage codes inside the kernels: BH_STORE_* in bits 0-1, plus kStoreWide for corner bricks of 2 GiB and
// more, whose records are read through 64-bit addresses; smaller bricks are read through a buffer
// resource (32-bit offsets, one cheaper address operation per sample), so they get kernels of their own.
constexpr int kStoreWide = 4;
__host__ __device__ constexpr int store_of(int s) { return s & 3; }
__host__ __device__ constexpr bool wide_of(int s) { return (s & kStoreWide) != 0; }

template <int STORE>
__device__ __forceinline__ float load_coeff(__amdgpu_buffer_rsrc_t rs, unsigned voff, unsigned soff) {
    if constexpr (store_of(STORE) == BH_STORE_F32) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, int(voff), int(soff), 0));
    } else if constexpr (store_of(STORE) == BH_STORE_BF16) {
        unsigned h = __builtin_amdgcn_raw_buffer_load_b16(rs, int(voff), int(soff), 0);
        return __builtin_bit_cast(float, h << 16);
    } else {
        unsigned short h = __builtin_amdgcn_raw_buffer_load_b16(rs, int(voff), int(soff), 0);
        return float(__builtin_bit_cast(_Float16, h));
    }
}

// Address of cell record `cell` (8 corners x 11 channels, corner-major) of the render's brick (wide bricks).
template <int STORE>
__device__ __forceinline__ const char *corner_ptr(const RenderArgs &A, unsigned cell) {
    constexpr size_t rec = 8u * kCoeffs * (store_of(STORE) == BH_STORE_F32 ? 4u : 2u);
    // the index stays 32-bit arithmetic (full-rate multiply-adds); only the byte address is 64-bit, one
    // multiply-add (folded, the whole index chain would become 64-bit multiply-adds)
    asm("" : "+v"(cell));
    const size_t base = (size_t(A.brick_base_hi) << 32) | A.brick_base;
    return static_cast<const char *>(A.coeffs) + base + size_t(cell) * rec;
}

// Index guess from the hint, clamped to [0, nmax]; NaN -> 0.  Only the r grid may be geometric
// (theta and phi hints are always linear), so only the r guess evaluates a logarithm.
template <bool MAYBE_LOG>
__device__ __forceinline__ int guess_index(float x, GridHint h, int nmax) {
    float u;
    if constexpr (MAYBE_LOG) u = h.log ? __builtin_amdgcn_logf(x) * h.scale + h.offset : x * h.scale + h.offset;
    else u = x * h.scale + h.offset;
    u = fminf(fmaxf(u, 0.f), float(nmax));
    return int(u);
}

// Exact bracket search from a guess i, for grids the hint does not fit: the i with
// (i == 0 || x >= g[i]) && (i == last || x < g[i + 1]), last = n - 2 for r/theta, n - 1 for phi.
template <typename R>
__device__ __forceinline__ int bracket_search(const R *g, int last, R x, int i) {
    while (i > 0 && x < g[i]) --i;
    while (i < last && x >= g[i + 1]) ++i;
    return i;
}

// Cell and weights of a point: native_kernels.py:137-146 (r, theta: searchsorted-right - 1
// clipped to [0, n-2], unit weight at the last grid value) and :149-161 (periodic phi with the
// seam cell [g[n-1], g[0] + 2 pi)).  The index comes from the hint; the six grid values that check
// it are read from LDS together, and only a point whose check fails searches.
template <typename R>
struct CellPos {
    int r0, t0, p0, p1;
    R wr, wt, wp;
};

// The first half of locate, which reads no grid: the r/theta range test, the wrapped phi and the
// three index guesses.
template <typename R>
struct CellGuess {
    R p;
    int ir, it, ip;
};

template <typename R>
__device__ __forceinline__ bool locate_guess(const RenderArgs &A, R x, R y, R z, CellGuess<R> &g) {
    const R r_first = R(A.hr.first), r_last = R(A.hr.last), t_first = R(A.ht.first), t_last = R(A.ht.last);
    if (x < r_first || x > r_last || y < t_first || y > t_last) return false;   // NaN passes, as in the reference
    g.p = Mth<R>::wrap(z, R(A.hp.first));
    g.ir = guess_index<true>(float(x), A.hr, A.nr - 2);
    g.it = guess_index<false>(float(y), A.ht, A.nt - 2);
    g.ip = guess_index<false>(float(g.p), A.hp, A.np - 1);
    return true;
}

// The second half: check the guesses against the grids (LDS), search where they miss, the seam cell
// and the weights.
template <typename R>
__device__ __forceinline__ void locate_finish(const RenderArgs &A, const R *gr, const R *gt, const R *gp, R x, R y,
                                              const CellGuess<R> &g, CellPos<R> &c) {
    const R r_last = R(A.hr.last), t_last = R(A.ht.last);
    const R p_first = R(A.hp.first);
    const R p = g.p;
    int ir = g.ir, it = g.it, ip = g.ip;
    R rl = gr[ir], rh = gr[ir + 1], tl = gt[it], th = gt[it + 1], pl = gp[ip], ph = gp[ip + 1];   // g[np] = +inf
    // bitwise, not short-circuit: all six reads are needed, so no branch may split them
    const bool bad = ((ir > 0) & (x < rl)) | ((ir < A.nr - 2) & (x >= rh)) | ((it > 0) & (y < tl)) |
                     ((it < A.nt - 2) & (y >= th)) | ((ip > 0) & (p < pl)) | ((ip < A.np - 1) & (p >= ph));
    if (bad) {
        ir = bracket_search(gr, A.nr - 2, x, ir);
        it = bracket_search(gt, A.nt - 2, y, it);
        ip = bracket_search(gp, A.np - 1, p, ip);
        rl = gr[ir];
        rh = gr[ir + 1];
        tl = gt[it];
        th = gt[it + 1];
        pl = gp[ip];
        ph = gp[ip + 1];
    }
    if (p < p_first) {   // wrap rounding: the seam cell
        ip = A.np - 1;
        pl = gp[ip];
    }
    const int jp = (ip + 1 == A.np) ? 0 : ip + 1;
    const R p_hi = (jp > ip) ? ph : p_first + Mth<R>::two_pi();
    const R pp = (p >= pl) ? p : p + Mth<R>::two_pi();
    c.r0 = ir;
    c.t0 = it;
    c.p0 = ip;
    c.p1 = jp;
    c.wr = (x == r_last) ? R(1) : Mth<R>::div(x - rl, rh - rl);
    c.wt = (y == t_last) ? R(1) : Mth<R>::div(y - tl, th - tl);
    c.wp = Mth<R>::div(pp - pl, p_hi - pl);
}

template <typename R>
__device__ __forceinline__ bool locate(const RenderArgs &A, const R *gr, const R *gt, const R *gp, R x, R y, R z,
                                       CellPos<R> &c) {
    CellGuess<R> g;
    if (!locate_guess(A, x, y, z, g)) return false;
    locate_finish(A, gr, gt, gp, x, y, g, c);
    return true;
}

// 1 - w for a trilinear weight, rounded on its own in FP32: contracted with the division that made w
// (which happens only when both sit in one block) it would round differently in the paths whose w
// comes from a register or from LDS, so every path sees the same weights.
template <typename R>
__device__ __forceinline__ R one_minus(R w) {
    if constexpr (sizeof(R) == 4) return __ocml_sub_rte_f32(1.0f, w);
    else return R(1) - w;
}

// Trilinear sample of all 11 channels at (r, theta, phi); false when outside r/theta.
// FP64 keeps the reference's nested (1-w)a + wb lerps (native_kernels.py:180-186);
// FP32 uses the equivalent eight-weight sum.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float half_lo(unsigned w, int store) {
    return store == BH_STORE_BF16 ? __builtin_bit_cast(float, w << 16)
                                  : float(__builtin_bit_cast(_Float16, (unsigned short)(w & 0xffffu)));
}
__device__ __forceinline__ float half_hi(unsigned w, int store) {
    return store == BH_STORE_BF16 ? __builtin_bit_cast(float, w & 0xffff0000u)
                                  : float(__builtin_bit_cast(_Float16, (unsigned short)(w >> 16)));
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// Two FP32 weights as bf16 pairs, a head and a tail: hi = bf16(w) (round to nearest even) and
// lo = bf16(w - hi), where w - hi is exact, so hi + lo holds w to ~2^-17 relative.
__device__ __forceinline__ void split_bf16_pair(float a, float b, bf16x2 &hi, bf16x2 &lo) {
    hi = __builtin_convertvector((f32x2){a, b}, bf16x2);
    const unsigned u = __builtin_bit_cast(unsigned, hi);
    lo = __builtin_convertvector((f32x2){a - __builtin_bit_cast(float, u << 16), b - __builtin_bit_cast(float, u & 0xffff0000u)},
                                 bf16x2);
}

template <typename R, int STORE, int LAYOUT>
__device__ __forceinline__ bool sample_brick(const RenderArgs &A, __amdgpu_buffer_rsrc_t rs, const R *gr, const R *gt,
                                             const R *gp, R x, R y, R z, R out[kCoeffs]) {
    CellPos<R> cp;
    if (!locate(A, gr, gt, gp, x, y, z, cp)) return false;
    const int r0 = cp.r0, t0 = cp.t0, p0 = cp.p0, p1 = cp.p1;
    const R wr = cp.wr, wt = cp.wt, wp = cp.wp;
    constexpr unsigned esz = store_of(STORE) == BH_STORE_F32 ? 4u : 2u;
    if constexpr (LAYOUT == BH_LAYOUT_CORNER) {
        // one cell record: corners (dr, dt, dp) in order dr*4 + dt*2 + dp, 11 channels each (bf16: channel-major)
        const unsigned cell = (unsigned(r0) * unsigned(A.nt - 1) + unsigned(t0)) * unsigned(A.np) + unsigned(p0);
        constexpr int kWords = 8 * kCoeffs * int(esz) / 4;   // 88 (f32) or 44 (16-bit)
        constexpr int kVec = kWords / 4;                      // 22 or 11 loads of 16 bytes
        unsigned w[kWords];
        const u32x4 *rec = nullptr;
        unsigned base = 0;
        if constexpr (wide_of(STORE)) rec = reinterpret_cast<const u32x4 *>(corner_ptr<STORE>(A, cell));
        else base = A.brick_base + cell * unsigned(8 * kCoeffs * esz);
#pragma unroll
        for (int j = 0; j < kVec; ++j) {
            u32x4 q;
            if constexpr (wide_of(STORE)) q = rec[j];
            else q = __builtin_amdgcn_raw_buffer_load_b128(rs, int(base + 16u * unsigned(j)), 0, 0);
            w[4 * j] = q.x;
            w[4 * j + 1] = q.y;
            w[4 * j + 2] = q.z;
            w[4 * j + 3] = q.w;
        }
        // all loads of the record leave before the first use, so a lone ray waits for one round trip
        __builtin_amdgcn_sched_barrier(0);
        const R ar = one_minus(wr), at = one_minus(wt), ap = one_minus(wp);
        const R q00 = ar * at, q01 = ar * wt, q10 = wr * at, q11 = wr * wt;
        const R w8[8] = {q00 * ap, q00 * wp, q01 * ap, q01 * wp, q10 * ap, q10 * wp, q11 * ap, q11 * wp};
        if constexpr (store_of(STORE) == BH_STORE_BF16 && sizeof(R) == 4) {
            // bf16 records are channel-major: load k holds channel k's 8 corners, corners 2m and 2m + 1 in
            // word m, so one v_dot2_f32_bf16 takes a corner pair against a pair of split weights (tails
            // first, then heads): 8 dot products per channel instead of 8 unpacks and 8 FMAs
            bf16x2 hi[4], lo[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) split_bf16_pair(w8[2 * m], w8[2 * m + 1], hi[m], lo[m]);
#pragma unroll
            for (int k = 0; k < kCoeffs; ++k) {
                float acc = 0.f;
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    acc = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, w[4 * k + m]), lo[m], acc, false);
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    acc = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, w[4 * k + m]), hi[m], acc, false);
                out[k] = acc;
            }
            return true;
        }
#pragma unroll
        for (int k = 0; k < kCoeffs; ++k) {
            R acc = R(0);
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int e = q * kCoeffs + k;
                float v;
                if constexpr (store_of(STORE) == BH_STORE_F32) v = __builtin_bit_cast(float, w[e]);
                else v = (e & 1) ? half_hi(w[e >> 1], store_of(STORE)) : half_lo(w[e >> 1], store_of(STORE));
                acc = (q == 0) ? w8[0] * R(v) : fma_hd(w8[q], R(v), acc);
            }
            out[k] = acc;
        }
        return true;
    }
    const unsigned row00 = A.brick_base / esz + (unsigned(r0) * unsigned(A.nt) + unsigned(t0)) * unsigned(A.np);
    const unsigned row01 = row00 + unsigned(A.np);
    const unsigned row10 = row00 + unsigned(A.nt) * unsigned(A.np);
    const unsigned row11 = row10 + unsigned(A.np);
    const unsigned o[8] = {(row00 + p0) * esz, (row00 + p1) * esz, (row01 + p0) * esz, (row01 + p1) * esz,
                           (row10 + p0) * esz, (row10 + p1) * esz, (row11 + p0) * esz, (row11 + p1) * esz};
    if constexpr (sizeof(R) == 8) {
        // FP64 keeps the reference's nested lerps; its loads are left to the scheduler (holding all 88
        // in flight would spill the FP64 kernel's registers)
        const R ap = R(1) - wp, at = R(1) - wt, ar = R(1) - wr;
#pragma unroll
        for (int k = 0; k < kCoeffs; ++k) {
            const unsigned so = unsigned(k) * A.plane_bytes;
            R v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = R(load_coeff<STORE>(rs, o[q], so));
            const R c00 = ap * v[0] + wp * v[1];
            const R c01 = ap * v[2] + wp * v[3];
            const R c10 = ap * v[4] + wp * v[5];
            const R c11 = ap * v[6] + wp * v[7];
            const R c0 = at * c00 + wt * c01;
            const R c1 = at * c10 + wt * c11;
            out[k] = ar * c0 + wr * c1;
        }
    } else {
        // all 88 loads leave before any arithmetic, so they are in flight together (one round trip)
        float raw[kCoeffs][8];
#pragma unroll
        for (int k = 0; k < kCoeffs; ++k) {
            const unsigned so = unsigned(k) * A.plane_bytes;
#pragma unroll
            for (int q = 0; q < 8; ++q) raw[k][q] = load_coeff<STORE>(rs, o[q], so);
        }
        __builtin_amdgcn_sched_barrier(0);
        const R ar = one_minus(wr), at = one_minus(wt), ap = one_minus(wp);
        const R q00 = ar * at, q01 = ar * wt, q10 = wr * at, q11 = wr * wt;
        const R w8[8] = {q00 * ap, q00 * wp, q01 * ap, q01 * wp, q10 * ap, q10 * wp, q11 * ap, q11 * wp};
#pragma unroll
        for (int k = 0; k < kCoeffs; ++k) {
            R acc = w8[0] * raw[k][0];
#pragma unroll
            for (int q = 1; q < 8; ++q) acc = fma_hd(w8[q], raw[k][q], acc);
            out[k] = acc;
        }
    }
    return true;
}

// The reference renderer's sample (accelerated_renderer.py:73-97 with the brackets of grmhd.py:229-257)
// as NumPy evaluates it on a float32 brick: its Python-float coordinates meet float32 grid scalars and
// float32 coefficient arrays, and NumPy casts such Python scalars to float32, so the range tests, the
// searchsorted brackets, the weights and all three lerp levels run in float32 arithmetic, each
// (1 - w) a + w b as two rounded products and a rounded sum; only the phi wrap and the seam shift
// p + 2 pi are Python (double) arithmetic.  Every float32 operation below is rounded on its own
// (contraction off: a fused multiply-add would round (1 - w) a + w b once instead of three times), so
// the eleven values equal NumPy's bit for bit.  FP64 render, float32 SoA brick.
__device__ __forceinline__ float lerp_numpy32(float a, float wa, float b, float wb) {
#pragma clang fp contract(off)
    return wa * a + wb * b;
}

__device__ __forceinline__ bool sample_brick_numpy32(const RenderArgs &A, __amdgpu_buffer_rsrc_t rs, const double *gr,
                                                     const double *gt, const double *gp, double x, double y, double z,
                                                     double out[kCoeffs]) {
#pragma clang fp contract(off)
    const float xf = float(x), yf = float(y);
    if (xf < A.hr.first || xf > A.hr.last || yf < A.ht.first || yf > A.ht.last) return false;   // NaN passes
    const double p = Mth<double>::wrap(z, double(A.hp.first));
    const float pf = float(p);
    // searchsorted(side='right') - 1 on float32 values (the double grids hold them exactly), clipped
    const int ir = bracket_search(gr, A.nr - 2, double(xf), guess_index<true>(xf, A.hr, A.nr - 2));
    const int it = bracket_search(gt, A.nt - 2, double(yf), guess_index<false>(yf, A.ht, A.nt - 2));
    int ip = bracket_search(gp, A.np - 1, double(pf), guess_index<false>(pf, A.hp, A.np - 1));
    if (pf < A.hp.first) ip = A.np - 1;   // searchsorted gave -1: the seam cell
    const int jp = (ip + 1 == A.np) ? 0 : ip + 1;
    const float rl = float(gr[ir]), rh = float(gr[ir + 1]), tl = float(gt[it]), th = float(gt[it + 1]);
    const float pl = float(gp[ip]);
    const float p_hi = (jp > ip) ? float(gp[jp]) : A.hp.first + 6.28318548f;   // grid[0] + 2 pi in float32
    const double pp = (pf >= pl) ? p : p + kTwoPi;
    const float wr = (xf == A.hr.last) ? 1.0f : (xf - rl) / (rh - rl);
    const float wt = (yf == A.ht.last) ? 1.0f : (yf - tl) / (th - tl);
    const float wp = (float(pp) - pl) / (p_hi - pl);
    const float ar = 1.0f - wr, at = 1.0f - wt, ap = 1.0f - wp;
    const unsigned row00 = A.brick_base / 4u + (unsigned(ir) * unsigned(A.nt) + unsigned(it)) * unsigned(A.np);
    const unsigned row01 = row00 + unsigned(A.np);
    const unsigned row10 = row00 + unsigned(A.nt) * unsigned(A.np);
    const unsigned row11 = row10 + unsigned(A.np);
    const unsigned o[8] = {(row00 + ip) * 4u, (row00 + jp) * 4u, (row01 + ip) * 4u, (row01 + jp) * 4u,
                           (row10 + ip) * 4u, (row10 + jp) * 4u, (row11 + ip) * 4u, (row11 + jp) * 4u};
#pragma unroll
    for (int k = 0; k < kCoeffs; ++k) {
        const unsigned so = unsigned(k) * A.plane_bytes;
        float v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = load_coeff<BH_STORE_F32>(rs, o[q], so);
        const float c00 = lerp_numpy32(v[0], ap, v[1], wp), c01 = lerp_numpy32(v[2], ap, v[3], wp);
        const float c10 = lerp_numpy32(v[4], ap, v[5], wp), c11 = lerp_numpy32(v[6], ap, v[7], wp);
        const float c0 = lerp_numpy32(c00, at, c01, wt), c1 = lerp_numpy32(c10, at, c11, wt);
        out[k] = double(lerp_numpy32(c0, ar, c1, wr));
    }
    return true;
}

// lib.rs:37-59 and 210-225.
template <typename R>
__device__ __forceinline__ void stokes_rhs(const R s[4], const R c[kCoeffs], R o[4]) {
#pragma clang fp contract(off)
    o[0] = c[0] - dot4_hd(c[4], s[0], c[5], s[1], c[6], s[2], c[7], s[3], R(0));
    o[1] = c[1] - dot4_hd(c[5], s[0], c[4], s[1], c[8], s[2], -c[10], s[3], R(0));
    o[2] = c[2] - dot4_hd(c[6], s[0], -c[8], s[1], c[4], s[2], c[9], s[3], R(0));
    o[3] = c[3] - dot4_hd(c[7], s[0], c[10], s[1], -c[9], s[2], c[4], s[3], R(0));
}

template <typename R>
__device__ __forceinline__ void stokes_rk2(R s[4], const R c[kCoeffs], R ds) {
#pragma clang fp contract(off)
    R k1[4], mid[4], k2[4];
    stokes_rhs(s, c, k1);
    const R hh = R(0.5) * ds;
#pragma unroll
    for (int i = 0; i < 4; ++i) mid[i] = fma_hd(hh, k1[i], s[i]);
    stokes_rhs(mid, c, k2);
#pragma unroll
    for (int i = 0; i < 4; ++i) s[i] = fma_hd(ds, k2[i], s[i]);
}

// ---- camera (kerr.py:188-233), FP64 per ray --------------------------------

__device__ __forceinline__ void camera_ray(const RenderArgs &A, int px, int py, double &pr, double &pth, double &pt,
                                           double &pph) {
    const double ndc_x = 2.0 * (double(px) + 0.5) / double(A.W) - 1.0;
    const double ndc_y = 1.0 - 2.0 * (double(py) + 0.5) / double(A.H);
    const CamConst &C = A.cam;
    double n_r = -1.0, n_th = -ndc_y * C.tan_y, n_ph = ndc_x * C.tan_x;
    if (C.roll) {
        const double t0 = C.cos_roll * n_th - C.sin_roll * n_ph;
        const double t1 = C.sin_roll * n_th + C.cos_roll * n_ph;
        n_th = t0;
        n_ph = t1;
    }
    const double nrm = ::sqrt(n_r * n_r + n_th * n_th + n_ph * n_ph);
    n_r /= nrm;
    n_th /= nrm;
    n_ph /= nrm;
    const double pc_t = C.e_t0;
    const double pc_r = n_r * C.e_r1;
    const double pc_th = n_th * C.e_th2;
    const double pc_ph = C.e_t3 + n_ph * C.e_ph3;
    pt = C.g00 * pc_t + C.g03 * pc_ph;
    pr = C.g11 * pc_r;
    pth = C.g22 * pc_th;
    pph = C.g03 * pc_t + C.g33 * pc_ph;
}

template <typename R>
struct RayState {
    R r, th, ph, pr, pth;
    R S[4];
    R tau;
    RayK<R> k;
    double Ed, Ld;   // E = -p_t and L = p_phi in FP64 (mixed-precision islands)
    int steps;
    int pix;
    bool clamped;
};

template <typename R>
__device__ __forceinline__ void ray_init(const RenderArgs &A, int px, int py, RayState<R> &s) {
    double pr, pth, pt, pph;
    camera_ray(A, px, py, pr, pth, pt, pph);
    s.r = R(A.cam.r);
    s.th = R(A.cam.th);
    s.ph = R(A.cam.ph);
    s.pr = R(pr);
    s.pth = R(pth);
    const double E = -pt, L = pph, a = A.a;
    s.k.a = R(a);
    s.k.a2 = R(a * a);
    s.k.E = R(E);
    s.k.L = R(L);
    s.k.L2 = R(L * L);
    s.k.a2E2 = R(a * a * E * E);
    s.k.twoaEL = R(2.0 * a * E * L);
    s.k.aE = R(a * E);
    s.k.aL = R(a * L);
    s.k.fourE = R(4.0 * E);
    s.Ed = E;
    s.Ld = L;
    s.S[0] = s.S[1] = s.S[2] = s.S[3] = R(0);
    s.tau = R(0);
    s.steps = 0;
    s.pix = py * A.W + px;
    s.clamped = false;
}

__device__ __forceinline__ bool finite_(float x) { return __builtin_isfinite(x); }
__device__ __forceinline__ bool finite_(double x) { return __builtin_isfinite(x); }

template <typename R>
__device__ __forceinline__ R nan_to_zero(R x) { return __builtin_isfinite(x) ? x : R(0); }

template <typename G>
__device__ __forceinline__ RayK<G> make_k(double a, double E, double L) {
    RayK<G> k;
    k.a = G(a);
    k.a2 = G(a * a);
    k.E = G(E);
    k.L = G(L);
    k.L2 = G(L * L);
    k.a2E2 = G(a * a * E * E);
    k.twoaEL = G(2.0 * a * E * L);
    k.aE = G(a * E);
    k.aL = G(a * L);
    k.fourE = G(4.0 * E);
    return k;
}

// RK2 midpoint step of the geodesic in precision G: k1 = f(y), k2 = f(y + h/2 k1), y + h k2.
// Returns the raw (unclamped) end theta in nt_raw and the midpoint theta in mt.  _free: the FP64 reference
// kernels (kerr_rhs_free); _pinned: the FP32 kernels and the islands, rounded alike in every code path.
template <typename G>
__device__ __forceinline__ void geo_step_free(const RayK<G> &k, G r, G th, G ph, G pr, G pth, G h, G &nr, G &nt_raw, G &np,
                                              G &npr, G &npt, G &mt) {
    G k1r, k1t, k1p, k1pr, k1pt;
    kerr_rhs_free<G>(k, r, th, pr, pth, k1r, k1t, k1p, k1pr, k1pt);
    const G hh = G(0.5) * h;
    mt = th + hh * k1t;
    G k2r, k2t, k2p, k2pr, k2pt;
    kerr_rhs_free<G>(k, r + hh * k1r, mt, pr + hh * k1pr, pth + hh * k1pt, k2r, k2t, k2p, k2pr, k2pt);
    nr = r + h * k2r;
    nt_raw = th + h * k2t;
    np = ph + h * k2p;
    npr = pr + h * k2pr;
    npt = pth + h * k2pt;
}

template <typename G, bool ISLAND = false>
__device__ __forceinline__ void geo_step_pinned(const RayK<G> &k, G r, G th, G ph, G pr, G pth, G h, G &nr, G &nt_raw, G &np,
                                                G &npr, G &npt, G &mt) {
#pragma clang fp contract(off)
    G k1r, k1t, k1p, k1pr, k1pt;
    kerr_rhs_pinned<G, ISLAND>(k, r, th, pr, pth, k1r, k1t, k1p, k1pr, k1pt);
    const G hh = G(0.5) * h;
    mt = fma_hd(hh, k1t, th);
    G k2r, k2t, k2p, k2pr, k2pt;
    kerr_rhs_pinned<G, ISLAND>(k, fma_hd(hh, k1r, r), mt, fma_hd(hh, k1pr, pr), fma_hd(hh, k1pt, pth), k2r, k2t, k2p, k2pr,
                        k2pt);
    nr = fma_hd(h, k2r, r);
    nt_raw = fma_hd(h, k2t, th);
    np = fma_hd(h, k2p, ph);
    npr = fma_hd(h, k2pr, pr);
    npt = fma_hd(h, k2pt, pth);
}

template <typename G, bool ISLAND = false>
__device__ __forceinline__ void geo_step(const RayK<G> &k, G r, G th, G ph, G pr, G pth, G h, G &nr, G &nt_raw, G &np,
                                         G &npr, G &npt, G &mt) {
    if constexpr (sizeof(G) == 8 && !ISLAND) geo_step_free<G>(k, r, th, ph, pr, pth, h, nr, nt_raw, np, npr, npt, mt);
    else geo_step_pinned<G, ISLAND>(k, r, th, ph, pr, pth, h, nr, nt_raw, np, npr, npt, mt);
}

// Classical RK4 step of the reference tracer (kerr.py:306-313): k1 = f(y), k2 = f(y + h/2 k1),
// k3 = f(y + h/2 k2), k4 = f(y + h k3), y + (h/6)(k1 + 2 k2 + 2 k3 + k4), summed in that order.
// Returns the raw end theta in nt_raw and the lowest and highest theta of the three inner stages.
template <typename G>
__device__ __forceinline__ void geo_step4_free(const RayK<G> &k, G r, G th, G ph, G pr, G pth, G h, G &nr, G &nt_raw,
                                               G &np, G &npr, G &npt, G &mlo, G &mhi) {
    G k1r, k1t, k1p, k1pr, k1pt;
    kerr_rhs_free<G>(k, r, th, pr, pth, k1r, k1t, k1p, k1pr, k1pt);
    const G hh = G(0.5) * h;
    const G t2 = th + hh * k1t;
    G k2r, k2t, k2p, k2pr, k2pt;
    kerr_rhs_free<G>(k, r + hh * k1r, t2, pr + hh * k1pr, pth + hh * k1pt, k2r, k2t, k2p, k2pr, k2pt);
    const G t3 = th + hh * k2t;
    G k3r, k3t, k3p, k3pr, k3pt;
    kerr_rhs_free<G>(k, r + hh * k2r, t3, pr + hh * k2pr, pth + hh * k2pt, k3r, k3t, k3p, k3pr, k3pt);
    const G t4 = th + h * k3t;
    G k4r, k4t, k4p, k4pr, k4pt;
    kerr_rhs_free<G>(k, r + h * k3r, t4, pr + h * k3pr, pth + h * k3pt, k4r, k4t, k4p, k4pr, k4pt);
    const G w = h / G(6);
    nr = r + w * (k1r + G(2) * k2r + G(2) * k3r + k4r);
    nt_raw = th + w * (k1t + G(2) * k2t + G(2) * k3t + k4t);
    np = ph + w * (k1p + G(2) * k2p + G(2) * k3p + k4p);
    npr = pr + w * (k1pr + G(2) * k2pr + G(2) * k3pr + k4pr);
    npt = pth + w * (k1pt + G(2) * k2pt + G(2) * k3pt + k4pt);
    mlo = t2 < t3 ? (t2 < t4 ? t2 : t4) : (t3 < t4 ? t3 : t4);
    mhi = t2 > t3 ? (t2 > t4 ? t2 : t4) : (t3 > t4 ? t3 : t4);
}

template <typename G, bool ISLAND = false>
__device__ __forceinline__ void geo_step4_pinned(const RayK<G> &k, G r, G th, G ph, G pr, G pth, G h, G &nr, G &nt_raw,
                                          G &np, G &npr, G &npt, G &mlo, G &mhi) {
#pragma clang fp contract(off)
    G k1r, k1t, k1p, k1pr, k1pt;
    kerr_rhs_pinned<G, ISLAND>(k, r, th, pr, pth, k1r, k1t, k1p, k1pr, k1pt);
    const G hh = G(0.5) * h;
    // the stages and the update are rounded as NumPy rounds them (kerr.py:306-313): a product, then a sum
    const G t2 = th + hh * k1t;
    G k2r, k2t, k2p, k2pr, k2pt;
    kerr_rhs_pinned<G, ISLAND>(k, r + hh * k1r, t2, pr + hh * k1pr, pth + hh * k1pt, k2r, k2t, k2p, k2pr, k2pt);
    const G t3 = th + hh * k2t;
    G k3r, k3t, k3p, k3pr, k3pt;
    kerr_rhs_pinned<G, ISLAND>(k, r + hh * k2r, t3, pr + hh * k2pr, pth + hh * k2pt, k3r, k3t, k3p, k3pr, k3pt);
    const G t4 = th + h * k3t;
    G k4r, k4t, k4p, k4pr, k4pt;
    kerr_rhs_pinned<G, ISLAND>(k, r + h * k3r, t4, pr + h * k3pr, pth + h * k3pt, k4r, k4t, k4p, k4pr, k4pt);
    const G w = h / G(6);
    nr = r + w * (k1r + G(2) * k2r + G(2) * k3r + k4r);
    nt_raw = th + w * (k1t + G(2) * k2t + G(2) * k3t + k4t);
    np = ph + w * (k1p + G(2) * k2p + G(2) * k3p + k4p);
    npr = pr + w * (k1pr + G(2) * k2pr + G(2) * k3pr + k4pr);
    npt = pth + w * (k1pt + G(2) * k2pt + G(2) * k3pt + k4pt);
    mlo = t2 < t3 ? (t2 < t4 ? t2 : t4) : (t3 < t4 ? t3 : t4);
    mhi = t2 > t3 ? (t2 > t4 ? t2 : t4) : (t3 > t4 ? t3 : t4);
}

template <typename G, bool ISLAND = false>
__device__ __forceinline__ void geo_step4(const RayK<G> &k, G r, G th, G ph, G pr, G pth, G h, G &nr, G &nt_raw,
                                          G &np, G &npr, G &npt, G &mlo, G &mhi) {
    if constexpr (sizeof(G) == 8 && !ISLAND)
        geo_step4_free<G>(k, r, th, ph, pr, pth, h, nr, nt_raw, np, npr, npt, mlo, mhi);
    else geo_step4_pinned<G, ISLAND>(k, r, th, ph, pr, pth, h, nr, nt_raw, np, npr, npt, mlo, mhi);
}

// The end point of one geodesic step: the state after it, whether it met the pole pathology (the
// theta clamp fired, or the step swung phi by more than 1 rad) and whether all five values are finite.
template <typename R>
struct GeoPoint {
    R r, th, ph, pr, pth;
    bool hit, fin;
};

// One geodesic step (RK2 midpoint, or RK4) of the ray from (r, th, ph, pr, pth); the ray's constants
// of motion come from s.  MIXED: the step runs in FP64 inside the islands (r < 1.5 r_+, within th_island of a pole,
// |p_theta| < pth_island near a theta turning point).
template <typename R, bool MIXED, bool RK4>
__device__ __forceinline__ GeoPoint<R> geo_next(const RenderArgs &A, const HotParams<R> &P, const RayState<R> &s, R r,
                                                R th, R ph, R pr, R pth) {
    const R lo = R(kEpsTheta), hi = Mth<R>::pi_minus_eps();
    R nr, nt_raw, np, npr, npt, mlo, mhi;
    bool island = false;
    if constexpr (MIXED) {
        island = (double(r) < A.r_island) || (double(th) < A.th_island) ||
                 (double(th) > 3.14159265358979323846 - A.th_island) || (fabs(double(pth)) < A.pth_island);
    }
    if (MIXED && island) {
        const double hd = A.adaptive ? A.step * (double(r) - A.r_plus) : A.step;
        const RayK<double> kd = make_k<double>(A.a, s.Ed, s.Ld);
        double dr_, dt_, dp_, dpr_, dpt_, dlo_, dhi_;
        if constexpr (RK4) {
            geo_step4<double, true>(kd, double(r), double(th), double(ph), double(pr), double(pth), hd, dr_, dt_, dp_,
                                    dpr_, dpt_, dlo_, dhi_);
        } else {
            geo_step<double, true>(kd, double(r), double(th), double(ph), double(pr), double(pth), hd, dr_, dt_, dp_,
                                   dpr_, dpt_, dlo_);
            dhi_ = dlo_;
        }
        nr = R(dr_);
        nt_raw = R(dt_);
        np = R(dp_);
        npr = R(dpr_);
        npt = R(dpt_);
        mlo = R(dlo_);
        mhi = R(dhi_);
    } else {
        const R h = P.adaptive ? P.step * (r - P.r_plus) : P.step;
        if constexpr (RK4) {
            geo_step4<R>(s.k, r, th, ph, pr, pth, h, nr, nt_raw, np, npr, npt, mlo, mhi);
        } else {
            geo_step<R>(s.k, r, th, ph, pr, pth, h, nr, nt_raw, np, npr, npt, mlo);
            mhi = mlo;
        }
    }
    GeoPoint<R> b;
    b.r = nr;
    b.th = clamp_theta(nt_raw);
    b.ph = np;
    b.pr = npr;
    b.pth = npt;
    b.hit = (mlo < lo) || (mhi > hi) || (nt_raw < lo) || (nt_raw > hi) || (np - ph > R(1)) || (ph - np > R(1));
    b.fin = finite_(nr) && finite_(b.th) && finite_(np) && finite_(npr) && finite_(npt);
    return b;
}

// ds of the segment from (r, th, ph) to b (accelerated_renderer.py:119-121).  The sum of squares is
// written as explicit fused multiply-adds with contraction off, so every code path that takes a step
// (the ray's own lane, a lone ray's batch, the helper of a split lone ray) rounds ds the same way.
template <typename R>
__device__ __forceinline__ R segment_ds_at(const HotParams<R> &P, R r, R th, R ph, const GeoPoint<R> &b) {
#pragma clang fp contract(off)
    const R dx = b.r - r, dy = b.th - th, dz = b.ph - ph;
    const R dlam = Mth<R>::sqrt(fma_hd(dz, dz, fma_hd(dy, dy, dx * dx)));
    return pymax(P.L * pymax(dlam, R(1.0e-14)), R(0));
}

// ds of the segment from the ray's position to b.
template <typename R>
__device__ __forceinline__ R segment_ds(const HotParams<R> &P, const RayState<R> &s, const GeoPoint<R> &b) {
    return segment_ds_at(P, s.r, s.th, s.ph, b);
}

// Transfer over a segment of length ds with the coefficients c sampled at its midpoint
// (accelerated_renderer.py:113-125).
template <typename R, bool EXACT>
__device__ __forceinline__ void segment_transfer(RayState<R> &s, R ds, const R c[kCoeffs]) {
    if constexpr (EXACT) stokes_exact(s.S, c, ds);
    else stokes_rk2(s.S, c, ds);
    s.tau = fma_hd(pymax(c[4], R(0)), ds, s.tau);
}

// The geodesic tests of a step that ends at b from radius r_prev: captured, escaped (MIXED: in FP64).
template <typename R, bool MIXED>
__device__ __forceinline__ void end_tests(const RenderArgs &A, const HotParams<R> &P, R r_prev, R r, bool &captured,
                                          bool &escaped) {
    if constexpr (MIXED) {
        captured = double(r) <= A.r_cap;
        escaped = (double(r) > A.r_esc) && (r > r_prev);
    } else {
        captured = r <= P.r_cap;
        escaped = (r > P.r_esc) && (r > r_prev);
    }
}

// First match wins: invalid, captured, escaped, opaque, out of steps (selects, no branches).
template <typename R>
__device__ __forceinline__ int end_code(const HotParams<R> &P, const RayState<R> &s, bool captured, bool escaped,
                                        bool fin) {
    int code = (s.steps >= P.max_steps) ? 0 : -1;
    code = (s.tau >= P.tau_max) ? 4 : code;
    code = escaped ? 2 : code;
    code = captured ? 1 : code;
    code = fin ? code : 3;
    return code;
}

// One RK2 step plus its transfer segment.  Returns the status code when the ray ends, -1 otherwise.
template <typename R, int STORE, int LAYOUT, bool MIXED, bool EXACT, bool RK4>
__device__ __forceinline__ int ray_step(const RenderArgs &A, const HotParams<R> &P, __amdgpu_buffer_rsrc_t rs,
                                        const R *gr, const R *gt, const R *gp, RayState<R> &s, unsigned &n_samples) {
    const GeoPoint<R> b = geo_next<R, MIXED, RK4>(A, P, s, s.r, s.th, s.ph, s.pr, s.pth);
    s.clamped = s.clamped || b.hit;
    R c[kCoeffs];
    bool inside;
    if constexpr (sizeof(R) == 8 && store_of(STORE) == BH_STORE_F32 && LAYOUT == BH_LAYOUT_SOA) {
        if (A.lerp32)
            inside = sample_brick_numpy32(A, rs, gr, gt, gp, midpoint(s.r, b.r), midpoint(s.th, b.th), midpoint(s.ph, b.ph), c);
        else
            inside = sample_brick<R, STORE, LAYOUT>(A, rs, gr, gt, gp, midpoint(s.r, b.r), midpoint(s.th, b.th),
                                                          midpoint(s.ph, b.ph), c);
    } else {
        inside = sample_brick<R, STORE, LAYOUT>(A, rs, gr, gt, gp, midpoint(s.r, b.r), midpoint(s.th, b.th),
                                                      midpoint(s.ph, b.ph), c);
    }
    if (inside) {
        segment_transfer<R, EXACT>(s, segment_ds(P, s, b), c);
        ++n_samples;
    }
    const R r_prev = s.r;
    s.r = b.r;
    s.th = b.th;
    s.ph = b.ph;
    s.pr = b.pr;
    s.pth = b.pth;
    s.steps += 1;
    bool captured, escaped;
    end_tests<R, MIXED>(A, P, r_prev, b.r, captured, escaped);
    return end_code(P, s, captured, escaped, b.fin);
}

template <typename T>
__device__ __forceinline__ T lane_value(T x, int lane) {
    if constexpr (sizeof(T) == 8) {
        const long long v = __builtin_bit_cast(long long, x);
        const int lo = __builtin_amdgcn_readlane(int(v & 0xffffffffll), lane);
        const int hi = __builtin_amdgcn_readlane(int(v >> 32), lane);
        return __builtin_bit_cast(T, (long long)(unsigned(lo)) | ((long long)hi << 32));
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), lane));
    }
}

// Copy the state of the ray in lane `owner` to every lane of the wave.
template <typename R>
__device__ __forceinline__ void broadcast_ray(RayState<R> &s, int owner) {
    s.r = lane_value(s.r, owner);
    s.th = lane_value(s.th, owner);
    s.ph = lane_value(s.ph, owner);
    s.pr = lane_value(s.pr, owner);
    s.pth = lane_value(s.pth, owner);
#pragma unroll
    for (int i = 0; i < 4; ++i) s.S[i] = lane_value(s.S[i], owner);
    s.tau = lane_value(s.tau, owner);
    s.k.a = lane_value(s.k.a, owner);
    s.k.a2 = lane_value(s.k.a2, owner);
    s.k.E = lane_value(s.k.E, owner);
    s.k.L = lane_value(s.k.L, owner);
    s.k.L2 = lane_value(s.k.L2, owner);
    s.k.a2E2 = lane_value(s.k.a2E2, owner);
    s.k.twoaEL = lane_value(s.k.twoaEL, owner);
    s.k.aE = lane_value(s.k.aE, owner);
    s.k.aL = lane_value(s.k.aL, owner);
    s.k.fourE = lane_value(s.k.fourE, owner);
    s.Ed = lane_value(s.Ed, owner);
    s.Ld = lane_value(s.Ld, owner);
    s.steps = lane_value(s.steps, owner);
    s.pix = lane_value(s.pix, owner);
    s.clamped = lane_value(int(s.clamped), owner) != 0;
}

template <typename R>
__device__ __forceinline__ void ray_store(const RenderArgs &A, const RayState<R> &s, int code) {
    R *o = static_cast<R *>(A.iquv) + size_t(s.pix) * 4;
    o[0] = nan_to_zero(s.S[0]);
    o[1] = nan_to_zero(s.S[1]);
    o[2] = nan_to_zero(s.S[2]);
    o[3] = nan_to_zero(s.S[3]);
    if (A.steps) A.steps[s.pix] = s.steps;
    if (A.status) A.status[s.pix] = uint8_t(code | (s.clamped ? 8 : 0));
}

// ---- a lone ray in batches of steps ------------------------------------------------------
// A wave left with one ray (every lane holding a copy) takes its steps in batches: the geodesic runs
// up to kLoneBatch steps ahead, lane j keeping step j's segment (midpoint, ds, end flags); then every
// lane samples its own segment's midpoint, so the batch's brick loads are in flight together (one
// memory round trip per batch instead of one per step); then the transfer and the end tests run over
// the batch's steps in order, and the first step that ends the ray ends it (geodesic steps computed
// past it are dropped).  Every value goes through the functions the per-lane step uses, in the same
// order, so the image is the per-lane image bit for bit.  Batches grow from 8 steps to kLoneBatch as the
// ray goes on, so a ray that ends soon wastes few geodesic steps.
constexpr int kLoneBatch = 64;
constexpr int kLoneStride = 12;   // floats per step in the wave's LDS buffer (11 coefficients, 16-byte rows)
constexpr int kLoneCheck = 8;     // geodesic steps between looks at the end flags

// One step of a batch, as the lane that keeps it holds it.
struct AheadStep {
    float x, y, z, ds;           // midpoint and ds of the segment
    unsigned flags;              // bit 0 captured, 1 escaped, 2 finite, 3 pole pathology
    float r, th, ph, pr, pth;    // the ray after the step (KEEP_END)
};

// Up to nb geodesic steps from (r, th, ph, pr, pth); lane j keeps step j in `mine`.  The end flags are
// looked at every kLoneCheck steps: a step past the ray's end only computes values that are dropped.
// Returns the steps taken; (r, th, ph, pr, pth) become the ray after the last one, `ended` whether a
// step ended the geodesic (captured, escaped or not finite).
template <bool MIXED, bool RK4, bool KEEP_END>
__device__ __forceinline__ int geo_ahead(const RenderArgs &A, const HotParams<float> &P, const RayState<float> &s,
                                         int nb, int lane, float &r, float &th, float &ph, float &pr, float &pth,
                                         AheadStep &mine, bool &ended) {
    int n = 0;
    bool end = false;
    while (n < nb) {
        const int stop = n + kLoneCheck < nb ? n + kLoneCheck : nb;
        for (; n < stop; ++n) {
            const GeoPoint<float> b = geo_next<float, MIXED, RK4>(A, P, s, r, th, ph, pr, pth);
            bool captured, escaped;
            end_tests<float, MIXED>(A, P, r, b.r, captured, escaped);
            const float x = midpoint(r, b.r), y = midpoint(th, b.th), z = midpoint(ph, b.ph);
            const float ds = segment_ds_at(P, r, th, ph, b);
            const unsigned f = (captured ? 1u : 0u) | (escaped ? 2u : 0u) | (b.fin ? 4u : 0u) | (b.hit ? 8u : 0u);
            if (lane == n) {
                mine.x = x;
                mine.y = y;
                mine.z = z;
                mine.ds = ds;
                mine.flags = f;
                if constexpr (KEEP_END) {
                    mine.r = b.r;
                    mine.th = b.th;
                    mine.ph = b.ph;
                    mine.pr = b.pr;
                    mine.pth = b.pth;
                }
            }
            end = end || captured || escaped || !b.fin;
            r = b.r;
            th = b.th;
            ph = b.ph;
            pr = b.pr;
            pth = b.pth;
        }
        if (__builtin_amdgcn_readfirstlane(int(end))) break;
    }
    ended = __builtin_amdgcn_readfirstlane(int(end)) != 0;
    return n;
}

// The transfer and the end tests of one step, in the per-lane step's order (flags as AheadStep's, plus
// bit 4 when the midpoint is in the brick and c holds its coefficients).  Returns the end code, -1 when
// the ray goes on.
template <bool EXACT>
__device__ __forceinline__ int take_step(const HotParams<float> &P, RayState<float> &s, unsigned f, float ds,
                                         const float c[kCoeffs], unsigned &steps, unsigned &samples) {
    if (f & 16u) {
        segment_transfer<float, EXACT>(s, ds, c);
        ++samples;
    }
    s.clamped = s.clamped || (f & 8u);
    s.steps += 1;
    ++steps;
    return end_code(P, s, (f & 1u) != 0, (f & 2u) != 0, (f & 4u) != 0);
}

// The lone ray for up to `budget` steps; returns its end code, or -1 when the budget ran out first (s then
// holds the ray after the last step).  cb: this wave's LDS buffer, kLoneBatch x kLoneStride floats.
template <int STORE, bool MIXED, bool EXACT, bool RK4>
__device__ __forceinline__ int lone_batched(const RenderArgs &A, const HotParams<float> &P, __amdgpu_buffer_rsrc_t rs,
                                            const float *gr, const float *gt, const float *gp, RayState<float> &s,
                                            int lane, float *cb, unsigned budget, unsigned &steps, unsigned &samples) {
    int batch = 8;
    while (budget > 0) {
        float r = s.r, th = s.th, ph = s.ph, pr = s.pr, pth = s.pth;
        AheadStep mine{};
        const int room = __builtin_amdgcn_readfirstlane(P.max_steps - s.steps);
        int nb = batch < room ? batch : room;
        nb = unsigned(nb) < budget ? nb : int(budget);
        bool ended;
        const int n = geo_ahead<MIXED, RK4, false>(A, P, s, nb, lane, r, th, ph, pr, pth, mine, ended);
        // the batch's samples, one per lane
        float c[kCoeffs];
        bool inside = false;
        if (lane < n) inside = sample_brick<float, STORE, BH_LAYOUT_CORNER>(A, rs, gr, gt, gp, mine.x, mine.y, mine.z, c);
        if (inside) {
#pragma unroll
            for (int k = 0; k < kCoeffs; ++k) cb[lane * kLoneStride + k] = c[k];
        }
        __builtin_amdgcn_wave_barrier();
        const unsigned mf = mine.flags | (inside ? 16u : 0u);
        // transfer and end tests in step order; step j + 1's coefficients are read while step j integrates
        float cn[kCoeffs];
#pragma unroll
        for (int k = 0; k < kCoeffs; ++k) cn[k] = cb[k];
        for (int j = 0; j < n; ++j) {
            float cj[kCoeffs];
#pragma unroll
            for (int k = 0; k < kCoeffs; ++k) cj[k] = cn[k];
            if (j + 1 < n) {
#pragma unroll
                for (int k = 0; k < kCoeffs; ++k) cn[k] = cb[(j + 1) * kLoneStride + k];
            }
            const int code = take_step<EXACT>(P, s, __builtin_amdgcn_readlane(mf, j),
                                              __builtin_bit_cast(float, __builtin_amdgcn_readlane(
                                                                            __builtin_bit_cast(int, mine.ds), j)),
                                              cj, steps, samples);
            if (code >= 0) return code;
        }
        __builtin_amdgcn_wave_barrier();   // the next batch rewrites this wave's LDS buffer
        s.r = r;
        s.th = th;
        s.ph = ph;
        s.pr = pr;
        s.pth = pth;
        budget -= unsigned(n);
        batch = batch * 2 < kLoneBatch ? batch * 2 : kLoneBatch;
    }
    return -1;
}

// ---- a lone ray split over two waves -----------------------------------------------------
// A ray that stays alone in its wave for coop_steps steps asks for help: a wave of the same
// workgroup that has no rays of its own takes the geodesic and the sampling, in the batches above
// (geodesic ahead, then every lane sampling its own step), and writes each step into a ring in LDS:
// the coefficients at the midpoint, ds, the end flags and the ray after the step.  The lone wave takes
// the steps in order and does the transfer and the end tests, the same functions in the same order as
// the per-lane step, so the image is the per-lane image bit for bit.  The two waves work at the same
// time, the helper a batch or two ahead.  Every wait is bounded: a helper that stops answering makes the
// lone wave finish the ray alone from the last step it took (counted in work[4]).
constexpr int kCoopRing = 2 * kLoneBatch;
constexpr int kCoopFirstBatch = 16;
constexpr unsigned kCoopSpin = 1u << 21;   // ~0.1 s of s_sleep(1) polls
// The hand-over is one word, phase | requesting wave << 8, so every transition is a single
// compare-and-swap on the exact value read: Idle -> Asked (lone wave), Asked -> Accepted (helper),
// Accepted -> Go (lone wave, after writing the ray) or Accepted -> Cancel (the ray ended first),
// Go / Cancel -> Free (helper, leaving), Free or Asked -> Idle (lone wave).
enum : int { kCoopIdle = 0, kCoopAsked = 1, kCoopAccepted = 2, kCoopGo = 3, kCoopCancel = 4, kCoopFree = 5 };
__device__ __forceinline__ int coop_word(int phase, int wave) { return phase | (wave << 8); }

struct alignas(16) CoopSample {
    float c[kCoeffs];            // the coefficients at the step's midpoint (when flags bit 4 is set)
    float ds;
    unsigned flags;              // AheadStep's, plus bit 4: the midpoint is in the brick
    float r, th, ph, pr, pth;    // the ray after the step (restart point if the helper stops answering)
    unsigned pad[2];
};

struct CoopLds {
    int state;       // coop_word(phase, requesting wave)
    int tracing;     // waves of the workgroup that may still trace rays
    unsigned head;   // steps written by the helper
    unsigned tail;   // steps taken by the lone wave
    int stop;        // the lone wave is done: the helper stops
    RayState<float> ray;
    CoopSample ring[kCoopRing];
};

// Ring indices: a wave's LDS operations execute in order, so a relaxed store of the index after the
// slots' stores (kept in order by a compiler fence) publishes the slots without waiting for the
// stores to complete, and a reader that sees the index sees the slots.
template <typename T>
__device__ __forceinline__ T lds_peek(T *p) {
    const T v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    return v;
}
template <typename T>
__device__ __forceinline__ void lds_publish(T *p, T v) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T, typename F>
__device__ __forceinline__ T lds_wait_relaxed(T *p, F pred, bool &timed_out) {
    T v = lds_peek(p);
    unsigned spins = 0;
    while (!pred(v)) {
        if (++spins > kCoopSpin) {
            timed_out = true;
            return v;
        }
        __builtin_amdgcn_s_sleep(1);
        v = lds_peek(p);
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T lds_load(T *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ void lds_store(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ bool lds_cas(int *p, int expect, int v) {
    return __hip_atomic_compare_exchange_strong(p, &expect, v, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
}
// one lane does it, the wave learns the result
__device__ __forceinline__ bool wave_cas(int *p, int expect, int v, int lane) {
    int ok = 0;
    if (lane == 0) ok = lds_cas(p, expect, v) ? 1 : 0;
    return __builtin_amdgcn_readfirstlane(ok) != 0;
}
// poll *p until pred(value) holds or the spin bound passes; returns the last value read
template <typename T, typename F>
__device__ __forceinline__ T lds_wait(T *p, F pred, bool &timed_out) {
    T v = lds_load(p);
    unsigned spins = 0;
    while (!pred(v)) {
        if (++spins > kCoopSpin) {
            timed_out = true;
            return v;
        }
        __builtin_amdgcn_s_sleep(1);
        v = lds_load(p);
    }
    return v;
}

// Helper side: the handed-over ray's geodesic and samples in batches, written to the ring, until the
// geodesic ends the ray, the step limit, the lone wave says stop, or the ring stays full past the spin
// bound.  Returns the steps written.
template <int STORE, bool MIXED, bool RK4>
__device__ __forceinline__ unsigned coop_produce(const RenderArgs &A, const HotParams<float> &P, __amdgpu_buffer_rsrc_t rs,
                                                 const float *gr, const float *gt, const float *gp, CoopLds *c,
                                                 int lane) {
    RayState<float> s = c->ray;
    __builtin_amdgcn_s_setprio(2);   // the lone wave waits on every batch of this wave
    unsigned head = 0, tail = 0;
    int batch = kCoopFirstBatch;
    for (;;) {
        const int room = __builtin_amdgcn_readfirstlane(P.max_steps - s.steps);
        const int nb = batch < room ? batch : room;
        if (nb <= 0) break;
        if (head + unsigned(nb) - tail > unsigned(kCoopRing)) {   // no room for the batch: wait for the lone wave
            bool timed_out = false;
            tail = lds_wait_relaxed(
                &c->tail, [&](unsigned v) { return head + unsigned(nb) - v <= unsigned(kCoopRing) || lds_peek(&c->stop); },
                timed_out);
            if (timed_out || head + unsigned(nb) - tail > unsigned(kCoopRing)) break;
        }
        if (lds_peek(&c->stop)) break;
        float r = s.r, th = s.th, ph = s.ph, pr = s.pr, pth = s.pth;
        AheadStep mine{};
        bool ended;
        const int n = geo_ahead<MIXED, RK4, true>(A, P, s, nb, lane, r, th, ph, pr, pth, mine, ended);
        float cf[kCoeffs];
        bool inside = false;
        if (lane < n) inside = sample_brick<float, STORE, BH_LAYOUT_CORNER>(A, rs, gr, gt, gp, mine.x, mine.y, mine.z, cf);
        if (lane < n) {
            CoopSample &sl = c->ring[(head + unsigned(lane)) & unsigned(kCoopRing - 1)];
            if (inside) {
#pragma unroll
                for (int k = 0; k < kCoeffs; ++k) sl.c[k] = cf[k];
            }
            sl.ds = mine.ds;
            sl.flags = mine.flags | (inside ? 16u : 0u);
            sl.r = mine.r;
            sl.th = mine.th;
            sl.ph = mine.ph;
            sl.pr = mine.pr;
            sl.pth = mine.pth;
        }
        head += unsigned(n);
        if (lane == 0) lds_publish(&c->head, head);
        s.r = r;
        s.th = th;
        s.ph = ph;
        s.pr = pr;
        s.pth = pth;
        s.steps += n;
        if (ended) break;   // the geodesic ends the ray within the steps written
        batch = batch * 2 < kLoneBatch ? batch * 2 : kLoneBatch;
    }
    return head;
}

// Lone-wave side: take the helper's steps in order until the ray ends.  Returns the end code, or -1 after
// a time-out (s then holds the ray after the last step taken, and the lone wave goes on alone).
template <bool EXACT>
__device__ __forceinline__ int coop_consume(const HotParams<float> &P, RayState<float> &s, CoopLds *c, int lane,
                                            unsigned &steps, unsigned &samples) {
    unsigned next = 0;   // steps taken
    for (;;) {
        bool timed_out = false;
        const unsigned h = lds_wait_relaxed(&c->head, [next](unsigned v) { return v != next; }, timed_out);
        if (timed_out) {
            if (lane == 0) lds_store(&c->stop, 1);
            return -1;
        }
        for (; next != h; ++next) {
            const CoopSample &sl = c->ring[next & unsigned(kCoopRing - 1)];
            float cf[kCoeffs];
#pragma unroll
            for (int k = 0; k < kCoeffs; ++k) cf[k] = sl.c[k];
            const int code = take_step<EXACT>(P, s, sl.flags, sl.ds, cf, steps, samples);
            s.r = sl.r;
            s.th = sl.th;
            s.ph = sl.ph;
            s.pr = sl.pr;
            s.pth = sl.pth;
            if (code >= 0) {
                if (lane == 0) lds_store(&c->stop, 1);
                return code;
            }
        }
        if (lane == 0) lds_publish(&c->tail, next);   // the slots taken may be rewritten
    }
}

// A wave left with one live ray (FP32 kernels, corner layout): every lane takes a copy of it and the ray runs
// to its end in batches; then its owner lane writes the result.  With coop != null and coop_steps > 0, a ray
// still alone after coop_steps steps asks an idle wave of the workgroup to take its geodesic and samples
// (see above).
template <typename R, int STORE, int LAYOUT, bool MIXED, bool EXACT, bool RK4>
__device__ __forceinline__ void lone_ray(const RenderArgs &A, const HotParams<R> &P, __amdgpu_buffer_rsrc_t rs,
                                         const R *gr, const R *gt, const R *gp, RayState<R> &s, int lane, int owner,
                                         unsigned &n_steps, unsigned &n_samples, unsigned long long &iters,
                                         CoopLds *coop, int wave, float *cb) {
    static_assert(sizeof(R) == 4 && LAYOUT == BH_LAYOUT_CORNER, "lone rays are FP32, corner layout");
    broadcast_ray(s, owner);
    __builtin_amdgcn_s_setprio(2);
    unsigned steps = 0, samples = 0;
    bool asked = false, helped = false, timed_out = false;
    const bool coop_on = coop != nullptr && A.coop_steps > 0;
    // alone until coop_steps steps; then, while the request waits, a few steps at a time
    int code = lone_batched<STORE, MIXED, EXACT, RK4>(A, P, rs, gr, gt, gp, s, lane, cb,
                                                      coop_on ? unsigned(A.coop_steps) : ~0u, steps, samples);
    unsigned wait = 8;   // steps between looks at the request, doubling while nobody answers
    while (code < 0) {
        if (!asked) {
            asked = wave_cas(&coop->state, kCoopIdle, coop_word(kCoopAsked, wave), lane);
        } else if (!timed_out && lds_load(&coop->state) == coop_word(kCoopAccepted, wave)) {
            if (lane == 0) {
                coop->ray = s;
                coop->head = 0u;
                coop->tail = 0u;
                coop->stop = 0;
                lds_store(&coop->state, coop_word(kCoopGo, wave));
            }
            helped = true;
            code = coop_consume<EXACT>(P, s, coop, lane, steps, samples);
            if (code >= 0) break;
            timed_out = true;   // finish alone from the last step received
        }
        code = lone_batched<STORE, MIXED, EXACT, RK4>(A, P, rs, gr, gt, gp, s, lane, cb, timed_out ? ~0u : wait, steps,
                                                      samples);
        wait = wait * 2 < unsigned(kLoneBatch) ? wait * 2 : unsigned(kLoneBatch);
    }
    if (asked && lane == 0) {
        // withdraw the request (nobody accepted it), or cancel it / wait for the helper to leave
        if (!lds_cas(&coop->state, coop_word(kCoopAsked, wave), kCoopIdle)) {
            lds_cas(&coop->state, coop_word(kCoopAccepted, wave), coop_word(kCoopCancel, wave));
            bool t2 = false;
            const int free_word = coop_word(kCoopFree, wave);
            lds_wait(&coop->state, [free_word](int v) { return v == free_word; }, t2);
            timed_out = timed_out || t2;
            lds_cas(&coop->state, free_word, kCoopIdle);   // on a time-out the word stays taken: no more hand-overs here
        }
    }
    if (lane == owner) {
        ray_store(A, s, code);
        n_steps += steps;
        n_samples += samples;
        if (helped) atomicAdd(&A.work[3], 1ull);
        if (timed_out) atomicAdd(&A.work[4], 1ull);
    }
    iters += steps;
}

// Helper side of the handshake, for a wave with no rays of its own.  Returns true when it helped (or
// took a request that was then withdrawn).
template <int STORE, bool MIXED, bool RK4>
__device__ __forceinline__ bool coop_help(const RenderArgs &A, const HotParams<float> &P, __amdgpu_buffer_rsrc_t rs,
                                          const float *gr, const float *gt, const float *gp, CoopLds *c, int lane,
                                          int wave, unsigned long long &iters) {
    const int word = lds_load(&c->state);
    const int requester = word >> 8;
    if ((word & 0xff) != kCoopAsked || requester == wave) return false;
    if (!wave_cas(&c->state, word, coop_word(kCoopAccepted, requester), lane)) return false;
    bool timed_out = false;
    const int go = coop_word(kCoopGo, requester), cancel = coop_word(kCoopCancel, requester);
    const int st = lds_wait(&c->state, [go, cancel](int v) { return v == go || v == cancel; }, timed_out);
    if (!timed_out && st == go) iters += coop_produce<STORE, MIXED, RK4>(A, P, rs, gr, gt, gp, c, lane);
    // leave: Go / Cancel -> Free (after a time-out the lone wave sees no Free and stops using the slot)
    if (lane == 0 && !timed_out) lds_store(&c->state, coop_word(kCoopFree, requester));
    return true;
}

__device__ __forceinline__ unsigned lane_rank(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi(unsigned(mask >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(mask), 0u));
}

#ifndef BH_RENDER_WAVES_PER_EU
#define BH_RENDER_WAVES_PER_EU 2
#endif

// The mixed kernels with 16-bit bricks hold the most wave-uniform values in their step loops (the FP64
// island constants besides the FP32 ones) and spill SGPRs into VGPR lanes.  Their copy of the launch
// arguments keeps the cell-lookup constants (grid ends, index-guess scales and offsets, n_theta, n_phi)
// in VGPRs instead: 14 more VGPRs (at most 253 of 256, no scratch), which removes ~3/4 of the v_readlane reloads of
// the sampler's loops and makes these kernels 5-7 % faster.  The FP32 kernels measured no gain, and
// the mixed kernels on f32 bricks would spill, so they keep the launch arguments as they are.
template <typename R, int STORE, bool MIXED>
__device__ __forceinline__ RenderArgs step_args(const RenderArgs &A0) {
    RenderArgs A = A0;
    if constexpr (sizeof(R) == 4 && MIXED && store_of(STORE) != BH_STORE_F32) {
        A.hr.first = in_vgpr(A.hr.first);
        A.hr.last = in_vgpr(A.hr.last);
        A.ht.first = in_vgpr(A.ht.first);
        A.ht.last = in_vgpr(A.ht.last);
        A.hp.first = in_vgpr(A.hp.first);
        A.hp.last = in_vgpr(A.hp.last);
        A.hr.scale = in_vgpr(A.hr.scale);
        A.hr.offset = in_vgpr(A.hr.offset);
        A.ht.scale = in_vgpr(A.ht.scale);
        A.ht.offset = in_vgpr(A.ht.offset);
        A.hp.scale = in_vgpr(A.hp.scale);
        A.hp.offset = in_vgpr(A.hp.offset);
        A.nt = in_vgpr(A.nt);
        A.np = in_vgpr(A.np);
    }
    return A;
}

template <typename R, int STORE, int LAYOUT, bool MIXED, bool EXACT, bool RK4>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BH_RENDER_WAVES_PER_EU))) void render_kernel(
    RenderArgs A_launch) {
    const RenderArgs A = step_args<R, STORE, MIXED>(A_launch);
    __shared__ R s_grid[kMaxGrid + 1];
    const int ng = A.nr + A.nt + A.np;
    for (int i = threadIdx.x; i < ng; i += blockDim.x) s_grid[i] = R(A.grids[i]);
    if (threadIdx.x == 0) s_grid[ng] = R(__builtin_huge_valf());   // sentinel g[np] of the phi grid
    __shared__ CoopLds coop;
    // per wave: a lone ray's batch of sampled coefficients (FP32 kernels)
    __shared__ __align__(16) float lone_c[sizeof(R) == 4 ? 4 * kLoneBatch * kLoneStride : 1];
    if (threadIdx.x == 0) {
        coop.state = kCoopIdle;
        coop.tracing = int((blockDim.x + 63) / 64);
    }
    __syncthreads();
    const R *gr = s_grid;
    const R *gt = s_grid + A.nr;
    const R *gp = gt + A.nt;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(A.coeffs), (short)0, int(A.total_bytes), int(kRsrcWord3));

    const HotParams<R> P = loop_params<R>(A);
    const int lane = threadIdx.x & 63;
    const int q_own = blockIdx.x % A.queues;
    const unsigned all_q = (1u << A.queues) - 1u;
    unsigned exhausted = 0;   // wave-uniform
    bool active = false;
    RayState<R> s;
    unsigned n_steps = 0, n_samples = 0;   // this lane's work over all its rays
    unsigned long long iters = 0;          // wave loop iterations
    int n_ids = A.n_ids, q_len = A.q_len;
    if (A.ray_list) {   // list mode: the ids are the listed pixels, split evenly over the queues
        n_ids = int(*A.n_list);
        q_len = (n_ids + A.queues - 1) / A.queues;
    }

    const int wave = int(threadIdx.x >> 6);
    for (;;) {
        unsigned long long idle = __ballot(!active);
        const int n_idle = __popcll(idle);
        if constexpr (sizeof(R) == 4 && LAYOUT == BH_LAYOUT_CORNER) {
            // a wave with no rays first serves a lone ray of its workgroup that asks for help
            if (n_idle == 64 && A.coop_steps > 0 && coop_help<STORE, MIXED, RK4>(A, P, rs, gr, gt, gp, &coop, lane, wave, iters)) continue;
        }
        if (exhausted != all_q && (n_idle >= A.refill_min || n_idle == 64)) {
            for (int t = 0; t < A.queues && idle; ++t) {
                const int q = (q_own + t) % A.queues;
                if (exhausted & (1u << q)) continue;
                const int need = __popcll(idle);
                const int leader = __builtin_ctzll(idle);
                unsigned got = 0;
                if (lane == leader) got = atomicAdd(&A.counters[q], unsigned(need));
                got = __builtin_amdgcn_readlane(got, leader);
                const int q_begin = q * q_len;
                const int q_end = min(q_begin + q_len, n_ids);
                const long long avail = (long long)q_end - ((long long)q_begin + (long long)got);
                if (avail < need) exhausted |= (1u << q);
                if (avail <= 0) continue;
                const unsigned rank = lane_rank(idle);
                if (!active && (long long)rank < avail) {
                    const int id = q_begin + int(got) + int(rank);
                    int px, py;
                    if (A.ray_list) {
                        const int pix = A.ray_list[id];
                        px = pix % A.W;
                        py = pix / A.W;
                    } else {
                        const int tile = A.tile_order ? A.tile_order[id >> 6] : (id >> 6), w = id & 63;
                        px = (tile % A.tiles_x) * 8 + (w & 7);
                        py = (tile / A.tiles_x) * 8 + (w >> 3);
                    }
                    if (px < A.W && py < A.H) {
                        ray_init(A, px, py, s);
                        active = true;
                    }
                }
                idle = __ballot(!active);
            }
        }
        const unsigned long long live_mask = __ballot(active);
        const int live = __popcll(live_mask);
        if (live == 0) {
            if (exhausted == all_q) {
                if constexpr (sizeof(R) == 4 && LAYOUT == BH_LAYOUT_CORNER) {
                    // out of rays for good: stay while another wave of the workgroup still traces, since its
                    // last ray may ask for help (bounded, like every wait of the hand-over)
                    if (lane == 0) __hip_atomic_fetch_add(&coop.tracing, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (A.coop_steps > 0) {
                        unsigned spins = 0;
                        while (lds_peek(&coop.tracing) > 0 && spins < kCoopSpin) {
                            if (coop_help<STORE, MIXED, RK4>(A, P, rs, gr, gt, gp, &coop, lane, wave, iters)) continue;
                            __builtin_amdgcn_s_sleep(2);
                            ++spins;
                        }
                    }
                }
                break;
            }
            continue;
        }
        if constexpr (sizeof(R) == 4 && LAYOUT == BH_LAYOUT_CORNER) {
            if (live == 1) {   // no refill happened above: the wave serves its last ray with all lanes
                lone_ray<R, STORE, LAYOUT, MIXED, EXACT, RK4>(A, P, rs, gr, gt, gp, s, lane, int(__builtin_ctzll(live_mask)),
                                                  n_steps, n_samples, iters, &coop, wave,
                                                  lone_c + (sizeof(R) == 4 ? wave * kLoneBatch * kLoneStride : 0));
                active = false;
                continue;
            }
        }
        if (live <= 8) __builtin_amdgcn_s_setprio(2);   // a few long rays hold this wave: let them issue first
        else __builtin_amdgcn_s_setprio(0);
        for (int k = 0; k < A.steps_per_sync; ++k) {
            if (active) {
                const int code = ray_step<R, STORE, LAYOUT, MIXED, EXACT, RK4>(A, P, rs, gr, gt, gp, s, n_samples);
                ++n_steps;
                if (code >= 0) {
                    ray_store(A, s, code);
                    active = false;
                }
            }
        }
        iters += unsigned(A.steps_per_sync);
    }
    // wave-reduce the work counters, one atomic per wave
    unsigned long long ws = n_steps, wsm = n_samples;
    for (int off = 32; off > 0; off >>= 1) {
        ws += __shfl_xor(ws, off);
        wsm += __shfl_xor(wsm, off);
    }
    if (lane == 0) {
        atomicAdd(&A.work[0], ws);
        atomicAdd(&A.work[1], wsm);
        atomicAdd(&A.work[2], iters * 64ull);
    }
}


// ---- batched persistent launch ---------------------------------------------------
// Many renders (own size, spin, camera and brick) in one launch.  Waves take whole 8x8
// tiles from per-XCD queues of a global longest-first schedule; a tile's render descriptor is
// loaded with scalar loads, so every render parameter is wave-uniform.  All bricks of a batch
// live in one coefficient buffer (one buffer resource) and one LDS grid table.

struct BatchArgs {
    const RenderArgs *descs;     // per render: scene, brick view, output offset
    const int *slots;            // queue slot -> (render << 20 | tile) or -1 for padding
    const float *grids;          // per brick: r | theta | phi | +inf
    int ng;
    const void *coeffs;
    unsigned total_bytes;
    void *iquv;
    int *steps;
    uint8_t *status;
    unsigned *counters;          // per-queue tile counters (list mode: counters[0] counts taken chunks)
    unsigned long long *work;    // [0] RK2 steps, [1] brick samples, [2] wave iterations x 64
    int queues, q_tiles, steps_per_sync;
    // list mode (the kernels compiled with LIST): render r traces the pixels list[out_off_r .. + list_count[r]),
    // in chunks of 64 rays; chunk_prefix (n_renders + 1 entries, on the device) is the exclusive prefix of the chunks
    const int *list;
    const unsigned *list_count, *chunk_prefix;
    int n_renders;
};

template <typename R, int STORE, int LAYOUT, bool MIXED, bool EXACT, bool RK4, bool LIST = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BH_RENDER_WAVES_PER_EU))) void batch_kernel(
    BatchArgs B) {
    __shared__ R s_grid[kMaxGrid + 8];
    // per wave: a lone ray's batch of sampled coefficients (FP32 kernels)
    __shared__ __align__(16) float lone_c[sizeof(R) == 4 ? 4 * kLoneBatch * kLoneStride : 1];
    for (int i = threadIdx.x; i < B.ng; i += blockDim.x) s_grid[i] = R(B.grids[i]);
    __syncthreads();
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(B.coeffs), (short)0, int(B.total_bytes), int(kRsrcWord3));
    const int lane = threadIdx.x & 63;
    const int q_own = blockIdx.x % B.queues;
    const unsigned all_q = (1u << B.queues) - 1u;
    unsigned exhausted = 0;
    bool active = false;
    RayState<R> s;
    RenderArgs A;
    unsigned n_steps = 0, n_samples = 0;
    unsigned long long iters = 0;
    for (;;) {
        if (!__any(active)) {
            bool took = false;
            if constexpr (LIST) {   // the next chunk of 64 listed rays of some render
                unsigned slot = 0;
                if (lane == 0) slot = atomicAdd(&B.counters[0], 1u);
                slot = __builtin_amdgcn_readfirstlane(slot);
                if (slot < B.chunk_prefix[B.n_renders]) {
                    int lo = 0, hi = B.n_renders - 1;   // the last render whose chunks start at or before slot
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (B.chunk_prefix[mid] <= slot) lo = mid;
                        else hi = mid - 1;
                    }
                    A = B.descs[lo];
                    const size_t esz = sizeof(R);
                    A.iquv = static_cast<char *>(B.iquv) + size_t(A.out_off) * 4 * esz;
                    A.steps = B.steps ? B.steps + A.out_off : nullptr;
                    A.status = B.status ? B.status + A.out_off : nullptr;
                    const unsigned e = (slot - B.chunk_prefix[lo]) * 64u + unsigned(lane);
                    if (e < B.list_count[lo]) {
                        const int pix = B.list[A.out_off + e];
                        ray_init(A, pix % A.W, pix / A.W, s);
                        active = true;
                    }
                    took = true;
                }
                exhausted = all_q;
            }
            while (!took && exhausted != all_q) {
                for (int t = 0; t < B.queues && !took; ++t) {
                    const int q = (q_own + t) % B.queues;
                    if (exhausted & (1u << q)) continue;
                    unsigned slot = 0;
                    if (lane == 0) slot = atomicAdd(&B.counters[q], 1u);
                    slot = __builtin_amdgcn_readfirstlane(slot);
                    if (slot >= unsigned(B.q_tiles)) {
                        exhausted |= (1u << q);
                        continue;
                    }
                    const int e = B.slots[q * B.q_tiles + int(slot)];
                    if (e < 0) continue;
                    const int rix = e >> 20, tile = e & 0xfffff;
                    A = B.descs[rix];
                    const size_t esz = sizeof(R);
                    A.iquv = static_cast<char *>(B.iquv) + size_t(A.out_off) * 4 * esz;
                    A.steps = B.steps ? B.steps + A.out_off : nullptr;
                    A.status = B.status ? B.status + A.out_off : nullptr;
                    const int px = (tile % A.tiles_x) * 8 + (lane & 7);
                    const int py = (tile / A.tiles_x) * 8 + (lane >> 3);
                    if (px < A.W && py < A.H) {
                        ray_init(A, px, py, s);
                        active = true;
                    }
                    took = true;
                }
            }
            if (!took) break;
        }
        const unsigned long long live_mask = __ballot(active);
        const int live = __popcll(live_mask);
        const R *gr = s_grid + A.grid_off;
        const R *gt = gr + A.nr;
        const R *gp = gt + A.nt;
        const HotParams<R> P = loop_params<R>(A);
        if constexpr (sizeof(R) == 4 && LAYOUT == BH_LAYOUT_CORNER) {
            if (live == 1) {
                lone_ray<R, STORE, LAYOUT, MIXED, EXACT, RK4>(A, P, rs, gr, gt, gp, s, lane, int(__builtin_ctzll(live_mask)),
                                                  n_steps, n_samples, iters, nullptr, 0,
                                                  lone_c + (sizeof(R) == 4 ? int(threadIdx.x >> 6) * kLoneBatch * kLoneStride : 0));
                active = false;
                continue;
            }
        }
        if (live <= 8) __builtin_amdgcn_s_setprio(2);
        else __builtin_amdgcn_s_setprio(0);
        for (int k = 0; k < B.steps_per_sync; ++k) {
            if (active) {
                const int code = ray_step<R, STORE, LAYOUT, MIXED, EXACT, RK4>(A, P, rs, gr, gt, gp, s, n_samples);
                ++n_steps;
                if (code >= 0) {
                    ray_store(A, s, code);
                    active = false;
                }
            }
        }
        iters += unsigned(B.steps_per_sync);
    }
    unsigned long long ws = n_steps, wsm = n_samples;
    for (int off = 32; off > 0; off >>= 1) {
        ws += __shfl_xor(ws, off);
        wsm += __shfl_xor(wsm, off);
    }
    if (lane == 0) {
        atomicAdd(&B.work[0], ws);
        atomicAdd(&B.work[1], wsm);
        atomicAdd(&B.work[2], iters * 64ull);
    }
}

// ---- host side ------------------------------------------------------------------

static int camera_constants(const bh_scene_desc &d, CamConst &C) {
    const double a = d.spin;
    const double th = d.cam_theta < kEpsTheta ? kEpsTheta
                      : (d.cam_theta > 3.14159265358979323846 - kEpsTheta ? 3.14159265358979323846 - kEpsTheta
                                                                           : d.cam_theta);
    const double r = d.cam_r;
    const double s = std::sin(th), c = std::cos(th);
    const double sig = r * r + a * a * c * c;
    const double dlt = r * r - 2.0 * r + a * a;
    const double A = (r * r + a * a) * (r * r + a * a) - a * a * dlt * s * s;
    if (!(dlt > 0.0)) return set_error(BH_ERR_VALUE, "ZAMO tetrad must be outside the outer horizon");
    if (std::fabs(s) < kEpsTheta) return set_error(BH_ERR_VALUE, "ZAMO tetrad is ill-conditioned at the coordinate pole");
    const double s2 = s * s;
    C.g00 = -(1.0 - 2.0 * r / sig);
    C.g03 = -2.0 * a * r * s2 / sig;
    C.g11 = sig / dlt;
    C.g22 = sig;
    C.g33 = A * s2 / sig;
    const double lapse = std::sqrt(sig * dlt / A);
    const double omega = 2.0 * a * r / A;
    C.e_t0 = 1.0 / lapse;
    C.e_t3 = omega / lapse;
    C.e_r1 = std::sqrt(dlt / sig);
    C.e_th2 = 1.0 / std::sqrt(sig);
    C.e_ph3 = 1.0 / std::sqrt(C.g33);
    const double deg = 3.14159265358979323846 / 180.0;   // math.radians multiplies by pi/180
    C.tan_y = std::tan(0.5 * (d.fov_y_degrees * deg));
    C.tan_x = (double(d.width) / double(d.height)) * C.tan_y;
    C.roll = d.roll_degrees != 0.0;
    const double roll = d.roll_degrees * deg;
    C.cos_roll = std::cos(roll);
    C.sin_roll = std::sin(roll);
    C.r = r;
    C.th = th;
    C.ph = d.cam_phi;
    return BH_OK;
}

// Bardeen's critical curve of a distant observer at inclination inc: spherical photon orbits
// r_p in [r_ph+, r_ph-] give xi = L/E and eta = Q/E^2; alpha = -xi / sin i,
// beta = +-sqrt(eta + a^2 cos^2 i - xi^2 cot^2 i).
static void critical_curve(double a, double inc, std::vector<double> &al, std::vector<double> &be) {
    const double aa = std::fabs(a) > 1e-9 ? a : 1e-9;
    const double r1 = 2.0 * (1.0 + std::cos(2.0 / 3.0 * std::acos(-std::fabs(aa))));
    const double r2 = 2.0 * (1.0 + std::cos(2.0 / 3.0 * std::acos(std::fabs(aa))));
    const double si = std::sin(inc), ci = std::cos(inc);
    const int n = 1024;
    for (int k = 0; k < n; ++k) {
        const double r = r2 + (r1 - r2) * (k + 0.5) / n;
        const double xi = (r * r * (3.0 - r) - aa * aa * (r + 1.0)) / (aa * (r - 1.0));
        const double eta = r * r * r * (4.0 * aa * aa - r * (r - 3.0) * (r - 3.0)) / (aa * aa * (r - 1.0) * (r - 1.0));
        const double b2 = eta + aa * aa * ci * ci - xi * xi * (ci / si) * (ci / si);
        if (b2 < 0.0) continue;
        al.push_back(-xi / si);
        be.push_back(std::sqrt(b2));
    }
}

// Tile schedule: 8x8 tiles sorted by the distance of their central ray from the critical curve
// (rays there circle the hole and need the most steps), dealt round-robin over the queues so every
// queue starts with its longest rays.
// Squared distance of each 8x8 tile's central ray from the critical curve in Bardeen's image plane.
static std::vector<float> tile_distances(const bh_scene_desc &d, const CamConst &C, int tiles_x, int tiles_y) {
    const int n_tiles = tiles_x * tiles_y;
    std::vector<double> al, be;
    const double inc = C.th;
    if (std::sin(inc) > 1e-3) critical_curve(d.spin, inc, al, be);
    std::vector<float> dist(n_tiles, 0.f);
    const double a = d.spin, s0 = std::sin(C.th), c0 = std::cos(C.th);
    for (int t = 0; t < n_tiles; ++t) {
        const double px = (t % tiles_x) * 8 + 3.5, py = (t / tiles_x) * 8 + 3.5;
        const double ndc_x = 2.0 * (px + 0.5) / d.width - 1.0, ndc_y = 1.0 - 2.0 * (py + 0.5) / d.height;
        double n_r = -1.0, n_th = -ndc_y * C.tan_y, n_ph = ndc_x * C.tan_x;
        if (C.roll) {
            const double t0 = C.cos_roll * n_th - C.sin_roll * n_ph, t1 = C.sin_roll * n_th + C.cos_roll * n_ph;
            n_th = t0;
            n_ph = t1;
        }
        const double nrm = std::sqrt(n_r * n_r + n_th * n_th + n_ph * n_ph);
        const double pc_t = C.e_t0, pc_th = n_th / nrm * C.e_th2, pc_ph = C.e_t3 + n_ph / nrm * C.e_ph3;
        const double E = -(C.g00 * pc_t + C.g03 * pc_ph), L = C.g03 * pc_t + C.g33 * pc_ph, pth = C.g22 * pc_th;
        const double Q = pth * pth + c0 * c0 * (-a * a * E * E + L * L / (s0 * s0));
        const double xi = L / E, bb = std::sqrt(std::fmax(Q / (E * E) + a * a * c0 * c0 - xi * xi * (c0 / s0) * (c0 / s0), 0.0));
        const double alpha = -xi / std::sin(inc), beta = bb;
        double best = 1e300;
        for (size_t k = 0; k < al.size(); ++k) {
            const double dx = alpha - al[k], dy = beta - be[k];
            best = std::fmin(best, dx * dx + dy * dy);
        }
        dist[t] = al.empty() ? float(t) : float(best);
    }
    return dist;
}

// Queue slots -> tiles: queue q holds slots [q * q_tiles, (q + 1) * q_tiles).
static std::vector<int> schedule_of(const bh_scene_desc &d, const CamConst &C, int tiles_x, int tiles_y, int queues,
                                    int q_tiles) {
    const int n_tiles = tiles_x * tiles_y;
    const std::vector<float> dist = tile_distances(d, C, tiles_x, tiles_y);
    std::vector<int> idx(n_tiles);
    for (int t = 0; t < n_tiles; ++t) idx[t] = t;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return dist[x] < dist[y]; });
    std::vector<int> order(size_t(queues) * q_tiles, 0);
    std::vector<int> fill(queues, 0);
    for (int k = 0; k < n_tiles; ++k) {
        const int q = k % queues;
        order[size_t(q) * q_tiles + fill[q]++] = idx[k];
    }
    // queue slots past the dealt tiles repeat a tile index that lies outside the image
    for (int q = 0; q < queues; ++q)
        for (int j = fill[q]; j < q_tiles; ++j) order[size_t(q) * q_tiles + j] = n_tiles;
    return order;
}

static int build_schedule(bh_ctx *ctx, const bh_scene_desc &d, const CamConst &C, int tiles_x, int tiles_y,
                          int queues, int q_tiles) {
    std::vector<double> key = {d.spin, d.cam_r, d.cam_theta, d.cam_phi, d.fov_y_degrees, d.roll_degrees,
                               double(d.width), double(d.height), double(queues)};
    if (ctx->d_order && ctx->order_key == key) return BH_OK;
    const std::vector<int> order = schedule_of(d, C, tiles_x, tiles_y, queues, q_tiles);
    const size_t bytes = order.size() * sizeof(int);
    if (bytes > ctx->order_cap) {
        if (ctx->d_order) (void)hipFree(ctx->d_order);
        ctx->d_order = nullptr;
        ctx->order_cap = 0;
        if (hipMalloc(&ctx->d_order, bytes) != hipSuccess) return set_error(BH_ERR_MEMORY, "schedule allocation failed");
        ctx->order_cap = bytes;
    }
    BH_HIP(hipMemcpyAsync(ctx->d_order, order.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    ctx->order_key = key;
    return BH_OK;
}

static int validate_scene(const bh_scene_desc &d) {
    if (!std::isfinite(d.spin) || std::fabs(d.spin) >= 1.0)
        return set_error(BH_ERR_VALUE, "Kerr spin must be finite and satisfy |a| < 1 for this integrator");
    if (d.width <= 0 || d.height <= 0) return set_error(BH_ERR_VALUE, "width and height must be positive");
    if (int64_t(d.width) * d.height > (int64_t(1) << 30)) return set_error(BH_ERR_VALUE, "image is too large");
    if (d.max_steps < 1) return set_error(BH_ERR_VALUE, "max_steps must be at least 1");
    if (!std::isfinite(d.step) || d.step <= 0.0) return set_error(BH_ERR_VALUE, "step must be finite and positive");
    if (d.step_mode != 0 && d.step_mode != 1) return set_error(BH_ERR_VALUE, "step_mode must be 0 (fixed) or 1 (adaptive)");
    if (!std::isfinite(d.cam_r) || !std::isfinite(d.cam_theta) || !std::isfinite(d.cam_phi) ||
        !std::isfinite(d.fov_y_degrees) || !std::isfinite(d.roll_degrees))
        return set_error(BH_ERR_VALUE, "camera parameters must be finite");
    if (!(d.fov_y_degrees > 0.0 && d.fov_y_degrees < 180.0))
        return set_error(BH_ERR_VALUE, "fov_y_degrees must lie in (0, 180)");
    if (std::isnan(d.escape_radius) || std::isnan(d.affine_to_length_cm) || std::isnan(d.max_optical_depth))
        return set_error(BH_ERR_VALUE, "escape_radius, affine_to_length_cm and max_optical_depth must not be NaN");
    if (d.transfer != BH_TRANSFER_RK2 && d.transfer != BH_TRANSFER_EXACT)
        return set_error(BH_ERR_VALUE, "transfer must be BH_TRANSFER_RK2 or BH_TRANSFER_EXACT");
    if (d.integrator != BH_INTEGRATOR_RK2 && d.integrator != BH_INTEGRATOR_RK4)
        return set_error(BH_ERR_VALUE, "integrator must be BH_INTEGRATOR_RK2 or BH_INTEGRATOR_RK4");
    if (d.sampling != BH_SAMPLING_NATIVE && d.sampling != BH_SAMPLING_NUMPY32)
        return set_error(BH_ERR_VALUE, "sampling must be BH_SAMPLING_NATIVE or BH_SAMPLING_NUMPY32");
    return BH_OK;
}

static GridHint make_hint(const float *g, int n, bool periodic, bool allow_log) {
    GridHint h{};
    const double g0 = g[0], gl = g[n - 1];
    bool geometric = allow_log && g0 > 0.0;
    if (geometric && n > 2) {
        const double q = std::log2(double(g[1]) / g0);
        for (int i = 1; i < n && geometric; ++i) {
            double qi = std::log2(double(g[i]) / double(g[i - 1]));
            if (std::fabs(qi - q) > 1e-3 * std::fabs(q)) geometric = false;
        }
        if (std::fabs(q) < 1e-12) geometric = false;
    }
    if (periodic) geometric = false;
    h.first = g[0];
    h.last = g[n - 1];
    if (geometric) {
        const double l0 = std::log2(g0), l1 = std::log2(gl);
        h.log = 1;
        h.scale = float((n - 1) / (l1 - l0));
        h.offset = float(-l0 * (n - 1) / (l1 - l0));
    } else {
        h.log = 0;
        const double span = periodic ? (gl - g0) * double(n) / double(n - 1) : (gl - g0);
        const double cells = periodic ? double(n) : double(n - 1);
        h.scale = float(cells / span);
        h.offset = float(-g0 * cells / span);
    }
    return h;
}

template <typename T>
static void transpose_soa(const float *aos, int nr, int nt, int np, T *soa, T (*cvt)(float)) {
    const size_t plane = size_t(nr) * nt * np;
    for (size_t e = 0; e < plane; ++e)
        for (int k = 0; k < kCoeffs; ++k) soa[size_t(k) * plane + e] = cvt(aos[e * kCoeffs + k]);
}

// Corner records: cell (ir, it, ip) holds corners (ir+dr, it+dt, (ip+dp) mod np), corner q = dr*4 + dt*2 + dp,
// corner-major (value (q, k) at q * 11 + k) or, for bf16 records, channel-major (at k * 8 + q).
template <typename T>
static void pack_corners(const float *aos, int nr, int nt, int np, T *out, T (*cvt)(float), bool channel_major) {
    size_t o = 0;
    for (int ir = 0; ir < nr - 1; ++ir)
        for (int it = 0; it < nt - 1; ++it)
            for (int ip = 0; ip < np; ++ip, o += 8 * kCoeffs)
                for (int q = 0; q < 8; ++q) {
                    const int jr = ir + (q >> 2), jt = it + ((q >> 1) & 1), jp = (ip + (q & 1)) % np;
                    const float *src = aos + ((size_t(jr) * nt + jt) * np + jp) * kCoeffs;
                    for (int k = 0; k < kCoeffs; ++k) out[o + (channel_major ? k * 8 + q : q * kCoeffs + k)] = cvt(src[k]);
                }
}

static float cvt_f32(float x) { return x; }
static uint16_t cvt_bf16(float x) {   // round to nearest even
    uint32_t u;
    std::memcpy(&u, &x, 4);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return uint16_t(u >> 16);
}
static uint16_t cvt_f16(float x) {
    _Float16 h = (_Float16)x;
    uint16_t u;
    std::memcpy(&u, &h, 2);
    return u;
}

typedef void (*RenderFn)(RenderArgs);
typedef void (*BatchFn)(BatchArgs);

// Kernel table: precision x storage x layout x transfer rule.  The exact Stokes step is a template
// parameter, so the RK2 kernels carry none of its code or registers.
template <typename R, int STORE, int LAYOUT, bool MIXED, bool EXACT, bool RK4>
struct KernelPair {
    static RenderFn render() { return render_kernel<R, STORE, LAYOUT, MIXED, EXACT, RK4>; }
    static BatchFn batch() { return batch_kernel<R, STORE, LAYOUT, MIXED, EXACT, RK4>; }
};

template <bool BATCH, typename R, int STORE, int LAYOUT, bool MIXED, bool EXACT, bool RK4>
static void *kernel_of() {
    using K = KernelPair<R, STORE, LAYOUT, MIXED, EXACT, RK4>;
    return BATCH ? reinterpret_cast<void *>(K::batch()) : reinterpret_cast<void *>(K::render());
}

// W: kStoreWide for the 64-bit-addressed corner kernels (corner bricks of 2 GiB and more), else 0.
template <bool BATCH, int LAYOUT, bool EXACT, bool RK4, int W = 0>
static void *pick_fp32(int storage, bool mixed) {
    switch (storage) {
        case BH_STORE_F32:
            return mixed ? kernel_of<BATCH, float, BH_STORE_F32 | W, LAYOUT, true, EXACT, RK4>()
                         : kernel_of<BATCH, float, BH_STORE_F32 | W, LAYOUT, false, EXACT, RK4>();
        case BH_STORE_BF16:
            return mixed ? kernel_of<BATCH, float, BH_STORE_BF16 | W, LAYOUT, true, EXACT, RK4>()
                         : kernel_of<BATCH, float, BH_STORE_BF16 | W, LAYOUT, false, EXACT, RK4>();
        case BH_STORE_F16:
            return mixed ? kernel_of<BATCH, float, BH_STORE_F16 | W, LAYOUT, true, EXACT, RK4>()
                         : kernel_of<BATCH, float, BH_STORE_F16 | W, LAYOUT, false, EXACT, RK4>();
    }
    return nullptr;
}

// Kernels are compiled per integrator: the RK4 step's registers would otherwise weigh on every RK2
// kernel (a uniform branch on the rule cost the mixed batch 14 %).  RK4 is built for the FP64 render
// and the corner-packed FP32 / mixed renders (the layouts the reference renderer's drop-in uploads).
template <bool BATCH, bool EXACT, bool RK4>
static void *pick_any(int precision, int storage, int layout, bool wide) {
    if (precision == BH_PREC_FP64)
        return (storage == BH_STORE_F32 && layout == BH_LAYOUT_SOA)
                   ? kernel_of<BATCH, double, BH_STORE_F32, BH_LAYOUT_SOA, false, EXACT, RK4>()
                   : nullptr;
    const bool mixed = precision == BH_PREC_MIXED;
    if (layout == BH_LAYOUT_CORNER)
        return wide ? pick_fp32<BATCH, BH_LAYOUT_CORNER, EXACT, RK4, kStoreWide>(storage, mixed)
                    : pick_fp32<BATCH, BH_LAYOUT_CORNER, EXACT, RK4>(storage, mixed);
    if constexpr (RK4) return nullptr;
    else return pick_fp32<BATCH, BH_LAYOUT_SOA, EXACT, false>(storage, mixed);
}

template <bool BATCH>
static void *pick_rule(int precision, int storage, int layout, bool exact, bool rk4, bool wide) {
    if (rk4) return exact ? pick_any<BATCH, true, true>(precision, storage, layout, wide) : pick_any<BATCH, false, true>(precision, storage, layout, wide);
    return exact ? pick_any<BATCH, true, false>(precision, storage, layout, wide) : pick_any<BATCH, false, false>(precision, storage, layout, wide);
}

static RenderFn pick_kernel(int precision, int storage, int layout, bool exact, bool rk4, bool wide) {
    return reinterpret_cast<RenderFn>(pick_rule<false>(precision, storage, layout, exact, rk4, wide));
}

static BatchFn pick_batch_kernel(int precision, int storage, int layout, bool exact, bool rk4, bool wide) {
    return reinterpret_cast<BatchFn>(pick_rule<true>(precision, storage, layout, exact, rk4, wide));
}

// Corner bricks (or a batch's merged corner bricks) of 2 GiB and more take the 64-bit-addressed kernels.
static bool wide_corner(int layout, size_t bytes) { return layout == BH_LAYOUT_CORNER && bytes >= (size_t(1) << 31); }

static int setup_args(const bh_scene_desc &sc, const BrickDev &b, RenderArgs &A) {
    int st;
    if ((st = camera_constants(sc, A.cam))) return st;
    const double a = sc.spin;
    A.a = a;
    A.r_plus = 1.0 + std::sqrt(1.0 - a * a);
    A.r_cap = A.r_plus * (1.0 + 2.0e-4);
    A.r_island = 1.5 * A.r_plus;
    A.th_island = kIslandTheta;
    A.pth_island = kIslandPtheta;
    A.step = sc.step;
    A.adaptive = sc.step_mode;
    A.r_esc = sc.escape_radius;
    A.L = sc.affine_to_length_cm;
    A.tau_max = sc.max_optical_depth;
    A.W = sc.width;
    A.H = sc.height;
    A.max_steps = sc.max_steps;
    A.lerp32 = sc.sampling == BH_SAMPLING_NUMPY32;
    A.step_f = float(A.step);
    A.r_plus_f = float(A.r_plus);
    A.r_cap_f = float(A.r_cap);
    A.r_esc_f = float(A.r_esc);
    A.L_f = float(A.L);
    A.tau_max_f = float(A.tau_max);
    A.coeffs = b.coeffs;
    A.nr = b.nr;
    A.nt = b.nt;
    A.np = b.np;
    const unsigned esz = b.storage == BH_STORE_F32 ? 4u : 2u;
    A.plane_bytes = unsigned(b.plane * esz);
    A.total_bytes = unsigned(std::min(b.bytes, size_t(0xffffffffu)));   // the SoA buffer resource's bound
    A.grids = b.grids;
    A.hr = b.hr;
    A.ht = b.ht;
    A.hp = b.hp;
    A.tiles_x = (A.W + 7) / 8;
    return BH_OK;
}

// ---- memory of the no-allocation render ------------------------------------------------------
// bh_render_device_noalloc keeps a launch's ray-queue counters (kCounterWords words) and its tile schedule
// (at most kScratchTiles tiles, padded to a multiple of the queue count) in this array instead of memory
// the context allocates.  Each GPU holds its own copy for the life of the process (257 KiB).  In host
// code `scratch` names a host-side stand-in, never the device memory: its device address comes from
// hipGetSymbolAddress, once per device, and only that address is handed to kernels and copies.
constexpr int kScratchTiles = 65536;   // 8x8 tiles: 2048 x 2048 pixels
constexpr int kScratchWords = kCounterWords + kScratchTiles + kMaxQueues;
static __device__ __align__(256) unsigned int scratch[kScratchWords];
static_assert(kScratchTiles == BH_NOALLOC_MAX_TILES && sizeof(scratch) == BH_NOALLOC_SCRATCH_BYTES,
              "bhgrrt.h states the size of the no-allocation scratch array");

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static unsigned *g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the context's device (which ctx_check has made current).
static int scratch_address(bh_ctx *ctx, unsigned **out) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return set_error(BH_ERR_DEVICE, "device %d out of range", ctx->device);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_addr[ctx->device]) {
        void *ptr = nullptr;
        BH_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return set_error(BH_ERR_DEVICE, "no device address for the render scratch array");
        g_scratch_addr[ctx->device] = static_cast<unsigned *>(ptr);
    }
    *out = g_scratch_addr[ctx->device];
    return BH_OK;
}

// Validate, set up and enqueue one render launch.  List mode (ray_list non-null): only the pixels
// ray_list[0 .. *n_list) are traced, n_list read on the device; max_list bounds it for the grid size.
// scratch (non-null): the launch's counters go to scratch[0, kCounterWords) and its tile schedule right after
// them, copied from the host, instead of the context's counters and its cached device schedule.
static int render_enqueue(bh_ctx *ctx, const bh_scene_desc *scene, int32_t brick_id, int32_t precision, void *d_iquv,
                          int32_t *d_steps, uint8_t *d_status, const int *ray_list, const unsigned *n_list,
                          int max_list, bool timed, unsigned *scratch = nullptr) {
    int st;
    if (!scene || !d_iquv) return set_error(BH_ERR_VALUE, "scene and iquv must not be NULL");
    if ((st = validate_scene(*scene))) return st;
    if (brick_id < 0 || size_t(brick_id) >= ctx->bricks.size() || !ctx->bricks[brick_id].live)
        return set_error(BH_ERR_STATE, "unknown brick id %d", brick_id);
    if (precision != BH_PREC_FP64 && precision != BH_PREC_FP32 && precision != BH_PREC_MIXED)
        return set_error(BH_ERR_VALUE, "precision must be BH_PREC_FP64, BH_PREC_FP32 or BH_PREC_MIXED");
    const BrickDev &b = ctx->bricks[brick_id];
    const bool rk4 = scene->integrator == BH_INTEGRATOR_RK4;
    RenderFn fn = pick_kernel(precision, b.storage, b.layout, scene->transfer == BH_TRANSFER_EXACT, rk4,
                              wide_corner(b.layout, b.bytes));
    if (!fn)
        return set_error(BH_ERR_VALUE, precision == BH_PREC_FP64
                                           ? "the FP64 render needs a float32 brick in the SoA layout"
                                           : "the RK4 integrator of the FP32 and mixed renders needs the corner layout");
    RenderArgs A{};
    if ((st = setup_args(*scene, b, A))) return st;
    A.iquv = d_iquv;
    A.steps = d_steps;
    A.status = d_status;
    unsigned *counters = scratch ? scratch : ctx->d_counters;
    A.counters = counters;
    A.work = reinterpret_cast<unsigned long long *>(counters + kWorkOffset);
    const Tuning &u = ctx->tuning[precision];
    A.tiles_x = (A.W + 7) / 8;
    const int tiles_y = (A.H + 7) / 8;
    const int n_tiles = A.tiles_x * tiles_y;
    if (scratch && n_tiles > kScratchTiles)
        return set_error(BH_ERR_VALUE,
                         "a %d x %d image has %d 8x8 tiles; the no-allocation render holds at most %d (2048 x 2048 "
                         "pixels)", A.W, A.H, n_tiles, kScratchTiles);
    A.tile_order = nullptr;
    if (ray_list) {
        A.ray_list = ray_list;
        A.n_list = n_list;
        A.queues = u.queues;
        A.n_ids = max_list;   // bound only: the kernel reads the count
        A.q_len = 0;
    } else {
        A.queues = u.queues < n_tiles ? u.queues : n_tiles;
        const int q_tiles = (n_tiles + A.queues - 1) / A.queues;
        A.q_len = q_tiles * 64;
        A.n_ids = A.queues * A.q_len;   // padded id space: every queue slot is issued; padding maps outside the image
        if (u.order == 1 && scratch) {
            const std::vector<int> order = schedule_of(*scene, A.cam, A.tiles_x, tiles_y, A.queues, q_tiles);
            int *d_sched = reinterpret_cast<int *>(scratch + kCounterWords);
            BH_HIP(hipMemcpyAsync(d_sched, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice,
                                  ctx->stream));
            BH_HIP(hipStreamSynchronize(ctx->stream));   // the host vector dies with this scope
            A.tile_order = d_sched;
        } else if (u.order == 1) {
            if ((st = build_schedule(ctx, *scene, A.cam, A.tiles_x, tiles_y, A.queues, q_tiles))) return st;
            A.tile_order = ctx->d_order;
        }
    }
    A.steps_per_sync = u.steps_per_sync;
    A.refill_min = u.refill_min;
    A.coop_steps = u.coop_steps;
    int per_cu = u.blocks_per_cu;
    if (per_cu <= 0) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, u.block_threads, 0) != hipSuccess || nb <= 0) nb = 1;
        per_cu = nb;
    }
    int grid = ctx->cus * per_cu;
    const int max_useful = (A.n_ids + u.block_threads - 1) / u.block_threads;
    if (grid > max_useful) grid = max_useful;
    if (grid < 1) grid = 1;
    if ((st = fill_u32(ctx, counters, 0u, kCounterWords))) return st;
    ctx->last_counters = counters;
    if (timed) begin_timing(ctx);
    void *args[] = {&A};
    BH_HIP(hipLaunchKernel(reinterpret_cast<const void *>(fn), dim3(grid), dim3(u.block_threads), args, 0, ctx->stream));
    if (timed) end_timing(ctx, 1);
    return BH_OK;
}

}  // namespace bh

using namespace bh;

extern "C" {

int bh_brick_upload(bh_ctx *ctx, const float *coeffs_aos, const float *r_grid, int32_t nr, const float *theta_grid,
                    int32_t nt, const float *phi_grid, int32_t np, int32_t storage, int32_t layout, int32_t *brick_id) {
    clear_error();
    int st = ctx_check(ctx);
    if (st) return st;
    if (!coeffs_aos || !r_grid || !theta_grid || !phi_grid || !brick_id)
        return set_error(BH_ERR_VALUE, "brick arrays must not be NULL");
    if (nr < 2 || nt < 2 || np < 2) return set_error(BH_ERR_VALUE, "brick grids must contain at least 2 values");
    if (nr + nt + np > kMaxGrid) return set_error(BH_ERR_VALUE, "brick grids hold more than %d values", kMaxGrid);
    if (storage != BH_STORE_F32 && storage != BH_STORE_BF16 && storage != BH_STORE_F16)
        return set_error(BH_ERR_VALUE, "storage must be BH_STORE_F32, BH_STORE_BF16 or BH_STORE_F16");
    if (layout != BH_LAYOUT_SOA && layout != BH_LAYOUT_CORNER)
        return set_error(BH_ERR_VALUE, "layout must be BH_LAYOUT_SOA or BH_LAYOUT_CORNER");
    const float *gs[3] = {r_grid, theta_grid, phi_grid};
    const int ns[3] = {nr, nt, np};
    const char *names[3] = {"r", "theta", "phi"};
    for (int k = 0; k < 3; ++k) {
        for (int i = 0; i < ns[k]; ++i) {
            if (!std::isfinite(gs[k][i])) return set_error(BH_ERR_VALUE, "%s grid must contain only finite values", names[k]);
            if (i && !(gs[k][i] > gs[k][i - 1])) return set_error(BH_ERR_VALUE, "%s grid must be strictly increasing", names[k]);
        }
    }
    const size_t plane = size_t(nr) * nt * np;
    const size_t esz = storage == BH_STORE_F32 ? 4 : 2;
    const size_t cells = size_t(nr - 1) * (nt - 1) * np;
    const size_t nbytes = layout == BH_LAYOUT_SOA ? plane * kCoeffs * esz : cells * 8 * kCoeffs * esz;
    // SoA planes are read through one buffer resource (32-bit offsets); corner records through 64-bit addresses
    if (layout == BH_LAYOUT_SOA && nbytes >= (size_t(1) << 31))
        return set_error(BH_ERR_VALUE, "SoA coefficient brick exceeds 2 GiB (use the corner layout)");
    for (size_t i = 0; i < plane * kCoeffs; ++i)
        if (!std::isfinite(coeffs_aos[i])) return set_error(BH_ERR_VALUE, "coefficient brick contains non-finite values");
    std::vector<unsigned char> host(nbytes);
    if (layout == BH_LAYOUT_SOA) {
        if (storage == BH_STORE_F32) transpose_soa<float>(coeffs_aos, nr, nt, np, reinterpret_cast<float *>(host.data()), cvt_f32);
        else if (storage == BH_STORE_BF16) transpose_soa<uint16_t>(coeffs_aos, nr, nt, np, reinterpret_cast<uint16_t *>(host.data()), cvt_bf16);
        else transpose_soa<uint16_t>(coeffs_aos, nr, nt, np, reinterpret_cast<uint16_t *>(host.data()), cvt_f16);
    } else {
        if (storage == BH_STORE_F32) pack_corners<float>(coeffs_aos, nr, nt, np, reinterpret_cast<float *>(host.data()), cvt_f32, false);
        else if (storage == BH_STORE_BF16) pack_corners<uint16_t>(coeffs_aos, nr, nt, np, reinterpret_cast<uint16_t *>(host.data()), cvt_bf16, true);
        else pack_corners<uint16_t>(coeffs_aos, nr, nt, np, reinterpret_cast<uint16_t *>(host.data()), cvt_f16, false);
    }
    BrickDev b;
    b.storage = storage;
    b.layout = layout;
    b.bytes = nbytes;
    b.nr = nr;
    b.nt = nt;
    b.np = np;
    b.plane = plane;
    if (hipMalloc(&b.coeffs, host.size()) != hipSuccess) return set_error(BH_ERR_MEMORY, "brick allocation failed");
    if (hipMalloc(&b.grids, sizeof(float) * (nr + nt + np)) != hipSuccess) {
        (void)hipFree(b.coeffs);
        return set_error(BH_ERR_MEMORY, "brick grid allocation failed");
    }
    std::vector<float> g(nr + nt + np);
    std::memcpy(g.data(), r_grid, sizeof(float) * nr);
    std::memcpy(g.data() + nr, theta_grid, sizeof(float) * nt);
    std::memcpy(g.data() + nr + nt, phi_grid, sizeof(float) * np);
    BH_HIP(hipMemcpy(b.coeffs, host.data(), host.size(), hipMemcpyHostToDevice));
    BH_HIP(hipMemcpy(b.grids, g.data(), sizeof(float) * g.size(), hipMemcpyHostToDevice));
    b.hgrids = g;
    b.hr = make_hint(r_grid, nr, false, true);
    b.ht = make_hint(theta_grid, nt, false, false);   // the kernels' theta and phi guesses are linear
    b.hp = make_hint(phi_grid, np, true, false);
    b.live = true;
    ctx->bricks.push_back(b);
    *brick_id = int32_t(ctx->bricks.size() - 1);
    return BH_OK;
}

int bh_brick_release(bh_ctx *ctx, int32_t brick_id) {
    int st = ctx_check(ctx);
    if (st) return st;
    if (brick_id < 0 || size_t(brick_id) >= ctx->bricks.size() || !ctx->bricks[brick_id].live)
        return set_error(BH_ERR_STATE, "unknown brick id %d", brick_id);
    BH_HIP(hipStreamSynchronize(ctx->stream));
    BrickDev &b = ctx->bricks[brick_id];
    (void)hipFree(b.coeffs);
    (void)hipFree(b.grids);
    b = BrickDev();
    return BH_OK;
}

int bh_ctx_last_work(bh_ctx *ctx, uint64_t *steps, uint64_t *samples, uint64_t *lane_slots) {
    int st = ctx_check(ctx);
    if (st) return st;
    unsigned long long w[3] = {0, 0, 0};
    const unsigned *c = ctx->last_counters ? ctx->last_counters : ctx->d_counters;
    BH_HIP(hipMemcpyAsync(w, c + kWorkOffset, sizeof(w), hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    if (steps) *steps = w[0];
    if (samples) *samples = w[1];
    if (lane_slots) *lane_slots = w[2];
    return BH_OK;
}

int bh_ctx_last_coop(bh_ctx *ctx, uint64_t *split_rays, uint64_t *timeouts) {
    int st = ctx_check(ctx);
    if (st) return st;
    unsigned long long w[2] = {0, 0};
    const unsigned *c = ctx->last_counters ? ctx->last_counters : ctx->d_counters;
    BH_HIP(hipMemcpyAsync(w, c + kWorkOffset + 6, sizeof(w), hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    if (split_rays) *split_rays = w[0];
    if (timeouts) *timeouts = w[1];
    return BH_OK;
}

int bh_ctx_set_tuning(bh_ctx *ctx, int32_t precision, const bh_tuning *t) {
    int st = ctx_check(ctx);
    if (st) return st;
    if (precision < 0 || precision > 2 || !t) return set_error(BH_ERR_VALUE, "bad precision or tuning");
    Tuning u = ctx->tuning[precision];
    if (t->block_threads) u.block_threads = t->block_threads;
    if (t->blocks_per_cu >= 0) u.blocks_per_cu = t->blocks_per_cu;
    if (t->steps_per_sync) u.steps_per_sync = t->steps_per_sync;
    if (t->refill_min) u.refill_min = t->refill_min;
    if (t->queues) u.queues = t->queues;
    if (t->order >= 0) u.order = t->order;
    if (t->coop_steps >= 0) u.coop_steps = t->coop_steps;
    if (u.block_threads < 64 || u.block_threads > 256 || u.block_threads % 64)
        return set_error(BH_ERR_VALUE, "block_threads must be 64, 128, 192 or 256");
    if (u.steps_per_sync < 1 || u.steps_per_sync > 4096) return set_error(BH_ERR_VALUE, "steps_per_sync out of range");
    if (u.refill_min < 1 || u.refill_min > 64) return set_error(BH_ERR_VALUE, "refill_min must lie in 1..64");
    if (u.queues < 1 || u.queues > kMaxQueues) return set_error(BH_ERR_VALUE, "queues must lie in 1..%d", kMaxQueues);
    if (u.blocks_per_cu > 64) return set_error(BH_ERR_VALUE, "blocks_per_cu out of range");
    if (u.order != 0 && u.order != 1) return set_error(BH_ERR_VALUE, "order must be 0 or 1");
    if (u.coop_steps > (1 << 20)) return set_error(BH_ERR_VALUE, "coop_steps out of range");
    ctx->tuning[precision] = u;
    return BH_OK;
}

int bh_ctx_get_tuning(bh_ctx *ctx, int32_t precision, bh_tuning *t) {
    int st = ctx_check(ctx);
    if (st) return st;
    if (precision < 0 || precision > 2 || !t) return set_error(BH_ERR_VALUE, "bad precision or tuning");
    const Tuning &u = ctx->tuning[precision];
    t->block_threads = u.block_threads;
    t->blocks_per_cu = u.blocks_per_cu;
    t->steps_per_sync = u.steps_per_sync;
    t->refill_min = u.refill_min;
    t->queues = u.queues;
    t->order = u.order;
    t->coop_steps = u.coop_steps;
    return BH_OK;
}

int bh_render_device(bh_ctx *ctx, const bh_scene_desc *scene, int32_t brick_id, int32_t precision, void *d_iquv,
                     int32_t *d_steps, uint8_t *d_status) {
    clear_error();
    int st = ctx_check(ctx);
    if (st) return st;
    return render_enqueue(ctx, scene, brick_id, precision, d_iquv, d_steps, d_status, nullptr, nullptr, 0, true);
}

int bh_render_device_noalloc(bh_ctx *ctx, const bh_scene_desc *scene, int32_t brick_id, int32_t precision,
                             void *d_iquv, int32_t *d_steps, uint8_t *d_status) {
    clear_error();
    int st = ctx_check(ctx);
    if (st) return st;
    unsigned *buf = nullptr;
    if ((st = scratch_address(ctx, &buf))) return st;
    return render_enqueue(ctx, scene, brick_id, precision, d_iquv, d_steps, d_status, nullptr, nullptr, 0, true, buf);
}

int bh_render(bh_ctx *ctx, const bh_scene_desc *scene, int32_t brick_id, int32_t precision, void *iquv, int32_t *steps,
              uint8_t *status) {
    clear_error();
    int st = ctx_check(ctx);
    if (st) return st;
    if (!scene || !iquv) return set_error(BH_ERR_VALUE, "scene and iquv must not be NULL");
    if ((st = validate_scene(*scene))) return st;
    const size_t n = size_t(scene->width) * scene->height;
    const size_t esz = precision == BH_PREC_FP64 ? 8 : 4;
    void *d = nullptr;
    if (hipMalloc(&d, n * 4 * esz + n * 4 + n + 1024) != hipSuccess)
        return set_error(BH_ERR_MEMORY, "render output allocation failed");
    char *base = static_cast<char *>(d);
    void *d_iquv = base;
    int32_t *d_steps = reinterpret_cast<int32_t *>(base + ((n * 4 * esz + 255) & ~size_t(255)));
    uint8_t *d_status = reinterpret_cast<uint8_t *>(reinterpret_cast<char *>(d_steps) + ((n * 4 + 255) & ~size_t(255)));
    st = bh_render_device(ctx, scene, brick_id, precision, d_iquv, d_steps, d_status);
    if (st == BH_OK) {
        hipError_t e = hipMemcpyAsync(iquv, d_iquv, n * 4 * esz, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess && steps) e = hipMemcpyAsync(steps, d_steps, n * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess && status) e = hipMemcpyAsync(status, d_status, n, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = set_error(BH_ERR_DEVICE, "render readback failed: %s", hipGetErrorString(e));
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    return st;
}

}  // extern "C"

struct bh_batch {
    bh_ctx *ctx = nullptr;
    int precision = 0, storage = 0, layout = 0, n = 0;
    bh::RenderArgs *d_descs = nullptr;
    int *d_slots = nullptr;
    float *d_grids = nullptr;
    void *d_coeffs = nullptr;
    size_t coeff_bytes = 0;
    int ng = 0, queues = 1, q_tiles = 0;
    long long total_pixels = 0;
    bh::BatchFn fn = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *g_iquv = nullptr, *g_steps = nullptr, *g_status = nullptr;
    int g_grid = 0, g_block = 0, g_spc = 0;
};

namespace bh {

static int batch_args(bh_batch *bt, void *d_iquv, int32_t *d_steps, uint8_t *d_status, BatchArgs &B, int &grid,
                      int &block) {
    bh_ctx *ctx = bt->ctx;
    const Tuning &u = ctx->tuning[bt->precision];
    B.descs = bt->d_descs;
    B.slots = bt->d_slots;
    B.grids = bt->d_grids;
    B.ng = bt->ng;
    B.coeffs = bt->d_coeffs;
    B.total_bytes = unsigned(std::min(bt->coeff_bytes, size_t(0xffffffffu)));
    B.iquv = d_iquv;
    B.steps = d_steps;
    B.status = d_status;
    B.counters = ctx->d_counters;
    B.work = reinterpret_cast<unsigned long long *>(ctx->d_counters + kWorkOffset);
    B.queues = bt->queues;
    B.q_tiles = bt->q_tiles;
    B.steps_per_sync = u.steps_per_sync;
    int per_cu = u.blocks_per_cu;
    if (per_cu <= 0) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, bt->fn, u.block_threads, 0) != hipSuccess || nb <= 0) nb = 1;
        per_cu = nb;
    }
    block = u.block_threads;
    grid = ctx->cus * per_cu;
    const int waves_per_block = block / 64;
    const int max_useful = (bt->queues * bt->q_tiles + waves_per_block - 1) / waves_per_block;
    if (grid > max_useful) grid = max_useful;
    if (grid < 1) grid = 1;
    return BH_OK;
}

}  // namespace bh

extern "C" {

int bh_batch_create(bh_ctx *ctx, const bh_batch_item *items, int32_t n, int32_t precision, int64_t *offsets,
                    bh_batch **out) {
    clear_error();
    int st = ctx_check(ctx);
    if (st) return st;
    if (!items || !out || n < 1) return set_error(BH_ERR_VALUE, "batch needs at least one item");
    if (n >= (1 << 11)) return set_error(BH_ERR_VALUE, "batch holds at most 2047 renders");
    if (precision != BH_PREC_FP64 && precision != BH_PREC_FP32 && precision != BH_PREC_MIXED)
        return set_error(BH_ERR_VALUE, "precision must be BH_PREC_FP64, BH_PREC_FP32 or BH_PREC_MIXED");
    *out = nullptr;
    std::vector<int> uniq;
    std::vector<int> kmap(n);
    int storage = -1, layout = -1;
    for (int i = 0; i < n; ++i) {
        if ((st = validate_scene(items[i].scene))) return st;
        const int id = items[i].brick_id;
        if (id < 0 || size_t(id) >= ctx->bricks.size() || !ctx->bricks[id].live)
            return set_error(BH_ERR_STATE, "unknown brick id %d", id);
        const BrickDev &b = ctx->bricks[id];
        if (storage < 0) {
            storage = b.storage;
            layout = b.layout;
        } else if (storage != b.storage || layout != b.layout) {
            return set_error(BH_ERR_VALUE, "all bricks of a batch must share storage and layout");
        }
        int k = -1;
        for (size_t j = 0; j < uniq.size(); ++j)
            if (uniq[j] == id) k = int(j);
        if (k < 0) {
            k = int(uniq.size());
            uniq.push_back(id);
        }
        kmap[i] = k;
        const long long tiles = (long long)((items[i].scene.width + 7) / 8) * ((items[i].scene.height + 7) / 8);
        if (tiles >= (1 << 20)) return set_error(BH_ERR_VALUE, "render %d has too many 8x8 tiles", i);
        if (items[i].scene.transfer != items[0].scene.transfer)
            return set_error(BH_ERR_VALUE, "all renders of a batch must share the transfer rule");
        if (items[i].scene.integrator != items[0].scene.integrator)
            return set_error(BH_ERR_VALUE, "all renders of a batch must share the integrator");
    }
    // merged coefficient buffer and grid table
    std::vector<size_t> coff(uniq.size());
    std::vector<int> goff(uniq.size());
    size_t total = 0;
    int ng = 0;
    for (size_t k = 0; k < uniq.size(); ++k) {
        const BrickDev &b = ctx->bricks[uniq[k]];
        coff[k] = total;
        total += (b.bytes + 255) & ~size_t(255);
        goff[k] = ng;
        ng += b.nr + b.nt + b.np + 1;
    }
    // SoA bricks are read through one buffer resource (32-bit offsets); corner records of 2 GiB and more
    // through 64-bit addresses
    if (layout == BH_LAYOUT_SOA && total >= (size_t(1) << 31))
        return set_error(BH_ERR_VALUE, "the SoA bricks of a batch exceed 2 GiB (corner bricks have no such limit)");
    BatchFn fn = pick_batch_kernel(precision, storage, layout, items[0].scene.transfer == BH_TRANSFER_EXACT,
                                   items[0].scene.integrator == BH_INTEGRATOR_RK4, wide_corner(layout, total));
    if (!fn)
        return set_error(BH_ERR_VALUE, precision == BH_PREC_FP64
                                           ? "the FP64 batch needs float32 bricks in the SoA layout"
                                           : "the RK4 integrator of the FP32 and mixed batches needs the corner layout");
    if (ng > kMaxGrid) return set_error(BH_ERR_VALUE, "the bricks of a batch hold more than %d grid values", kMaxGrid);
    bh_batch *bt = new bh_batch();
    bt->ctx = ctx;
    bt->precision = precision;
    bt->storage = storage;
    bt->layout = layout;
    bt->n = n;
    bt->fn = fn;
    bt->ng = ng;
    bt->coeff_bytes = total;
    auto fail = [&](int code) {
        bh_batch_destroy(bt);
        return code;
    };
    if (hipMalloc(&bt->d_coeffs, total) != hipSuccess) return fail(set_error(BH_ERR_MEMORY, "batch brick buffer allocation failed"));
    std::vector<float> grids(ng);
    for (size_t k = 0; k < uniq.size(); ++k) {
        const BrickDev &b = ctx->bricks[uniq[k]];
        if (hipMemcpyAsync(static_cast<char *>(bt->d_coeffs) + coff[k], b.coeffs, b.bytes, hipMemcpyDeviceToDevice,
                           ctx->stream) != hipSuccess)
            return fail(set_error(BH_ERR_DEVICE, "batch brick copy failed"));
        std::copy(b.hgrids.begin(), b.hgrids.end(), grids.begin() + goff[k]);
        grids[goff[k] + b.nr + b.nt + b.np] = __builtin_huge_valf();
    }
    if (hipMalloc(&bt->d_grids, sizeof(float) * ng) != hipSuccess) return fail(set_error(BH_ERR_MEMORY, "batch grid allocation failed"));
    if (hipMemcpyAsync(bt->d_grids, grids.data(), sizeof(float) * ng, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        return fail(set_error(BH_ERR_DEVICE, "batch grid copy failed"));
    // render descriptors
    std::vector<RenderArgs> descs(n);
    std::vector<std::pair<float, int>> order;   // (distance, render << 20 | tile)
    long long pix = 0;
    const Tuning &u = ctx->tuning[precision];
    for (int i = 0; i < n; ++i) {
        const BrickDev &b = ctx->bricks[uniq[kmap[i]]];
        RenderArgs &A = descs[i];
        A = RenderArgs{};
        if ((st = setup_args(items[i].scene, b, A))) return fail(st);
        A.coeffs = bt->d_coeffs;
        A.total_bytes = unsigned(std::min(total, size_t(0xffffffffu)));   // SoA batches stay below 2 GiB
        A.brick_base = unsigned(coff[kmap[i]]);
        A.brick_base_hi = unsigned(coff[kmap[i]] >> 32);
        A.grid_off = goff[kmap[i]];
        A.out_off = pix;
        if (offsets) offsets[i] = pix;
        pix += (long long)items[i].scene.width * items[i].scene.height;
        const int tx = A.tiles_x, ty = (A.H + 7) / 8;
        std::vector<float> dist;
        if (u.order == 1) dist = tile_distances(items[i].scene, A.cam, tx, ty);
        for (int t = 0; t < tx * ty; ++t)
            order.push_back({u.order == 1 ? dist[t] : float(order.size()), (i << 20) | t});
    }
    if (offsets) offsets[n] = pix;
    bt->total_pixels = pix;
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<float, int> &x, const std::pair<float, int> &y) { return x.first < y.first; });
    const int n_tiles = int(order.size());
    bt->queues = u.queues < n_tiles ? u.queues : n_tiles;
    bt->q_tiles = (n_tiles + bt->queues - 1) / bt->queues;
    std::vector<int> slots(size_t(bt->queues) * bt->q_tiles, -1);
    for (int k = 0; k < n_tiles; ++k) slots[size_t(k % bt->queues) * bt->q_tiles + k / bt->queues] = order[k].second;
    if (hipMalloc(&bt->d_descs, sizeof(RenderArgs) * n) != hipSuccess ||
        hipMalloc(&bt->d_slots, sizeof(int) * slots.size()) != hipSuccess)
        return fail(set_error(BH_ERR_MEMORY, "batch table allocation failed"));
    if (hipMemcpyAsync(bt->d_descs, descs.data(), sizeof(RenderArgs) * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(bt->d_slots, slots.data(), sizeof(int) * slots.size(), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return fail(set_error(BH_ERR_DEVICE, "batch table upload failed"));
    *out = bt;
    return BH_OK;
}

int bh_batch_launch(bh_batch *bt, void *d_iquv, int32_t *d_steps, uint8_t *d_status) {
    clear_error();
    if (!bt) return set_error(BH_ERR_STATE, "batch is NULL");
    int st = ctx_check(bt->ctx);
    if (st) return st;
    if (!d_iquv) return set_error(BH_ERR_VALUE, "iquv must not be NULL");
    BatchArgs B{};
    int grid = 0, block = 0;
    if ((st = batch_args(bt, d_iquv, d_steps, d_status, B, grid, block))) return st;
    if ((st = fill_u32(bt->ctx, bt->ctx->d_counters, 0u, kCounterWords))) return st;
    bt->ctx->last_counters = bt->ctx->d_counters;
    begin_timing(bt->ctx);
    void *args[] = {&B};
    BH_HIP(hipLaunchKernel(reinterpret_cast<const void *>(bt->fn), dim3(grid), dim3(block), args, 0, bt->ctx->stream));
    end_timing(bt->ctx, 1);
    return BH_OK;
}

int bh_batch_launch_graph(bh_batch *bt, void *d_iquv, int32_t *d_steps, uint8_t *d_status) {
    clear_error();
    if (!bt) return set_error(BH_ERR_STATE, "batch is NULL");
    int st = ctx_check(bt->ctx);
    if (st) return st;
    if (!d_iquv) return set_error(BH_ERR_VALUE, "iquv must not be NULL");
    bh_ctx *ctx = bt->ctx;
    BatchArgs B{};
    int grid = 0, block = 0;
    if ((st = batch_args(bt, d_iquv, d_steps, d_status, B, grid, block))) return st;
    ctx->last_counters = ctx->d_counters;
    const bool same = bt->exec && bt->g_iquv == d_iquv && bt->g_steps == (void *)d_steps && bt->g_status == (void *)d_status &&
                      bt->g_grid == grid && bt->g_block == block && bt->g_spc == B.steps_per_sync;
    if (!same) {
        if (bt->exec) (void)hipGraphExecDestroy(bt->exec);
        if (bt->graph) (void)hipGraphDestroy(bt->graph);
        bt->exec = nullptr;
        bt->graph = nullptr;
        BH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        // counter reset by a fill kernel (no memset node in the graph), then the persistent kernel
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(kCounterWords), 0, ctx->stream, ctx->d_counters, 0u, kCounterWords);
        void *args[] = {&B};
        hipError_t e = hipLaunchKernel(reinterpret_cast<const void *>(bt->fn), dim3(grid), dim3(block), args, 0, ctx->stream);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
        if (e != hipSuccess || e2 != hipSuccess || !g)
            return set_error(BH_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
        bt->graph = g;
        BH_HIP(hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0));
        bt->g_iquv = d_iquv;
        bt->g_steps = d_steps;
        bt->g_status = d_status;
        bt->g_grid = grid;
        bt->g_block = block;
        bt->g_spc = B.steps_per_sync;
    }
    begin_timing(ctx);
    BH_HIP(hipGraphLaunch(bt->exec, ctx->stream));
    end_timing(ctx, 1);
    return BH_OK;
}

int bh_batch_destroy(bh_batch *bt) {
    if (!bt) return BH_OK;
    if (bt->ctx) {
        (void)hipSetDevice(bt->ctx->device);
        (void)hipStreamSynchronize(bt->ctx->stream);
    }
    if (bt->exec) (void)hipGraphExecDestroy(bt->exec);
    if (bt->graph) (void)hipGraphDestroy(bt->graph);
    if (bt->d_descs) (void)hipFree(bt->d_descs);
    if (bt->d_slots) (void)hipFree(bt->d_slots);
    if (bt->d_grids) (void)hipFree(bt->d_grids);
    if (bt->d_coeffs) (void)hipFree(bt->d_coeffs);
    delete bt;
    return BH_OK;
}

}  // extern "C"

// ---- learned surrogate: geodesic knots and the transfer over predicted knots -----------

namespace bh {

constexpr int kCurveMax = 512;   // critical-curve points staged in LDS
constexpr int kMaxKnots = 512;
constexpr int kHeads = 3;        // captured, flag, recorded states

// ZAMO tetrad and metric of a camera at (r, th), phi 0, no roll (kerr.py:67-83, 144-148), FP64.
__device__ __forceinline__ void cam_tetrad(double a, double r, double th, CamConst &C) {
    double s, c;
    ::sincos(th, &s, &c);
    const double sig = r * r + a * a * c * c;
    const double dlt = r * r - 2.0 * r + a * a;
    const double Aq = (r * r + a * a) * (r * r + a * a) - a * a * dlt * s * s;
    const double s2 = s * s;
    C.g00 = -(1.0 - 2.0 * r / sig);
    C.g03 = -2.0 * a * r * s2 / sig;
    C.g11 = sig / dlt;
    C.g22 = sig;
    C.g33 = Aq * s2 / sig;
    const double lapse = ::sqrt(sig * dlt / Aq);
    const double omega = 2.0 * a * r / Aq;
    C.e_t0 = 1.0 / lapse;
    C.e_t3 = omega / lapse;
    C.e_r1 = ::sqrt(dlt / sig);
    C.e_th2 = 1.0 / ::sqrt(sig);
    C.e_ph3 = 1.0 / ::sqrt(C.g33);
    C.tan_x = C.tan_y = 0.0;
    C.cos_roll = 1.0;
    C.sin_roll = 0.0;
    C.roll = 0;
    C.r = r;
    C.th = th;
    C.ph = 0.0;
}

// Redshift seen by a zero-angular-momentum observer at (r, th): g = E alpha / (E - omega L).
__device__ __forceinline__ double zamo_redshift(double a, double r, double th, double E, double L) {
    double s, c;
    ::sincos(th, &s, &c);
    const double sig = r * r + a * a * c * c, dlt = r * r - 2.0 * r + a * a;
    const double Aq = (r * r + a * a) * (r * r + a * a) - a * a * dlt * s * s;
    const double lapse = ::sqrt(fmax(sig * dlt / Aq, 0.0));
    const double omega = 2.0 * a * r / Aq;
    return E * lapse / (E - omega * L);
}

struct KnotArgs {
    const float *params;   // n x 4: alpha, beta, a, cos i
    long long n;
    double cam_r, step, r_esc, r_lo, r_hi;
    int adaptive, max_steps, K;
    float *rows;           // n x (4K + kHeads)
    int *steps;
    float *weight;         // n: 1 for an unflagged ray, else 0
};

// One thread per ray: the FP64 RK2 geodesic of ray_step (same step rule, theta clamp, pole test and
// termination order, without the transfer), recording knots from the first step whose midpoint
// radius lies in the brick's radial range.
__global__ __launch_bounds__(256) void knots_kernel(KnotArgs A) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const float *p = A.params + i * 4;
    const double alpha = p[0], beta = p[1], a = p[2];
    const double ci = fmin(fmax(double(p[3]), -1.0), 1.0);
    const double th0 = clamp_theta<double>(::acos(ci));
    CamConst C;
    cam_tetrad(a, A.cam_r, th0, C);
    double n_r = -1.0, n_th = -beta / A.cam_r, n_ph = alpha / A.cam_r;
    const double nrm = ::sqrt(n_r * n_r + n_th * n_th + n_ph * n_ph);
    n_r /= nrm;
    n_th /= nrm;
    n_ph /= nrm;
    const double pc_t = C.e_t0, pc_r = n_r * C.e_r1, pc_th = n_th * C.e_th2, pc_ph = C.e_t3 + n_ph * C.e_ph3;
    const double E = -(C.g00 * pc_t + C.g03 * pc_ph), L = C.g03 * pc_t + C.g33 * pc_ph;
    double r = A.cam_r, th = th0, ph = 0.0, pr = C.g11 * pc_r, pth = C.g22 * pc_th;
    const RayK<double> k = make_k<double>(a, E, L);
    const double r_plus = 1.0 + ::sqrt(1.0 - a * a), r_cap = r_plus * (1.0 + 2.0e-4);
    const double lo = kEpsTheta, hi = 3.14159265358979323846 - kEpsTheta;
    const int K = A.K;
    float *row = A.rows + i * (4LL * K + kHeads);
    int rec = 0, steps = 0, status = 0;
    bool entered = false, overflow = false, clamped = false;
    auto put = [&](double kr, double kt, double kp) {
        float *o = row + 4 * rec;
        o[0] = float(kr);
        o[1] = float(kt);
        o[2] = float(kp);
        o[3] = float(zamo_redshift(a, kr, kt, E, L));
        ++rec;
    };
    for (;;) {
        const double h = A.adaptive ? A.step * (r - r_plus) : A.step;
        double nr, nt_raw, np, npr, npt, mt;
        geo_step<double>(k, r, th, ph, pr, pth, h, nr, nt_raw, np, npr, npt, mt);
        const double nt = clamp_theta(nt_raw);
        clamped = clamped || (mt < lo) || (mt > hi) || (nt_raw < lo) || (nt_raw > hi) || (np - ph > 1.0) ||
                  (ph - np > 1.0);
        const bool fin = finite_(nr) && finite_(nt) && finite_(np) && finite_(npr) && finite_(npt);
        const double rm = 0.5 * (r + nr);
        const bool inb = rm >= A.r_lo && rm <= A.r_hi;
        if (inb && !entered) {
            entered = true;
            put(r, th, ph);
        }
        if (entered && fin && !clamped) {   // a pole-clamped ray is flagged: its later states are not recorded
            if (rec < K) put(nr, nt, np);
            else if (inb) overflow = true;
        }
        const double r_prev = r;
        r = nr;
        th = nt;
        ph = np;
        pr = npr;
        pth = npt;
        ++steps;
        if (!fin) { status = 3; break; }
        if (nr <= r_cap) { status = 1; break; }
        if (nr > A.r_esc && nr > r_prev) { status = 2; break; }
        if (steps >= A.max_steps) { status = 0; break; }
    }
    const int n_rec = rec;
    if (rec == 0) put(r, th, ph);   // never reached the brick: the end state stands for every knot
    for (int j = rec; j < K; ++j) {
        row[4 * j] = row[4 * (rec - 1)];
        row[4 * j + 1] = row[4 * (rec - 1) + 1];
        row[4 * j + 2] = row[4 * (rec - 1) + 2];
        row[4 * j + 3] = row[4 * (rec - 1) + 3];
    }
    const bool flag = overflow || clamped || status == 3 || status == 0;
    row[4 * K] = status == 1 ? 1.0f : 0.0f;
    row[4 * K + 1] = flag ? 1.0f : 0.0f;
    row[4 * K + 2] = float(n_rec);
    if (A.steps) A.steps[i] = steps;
    if (A.weight) A.weight[i] = flag ? 0.0f : 1.0f;
}

// Network input of every pixel: ((alpha, beta, a, cos i) - add) * mul, alpha/beta the image-plane
// offsets (M) of the pixel's camera direction.
// Network input of pixel pix: ((alpha, beta, a, cos i) - add) * mul, alpha and beta the image-plane offsets
// (M) of the pixel's camera direction.
__device__ __forceinline__ void surrogate_input(const RenderArgs &A, int pix, float4 mul, float4 add, float *o) {
    const int px = pix % A.W, py = pix / A.W;
    const double ndc_x = 2.0 * (double(px) + 0.5) / double(A.W) - 1.0;
    const double ndc_y = 1.0 - 2.0 * (double(py) + 0.5) / double(A.H);
    const CamConst &C = A.cam;
    double n_th = -ndc_y * C.tan_y, n_ph = ndc_x * C.tan_x;
    if (C.roll) {
        const double t0 = C.cos_roll * n_th - C.sin_roll * n_ph;
        const double t1 = C.sin_roll * n_th + C.cos_roll * n_ph;
        n_th = t0;
        n_ph = t1;
    }
    o[0] = (float(n_ph * C.r) - add.x) * mul.x;
    o[1] = (float(-n_th * C.r) - add.y) * mul.y;
    o[2] = (float(A.a) - add.z) * mul.z;
    o[3] = (float(::cos(C.th)) - add.w) * mul.w;
}

__global__ void surrogate_input_kernel(RenderArgs A, float4 mul, float4 add, float *x) {
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= A.W * A.H) return;
    surrogate_input(A, pix, mul, add, x + size_t(pix) * 4);
}

// The inputs of every render of a batch, row g of the batch belonging to the last render with out_off <= g.
__global__ void surrogate_input_batch_kernel(const RenderArgs *descs, int n, long long total, float4 mul, float4 add,
                                             float *x) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (descs[mid].out_off <= g) lo = mid;
        else hi = mid - 1;
    }
    surrogate_input(descs[lo], int(g - descs[lo].out_off), mul, add, x + size_t(g) * 4);
}

struct SurArgs {
    RenderArgs R;            // scene, brick view and outputs (FP32 IQUV)
    const void *y;           // network outputs, [n_out][ld], FP32 or FP16 (the kernel's KT)
    long long ld;
    const float *mul, *add;  // denormalisation per output (device)
    const float *curve;      // critical curve (alpha, beta) pairs
    int n_curve, K;
    float crit2, flag_max, cap_lo, cap_hi;
    int *list;               // fallback pixels
    unsigned *n_list;
};

// The per-render settings of the surrogate, shared by every render of a batch.
struct SurShared {
    const float *mul, *add;
    int K;
    float crit2, flag_max, cap_lo, cap_hi;
};

// One 8x8 tile of a surrogate render, one lane per pixel.  Pixels near the critical curve or flagged by
// the heads are appended to the fallback list (their pixel index in the render); the others integrate
// the Stokes transfer over the predicted knots with the RK2 transfer rule (midpoint sample, Euclidean
// ds, accelerated_renderer.py:113-125).  y points at the knot column of the render's pixel 0, gr at the
// render's brick grids in LDS (+inf after phi), curve at the render's critical curve; A's outputs are
// the render's own.
template <int STORE, int LAYOUT, typename KT>
__device__ __forceinline__ void surrogate_tile(const RenderArgs &A, const float *gr, const float *s_mul,
                                               const float *s_add, const KT *y, long long ld, const float *curve,
                                               int n_curve, const SurShared &S, int *list, unsigned *n_list,
                                               int tile, int lane) {
    const float *gt = gr + A.nr, *gp = gt + A.nt;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(A.coeffs), (short)0, int(A.total_bytes), int(kRsrcWord3));
    const int px = (tile % A.tiles_x) * 8 + (lane & 7), py = (tile / A.tiles_x) * 8 + (lane >> 3);
    const bool inside = px < A.W && py < A.H;
    const int pix = py * A.W + px;
    const KT *yc = y + pix;
    bool fb = false;
    float cap = 0.f;
    if (inside) {
        // Bardeen image coordinates of the ray (as the tile schedule) and distance to the critical curve
        double pr, pth, pt, pph;
        camera_ray(A, px, py, pr, pth, pt, pph);
        const double E = -pt, L = pph, a = A.a;
        double s0, c0;
        ::sincos(A.cam.th, &s0, &c0);
        const double Q = pth * pth + c0 * c0 * (-a * a * E * E + L * L / (s0 * s0));
        const double xi = L / E;
        const float al = float(-xi / s0);
        const float be = float(::sqrt(fmax(Q / (E * E) + a * a * c0 * c0 - xi * xi * (c0 / s0) * (c0 / s0), 0.0)));
        float d2 = __builtin_huge_valf();
        for (int j = 0; j < n_curve; ++j) {
            const float dx = al - curve[2 * j], dy = be - curve[2 * j + 1];
            d2 = fminf(d2, dx * dx + dy * dy);
        }
        const int hc = 4 * S.K;
        cap = float(yc[hc * ld]) * s_mul[hc] + s_add[hc];
        const float flag = float(yc[(hc + 1) * ld]) * s_mul[hc + 1] + s_add[hc + 1];
        fb = d2 < S.crit2 || flag > S.flag_max || (cap > S.cap_lo && cap < S.cap_hi) || !(cap == cap);
    }
    const unsigned long long m = __ballot(fb);
    if (m) {
        const int leader = __builtin_ctzll(m);
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(n_list, unsigned(__popcll(m)));
        base = __builtin_amdgcn_readlane(base, leader);
        if (fb) list[base + lane_rank(m)] = pix;
    }
    if (!inside || fb) return;

    const HotParams<float> P = hot_params<float>(A);
    const float r_lo = gr[0], r_hi = gr[A.nr - 1], ph0 = float(A.cam.ph);
    // knot j's raw outputs; every row holds K knots, so any j < K is in bounds
    auto load_knot = [&](int j, KT &hr, KT &ht, KT &hp) {
        const int c = 4 * j;
        hr = yc[c * ld];
        ht = yc[(c + 1) * ld];
        hp = yc[(c + 2) * ld];
    };
    auto decode_knot = [&](int j, KT hr, KT ht, KT hp, float &kr, float &kt, float &kp) {
        const int c = 4 * j;
        kr = float(hr) * s_mul[c] + s_add[c];
        kt = clamp_theta(float(ht) * s_mul[c + 1] + s_add[c + 1]);
        kp = float(hp) * s_mul[c + 2] + s_add[c + 2] + ph0;
    };
    float St[4] = {0.f, 0.f, 0.f, 0.f}, tau = 0.f;
    float r0, t0, p0;
    KT hr, ht, hp;
    load_knot(0, hr, ht, hp);
    decode_knot(0, hr, ht, hp, r0, t0, p0);
    if (S.K > 1) load_knot(1, hr, ht, hp);
    int code = -1, seg = 0;
    for (int j = 1; j < S.K; ++j) {
        float r1, t1, p1;
        decode_knot(j, hr, ht, hp, r1, t1, p1);
        // the next knot's loads leave before this segment's brick loads, so a segment waits for one
        // memory round trip instead of two
        if (j + 1 < S.K) load_knot(j + 1, hr, ht, hp);
        const float rm = 0.5f * (r0 + r1);
        // outside the brick's radial range and moving away from it: no later segment samples it
        if ((rm > r_hi && r1 > r0) || (rm < r_lo && r1 < r0)) {
            code = r1 < r0 ? 1 : 2;
            break;
        }
        float c[kCoeffs];
        if (sample_brick<float, STORE, LAYOUT>(A, rs, gr, gt, gp, rm, 0.5f * (t0 + t1), 0.5f * (p0 + p1), c)) {
            const float dx = r1 - r0, dy = t1 - t0, dz = p1 - p0;
            const float dlam = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
            const float ds = pymax(P.L * pymax(dlam, 1.0e-14f), 0.0f);
            stokes_rk2(St, c, ds);
            tau = tau + pymax(c[4], 0.0f) * ds;
        }
        ++seg;
        if (r1 <= P.r_cap) { code = 1; break; }
        if (tau >= P.tau_max) { code = 4; break; }
        r0 = r1;
        t0 = t1;
        p0 = p1;
    }
    if (code < 0) code = cap >= 0.5f ? 1 : 2;
    float *o = static_cast<float *>(A.iquv) + size_t(pix) * 4;
    o[0] = nan_to_zero(St[0]);
    o[1] = nan_to_zero(St[1]);
    o[2] = nan_to_zero(St[2]);
    o[3] = nan_to_zero(St[3]);
    if (A.steps) A.steps[pix] = seg;
    if (A.status) A.status[pix] = uint8_t(code | 16);
}

// One render: one wave per 8x8 tile.
template <int STORE, int LAYOUT, typename KT>
__global__ __launch_bounds__(256) void surrogate_kernel(SurArgs S) {
    __shared__ float s_grid[kMaxGrid + 1];
    __shared__ float s_curve[2 * kCurveMax];
    __shared__ float s_mul[4 * kMaxKnots + kHeads], s_add[4 * kMaxKnots + kHeads];
    const RenderArgs &A = S.R;
    const int ng = A.nr + A.nt + A.np;
    for (int j = threadIdx.x; j < ng; j += blockDim.x) s_grid[j] = A.grids[j];
    if (threadIdx.x == 0) s_grid[ng] = __builtin_huge_valf();
    for (int j = threadIdx.x; j < 2 * S.n_curve; j += blockDim.x) s_curve[j] = S.curve[j];
    const int n_out = 4 * S.K + kHeads;
    for (int j = threadIdx.x; j < n_out; j += blockDim.x) {
        s_mul[j] = S.mul[j];
        s_add[j] = S.add[j];
    }
    __syncthreads();
    const SurShared sh{S.mul, S.add, S.K, S.crit2, S.flag_max, S.cap_lo, S.cap_hi};
    const int tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    surrogate_tile<STORE, LAYOUT, KT>(A, s_grid, s_mul, s_add, static_cast<const KT *>(S.y), S.ld, s_curve, S.n_curve, sh,
                                      S.list, S.n_list, tile, threadIdx.x & 63);
}

// Many renders: wave w takes tile tiles[w] = render << 20 | tile.  The renders share one merged brick
// buffer and grid table (as the batched RK2 kernel), their network outputs are columns out_off .. of y,
// their fallback pixels go to list[out_off + counts[render]++], and their critical curves (curve_at[r],
// curve_n[r] points) are read from global memory.
struct SurBatchArgs {
    const RenderArgs *descs;
    const int *tiles;
    int n_tiles;
    const float *grids;
    int ng;
    const _Float16 *y;
    long long ld;
    const float *curves;
    const int *curve_at, *curve_n;
    SurShared sh;
    void *iquv;
    int *steps;
    uint8_t *status;
    int *list;
    unsigned *counts;
};

template <int STORE, int LAYOUT>
__global__ __launch_bounds__(256) void surrogate_batch_kernel(SurBatchArgs B) {
    __shared__ float s_grid[kMaxGrid + 8];
    __shared__ float s_mul[4 * kMaxKnots + kHeads], s_add[4 * kMaxKnots + kHeads];
    for (int j = threadIdx.x; j < B.ng; j += blockDim.x) s_grid[j] = B.grids[j];
    const int n_out = 4 * B.sh.K + kHeads;
    for (int j = threadIdx.x; j < n_out; j += blockDim.x) {
        s_mul[j] = B.sh.mul[j];
        s_add[j] = B.sh.add[j];
    }
    __syncthreads();
    // the wave's tile and render are wave-uniform: read them into SGPRs so the render descriptor is a
    // scalar load and the brick's buffer resource needs no per-load waterfall
    const int w = __builtin_amdgcn_readfirstlane(int(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
    if (w >= B.n_tiles) return;
    const int e = __builtin_amdgcn_readfirstlane(B.tiles[w]);
    const int rix = e >> 20, tile = e & 0xfffff;
    RenderArgs A = B.descs[rix];
    A.iquv = static_cast<float *>(B.iquv) + size_t(A.out_off) * 4;
    A.steps = B.steps ? B.steps + A.out_off : nullptr;
    A.status = B.status ? B.status + A.out_off : nullptr;
    surrogate_tile<STORE, LAYOUT, _Float16>(A, s_grid + A.grid_off, s_mul, s_add, B.y + A.out_off, B.ld,
                                            B.curves + 2 * B.curve_at[rix], B.curve_n[rix], B.sh, B.list + A.out_off,
                                            B.counts + rix, tile, threadIdx.x & 63);
}

// Chunks of 64 fallback rays per render, as an exclusive prefix: prefix[r] .. prefix[r + 1] - 1 are the
// batched RK2 kernel's slots for render r, prefix[n] their number.
__global__ void fallback_chunks_kernel(const unsigned *counts, int n, unsigned *prefix) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned acc = 0;
    for (int r = 0; r < n; ++r) {
        prefix[r] = acc;
        acc += (counts[r] + 63u) / 64u;
    }
    prefix[n] = acc;
}

__global__ void fill_u32_once(unsigned *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0u;
}

typedef void (*SurFn)(SurArgs);
typedef void (*SurBatchFn)(SurBatchArgs);

// The batched RK2 kernel in list mode, for the fallback rays of batched surrogate renders (FP32 or mixed
// precision, RK2 transfer and integrator); W as in pick_fp32.
template <int W>
static BatchFn pick_list_kernel_w(int storage, int layout, bool mixed) {
#define BH_LIST(S, L) (mixed ? batch_kernel<float, S, L, true, false, false, true> : batch_kernel<float, S, L, false, false, false, true>)
    if (layout == BH_LAYOUT_CORNER) {
        if (storage == BH_STORE_F32) return BH_LIST(BH_STORE_F32 | W, BH_LAYOUT_CORNER);
        if (storage == BH_STORE_BF16) return BH_LIST(BH_STORE_BF16 | W, BH_LAYOUT_CORNER);
        return BH_LIST(BH_STORE_F16 | W, BH_LAYOUT_CORNER);
    }
    if constexpr (W == 0) {
        if (storage == BH_STORE_F32) return BH_LIST(BH_STORE_F32, BH_LAYOUT_SOA);
        if (storage == BH_STORE_BF16) return BH_LIST(BH_STORE_BF16, BH_LAYOUT_SOA);
        return BH_LIST(BH_STORE_F16, BH_LAYOUT_SOA);
    }
    return nullptr;
#undef BH_LIST
}
static BatchFn pick_list_kernel(int storage, int layout, bool mixed, bool wide) {
    return wide ? pick_list_kernel_w<kStoreWide>(storage, layout, mixed) : pick_list_kernel_w<0>(storage, layout, mixed);
}

template <typename KT, int W>
static SurFn pick_surrogate_t(int storage, int layout) {
    if (layout == BH_LAYOUT_CORNER) {
        if (storage == BH_STORE_F32) return surrogate_kernel<BH_STORE_F32 | W, BH_LAYOUT_CORNER, KT>;
        if (storage == BH_STORE_BF16) return surrogate_kernel<BH_STORE_BF16 | W, BH_LAYOUT_CORNER, KT>;
        return surrogate_kernel<BH_STORE_F16 | W, BH_LAYOUT_CORNER, KT>;
    }
    if constexpr (W == 0) {
        if (storage == BH_STORE_F32) return surrogate_kernel<BH_STORE_F32, BH_LAYOUT_SOA, KT>;
        if (storage == BH_STORE_BF16) return surrogate_kernel<BH_STORE_BF16, BH_LAYOUT_SOA, KT>;
        return surrogate_kernel<BH_STORE_F16, BH_LAYOUT_SOA, KT>;
    }
    return nullptr;
}

// half_knots: the knots come from the network as FP16; otherwise FP32 rows the caller supplies.
static SurFn pick_surrogate(int storage, int layout, bool half_knots, bool wide) {
    if (wide) return half_knots ? pick_surrogate_t<_Float16, kStoreWide>(storage, layout) : pick_surrogate_t<float, kStoreWide>(storage, layout);
    return half_knots ? pick_surrogate_t<_Float16, 0>(storage, layout) : pick_surrogate_t<float, 0>(storage, layout);
}

static SurBatchFn pick_surrogate_batch(int storage, int layout, bool wide) {
    if (layout == BH_LAYOUT_CORNER) {
        if (wide) {
            if (storage == BH_STORE_F32) return surrogate_batch_kernel<BH_STORE_F32 | kStoreWide, BH_LAYOUT_CORNER>;
            if (storage == BH_STORE_BF16) return surrogate_batch_kernel<BH_STORE_BF16 | kStoreWide, BH_LAYOUT_CORNER>;
            return surrogate_batch_kernel<BH_STORE_F16 | kStoreWide, BH_LAYOUT_CORNER>;
        }
        if (storage == BH_STORE_F32) return surrogate_batch_kernel<BH_STORE_F32, BH_LAYOUT_CORNER>;
        if (storage == BH_STORE_BF16) return surrogate_batch_kernel<BH_STORE_BF16, BH_LAYOUT_CORNER>;
        return surrogate_batch_kernel<BH_STORE_F16, BH_LAYOUT_CORNER>;
    }
    if (storage == BH_STORE_F32) return surrogate_batch_kernel<BH_STORE_F32, BH_LAYOUT_SOA>;
    if (storage == BH_STORE_BF16) return surrogate_batch_kernel<BH_STORE_BF16, BH_LAYOUT_SOA>;
    return surrogate_batch_kernel<BH_STORE_F16, BH_LAYOUT_SOA>;
}

static size_t surrogate_work_bytes(int64_t rows, int n_knots) {
    const size_t n_out = size_t(4 * n_knots + kHeads);
    const size_t al = 256;
    auto up = [&](size_t b) { return (b + al - 1) / al * al; };
    return 2 * up(n_out * 4) + up(2 * kCurveMax * 4) + up(64 * 4) + up(size_t(rows) * 4) + up(size_t(rows) * 16) +
           up(size_t(rows) * n_out * 4);
}


// Shared body of the surrogate entry points: knot rows come from the MLP (mlp non-null; network
// input and output live in the work buffer) or from the caller (rows, [n_out][ld]).
// Critical curve of a spin and inclination as (alpha, beta) pairs, thinned to kCurveMax points.
static std::vector<float> thinned_curve(double spin, double theta) {
    std::vector<double> al, be;
    if (std::sin(theta) > 1e-3) critical_curve(spin, theta, al, be);
    std::vector<float> curve;
    const size_t stride = al.empty() ? 1 : (al.size() + kCurveMax - 1) / kCurveMax;
    for (size_t j = 0; j < al.size(); j += stride) {
        curve.push_back(float(al[j]));
        curve.push_back(float(be[j]));
    }
    return curve;
}

static int check_surrogate_desc(const bh_surrogate_desc *sd, bh_mlp *mlp, bh_ctx *ctx) {
    if (!sd || !sd->out_mul || !sd->out_add) return set_error(BH_ERR_VALUE, "NULL surrogate argument");
    const int K = sd->n_knots;
    if (K < 2 || K > kMaxKnots) return set_error(BH_ERR_VALUE, "n_knots must lie in [2, %d]", kMaxKnots);
    if (mlp) {
        if (mlp_context(mlp) != ctx) return set_error(BH_ERR_VALUE, "the MLP belongs to another context");
        if (mlp_in_cols(mlp) != 4 || mlp_out_cols(mlp) != 4 * K + kHeads)
            return set_error(BH_ERR_VALUE, "the MLP must map 4 inputs to 4 * n_knots + %d outputs", kHeads);
    }
    if (sd->fallback_precision != BH_PREC_FP32 && sd->fallback_precision != BH_PREC_MIXED)
        return set_error(BH_ERR_VALUE, "fallback_precision must be BH_PREC_FP32 or BH_PREC_MIXED");
    return BH_OK;
}

// Host floats into device memory through kernel arguments, which the launch copies: the host array need not
// outlive the call and the stream is not synchronised, so a caller may capture the sequence in a graph.
constexpr int kArgFloats = 896;   // 3.5 KB of values per launch, inside the kernel-argument limit
struct FloatChunk {
    float v[kArgFloats];
    float *dst;
    int n;
};

__global__ __launch_bounds__(256) void upload_floats_kernel(FloatChunk c) {
    for (int i = threadIdx.x; i < c.n; i += blockDim.x) c.dst[i] = c.v[i];
}

static int upload_floats(hipStream_t stream, float *dst, const float *src, size_t n) {
    for (size_t o = 0; o < n; o += kArgFloats) {
        FloatChunk c;
        c.dst = dst + o;
        c.n = int(n - o < size_t(kArgFloats) ? n - o : size_t(kArgFloats));
        std::memcpy(c.v, src + o, size_t(c.n) * sizeof(float));
        hipLaunchKernelGGL(upload_floats_kernel, dim3(1), dim3(256), 0, stream, c);
    }
    BH_HIP(hipGetLastError());
    return BH_OK;
}

static int knot_render(bh_ctx *ctx, const bh_scene_desc *scene, int32_t brick_id, bh_mlp *mlp, const float *rows_in,
                       int64_t ld_in, const bh_surrogate_desc *sd, float *d_work, size_t work_bytes, void *d_iquv,
                       int32_t *d_steps, uint8_t *d_status) {
    int st;
    if (!scene || !sd || !d_work || !d_iquv || !sd->out_mul || !sd->out_add)
        return set_error(BH_ERR_VALUE, "NULL surrogate argument");
    if ((st = validate_scene(*scene))) return st;
    if (scene->transfer != BH_TRANSFER_RK2)
        return set_error(BH_ERR_VALUE, "the surrogate's transfer over knots takes the RK2 Stokes step only");
    if (brick_id < 0 || size_t(brick_id) >= ctx->bricks.size() || !ctx->bricks[brick_id].live)
        return set_error(BH_ERR_STATE, "unknown brick id %d", brick_id);
    const int K = sd->n_knots;
    if (K < 2 || K > kMaxKnots) return set_error(BH_ERR_VALUE, "n_knots must lie in [2, %d]", kMaxKnots);
    const int n_out = 4 * K + kHeads;
    if (mlp) {
        if (mlp_context(mlp) != ctx) return set_error(BH_ERR_VALUE, "the MLP belongs to another context");
        if (mlp_in_cols(mlp) != 4 || mlp_out_cols(mlp) != n_out)
            return set_error(BH_ERR_VALUE, "the MLP must map 4 inputs to 4 * n_knots + %d outputs", kHeads);
    }
    if (sd->fallback_precision != BH_PREC_FP32 && sd->fallback_precision != BH_PREC_MIXED)
        return set_error(BH_ERR_VALUE, "fallback_precision must be BH_PREC_FP32 or BH_PREC_MIXED");
    const int64_t rows = int64_t(scene->width) * scene->height;
    const size_t need = surrogate_work_bytes(mlp ? rows : 0, K) + (mlp ? 0 : size_t(rows) * 4 + 256);
    if (work_bytes < need) return set_error(BH_ERR_VALUE, "work buffer holds %zu bytes, %zu needed", work_bytes, need);
    const BrickDev &b = ctx->bricks[brick_id];
    SurArgs S{};
    if ((st = setup_args(*scene, b, S.R))) return st;
    S.R.iquv = d_iquv;
    S.R.steps = d_steps;
    S.R.status = d_status;
    // work layout: mul | add | curve | counter | fallback list | network input | network output
    char *w = reinterpret_cast<char *>(d_work);
    auto take = [&](size_t bytes) {
        char *p = w;
        w += (bytes + 255) / 256 * 256;
        return p;
    };
    float *d_mul = reinterpret_cast<float *>(take(size_t(n_out) * 4));
    float *d_add = reinterpret_cast<float *>(take(size_t(n_out) * 4));
    float *d_curve = reinterpret_cast<float *>(take(2 * kCurveMax * 4));
    unsigned *d_count = reinterpret_cast<unsigned *>(take(64 * 4));
    int *d_list = reinterpret_cast<int *>(take(size_t(rows) * 4));
    float *d_x = mlp ? reinterpret_cast<float *>(take(size_t(rows) * 16)) : nullptr;
    float *d_y = mlp ? reinterpret_cast<float *>(take(size_t(rows) * n_out * 4)) : nullptr;
    const std::vector<float> curve = thinned_curve(scene->spin, S.R.cam.th);
    S.n_curve = int(curve.size() / 2);
    if ((st = upload_floats(ctx->stream, d_mul, sd->out_mul, size_t(n_out))) ||
        (st = upload_floats(ctx->stream, d_add, sd->out_add, size_t(n_out))) ||
        (st = upload_floats(ctx->stream, d_curve, curve.data(), curve.size())))
        return st;
    // the fused inference writes the knots as FP16 (half the bytes to write and to read back)
    const bool half = mlp && mlp_half_outputs(mlp);
    S.y = mlp ? static_cast<const void *>(d_y) : static_cast<const void *>(rows_in);
    S.ld = mlp ? rows : ld_in;
    S.mul = d_mul;
    S.add = d_add;
    S.curve = d_curve;
    S.K = K;
    S.crit2 = sd->crit_distance * sd->crit_distance;
    S.flag_max = sd->flag_max;
    S.cap_lo = sd->cap_lo;
    S.cap_hi = sd->cap_hi;
    S.list = d_list;
    S.n_list = d_count;
    ctx->sur_count = d_count;

    begin_timing(ctx);
    hipLaunchKernelGGL(fill_u32_once, dim3(1), dim3(64), 0, ctx->stream, d_count);
    if (mlp) {
        const float4 imul = make_float4(sd->in_mul[0], sd->in_mul[1], sd->in_mul[2], sd->in_mul[3]);
        const float4 iadd = make_float4(sd->in_add[0], sd->in_add[1], sd->in_add[2], sd->in_add[3]);
        hipLaunchKernelGGL(surrogate_input_kernel, dim3(unsigned((rows + 255) / 256)), dim3(256), 0, ctx->stream, S.R,
                           imul, iadd, d_x);
        if (half) {
            if ((st = mlp_forward_enqueue_f16(mlp, d_x, rows, reinterpret_cast<_Float16 *>(d_y), rows))) return st;
        } else if ((st = mlp_forward_enqueue(mlp, d_x, rows, d_y, 1))) {
            return st;
        }
    }
    const int tiles_y = (S.R.H + 7) / 8;
    const int n_tiles = S.R.tiles_x * tiles_y;
    hipLaunchKernelGGL(pick_surrogate(b.storage, b.layout, half, wide_corner(b.layout, b.bytes)),
                       dim3(unsigned((n_tiles + 3) / 4)), dim3(256), 0,
                       ctx->stream, S);
    BH_HIP(hipGetLastError());
    if ((st = render_enqueue(ctx, scene, brick_id, sd->fallback_precision, d_iquv, d_steps, d_status, d_list, d_count,
                             int(rows), false)))
        return st;
    end_timing(ctx, 1);
    return BH_OK;
}

}  // namespace bh

extern "C" {

int bh_geodesic_knots(bh_ctx *ctx, const bh_knot_desc *desc, const float *d_params, int64_t n, float *d_rows,
                      int32_t *d_steps, float *d_weight) {
    clear_error();
    int st = ctx_check(ctx);
    if (st) return st;
    if (!desc || !d_params || !d_rows || n < 0) return set_error(BH_ERR_VALUE, "bad knot arguments");
    const bh_knot_desc &d = *desc;
    if (d.n_knots < 2 || d.n_knots > kMaxKnots) return set_error(BH_ERR_VALUE, "n_knots must lie in [2, %d]", kMaxKnots);
    if (!(d.cam_r > 2.0) || !(d.step > 0.0) || d.max_steps < 1 || !(d.r_hi > d.r_lo))
        return set_error(BH_ERR_VALUE, "bad knot scene (cam_r > 2, step > 0, max_steps >= 1, r_hi > r_lo)");
    KnotArgs A{};
    A.params = d_params;
    A.n = n;
    A.cam_r = d.cam_r;
    A.step = d.step;
    A.r_esc = d.escape_radius;
    A.r_lo = d.r_lo;
    A.r_hi = d.r_hi;
    A.adaptive = d.step_mode;
    A.max_steps = d.max_steps;
    A.K = d.n_knots;
    A.rows = d_rows;
    A.steps = d_steps;
    A.weight = d_weight;
    if (n == 0) return BH_OK;
    begin_timing(ctx);
    hipLaunchKernelGGL(knots_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, ctx->stream, A);
    BH_HIP(hipGetLastError());
    end_timing(ctx, 1);
    return BH_OK;
}

int bh_surrogate_work_bytes(int64_t rows, int32_t n_knots, size_t *bytes) {
    if (!bytes || rows < 0 || n_knots < 2 || n_knots > kMaxKnots) return set_error(BH_ERR_VALUE, "bad arguments");
    *bytes = surrogate_work_bytes(rows, n_knots);
    return BH_OK;
}

int bh_surrogate_render_device(bh_ctx *ctx, const bh_scene_desc *scene, int32_t brick_id, bh_mlp *mlp,
                               const bh_surrogate_desc *sd, float *d_work, size_t work_bytes, void *d_iquv,
                               int32_t *d_steps, uint8_t *d_status) {
    clear_error();
    int st = ctx_check(ctx);
    if (st) return st;
    if (!mlp) return set_error(BH_ERR_VALUE, "NULL MLP");
    return knot_render(ctx, scene, brick_id, mlp, nullptr, 0, sd, d_work, work_bytes, d_iquv, d_steps, d_status);
}

int bh_knot_render_device(bh_ctx *ctx, const bh_scene_desc *scene, int32_t brick_id, const float *d_rows, int64_t ld,
                          const bh_surrogate_desc *sd, float *d_work, size_t work_bytes, void *d_iquv, int32_t *d_steps,
                          uint8_t *d_status) {
    clear_error();
    int st = ctx_check(ctx);
    if (st) return st;
    if (!d_rows) return set_error(BH_ERR_VALUE, "NULL knot rows");
    if (scene && ld < int64_t(scene->width) * scene->height) return set_error(BH_ERR_VALUE, "ld must be >= width * height");
    return knot_render(ctx, scene, brick_id, nullptr, d_rows, ld, sd, d_work, work_bytes, d_iquv, d_steps, d_status);
}

int bh_surrogate_last_fallback(bh_ctx *ctx, int64_t *n_fallback) {
    clear_error();
    int st = ctx_check(ctx);
    if (st) return st;
    if (!n_fallback) return set_error(BH_ERR_VALUE, "NULL output");
    if (!ctx->sur_count) return set_error(BH_ERR_STATE, "no surrogate render on this context");
    unsigned v = 0;
    BH_HIP(hipMemcpyAsync(&v, ctx->sur_count, 4, hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    *n_fallback = int64_t(v);
    return BH_OK;
}

}  // extern "C"

// ---- batched surrogate renders --------------------------------------------------------------
// Many renders through the learned surrogate in one sequence: the network inputs of every render's rays,
// one fused FP16 inference over all of them, the knot transfer of every render's tiles in one launch, and
// every render's fallback rays pooled into one persistent batched RK2 launch (list mode).  The critical
// curves are computed once at creation; the launch needs no host synchronisation and is captured in a
// hipGraph on first use, then replayed.

struct bh_surrogate_batch {
    bh_ctx *ctx = nullptr;
    bh_batch *fb = nullptr;              // merged bricks, per-render descriptors, the fallback RK2 kernel
    bh_mlp *mlp = nullptr;
    int n = 0, n_tiles = 0;
    long long total = 0;
    float4 in_mul{}, in_add{};
    bh::SurShared sh{};
    bh::SurBatchFn sfn = nullptr;
    bh::BatchFn list_fn = nullptr;       // the fallback kernel in list mode
    float *d_mul = nullptr, *d_add = nullptr, *d_curves = nullptr, *d_x = nullptr;
    int *d_curve_at = nullptr, *d_curve_n = nullptr, *d_tiles = nullptr, *d_list = nullptr;
    _Float16 *d_y = nullptr;
    unsigned *d_counts = nullptr;        // n per-render fallback counts, then the fallback chunk counter
    unsigned *d_prefix = nullptr;        // n + 1
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *g_iquv = nullptr, *g_steps = nullptr, *g_status = nullptr;
    int g_grid = 0, g_block = 0, g_spc = 0;
};

namespace bh {

// Enqueue the whole batched surrogate render on the context stream.
static int surrogate_batch_enqueue(bh_surrogate_batch *sb, void *d_iquv, int32_t *d_steps, uint8_t *d_status,
                                   const BatchArgs &B, int grid, int block) {
    bh_ctx *ctx = sb->ctx;
    int st;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(kCounterWords), 0, ctx->stream, ctx->d_counters, 0u, kCounterWords);
    if ((st = fill_u32(ctx, sb->d_counts, 0u, sb->n + 1))) return st;
    hipLaunchKernelGGL(surrogate_input_batch_kernel, dim3(unsigned((sb->total + 255) / 256)), dim3(256), 0, ctx->stream,
                       sb->fb->d_descs, sb->n, sb->total, sb->in_mul, sb->in_add, sb->d_x);
    if ((st = mlp_forward_enqueue_f16(sb->mlp, sb->d_x, sb->total, sb->d_y, sb->total))) return st;
    SurBatchArgs S{};
    S.descs = sb->fb->d_descs;
    S.tiles = sb->d_tiles;
    S.n_tiles = sb->n_tiles;
    S.grids = sb->fb->d_grids;
    S.ng = sb->fb->ng;
    S.y = sb->d_y;
    S.ld = sb->total;
    S.curves = sb->d_curves;
    S.curve_at = sb->d_curve_at;
    S.curve_n = sb->d_curve_n;
    S.sh = sb->sh;
    S.iquv = d_iquv;
    S.steps = d_steps;
    S.status = d_status;
    S.list = sb->d_list;
    S.counts = sb->d_counts;
    hipLaunchKernelGGL(sb->sfn, dim3(unsigned((sb->n_tiles + 3) / 4)), dim3(256), 0, ctx->stream, S);
    hipLaunchKernelGGL(fallback_chunks_kernel, dim3(1), dim3(64), 0, ctx->stream, sb->d_counts, sb->n, sb->d_prefix);
    void *args[] = {const_cast<BatchArgs *>(&B)};
    BH_HIP(hipLaunchKernel(reinterpret_cast<const void *>(sb->list_fn), dim3(grid), dim3(block), args, 0, ctx->stream));
    BH_HIP(hipGetLastError());
    return BH_OK;
}

}  // namespace bh

extern "C" {

int bh_surrogate_batch_destroy(bh_surrogate_batch *sb) {
    if (!sb) return BH_OK;
    if (sb->ctx) {
        (void)hipSetDevice(sb->ctx->device);
        (void)hipStreamSynchronize(sb->ctx->stream);
    }
    if (sb->exec) (void)hipGraphExecDestroy(sb->exec);
    if (sb->graph) (void)hipGraphDestroy(sb->graph);
    for (void *p : {(void *)sb->d_mul, (void *)sb->d_add, (void *)sb->d_curves, (void *)sb->d_x, (void *)sb->d_curve_at,
                    (void *)sb->d_curve_n, (void *)sb->d_tiles, (void *)sb->d_list, (void *)sb->d_y,
                    (void *)sb->d_counts, (void *)sb->d_prefix})
        if (p) (void)hipFree(p);
    if (sb->fb) bh_batch_destroy(sb->fb);
    if (sb->ctx && sb->ctx->sur_count == sb->d_counts) sb->ctx->sur_count = nullptr;
    delete sb;
    return BH_OK;
}

int bh_surrogate_batch_create(bh_ctx *ctx, const bh_batch_item *items, int32_t n, bh_mlp *mlp,
                              const bh_surrogate_desc *sd, int64_t *offsets, bh_surrogate_batch **out) {
    clear_error();
    int st = ctx_check(ctx);
    if (st) return st;
    if (!out || !items || n < 1) return set_error(BH_ERR_VALUE, "a surrogate batch needs at least one render");
    *out = nullptr;
    if (!mlp) return set_error(BH_ERR_VALUE, "NULL MLP");
    if ((st = check_surrogate_desc(sd, mlp, ctx))) return st;
    if (!mlp_half_outputs(mlp))
        return set_error(BH_ERR_VALUE, "the batched surrogate needs the fused width-256 inference (FP16 knots)");
    for (int i = 0; i < n; ++i) {
        if ((st = validate_scene(items[i].scene))) return st;
        if (items[i].scene.transfer != BH_TRANSFER_RK2)
            return set_error(BH_ERR_VALUE, "the surrogate's transfer over knots takes the RK2 Stokes step only");
        if (items[i].scene.integrator != BH_INTEGRATOR_RK2)
            return set_error(BH_ERR_VALUE, "the surrogate is trained on RK2 geodesics: integrator must be RK2");
    }
    bh_surrogate_batch *sb = new bh_surrogate_batch();
    sb->ctx = ctx;
    sb->mlp = mlp;
    sb->n = n;
    auto fail = [&](int code) {
        bh_surrogate_batch_destroy(sb);
        return code;
    };
    std::vector<int64_t> off(size_t(n) + 1);
    if ((st = bh_batch_create(ctx, items, n, sd->fallback_precision, off.data(), &sb->fb))) return fail(st);
    if (offsets) std::copy(off.begin(), off.end(), offsets);
    sb->total = off[n];
    if (sb->total >= (int64_t(1) << 31)) return fail(set_error(BH_ERR_VALUE, "a surrogate batch holds at most 2^31 - 1 rays"));
    const bool wide = wide_corner(sb->fb->layout, sb->fb->coeff_bytes);
    sb->sfn = pick_surrogate_batch(sb->fb->storage, sb->fb->layout, wide);
    sb->list_fn = pick_list_kernel(sb->fb->storage, sb->fb->layout, sd->fallback_precision == BH_PREC_MIXED, wide);
    const int K = sd->n_knots, n_out = 4 * K + kHeads;
    sb->in_mul = make_float4(sd->in_mul[0], sd->in_mul[1], sd->in_mul[2], sd->in_mul[3]);
    sb->in_add = make_float4(sd->in_add[0], sd->in_add[1], sd->in_add[2], sd->in_add[3]);
    // critical curves and the tile table
    std::vector<float> curves;
    std::vector<int> curve_at(n), curve_n(n), tiles;
    for (int i = 0; i < n; ++i) {
        CamConst C{};
        if ((st = camera_constants(items[i].scene, C))) return fail(st);
        const std::vector<float> c = thinned_curve(items[i].scene.spin, C.th);
        curve_at[i] = int(curves.size() / 2);
        curve_n[i] = int(c.size() / 2);
        curves.insert(curves.end(), c.begin(), c.end());
        const int tx = (items[i].scene.width + 7) / 8, ty = (items[i].scene.height + 7) / 8;
        if (tx * ty >= (1 << 20)) return fail(set_error(BH_ERR_VALUE, "render %d has too many 8x8 tiles", i));
        for (int t = 0; t < tx * ty; ++t) tiles.push_back((i << 20) | t);
    }
    if (curves.empty()) curves.assign(2, 0.f);
    sb->n_tiles = int(tiles.size());
    const size_t rows = size_t(sb->total);
    auto alloc = [&](void **p, size_t bytes) { return hipMalloc(p, bytes ? bytes : 4) == hipSuccess; };
    if (!alloc((void **)&sb->d_mul, n_out * 4) || !alloc((void **)&sb->d_add, n_out * 4) ||
        !alloc((void **)&sb->d_curves, curves.size() * 4) || !alloc((void **)&sb->d_curve_at, size_t(n) * 4) ||
        !alloc((void **)&sb->d_curve_n, size_t(n) * 4) || !alloc((void **)&sb->d_tiles, tiles.size() * 4) ||
        !alloc((void **)&sb->d_x, rows * 16) || !alloc((void **)&sb->d_y, rows * size_t(n_out) * 2) ||
        !alloc((void **)&sb->d_list, rows * 4) || !alloc((void **)&sb->d_counts, (size_t(n) + 1) * 4) ||
        !alloc((void **)&sb->d_prefix, (size_t(n) + 1) * 4))
        return fail(set_error(BH_ERR_MEMORY, "surrogate batch allocation failed"));
    hipError_t e = hipMemcpyAsync(sb->d_mul, sd->out_mul, size_t(n_out) * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sb->d_add, sd->out_add, size_t(n_out) * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sb->d_curves, curves.data(), curves.size() * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sb->d_curve_at, curve_at.data(), size_t(n) * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sb->d_curve_n, curve_n.data(), size_t(n) * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sb->d_tiles, tiles.data(), tiles.size() * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(set_error(BH_ERR_DEVICE, "surrogate batch upload failed: %s", hipGetErrorString(e)));
    sb->sh = SurShared{sb->d_mul, sb->d_add, K, sd->crit_distance * sd->crit_distance, sd->flag_max, sd->cap_lo,
                       sd->cap_hi};
    *out = sb;
    return BH_OK;
}

int bh_surrogate_batch_launch(bh_surrogate_batch *sb, void *d_iquv, int32_t *d_steps, uint8_t *d_status, int32_t graph) {
    clear_error();
    if (!sb) return set_error(BH_ERR_STATE, "NULL surrogate batch");
    int st = ctx_check(sb->ctx);
    if (st) return st;
    if (!d_iquv) return set_error(BH_ERR_VALUE, "iquv must not be NULL");
    bh_ctx *ctx = sb->ctx;
    BatchArgs B{};
    int grid = 0, block = 0;
    if ((st = batch_args(sb->fb, d_iquv, d_steps, d_status, B, grid, block))) return st;
    B.counters = sb->d_counts + sb->n;   // the fallback chunk counter
    B.queues = 1;
    B.list = sb->d_list;
    B.list_count = sb->d_counts;
    B.chunk_prefix = sb->d_prefix;
    B.n_renders = sb->n;
    ctx->sur_count = nullptr;
    ctx->last_counters = ctx->d_counters;
    begin_timing(ctx);
    if (!graph) {
        if ((st = surrogate_batch_enqueue(sb, d_iquv, d_steps, d_status, B, grid, block))) return st;
        end_timing(ctx, 1);
        return BH_OK;
    }
    const bool same = sb->exec && sb->g_iquv == d_iquv && sb->g_steps == (void *)d_steps &&
                      sb->g_status == (void *)d_status && sb->g_grid == grid && sb->g_block == block &&
                      sb->g_spc == B.steps_per_sync;
    if (!same) {
        if (sb->exec) (void)hipGraphExecDestroy(sb->exec);
        if (sb->graph) (void)hipGraphDestroy(sb->graph);
        sb->exec = nullptr;
        sb->graph = nullptr;
        BH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int est = surrogate_batch_enqueue(sb, d_iquv, d_steps, d_status, B, grid, block);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
        if (est != BH_OK) {
            if (g) (void)hipGraphDestroy(g);
            return est;
        }
        if (e2 != hipSuccess || !g) return set_error(BH_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e2));
        sb->graph = g;
        BH_HIP(hipGraphInstantiate(&sb->exec, sb->graph, nullptr, nullptr, 0));
        sb->g_iquv = d_iquv;
        sb->g_steps = d_steps;
        sb->g_status = d_status;
        sb->g_grid = grid;
        sb->g_block = block;
        sb->g_spc = B.steps_per_sync;
    }
    BH_HIP(hipGraphLaunch(sb->exec, ctx->stream));
    end_timing(ctx, 1);
    return BH_OK;
}

int bh_surrogate_batch_fallback(bh_surrogate_batch *sb, int64_t *per_render, int64_t *total) {
    clear_error();
    if (!sb) return set_error(BH_ERR_STATE, "NULL surrogate batch");
    int st = ctx_check(sb->ctx);
    if (st) return st;
    std::vector<unsigned> c(size_t(sb->n));
    BH_HIP(hipMemcpyAsync(c.data(), sb->d_counts, c.size() * 4, hipMemcpyDeviceToHost, sb->ctx->stream));
    BH_HIP(hipStreamSynchronize(sb->ctx->stream));
    int64_t t = 0;
    for (int i = 0; i < sb->n; ++i) {
        if (per_render) per_render[i] = int64_t(c[size_t(i)]);
        t += int64_t(c[size_t(i)]);
    }
    if (total) *total = t;
    return BH_OK;
}

}  // extern "C"
