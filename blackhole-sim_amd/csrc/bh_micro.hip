// FP64 drop-in micro-kernels: stokes_rk2_brick, sample_brick_trilinear,
// sample_and_step_stokes (+ invalid_policy) and sample_brick_valid_mask.
//
// Arithmetic follows native/core/src/lib.rs:37-366 and
// python/blackhole_sim/native_kernels.py:133-222: searchsorted-right brackets
// clamped to [0, n-2], the x == grid[-1] special case, the periodic phi seam
// cell anchored at phi_grid[0], trilinear lerps (1-w)a + wb in phi, theta, r
// order, and the RK2 Stokes step k1 = f(S), k2 = f(S + ds/2 k1), S + ds k2 with
// f = j - K S.  One thread per point/cell; the validation mirrors lib.rs
// (messages included) and runs on the host before anything is copied.

#include "bh_internal.h"

#include <algorithm>
#include <cmath>

namespace bh {

// ---- device arithmetic ----------------------------------------------------

__device__ __forceinline__ int search_right_d(const double *g, int n, double x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = lo + ((hi - lo) >> 1);
        if (x < g[mid]) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// native_kernels.py:137-146; false when outside the closed range.
__device__ __forceinline__ bool bracket_linear_d(const double *g, int n, double x, int &i0, double &w) {
    if (x < g[0] || x > g[n - 1]) return false;
    if (x == g[n - 1]) {
        i0 = n - 2;
        w = 1.0;
        return true;
    }
    int i = search_right_d(g, n, x) - 1;
    i = i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
    double den = g[i + 1] - g[i];
    den = (1.0e-300 > den) ? 1.0e-300 : den;
    i0 = i;
    w = (x - g[i]) / den;
    return true;
}

// Python float %: fmod, shifted into [0, 2 pi) when negative (native_kernels.py:133-134).
__device__ __forceinline__ double wrap_phi_d(double phi, double base) {
    double m = fmod(phi - base, kTwoPi);
    if (m < 0.0) m += kTwoPi;
    return m + base;
}

// native_kernels.py:149-161.
__device__ __forceinline__ void bracket_phi_d(const double *g, int n, double phi, int &i0, int &i1, double &w) {
    double p = wrap_phi_d(phi, g[0]);
    int i = search_right_d(g, n, p) - 1;
    if (i < 0) i = n - 1;
    if (i >= n) i = n - 1;
    int j = (i + 1) % n;
    double hi = (j > i) ? g[j] : g[0] + kTwoPi;
    double pp = (p >= g[i]) ? p : p + kTwoPi;
    double den = hi - g[i];
    den = (1.0e-300 > den) ? 1.0e-300 : den;
    i0 = i;
    i1 = j;
    w = (pp - g[i]) / den;
}

// One channel's trilinear value from its 8 corner values (o000 .. o111, phi fastest): the nested
// (1 - w) a + w b lerps of lib.rs:199-205 and native_kernels.py:180-186, every product and sum rounded
// on its own as Rust and NumPy evaluate them (no fused multiply-add), so all sampling kernels agree bit
// for bit.
__device__ __forceinline__ double trilerp_d(double v000, double v001, double v010, double v011, double v100,
                                            double v101, double v110, double v111, double wr, double wt, double wp) {
#pragma clang fp contract(off)
    const double c00 = (1.0 - wp) * v000 + wp * v001;
    const double c01 = (1.0 - wp) * v010 + wp * v011;
    const double c10 = (1.0 - wp) * v100 + wp * v101;
    const double c11 = (1.0 - wp) * v110 + wp * v111;
    const double c0 = (1.0 - wt) * c00 + wt * c01;
    const double c1 = (1.0 - wt) * c10 + wt * c11;
    return (1.0 - wr) * c0 + wr * c1;
}

// lib.rs:171-208 over an AoS FP64 brick.
__device__ __forceinline__ bool sample_aos_d(const double *__restrict__ c, const double *rg, int nr, const double *tg,
                                             int nt, const double *pg, int np, double r, double th, double ph,
                                             double out[kCoeffs]) {
    int r0, t0, p0, p1;
    double wr, wt, wp;
    if (!bracket_linear_d(rg, nr, r, r0, wr)) return false;
    if (!bracket_linear_d(tg, nt, th, t0, wt)) return false;
    bracket_phi_d(pg, np, ph, p0, p1, wp);
    const int r1 = r0 + 1, t1 = t0 + 1;
    const size_t o000 = ((size_t(r0) * nt + t0) * np + p0) * kCoeffs;
    const size_t o001 = ((size_t(r0) * nt + t0) * np + p1) * kCoeffs;
    const size_t o010 = ((size_t(r0) * nt + t1) * np + p0) * kCoeffs;
    const size_t o011 = ((size_t(r0) * nt + t1) * np + p1) * kCoeffs;
    const size_t o100 = ((size_t(r1) * nt + t0) * np + p0) * kCoeffs;
    const size_t o101 = ((size_t(r1) * nt + t0) * np + p1) * kCoeffs;
    const size_t o110 = ((size_t(r1) * nt + t1) * np + p0) * kCoeffs;
    const size_t o111 = ((size_t(r1) * nt + t1) * np + p1) * kCoeffs;
#pragma unroll
    for (int k = 0; k < kCoeffs; ++k)
        out[k] = trilerp_d(c[o000 + k], c[o001 + k], c[o010 + k], c[o011 + k], c[o100 + k], c[o101 + k], c[o110 + k],
                           c[o111 + k], wr, wt, wp);
    return true;
}

// lib.rs:37-59: j - K S.
__device__ __forceinline__ void stokes_rhs_d(const double s[4], const double *c, double o[4]) {
    o[0] = c[0] - (c[4] * s[0] + c[5] * s[1] + c[6] * s[2] + c[7] * s[3]);
    o[1] = c[1] - (c[5] * s[0] + c[4] * s[1] + c[8] * s[2] - c[10] * s[3]);
    o[2] = c[2] - (c[6] * s[0] - c[8] * s[1] + c[4] * s[2] + c[9] * s[3]);
    o[3] = c[3] - (c[7] * s[0] + c[10] * s[1] - c[9] * s[2] + c[4] * s[3]);
}

// lib.rs:210-225.
__device__ __forceinline__ void stokes_rk2_d(double s[4], const double *c, double ds) {
    double k1[4], mid[4], k2[4];
    stokes_rhs_d(s, c, k1);
#pragma unroll
    for (int i = 0; i < 4; ++i) mid[i] = s[i] + 0.5 * ds * k1[i];
    stokes_rhs_d(mid, c, k2);
#pragma unroll
    for (int i = 0; i < 4; ++i) s[i] = s[i] + ds * k2[i];
}

// ---- kernels ---------------------------------------------------------------

__global__ void stokes_rk2_kernel(const double *__restrict__ coeffs, int64_t cells, double ds,
                                  const double *__restrict__ init, int broadcast, double *__restrict__ out) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    const double *ip = init + (broadcast ? 0 : i * kStokes);
    double s[4] = {ip[0], ip[1], ip[2], ip[3]};
    stokes_rk2_d(s, coeffs + i * kCoeffs, ds);
    double *op = out + i * kStokes;
#pragma unroll
    for (int k = 0; k < 4; ++k) op[k] = s[k];
}

constexpr int kGatherLanes = 16;
constexpr int kGatherGrid = 2048;   // grid values staged in LDS (r + theta + phi)

__device__ __forceinline__ void gather_points(const double *__restrict__ coeffs, const double *rg, int nr,
                                              const double *tg, int nt, const double *pg, int np, const double *__restrict__ pts, int64_t n,
                                              double *__restrict__ out, double (*vals)[8 * kCoeffs]) {
    constexpr int kGroups = 256 / kGatherLanes;
    constexpr int kVals = 8 * kCoeffs;
    const int g = threadIdx.x / kGatherLanes, gl = threadIdx.x % kGatherLanes;
    const int64_t stride = int64_t(gridDim.x) * kGroups;
    for (int64_t i = int64_t(blockIdx.x) * kGroups + g; i < n; i += stride) {
        const double *p = pts + i * 3;
        const double r = p[0], th = p[1], ph = p[2];
        int r0 = 0, t0 = 0, p0 = 0, p1 = 0;
        double wr = 0.0, wt = 0.0, wp = 0.0;
        const bool ok = bracket_linear_d(rg, nr, r, r0, wr) && bracket_linear_d(tg, nt, th, t0, wt);
        if (ok) {
            bracket_phi_d(pg, np, ph, p0, p1, wp);
#pragma unroll
            for (int round = 0; round < (kVals + kGatherLanes - 1) / kGatherLanes; ++round) {
                const int f = round * kGatherLanes + gl;   // value index: corner q = f / 11, channel f % 11
                if (f < kVals) {
                    const int q = f / kCoeffs, k = f - q * kCoeffs;
                    const int ri = r0 + (q >> 2), ti = t0 + ((q >> 1) & 1), pi = (q & 1) ? p1 : p0;
                    vals[g][f] = coeffs[((size_t(ri) * nt + ti) * np + pi) * kCoeffs + k];
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (gl < kCoeffs) {
            double res = __builtin_nan("");
            if (ok) {
                const double *v = vals[g];
                const int k = gl;
                // corners in order (dr, dt, dp) = q >> 2, (q >> 1) & 1, q & 1, as o000 .. o111
                res = trilerp_d(v[0 * kCoeffs + k], v[1 * kCoeffs + k], v[2 * kCoeffs + k], v[3 * kCoeffs + k],
                                v[4 * kCoeffs + k], v[5 * kCoeffs + k], v[6 * kCoeffs + k], v[7 * kCoeffs + k], wr, wt, wp);
            }
            out[i * kCoeffs + gl] = res;
        }
        __builtin_amdgcn_wave_barrier();   // the next point reuses this group's LDS values
    }
}

// FP64 trilinear sampler over many points (the rule of sample_aos_d; NaN for a point outside the
// r/theta range) with a cooperative gather: a group of 16 lanes serves one point.  The group loads the
// point's 8 corner records (88 doubles) with consecutive lanes on consecutive doubles of a record, so
// one load instruction touches a few whole sectors instead of one sector per lane; the values go
// through LDS, and lane k < 11 interpolates channel k with the nested lerps of sample_aos_d and writes
// it, so a point's 11 outputs are one contiguous 88-byte store.  Grids of up to 2048 values in all are
// searched in LDS.
__global__ __launch_bounds__(256) void sample_gather_kernel(const double *__restrict__ coeffs,
                                                            const double *__restrict__ rg, int nr,
                                                            const double *__restrict__ tg, int nt,
                                                            const double *__restrict__ pg, int np,
                                                            const double *__restrict__ pts, int64_t n,
                                                            double *__restrict__ out) {
    __shared__ double vals[256 / kGatherLanes][8 * kCoeffs];
    __shared__ double sgrid[kGatherGrid];
    if (nr + nt + np <= kGatherGrid) {
        for (int j = threadIdx.x; j < nr + nt + np; j += blockDim.x)
            sgrid[j] = j < nr ? rg[j] : (j < nr + nt ? tg[j - nr] : pg[j - nr - nt]);
        __syncthreads();
        gather_points(coeffs, sgrid, nr, sgrid + nr, nt, sgrid + nr + nt, np, pts, n, out, vals);
    } else {
        gather_points(coeffs, rg, nr, tg, nt, pg, np, pts, n, out, vals);
    }
}

// ---- points bucketed by brick cell ------------------------------------------------------------
// Random points gather eight random 88-byte corner records each; sorted by the trilinear cell they fall
// in, the points of a cell and of its phi neighbours (which share four corners) sit in neighbouring lanes,
// so one lane per point loads its corners directly: the lanes of a load instruction hit a few records,
// which come from L1/L2 instead of the fabric, and a wave keeps 64 points in flight.  A counting sort:
// the histogram's atomic returns each point's rank inside its cell, an exclusive scan turns the counts
// into cell offsets, and a scatter writes each point (coordinates, index, cell) to offset + rank.  The
// gather takes the cell from the sort key instead of searching the grids; the weights and the lerps are
// those of the unsorted gather (same expressions on the same cell), so the results are the same bits.

constexpr int kScanBlock = 1024;   // threads of the scan kernels: 16 waves of 64

// Inclusive prefix sum over the 1024 threads of a block; lds holds 16 words.
__device__ __forceinline__ unsigned block_inclusive_scan(unsigned v, unsigned *lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    if (lane == 63) lds[wave] = v;
    __syncthreads();
    if (wave == 0) {
        unsigned t = lane < 16 ? lds[lane] : 0u;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            const unsigned u = __shfl_up(t, d, 64);
            if (lane >= d) t += u;
        }
        if (lane < 16) lds[lane] = t;
    }
    __syncthreads();
    if (wave > 0) v += lds[wave - 1];
    __syncthreads();   // lds is free for the next call
    return v;
}

// Cell of each point (cells in the brick's memory order, phi fastest; points outside the r/theta range go
// to the extra cell n_cells) and its rank among the points of that cell.
__global__ __launch_bounds__(256) void bucket_count_kernel(const double *__restrict__ rg, int nr,
                                                           const double *__restrict__ tg, int nt,
                                                           const double *__restrict__ pg, int np,
                                                           const double *__restrict__ pts, int64_t n,
                                                           unsigned *__restrict__ key, unsigned *__restrict__ rank,
                                                           unsigned *__restrict__ count) {
    __shared__ double sgrid[kGatherGrid];
    const bool staged = nr + nt + np <= kGatherGrid;
    if (staged) {
        for (int j = threadIdx.x; j < nr + nt + np; j += blockDim.x)
            sgrid[j] = j < nr ? rg[j] : (j < nr + nt ? tg[j - nr] : pg[j - nr - nt]);
        __syncthreads();
    }
    const double *r_ = staged ? sgrid : rg;
    const double *t_ = staged ? sgrid + nr : tg;
    const double *p_ = staged ? sgrid + nr + nt : pg;
    const unsigned n_cells = unsigned(nr - 1) * unsigned(nt - 1) * unsigned(np);
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const double *p = pts + i * 3;
        int r0 = 0, t0 = 0, p0 = 0, p1 = 0;
        double w = 0.0;
        unsigned c = n_cells;
        if (bracket_linear_d(r_, nr, p[0], r0, w) && bracket_linear_d(t_, nt, p[1], t0, w)) {
            bracket_phi_d(p_, np, p[2], p0, p1, w);
            c = (unsigned(r0) * unsigned(nt - 1) + unsigned(t0)) * unsigned(np) + unsigned(p0);
        }
        key[i] = c;
        rank[i] = atomicAdd(&count[c], 1u);
    }
}

// Exclusive scan of a[0, n) in chunks of 1024; the chunk totals go to sums.
__global__ __launch_bounds__(kScanBlock) void scan_chunks_kernel(unsigned *__restrict__ a, unsigned n,
                                                                 unsigned *__restrict__ sums) {
    __shared__ unsigned lds[16];
    const unsigned i = blockIdx.x * kScanBlock + threadIdx.x;
    const unsigned v = i < n ? a[i] : 0u;
    const unsigned inc = block_inclusive_scan(v, lds);
    if (i < n) a[i] = inc - v;
    if (threadIdx.x == kScanBlock - 1) sums[blockIdx.x] = inc;
}

// Exclusive scan of the chunk totals, in place, by one block.
__global__ __launch_bounds__(kScanBlock) void scan_sums_kernel(unsigned *__restrict__ sums, unsigned nb) {
    __shared__ unsigned lds[16];
    __shared__ unsigned carry;
    if (threadIdx.x == 0) carry = 0u;
    __syncthreads();
    for (unsigned base = 0; base < nb; base += kScanBlock) {
        const unsigned i = base + threadIdx.x;
        const unsigned v = i < nb ? sums[i] : 0u;
        const unsigned inc = block_inclusive_scan(v, lds);
        const unsigned c = carry;
        if (i < nb) sums[i] = c + inc - v;
        __syncthreads();
        if (threadIdx.x == kScanBlock - 1) carry = c + inc;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kScanBlock) void scan_add_kernel(unsigned *__restrict__ a, unsigned n,
                                                              const unsigned *__restrict__ sums) {
    const unsigned i = blockIdx.x * kScanBlock + threadIdx.x;
    if (i < n) a[i] += sums[blockIdx.x];
}

struct alignas(16) SortedPoint {
    double r, th, ph;
    unsigned idx;    // position of the point in the caller's arrays
    unsigned cell;   // (r0 (nt - 1) + t0) np + p0, or the outside cell
};

__global__ __launch_bounds__(256) void bucket_scatter_kernel(const double *__restrict__ pts,
                                                             const unsigned *__restrict__ key,
                                                             const unsigned *__restrict__ rank,
                                                             const unsigned *__restrict__ offset, int64_t n,
                                                             SortedPoint *__restrict__ sorted) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const unsigned c = key[i];
        const double *p = pts + i * 3;
        sorted[offset[c] + rank[i]] = SortedPoint{p[0], p[1], p[2], unsigned(i), c};
    }
}

// One lane per point, the points in cell order; the weights repeat bracket_linear_d / bracket_phi_d on the
// cell the sort found, and the corners load straight from the brick.  A wave's 64 x 11 results pass
// through LDS so that the stores go out point by point: consecutive lanes write consecutive channels of one
// point, each point's 88 bytes in one piece, instead of every lane scattering 11 doubles.
__global__ __launch_bounds__(256) void sorted_gather_kernel(const double *__restrict__ coeffs,
                                                            const double *__restrict__ rg, int nr,
                                                            const double *__restrict__ tg, int nt,
                                                            const double *__restrict__ pg, int np,
                                                            const SortedPoint *__restrict__ sorted, int64_t n,
                                                            double *__restrict__ out) {
    __shared__ double sgrid[kGatherGrid];
    __shared__ double res_lds[4][64 * kCoeffs];
    __shared__ unsigned idx_lds[4][64];
    const bool staged = nr + nt + np <= kGatherGrid;
    if (staged) {
        for (int j = threadIdx.x; j < nr + nt + np; j += blockDim.x)
            sgrid[j] = j < nr ? rg[j] : (j < nr + nt ? tg[j - nr] : pg[j - nr - nt]);
        __syncthreads();
    }
    const double *r_ = staged ? sgrid : rg;
    const double *t_ = staged ? sgrid + nr : tg;
    const double *p_ = staged ? sgrid + nr + nt : pg;
    const unsigned n_cells = unsigned(nr - 1) * unsigned(nt - 1) * unsigned(np);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double *res = res_lds[wave];
    unsigned *idx = idx_lds[wave];
    // the wave's 64 points per round: a wave-uniform loop, so every lane takes part in the LDS exchange
    for (int64_t base = (int64_t(blockIdx.x) * 4 + wave) * 64; base < n; base += int64_t(gridDim.x) * 256) {
        const int64_t j = base + lane;
        unsigned my_idx = ~0u;
        if (j < n) {
            const SortedPoint sp = sorted[j];
            my_idx = sp.idx;
            if (sp.cell >= n_cells) {
#pragma unroll
                for (int k = 0; k < kCoeffs; ++k) res[lane * kCoeffs + k] = __builtin_nan("");
            } else {
                const int p0 = int(sp.cell % unsigned(np));
                const unsigned rt = sp.cell / unsigned(np);
                const int t0 = int(rt % unsigned(nt - 1)), r0 = int(rt / unsigned(nt - 1));
                double wr, wt;
                if (sp.r == r_[nr - 1]) {
                    wr = 1.0;
                } else {
                    double den = r_[r0 + 1] - r_[r0];
                    den = (1.0e-300 > den) ? 1.0e-300 : den;
                    wr = (sp.r - r_[r0]) / den;
                }
                if (sp.th == t_[nt - 1]) {
                    wt = 1.0;
                } else {
                    double den = t_[t0 + 1] - t_[t0];
                    den = (1.0e-300 > den) ? 1.0e-300 : den;
                    wt = (sp.th - t_[t0]) / den;
                }
                const double p = wrap_phi_d(sp.ph, p_[0]);
                const int p1 = (p0 + 1) % np;
                const double hi = (p1 > p0) ? p_[p1] : p_[0] + kTwoPi;
                const double pp = (p >= p_[p0]) ? p : p + kTwoPi;
                double den = hi - p_[p0];
                den = (1.0e-300 > den) ? 1.0e-300 : den;
                const double wp = (pp - p_[p0]) / den;
                const size_t row00 = (size_t(r0) * nt + t0) * np, row01 = row00 + np;
                const size_t row10 = row00 + size_t(nt) * np, row11 = row10 + np;
                const double *c000 = coeffs + (row00 + p0) * kCoeffs, *c001 = coeffs + (row00 + p1) * kCoeffs;
                const double *c010 = coeffs + (row01 + p0) * kCoeffs, *c011 = coeffs + (row01 + p1) * kCoeffs;
                const double *c100 = coeffs + (row10 + p0) * kCoeffs, *c101 = coeffs + (row10 + p1) * kCoeffs;
                const double *c110 = coeffs + (row11 + p0) * kCoeffs, *c111 = coeffs + (row11 + p1) * kCoeffs;
#pragma unroll
                for (int k = 0; k < kCoeffs; ++k)
                    res[lane * kCoeffs + k] =
                        trilerp_d(c000[k], c001[k], c010[k], c011[k], c100[k], c101[k], c110[k], c111[k], wr, wt, wp);
            }
        }
        idx[lane] = my_idx;
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int q = 0; q < kCoeffs; ++q) {   // value f of the wave's 64 x 11: point f / 11, channel f % 11
            const int f = q * 64 + lane;
            const int pt = f / kCoeffs;
            const unsigned ii = idx[pt];
            if (ii != ~0u) out[size_t(ii) * kCoeffs + (f - pt * kCoeffs)] = res[f];
        }
        __builtin_amdgcn_wave_barrier();   // the next round rewrites this wave's LDS
    }
}

__global__ void sample_step_kernel(const double *__restrict__ coeffs, const double *__restrict__ rg, int nr,
                                   const double *__restrict__ tg, int nt, const double *__restrict__ pg, int np,
                                   const double *__restrict__ pts, int64_t n, double ds,
                                   const double *__restrict__ init, int broadcast, int policy,
                                   double *__restrict__ out) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *p = pts + i * 3;
    const double *ip = init + (broadcast ? 0 : i * kStokes);
    double s[4] = {ip[0], ip[1], ip[2], ip[3]};
    double v[kCoeffs];
    double *o = out + i * kStokes;
    if (sample_aos_d(coeffs, rg, nr, tg, nt, pg, np, p[0], p[1], p[2], v)) {
        stokes_rk2_d(s, v, ds);
    } else if (policy == BH_POLICY_ZERO) {
        s[0] = s[1] = s[2] = s[3] = 0.0;
    } else if (policy == BH_POLICY_NAN) {
        s[0] = s[1] = s[2] = s[3] = __builtin_nan("");
    }  // BH_POLICY_INITIAL keeps the initial row
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = s[k];
}

__global__ void valid_mask_kernel(double r0, double r1, double t0, double t1, const double *__restrict__ pts,
                                  int64_t n, uint8_t *__restrict__ out) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r = pts[i * 3], th = pts[i * 3 + 1];
    out[i] = uint8_t((r >= r0) & (r <= r1) & (th >= t0) & (th <= t1));
}

// ---- host validation (lib.rs:112-169, 228-256, 318-334) --------------------

static bool all_finite(const double *v, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

static int validate_grid(const char *name, const double *g, size_t n) {
    if (!g || n < 2) return set_error(BH_ERR_VALUE, "%s grid must contain at least 2 values", name);
    if (!all_finite(g, n)) return set_error(BH_ERR_VALUE, "%s grid must contain only finite values", name);
    for (size_t i = 1; i < n; ++i)
        if (!(g[i] > g[i - 1])) return set_error(BH_ERR_VALUE, "%s grid must be strictly increasing", name);
    if (n > size_t(INT32_MAX)) return set_error(BH_ERR_VALUE, "%s grid is too large", name);
    return BH_OK;
}

static int validate_ds(double ds) {
    if (!std::isfinite(ds) || ds < 0.0) return set_error(BH_ERR_VALUE, "ds_cm must be finite and non-negative");
    return BH_OK;
}

static int validate_sampler(const double *coeffs, size_t n_coeffs, const double *rg, size_t nr, const double *tg,
                            size_t nt, const double *pg, size_t np, const double *pts, size_t n_pts_flat,
                            size_t *n_points) {
    int st;
    if ((st = validate_grid("r", rg, nr))) return st;
    if ((st = validate_grid("theta", tg, nt))) return st;
    if ((st = validate_grid("phi", pg, np))) return st;
    if (n_coeffs && (!coeffs || !all_finite(coeffs, n_coeffs)))
        return set_error(BH_ERR_VALUE, "coeffs must contain only finite values");
    size_t expected = 0;
    if (__builtin_mul_overflow(nr, nt, &expected) || __builtin_mul_overflow(expected, np, &expected) ||
        __builtin_mul_overflow(expected, size_t(kCoeffs), &expected))
        return set_error(BH_ERR_VALUE, "coefficient brick shape is too large");
    if (n_coeffs != expected)
        return set_error(BH_ERR_VALUE, "coeffs must contain exactly r*theta*phi*11 values; expected %zu, got %zu",
                         expected, n_coeffs);
    if (n_pts_flat % 3 != 0) return set_error(BH_ERR_VALUE, "points must contain flat r, theta, phi triples");
    if (n_pts_flat && (!pts || !all_finite(pts, n_pts_flat)))
        return set_error(BH_ERR_VALUE, "points must contain only finite values");
    *n_points = n_pts_flat / 3;
    return BH_OK;
}

static inline unsigned blocks_for(int64_t n, int t) { return unsigned((n + t - 1) / t); }

// Grid of sample_gather_kernel: one group of 16 lanes per point, grid-stride beyond 32 blocks per CU.
static inline unsigned gather_blocks(bh_ctx *ctx, int64_t n) {
    const int64_t want = (n + 15) / 16;
    const int64_t cap = int64_t(ctx->cus > 0 ? ctx->cus : 256) * 32;
    return unsigned(want < cap ? (want > 0 ? want : 1) : cap);
}

// Copies host arrays into one staging allocation; returns device pointers in order.
struct Stage {
    char *base = nullptr;
    size_t off = 0;
    template <typename T>
    T *put(bh_ctx *ctx, const T *src, size_t count, int *st) {
        T *d = reinterpret_cast<T *>(base + off);
        size_t bytes = count * sizeof(T);
        off += (bytes + 255) & ~size_t(255);
        if (count && src) {
            hipError_t e = hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, ctx->stream);
            if (e != hipSuccess) *st = set_error(BH_ERR_DEVICE, "hipMemcpyAsync failed: %s", hipGetErrorString(e));
        }
        return d;
    }
};

static size_t rounded(size_t bytes) { return (bytes + 255) & ~size_t(255); }

// The bucketed order pays for its sort when the points outnumber the cells (each record then serves
// several points) and there are enough of them to fill the chip.
constexpr int64_t kBucketMinPoints = int64_t(1) << 16;

// Bucketed large-N sample into device buffers; the workspace (cell counts and offsets, chunk sums, key and
// rank of every point, the points in cell order) is the context's, grown on demand.
static int sample_bucketed(bh_ctx *ctx, const double *d_coeffs, const double *d_r, int nr, const double *d_t, int nt,
                           const double *d_p, int np, const double *d_points, int64_t n, double *d_out) {
    const size_t n_cells = size_t(nr - 1) * size_t(nt - 1) * size_t(np) + 1;   // + the cell of outside points
    const size_t nb = (n_cells + kScanBlock - 1) / kScanBlock;
    const size_t need = rounded(n_cells * 4) + rounded(nb * 4) + 2 * rounded(size_t(n) * 4) +
                        rounded(size_t(n) * sizeof(SortedPoint));
    if (need > ctx->sampler_work_cap) {
        BH_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->d_sampler_work) (void)hipFree(ctx->d_sampler_work);
        ctx->d_sampler_work = nullptr;
        ctx->sampler_work_cap = 0;
        if (hipMalloc(&ctx->d_sampler_work, need) != hipSuccess)
            return set_error(BH_ERR_MEMORY, "sampler workspace allocation of %zu bytes failed", need);
        ctx->sampler_work_cap = need;
    }
    char *w = static_cast<char *>(ctx->d_sampler_work);
    unsigned *count = reinterpret_cast<unsigned *>(w);
    unsigned *sums = reinterpret_cast<unsigned *>(w + rounded(n_cells * 4));
    unsigned *key = reinterpret_cast<unsigned *>(w + rounded(n_cells * 4) + rounded(nb * 4));
    unsigned *rank = key + rounded(size_t(n) * 4) / 4;
    SortedPoint *sorted = reinterpret_cast<SortedPoint *>(rank + rounded(size_t(n) * 4) / 4);
    int st;
    if ((st = fill_u32(ctx, count, 0u, int(n_cells)))) return st;
    const int64_t cap = int64_t(ctx->cus > 0 ? ctx->cus : 256) * 16;
    const unsigned pblocks = unsigned(std::min<int64_t>((n + 255) / 256, cap));
    hipLaunchKernelGGL(bucket_count_kernel, dim3(pblocks), dim3(256), 0, ctx->stream, d_r, nr, d_t, nt, d_p, np, d_points, n,
                       key, rank, count);
    hipLaunchKernelGGL(scan_chunks_kernel, dim3(unsigned(nb)), dim3(kScanBlock), 0, ctx->stream, count, unsigned(n_cells),
                       sums);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kScanBlock), 0, ctx->stream, sums, unsigned(nb));
    hipLaunchKernelGGL(scan_add_kernel, dim3(unsigned(nb)), dim3(kScanBlock), 0, ctx->stream, count, unsigned(n_cells), sums);
    hipLaunchKernelGGL(bucket_scatter_kernel, dim3(pblocks), dim3(256), 0, ctx->stream, d_points, key, rank, count, n,
                       sorted);
    hipLaunchKernelGGL(sorted_gather_kernel, dim3(pblocks), dim3(256), 0, ctx->stream, d_coeffs, d_r, nr, d_t, nt, d_p, np,
                       sorted, n, d_out);
    BH_HIP(hipGetLastError());
    return BH_OK;
}

}  // namespace bh

using namespace bh;

extern "C" {

int bh_stokes_rk2_brick(const double *coeffs, size_t n_coeffs, double ds_cm, const double *initial, size_t n_initial,
                        double *out) {
    clear_error();
    int st;
    if ((st = validate_ds(ds_cm))) return st;
    if (n_coeffs == 0 || n_coeffs % kCoeffs != 0 || !coeffs)
        return set_error(BH_ERR_VALUE, "coeffs must contain one or more flat 11-coefficient cells");
    if (!all_finite(coeffs, n_coeffs)) return set_error(BH_ERR_VALUE, "coeffs must contain only finite values");
    size_t cells = n_coeffs / kCoeffs;
    bool bcast = n_initial == size_t(kStokes);
    if (!bcast && n_initial != cells * kStokes)
        return set_error(BH_ERR_VALUE, "initial must contain either 4 values or 4 values per coefficient cell");
    if (!initial || !all_finite(initial, n_initial))
        return set_error(BH_ERR_VALUE, "initial must contain only finite values");
    if (!out) return set_error(BH_ERR_VALUE, "out is NULL");
    std::lock_guard<std::mutex> lock(default_ctx_mutex());
    bh_ctx *ctx = default_ctx(&st);
    if (!ctx) return st;
    void *buf = nullptr;
    size_t need = rounded(n_coeffs * 8) + rounded(n_initial * 8) + rounded(cells * kStokes * 8);
    if ((st = staging(need, &buf))) return st;
    Stage sg{static_cast<char *>(buf)};
    st = BH_OK;
    double *dc = sg.put(ctx, coeffs, n_coeffs, &st);
    double *di = sg.put(ctx, initial, n_initial, &st);
    double *dout = sg.put<double>(ctx, nullptr, cells * kStokes, &st);
    if (st) return st;
    begin_timing(ctx);
    hipLaunchKernelGGL(stokes_rk2_kernel, dim3(blocks_for(int64_t(cells), 256)), dim3(256), 0, ctx->stream, dc,
                       int64_t(cells), ds_cm, di, int(bcast), dout);
    BH_HIP(hipGetLastError());
    end_timing(ctx, 1);
    BH_HIP(hipMemcpyAsync(out, dout, cells * kStokes * 8, hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    return BH_OK;
}

int bh_sample_brick_trilinear(const double *coeffs, size_t n_coeffs, const double *r_grid, size_t nr,
                              const double *theta_grid, size_t nt, const double *phi_grid, size_t np,
                              const double *points, size_t n_points_flat, double *out) {
    clear_error();
    size_t n = 0;
    int st = validate_sampler(coeffs, n_coeffs, r_grid, nr, theta_grid, nt, phi_grid, np, points, n_points_flat, &n);
    if (st) return st;
    if (n == 0) return BH_OK;
    if (!out) return set_error(BH_ERR_VALUE, "out is NULL");
    std::lock_guard<std::mutex> lock(default_ctx_mutex());
    bh_ctx *ctx = default_ctx(&st);
    if (!ctx) return st;
    void *buf = nullptr;
    size_t need = rounded(n_coeffs * 8) + rounded(nr * 8) + rounded(nt * 8) + rounded(np * 8) +
                  rounded(n_points_flat * 8) + rounded(n * kCoeffs * 8);
    if ((st = staging(need, &buf))) return st;
    Stage sg{static_cast<char *>(buf)};
    double *dc = sg.put(ctx, coeffs, n_coeffs, &st);
    double *dr = sg.put(ctx, r_grid, nr, &st);
    double *dt = sg.put(ctx, theta_grid, nt, &st);
    double *dp = sg.put(ctx, phi_grid, np, &st);
    double *dpts = sg.put(ctx, points, n_points_flat, &st);
    double *dout = sg.put<double>(ctx, nullptr, n * kCoeffs, &st);
    if (st) return st;
    begin_timing(ctx);
    hipLaunchKernelGGL(sample_gather_kernel, dim3(gather_blocks(ctx, int64_t(n))), dim3(256), 0, ctx->stream, dc, dr, int(nr),
                       dt, int(nt), dp, int(np), dpts, int64_t(n), dout);
    BH_HIP(hipGetLastError());
    end_timing(ctx, 1);
    BH_HIP(hipMemcpyAsync(out, dout, n * kCoeffs * 8, hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    return BH_OK;
}

int bh_sample_and_step_stokes(const double *coeffs, size_t n_coeffs, const double *r_grid, size_t nr,
                              const double *theta_grid, size_t nt, const double *phi_grid, size_t np,
                              const double *points, size_t n_points_flat, double ds_cm, const double *initial,
                              size_t n_initial, int invalid_policy, double *out) {
    clear_error();
    int st;
    if ((st = validate_ds(ds_cm))) return st;
    size_t n = 0;
    if ((st = validate_sampler(coeffs, n_coeffs, r_grid, nr, theta_grid, nt, phi_grid, np, points, n_points_flat, &n)))
        return st;
    bool bcast = n_initial == size_t(kStokes);
    if (!bcast && n_initial != n * kStokes)
        return set_error(BH_ERR_VALUE, "initial must contain either 4 values or 4 values per sample point");
    if (n_initial && (!initial || !all_finite(initial, n_initial)))
        return set_error(BH_ERR_VALUE, "initial must contain only finite values");
    if (invalid_policy != BH_POLICY_NAN && invalid_policy != BH_POLICY_INITIAL && invalid_policy != BH_POLICY_ZERO)
        return set_error(BH_ERR_VALUE, "invalid_policy must be 'nan', 'initial', or 'zero'");
    if (n == 0) return BH_OK;
    if (!out) return set_error(BH_ERR_VALUE, "out is NULL");
    std::lock_guard<std::mutex> lock(default_ctx_mutex());
    bh_ctx *ctx = default_ctx(&st);
    if (!ctx) return st;
    void *buf = nullptr;
    size_t need = rounded(n_coeffs * 8) + rounded(nr * 8) + rounded(nt * 8) + rounded(np * 8) +
                  rounded(n_points_flat * 8) + rounded(n_initial * 8) + rounded(n * kStokes * 8);
    if ((st = staging(need, &buf))) return st;
    Stage sg{static_cast<char *>(buf)};
    double *dc = sg.put(ctx, coeffs, n_coeffs, &st);
    double *dr = sg.put(ctx, r_grid, nr, &st);
    double *dt = sg.put(ctx, theta_grid, nt, &st);
    double *dp = sg.put(ctx, phi_grid, np, &st);
    double *dpts = sg.put(ctx, points, n_points_flat, &st);
    double *di = sg.put(ctx, initial, n_initial, &st);
    double *dout = sg.put<double>(ctx, nullptr, n * kStokes, &st);
    if (st) return st;
    begin_timing(ctx);
    hipLaunchKernelGGL(sample_step_kernel, dim3(blocks_for(int64_t(n), 256)), dim3(256), 0, ctx->stream, dc, dr,
                       int(nr), dt, int(nt), dp, int(np), dpts, int64_t(n), ds_cm, di, int(bcast), invalid_policy,
                       dout);
    BH_HIP(hipGetLastError());
    end_timing(ctx, 1);
    BH_HIP(hipMemcpyAsync(out, dout, n * kStokes * 8, hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    return BH_OK;
}

int bh_sample_brick_valid_mask(const double *r_grid, size_t nr, const double *theta_grid, size_t nt,
                               const double *points, size_t n_points_flat, uint8_t *out) {
    clear_error();
    int st;
    if ((st = validate_grid("r", r_grid, nr))) return st;
    if ((st = validate_grid("theta", theta_grid, nt))) return st;
    if (n_points_flat % 3 != 0) return set_error(BH_ERR_VALUE, "points must contain r/theta/phi triples");
    if (n_points_flat && (!points || !all_finite(points, n_points_flat)))
        return set_error(BH_ERR_VALUE, "points must contain only finite values");
    size_t n = n_points_flat / 3;
    if (n == 0) return BH_OK;
    if (!out) return set_error(BH_ERR_VALUE, "out is NULL");
    std::lock_guard<std::mutex> lock(default_ctx_mutex());
    bh_ctx *ctx = default_ctx(&st);
    if (!ctx) return st;
    void *buf = nullptr;
    if ((st = staging(rounded(n_points_flat * 8) + rounded(n), &buf))) return st;
    Stage sg{static_cast<char *>(buf)};
    double *dpts = sg.put(ctx, points, n_points_flat, &st);
    uint8_t *dout = sg.put<uint8_t>(ctx, nullptr, n, &st);
    if (st) return st;
    begin_timing(ctx);
    hipLaunchKernelGGL(valid_mask_kernel, dim3(blocks_for(int64_t(n), 256)), dim3(256), 0, ctx->stream, r_grid[0],
                       r_grid[nr - 1], theta_grid[0], theta_grid[nt - 1], dpts, int64_t(n), dout);
    BH_HIP(hipGetLastError());
    end_timing(ctx, 1);
    BH_HIP(hipMemcpyAsync(out, dout, n, hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    return BH_OK;
}

int bh_sample_brick_trilinear_device(bh_ctx *ctx, const double *d_coeffs, const double *d_r, int32_t nr,
                                     const double *d_theta, int32_t nt, const double *d_phi, int32_t np,
                                     const double *d_points, int64_t n, double *d_out) {
    clear_error();
    int st = ctx_check(ctx);
    if (st) return st;
    if (nr < 2 || nt < 2 || np < 2) return set_error(BH_ERR_VALUE, "grids must contain at least 2 values");
    if (n < 0) return set_error(BH_ERR_VALUE, "n must be non-negative");
    if (n == 0) return BH_OK;
    const int64_t n_cells = int64_t(nr - 1) * (nt - 1) * np + 1;
    const bool bucketed = n >= kBucketMinPoints && n >= n_cells && n < (int64_t(1) << 32) &&
                          n_cells < (int64_t(1) << 31);
    begin_timing(ctx);
    if (bucketed) {
        if ((st = sample_bucketed(ctx, d_coeffs, d_r, nr, d_theta, nt, d_phi, np, d_points, n, d_out))) return st;
    } else {
        hipLaunchKernelGGL(sample_gather_kernel, dim3(gather_blocks(ctx, n)), dim3(256), 0, ctx->stream, d_coeffs, d_r,
                           int(nr), d_theta, int(nt), d_phi, int(np), d_points, n, d_out);
        BH_HIP(hipGetLastError());
    }
    end_timing(ctx, 1);   // one timed call: the sort and the gather together
    return BH_OK;
}

}  // extern "C"
