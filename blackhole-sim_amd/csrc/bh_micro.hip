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
#include <cstring>

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
    const double d = phi - base;
    // fmod returns d itself inside [0, 2 pi); the library call (a loop) only runs outside it
    double m = (d >= 0.0 && d < kTwoPi) ? d : fmod(d, kTwoPi);
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

// ---- points partitioned by block of brick cells ------------------------------------------------
// Random points gather eight random 88-byte corner records each, from a brick (92 MB at 128x64x128)
// that does not fit an XCD's 4 MB L2.  The large-N form partitions the points by block of brick cells
// (br x bt x bp cells), and one workgroup per block copies the block's corner nodes into LDS once and
// samples all of the block's points from there: the brick is read about once per call, and the 88
// corner values of a point come from LDS.  The partition is a counting sort over blocks in two passes
// over the points, each workgroup counting a fixed chunk of points into an LDS histogram (no
// device-scope atomics): pass one writes the chunk's count per block, an exclusive scan over
// (block, chunk) gives each chunk its offset inside every block's segment, and pass two writes each
// point's 32-byte record there.  Cells are those of bracket_linear_d / bracket_phi_d; the gather repeats
// the weight expressions of sample_aos_d on the cell and the same lerps on the same values, so its
// results equal the unsorted gather's bit for bit.

constexpr int kScanBlock = 1024;      // threads of the scan kernels: 16 waves of 64
constexpr int kPartThreads = 1024;    // threads of the two partition passes
constexpr int kMaxBlocks = 8192;      // blocks of one partition; the LDS histogram holds kMaxBlocks + 1
constexpr int kBlockThreads = 256;    // threads of the block gather
constexpr int kOutsideGroups = 64;    // workgroups that write the NaN rows of points outside the r/theta range
constexpr int kMaxRegionBytes = 64 * 1024;

struct BlockGeom {
    int br, bt, bp;      // cells per block in r, theta, phi (powers of two)
    int lbr, lbt, lbp;   // their base-2 logarithms
    int nbr, nbt, nbp;   // blocks per dimension
    int nblocks;         // nbr nbt nbp; block nblocks holds the points outside the r/theta range
    int split;           // workgroups per block in the gather
};

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

// Exclusive scan of sums[0, nb), in place, by one block.
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

// Per-block offsets of the partition: h[c][b] (chunk-major, nb1 blocks per chunk row) holds chunk c's count
// of block b and becomes the exclusive prefix over chunks, sum over c' < c of h[c'][b]; tot[b] receives the
// block's total.  A workgroup takes 64 blocks; its 16 waves each sum a sixteenth of the chunks, an LDS pass
// turns those sums into the waves' starting offsets, and each wave rewrites its rows.
__global__ __launch_bounds__(1024) void column_scan_kernel(unsigned *__restrict__ h, int chunks, int nb1,
                                                           unsigned *__restrict__ tot) {
    __shared__ unsigned part[16][64];
    const int lane = threadIdx.x & 63, seg = threadIdx.x >> 6;
    const int b = int(blockIdx.x) * 64 + lane;
    const int per = (chunks + 15) / 16;
    const int c0 = seg * per < chunks ? seg * per : chunks;
    const int c1 = c0 + per < chunks ? c0 + per : chunks;
    unsigned s = 0u;
    if (b < nb1) {
#pragma unroll 8
        for (int c = c0; c < c1; ++c) s += h[size_t(c) * nb1 + b];
    }
    part[seg][lane] = s;
    __syncthreads();
    unsigned pre = 0u;
    for (int k = 0; k < seg; ++k) pre += part[k][lane];
    if (b < nb1) {
        for (int c = c0; c < c1; ++c) {
            const unsigned v = h[size_t(c) * nb1 + b];
            h[size_t(c) * nb1 + b] = pre;
            pre += v;
        }
        if (seg == 15) tot[b] = pre;
    }
}

struct alignas(16) SortedPoint {
    double r, th, ph;
    unsigned idx;    // position of the point in the caller's arrays
    unsigned cell;   // the cell inside its block: r0 % br | (t0 % bt) << 8 | (p0 % bp) << 16; ~0u outside
};

// Cell lookup of the partition passes: searchsorted-right - 1 from a guess.  Each grid gets a table of
// kGuessBins + 1 entries over its range ([g0, g_last], and [g0, g0 + 2 pi] for the wrapped phi), entry k the
// exact index of the bin's left edge; a point starts from its bin's entry and walks to the index where
// g[i] <= x < g[i + 1], which is searchsorted-right - 1 whatever the start, so the rounding of the bin only
// changes the length of the walk (a step or two for grids whose cells do not vary a lot in width).
constexpr int kGuessBins = 512;

struct Guess {
    const double *g;
    int n;
    double lo, scale;   // bin of x: (x - lo) * scale
    const int *t;       // kGuessBins + 1 start indices
};

__device__ __forceinline__ int guess_search(const Guess &q, double x) {
    const double u = (x - q.lo) * q.scale;
    int i = q.t[u > 0.0 ? (u < double(kGuessBins) ? int(u) : kGuessBins) : 0];
    while (i + 1 < q.n && q.g[i + 1] <= x) ++i;
    while (i > 0 && q.g[i] > x) --i;
    return i;
}

// Grids of up to kGatherGrid values in all are staged in LDS, and the three guess tables built in LDS
// (tab holds 3 (kGuessBins + 1) ints).
__device__ __forceinline__ void stage_guesses(const double *rg, int nr, const double *tg, int nt, const double *pg,
                                              int np, double *sgrid, int *tab, Guess q[3]) {
    const bool staged = nr + nt + np <= kGatherGrid;
    if (staged)
        for (int j = threadIdx.x; j < nr + nt + np; j += blockDim.x)
            sgrid[j] = j < nr ? rg[j] : (j < nr + nt ? tg[j - nr] : pg[j - nr - nt]);
    q[0] = Guess{staged ? sgrid : rg, nr, rg[0], double(kGuessBins) / (rg[nr - 1] - rg[0]), tab};
    q[1] = Guess{staged ? sgrid + nr : tg, nt, tg[0], double(kGuessBins) / (tg[nt - 1] - tg[0]), tab + kGuessBins + 1};
    q[2] = Guess{staged ? sgrid + nr + nt : pg, np, pg[0], double(kGuessBins) / kTwoPi, tab + 2 * (kGuessBins + 1)};
    for (int e = threadIdx.x; e < 3 * (kGuessBins + 1); e += blockDim.x) {
        const int w = e / (kGuessBins + 1), k = e - w * (kGuessBins + 1);
        const double *g = w == 0 ? rg : (w == 1 ? tg : pg);
        const int n = w == 0 ? nr : (w == 1 ? nt : np);
        const double edge = q[w].lo + double(k) / q[w].scale;
        const int i = search_right_d(g, n, edge) - 1;
        tab[e] = i < 0 ? 0 : i;
    }
    __syncthreads();
}

// Block of a point and its cell inside the block, the cell being the one bracket_linear_d / bracket_phi_d
// find (the r and theta index clamped to [0, n - 2], so x == g[n - 1] lands in the last cell; the phi index of
// the wrapped angle, n - 1 when out of [0, n)); false (block nblocks) outside the r/theta range or for a NaN
// coordinate.
__device__ __forceinline__ bool locate_point(const Guess q[3], const BlockGeom &g, double r, double th, double ph,
                                             unsigned &cell, unsigned &blk) {
    if (!(r >= q[0].g[0] && r <= q[0].g[q[0].n - 1] && th >= q[1].g[0] && th <= q[1].g[q[1].n - 1])) {
        cell = ~0u;
        blk = unsigned(g.nblocks);
        return false;
    }
    int r0 = guess_search(q[0], r), t0 = guess_search(q[1], th);
    r0 = r0 > q[0].n - 2 ? q[0].n - 2 : r0;
    t0 = t0 > q[1].n - 2 ? q[1].n - 2 : t0;
    int p0 = guess_search(q[2], wrap_phi_d(ph, q[2].g[0]));
    p0 = (p0 < 0 || p0 >= q[2].n) ? q[2].n - 1 : p0;
    cell = unsigned(r0 & (g.br - 1)) | unsigned(t0 & (g.bt - 1)) << 8 | unsigned(p0 & (g.bp - 1)) << 16;
    blk = (unsigned(r0 >> g.lbr) * unsigned(g.nbt) + unsigned(t0 >> g.lbt)) * unsigned(g.nbp) + unsigned(p0 >> g.lbp);
    return true;
}

// Pass one: workgroup c counts the points [c chunk, (c + 1) chunk) per block into hist[c][b].
__global__ __launch_bounds__(kPartThreads) void part_count_kernel(const double *__restrict__ rg, int nr,
                                                                  const double *__restrict__ tg, int nt,
                                                                  const double *__restrict__ pg, int np,
                                                                  const double *__restrict__ pts, int64_t n,
                                                                  int64_t chunk, BlockGeom g,
                                                                  unsigned *__restrict__ hist) {
    __shared__ double sgrid[kGatherGrid];
    __shared__ int tab[3 * (kGuessBins + 1)];
    __shared__ unsigned lh[kMaxBlocks + 1];
    for (int j = threadIdx.x; j <= g.nblocks; j += blockDim.x) lh[j] = 0u;
    Guess q[3];
    stage_guesses(rg, nr, tg, nt, pg, np, sgrid, tab, q);
    const int64_t lo = int64_t(blockIdx.x) * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    constexpr int kAhead = 4;   // points per thread whose coordinates load together
    for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += kAhead * int64_t(blockDim.x)) {
        double c[kAhead][3];
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
            const int64_t i = i0 + u * int64_t(blockDim.x);
            const double *p = pts + (i < hi ? i : i0) * 3;
            c[u][0] = p[0];
            c[u][1] = p[1];
            c[u][2] = p[2];
        }
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
            unsigned cell, blk;
            locate_point(q, g, c[u][0], c[u][1], c[u][2], cell, blk);
            if (i0 + u * int64_t(blockDim.x) < hi) atomicAdd(&lh[blk], 1u);
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j <= g.nblocks; j += blockDim.x) hist[size_t(blockIdx.x) * (g.nblocks + 1) + j] = lh[j];
}

// Pass two: the same chunks; each point's record goes to its block's segment (start[b]) at the chunk's offset
// inside it (offs[c][b], the column-scanned histogram) plus its rank among the chunk's points of that block,
// and inv[i] records where point i went.
__global__ __launch_bounds__(kPartThreads) void part_scatter_kernel(const double *__restrict__ rg, int nr,
                                                                    const double *__restrict__ tg, int nt,
                                                                    const double *__restrict__ pg, int np,
                                                                    const double *__restrict__ pts, int64_t n,
                                                                    int64_t chunk, BlockGeom g,
                                                                    const unsigned *__restrict__ offs,
                                                                    const unsigned *__restrict__ start,
                                                                    SortedPoint *__restrict__ sorted,
                                                                    unsigned *__restrict__ inv) {
    __shared__ double sgrid[kGatherGrid];
    __shared__ int tab[3 * (kGuessBins + 1)];
    __shared__ unsigned pos[kMaxBlocks + 1];
    for (int j = threadIdx.x; j <= g.nblocks; j += blockDim.x)
        pos[j] = start[j] + offs[size_t(blockIdx.x) * (g.nblocks + 1) + j];
    Guess q[3];
    stage_guesses(rg, nr, tg, nt, pg, np, sgrid, tab, q);
    const int64_t lo = int64_t(blockIdx.x) * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    constexpr int kAhead = 4;   // points per thread whose coordinates load together
    for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += kAhead * int64_t(blockDim.x)) {
        double c[kAhead][3];
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
            const int64_t i = i0 + u * int64_t(blockDim.x);
            const double *p = pts + (i < hi ? i : i0) * 3;
            c[u][0] = p[0];
            c[u][1] = p[1];
            c[u][2] = p[2];
        }
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
            const int64_t i = i0 + u * int64_t(blockDim.x);
            unsigned cell, blk;
            locate_point(q, g, c[u][0], c[u][1], c[u][2], cell, blk);
            if (i < hi) {
                const unsigned slot = atomicAdd(&pos[blk], 1u);
                sorted[slot] = SortedPoint{c[u][0], c[u][1], c[u][2], unsigned(i), cell};
                inv[i] = slot;
            }
        }
    }
}

// Last pass: out[i] = res[inv[i]], the 11 results of each point read from where the gather wrote them in
// block order; lanes on consecutive doubles of the output, so the stores are whole lines and the random
// accesses are reads (random 88-byte reads cost a little less than random 88-byte writes here, 0.78 vs
// ~0.85 ms per 2^24 points, and leave the gather's stores whole lines).
__global__ __launch_bounds__(256) void unpermute_kernel(const double *__restrict__ res, const unsigned *__restrict__ inv,
                                                        int64_t n, double *__restrict__ out) {
    const int64_t total = n * kCoeffs;
    for (int64_t f = int64_t(blockIdx.x) * 256 + threadIdx.x; f < total; f += int64_t(gridDim.x) * 256) {
        const int64_t i = f / kCoeffs;
        out[f] = res[size_t(inv[i]) * kCoeffs + size_t(f - i * kCoeffs)];
    }
}

// The gather.  Workgroup w < nblocks * split serves the (w % split)-th part of block w / split: it copies
// the block's corner nodes ([r][theta][phi][11] doubles, phi wrapping at the seam) and grid values into LDS,
// then each lane takes one point of the segment at a time, forms the weights from the point's cell as
// sample_aos_d does and interpolates the 11 channels from LDS.  The last kOutsideGroups workgroups write the
// NaN rows of the points outside the r/theta range.  Results are written in block order (res[j] for the
// point in sorted[j]); a wave's 64 x 11 values pass through LDS so that its stores cover whole lines.
__global__ __launch_bounds__(kBlockThreads) void block_gather_kernel(const double *__restrict__ coeffs,
                                                                     const double *__restrict__ rg, int nr,
                                                                     const double *__restrict__ tg, int nt,
                                                                     const double *__restrict__ pg, int np,
                                                                     const SortedPoint *__restrict__ sorted,
                                                                     const unsigned *__restrict__ start,
                                                                     int64_t n, BlockGeom g,
                                                                     double *__restrict__ out) {
    extern __shared__ double lds[];   // 4 waves x 64 x 11 results, then the block's corner nodes
    __shared__ double gloc[64];       // r nodes | theta nodes | phi nodes of the block
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double *res = lds + wave * 64 * kCoeffs;
    double *node = lds + (kBlockThreads / 64) * 64 * kCoeffs;
    const int groups = g.nblocks * g.split;
    const bool inside = int(blockIdx.x) < groups;
    const int b = inside ? int(blockIdx.x) / g.split : g.nblocks;
    const int part = inside ? int(blockIdx.x) % g.split : int(blockIdx.x) - groups;
    const int parts = inside ? g.split : kOutsideGroups;
    const int64_t seg_lo = start[b];
    const int64_t seg_hi = b < g.nblocks ? int64_t(start[b + 1]) : n;
    const int64_t len = seg_hi - seg_lo;
    const int64_t lo = seg_lo + len * part / parts, hi = seg_lo + len * (part + 1) / parts;
    if (lo >= hi) return;   // workgroup-uniform
    int r_lo = 0, t_lo = 0, p_lo = 0, nrr = 0, ntt = 0, npp = 0;
    double r_last = 0.0, t_last = 0.0, p_first = 0.0;
    if (inside) {
        const int pb = b % g.nbp, tb = (b / g.nbp) % g.nbt, rb = b / (g.nbp * g.nbt);
        r_lo = rb * g.br;
        t_lo = tb * g.bt;
        p_lo = pb * g.bp;
        nrr = (g.br < nr - 1 - r_lo ? g.br : nr - 1 - r_lo) + 1;
        ntt = (g.bt < nt - 1 - t_lo ? g.bt : nt - 1 - t_lo) + 1;
        npp = (g.bp < np - p_lo ? g.bp : np - p_lo) + 1;
        const int vals = nrr * ntt * npp * kCoeffs;
        for (int e = threadIdx.x; e < vals; e += blockDim.x) {
            const int ch = e % kCoeffs, q = e / kCoeffs;
            const int k = q % npp, j = (q / npp) % ntt, i = q / (npp * ntt);
            int pk = p_lo + k;
            pk = pk >= np ? pk - np : pk;
            node[e] = coeffs[((size_t(r_lo + i) * nt + (t_lo + j)) * np + pk) * kCoeffs + ch];
        }
        for (int e = threadIdx.x; e < nrr + ntt + npp; e += blockDim.x) {
            double v;
            if (e < nrr) v = rg[r_lo + e];
            else if (e < nrr + ntt) v = tg[t_lo + e - nrr];
            else {
                int pk = p_lo + e - nrr - ntt;
                v = pk >= np ? pg[0] + kTwoPi : pg[pk];   // the seam cell's upper node, as bracket_phi_d's hi
            }
            gloc[e] = v;
        }
        r_last = rg[nr - 1];
        t_last = tg[nt - 1];
        p_first = pg[0];
        __syncthreads();
    }
    const double *rl = gloc, *tl = gloc + nrr, *pl = gloc + nrr + ntt;
    const int sr = ntt * npp * kCoeffs, st = npp * kCoeffs;
    // wave-uniform rounds of 64 points, so that every lane takes part in the LDS exchange
    for (int64_t base = lo + wave * 64; base < hi; base += kBlockThreads) {
        const int64_t j = base + lane;
        if (j < hi) {
            const SortedPoint sp = sorted[j];
            if (!inside) {
#pragma unroll
                for (int k = 0; k < kCoeffs; ++k) res[lane * kCoeffs + k] = __builtin_nan("");
            } else {
                const int ri = int(sp.cell & 255u), ti = int((sp.cell >> 8) & 255u), pi = int(sp.cell >> 16);
                double wr, wt;
                if (sp.r == r_last) {
                    wr = 1.0;
                } else {
                    double den = rl[ri + 1] - rl[ri];
                    den = (1.0e-300 > den) ? 1.0e-300 : den;
                    wr = (sp.r - rl[ri]) / den;
                }
                if (sp.th == t_last) {
                    wt = 1.0;
                } else {
                    double den = tl[ti + 1] - tl[ti];
                    den = (1.0e-300 > den) ? 1.0e-300 : den;
                    wt = (sp.th - tl[ti]) / den;
                }
                const double p = wrap_phi_d(sp.ph, p_first);
                const double hi_p = pl[pi + 1];   // p_[p1], or p_[0] + 2 pi across the seam
                const double pp = (p >= pl[pi]) ? p : p + kTwoPi;
                double den = hi_p - pl[pi];
                den = (1.0e-300 > den) ? 1.0e-300 : den;
                const double wp = (pp - pl[pi]) / den;
                const double *c000 = node + ri * sr + ti * st + pi * kCoeffs;
                const double *c001 = c000 + kCoeffs, *c010 = c000 + st, *c011 = c010 + kCoeffs;
                const double *c100 = c000 + sr, *c101 = c100 + kCoeffs, *c110 = c100 + st, *c111 = c110 + kCoeffs;
#pragma unroll
                for (int k = 0; k < kCoeffs; ++k)
                    res[lane * kCoeffs + k] =
                        trilerp_d(c000[k], c001[k], c010[k], c011[k], c100[k], c101[k], c110[k], c111[k], wr, wt, wp);
            }
        }
        __builtin_amdgcn_wave_barrier();
        const int64_t lim = (hi - base) * kCoeffs;   // values of the round's points that exist
#pragma unroll
        for (int q = 0; q < kCoeffs; ++q) {   // value f of the wave's 64 x 11: point f / 11, channel f % 11
            const int f = q * 64 + lane;
            if (f < lim) out[size_t(base) * kCoeffs + f] = res[f];
        }
        __builtin_amdgcn_wave_barrier();   // the next round rewrites this wave's LDS
    }
}

// One lane per point; grids of up to kGatherGrid values in all are searched in LDS (the searches' dependent
// reads then stay on the CU, also when the arrays are in pinned host memory).
__global__ __launch_bounds__(256) void sample_step_kernel(const double *__restrict__ coeffs,
                                                          const double *__restrict__ rg_in, int nr,
                                                          const double *__restrict__ tg_in, int nt,
                                                          const double *__restrict__ pg_in, int np,
                                                          const double *__restrict__ pts, int64_t n, double ds,
                                                          const double *__restrict__ init, int broadcast, int policy,
                                                          double *__restrict__ out) {
    __shared__ double sgrid[kGatherGrid];
    const bool staged = nr + nt + np <= kGatherGrid;
    if (staged) {
        for (int j = threadIdx.x; j < nr + nt + np; j += blockDim.x)
            sgrid[j] = j < nr ? rg_in[j] : (j < nr + nt ? tg_in[j - nr] : pg_in[j - nr - nt]);
        __syncthreads();
    }
    const double *rg = staged ? sgrid : rg_in, *tg = staged ? sgrid + nr : tg_in;
    const double *pg = staged ? sgrid + nr + nt : pg_in;
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

static size_t rounded(size_t bytes) { return (bytes + 255) & ~size_t(255); }

// Host-buffer calls whose arrays take up to this many bytes in all (the reference's micro-benchmark sizes:
// an 8x6x8 brick and 64 points are 37 KB) run zero-copy: their kernels read the inputs from, and write
// the outputs to, pinned host memory, so the call is one launch and one wait.  Larger calls copy through
// device memory.
constexpr size_t kZeroCopyBytes = size_t(256) << 10;

// The host arrays of one call, laid out back to back (256-byte aligned) in a staging area: the pinned host
// buffer (zero-copy, the arrays copied in by the CPU) or the device staging buffer (copied by the stream).
// put() returns the device address of an array; fetch() brings an output back after the kernel and waits.
struct Stage {
    bh_ctx *ctx = nullptr;
    char *dev = nullptr;    // device address of the area
    char *host = nullptr;   // host address of the area when zero-copy, else null
    size_t off = 0;
    template <typename T>
    T *put(const T *src, size_t count, int *st) {
        T *d = reinterpret_cast<T *>(dev + off);
        const size_t bytes = count * sizeof(T);
        if (count && src) {
            if (host) {
                memcpy(host + off, src, bytes);
            } else {
                hipError_t e = hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, ctx->stream);
                if (e != hipSuccess) *st = set_error(BH_ERR_DEVICE, "hipMemcpyAsync failed: %s", hipGetErrorString(e));
            }
        }
        off += rounded(bytes);
        return d;
    }
    int fetch(void *dst, const void *d, size_t bytes) {
        if (host) {
            BH_HIP(hipStreamSynchronize(ctx->stream));
            memcpy(dst, host + (static_cast<const char *>(d) - dev), bytes);
        } else {
            BH_HIP(hipMemcpyAsync(dst, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
            BH_HIP(hipStreamSynchronize(ctx->stream));
        }
        return BH_OK;
    }
};

// Opens a staging area of `need` bytes for a call on the default context.
static int stage_open(bh_ctx *ctx, size_t need, Stage &sg) {
    sg.ctx = ctx;
    sg.off = 0;
    int st;
    if (need <= kZeroCopyBytes) {
        void *h = nullptr, *d = nullptr;
        if ((st = host_staging(need, &h, &d))) return st;
        sg.host = static_cast<char *>(h);
        sg.dev = static_cast<char *>(d);
        return BH_OK;
    }
    void *buf = nullptr;
    if ((st = staging(need, &buf))) return st;
    sg.host = nullptr;
    sg.dev = static_cast<char *>(buf);
    return BH_OK;
}

// Blocks of the partition: the first of these shapes with at most kMaxBlocks blocks whose corner nodes fit
// kMaxRegionBytes; false when none does (the call then gathers unsorted).
static bool choose_blocks(int nr, int nt, int np, int64_t n, BlockGeom &g) {
    static const int shapes[][3] = {{4, 4, 8}, {4, 8, 8}, {8, 8, 8}};
    for (const auto &s : shapes) {
        g.br = s[0];
        g.bt = s[1];
        g.bp = s[2];
        g.lbr = __builtin_ctz(unsigned(g.br));
        g.lbt = __builtin_ctz(unsigned(g.bt));
        g.lbp = __builtin_ctz(unsigned(g.bp));
        g.nbr = (nr - 1 + g.br - 1) / g.br;
        g.nbt = (nt - 1 + g.bt - 1) / g.bt;
        g.nbp = (np + g.bp - 1) / g.bp;
        const int64_t nb = int64_t(g.nbr) * g.nbt * g.nbp;
        const int64_t region = int64_t(g.br + 1) * (g.bt + 1) * (g.bp + 1) * kCoeffs * 8;
        if (nb <= kMaxBlocks && region <= kMaxRegionBytes) {
            g.nblocks = int(nb);
            // a workgroup per ~2048 points of a block on average
            g.split = int(std::min<int64_t>(64, std::max<int64_t>(1, n / nb / 2048)));
            return true;
        }
    }
    return false;
}

// The partition pays when the points outnumber the blocks well (a block's corner nodes, ~20 KB, are read
// once for all its points) and there are enough of them to fill the chip.
constexpr int64_t kBucketMinPoints = int64_t(1) << 16;
constexpr int64_t kMinPointsPerBlock = 32;

// Partitioned large-N sample into device buffers; the workspace (histogram and offsets per block and
// chunk, segment starts, the points and their results in block order, each point's place in that order) is
// the context's, grown on demand.
static int sample_partitioned(bh_ctx *ctx, const BlockGeom &g, const double *d_coeffs, const double *d_r, int nr,
                              const double *d_t, int nt, const double *d_p, int np, const double *d_points,
                              int64_t n, double *d_out) {
    const int cus = ctx->cus > 0 ? ctx->cus : 256;
    // two workgroups of the partition passes per CU (their LDS, 55 KB each, allows it)
    const int chunks = int(std::max<int64_t>(1, std::min<int64_t>(2 * cus, (n + 16383) / 16384)));
    const int64_t chunk = (n + chunks - 1) / chunks;
    const int nb1 = g.nblocks + 1;
    const size_t total = size_t(nb1) * size_t(chunks);
    const size_t need = rounded(total * 4) + rounded(size_t(nb1) * 4) + rounded(size_t(n) * sizeof(SortedPoint)) +
                        rounded(size_t(n) * 4) + rounded(size_t(n) * kCoeffs * 8);
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
    unsigned *hist = reinterpret_cast<unsigned *>(w);
    unsigned *start = reinterpret_cast<unsigned *>(w + rounded(total * 4));
    SortedPoint *sorted = reinterpret_cast<SortedPoint *>(w + rounded(total * 4) + rounded(size_t(nb1) * 4));
    unsigned *inv = reinterpret_cast<unsigned *>(reinterpret_cast<char *>(sorted) + rounded(size_t(n) * sizeof(SortedPoint)));
    double *res = reinterpret_cast<double *>(reinterpret_cast<char *>(inv) + rounded(size_t(n) * 4));
    hipLaunchKernelGGL(part_count_kernel, dim3(unsigned(chunks)), dim3(kPartThreads), 0, ctx->stream, d_r, nr, d_t, nt,
                       d_p, np, d_points, n, chunk, g, hist);
    hipLaunchKernelGGL(column_scan_kernel, dim3(unsigned((nb1 + 63) / 64)), dim3(1024), 0, ctx->stream, hist, chunks, nb1,
                       start);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kScanBlock), 0, ctx->stream, start, unsigned(nb1));
    hipLaunchKernelGGL(part_scatter_kernel, dim3(unsigned(chunks)), dim3(kPartThreads), 0, ctx->stream, d_r, nr, d_t, nt,
                       d_p, np, d_points, n, chunk, g, hist, start, sorted, inv);
    const size_t region = size_t(g.br + 1) * (g.bt + 1) * (g.bp + 1) * kCoeffs * 8;
    const size_t shmem = size_t(kBlockThreads / 64) * 64 * kCoeffs * 8 + region;
    const unsigned groups = unsigned(g.nblocks * g.split + kOutsideGroups);
    hipLaunchKernelGGL(block_gather_kernel, dim3(groups), dim3(kBlockThreads), shmem, ctx->stream, d_coeffs, d_r, nr, d_t,
                       nt, d_p, np, sorted, start, n, g, res);
    hipLaunchKernelGGL(unpermute_kernel, dim3(unsigned(cus * 16)), dim3(256), 0, ctx->stream, res, inv, n, d_out);
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
    const size_t need = rounded(n_coeffs * 8) + rounded(n_initial * 8) + rounded(cells * kStokes * 8);
    Stage sg;
    if ((st = stage_open(ctx, need, sg))) return st;
    double *dc = sg.put(coeffs, n_coeffs, &st);
    double *di = sg.put(initial, n_initial, &st);
    double *dout = sg.put<double>(nullptr, cells * kStokes, &st);
    if (st) return st;
    hipLaunchKernelGGL(stokes_rk2_kernel, dim3(blocks_for(int64_t(cells), 256)), dim3(256), 0, ctx->stream, dc,
                       int64_t(cells), ds_cm, di, int(bcast), dout);
    BH_HIP(hipGetLastError());
    return sg.fetch(out, dout, cells * kStokes * 8);
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
    const size_t need = rounded(n_coeffs * 8) + rounded(nr * 8) + rounded(nt * 8) + rounded(np * 8) +
                        rounded(n_points_flat * 8) + rounded(n * kCoeffs * 8);
    Stage sg;
    if ((st = stage_open(ctx, need, sg))) return st;
    double *dc = sg.put(coeffs, n_coeffs, &st);
    double *dr = sg.put(r_grid, nr, &st);
    double *dt = sg.put(theta_grid, nt, &st);
    double *dp = sg.put(phi_grid, np, &st);
    double *dpts = sg.put(points, n_points_flat, &st);
    double *dout = sg.put<double>(nullptr, n * kCoeffs, &st);
    if (st) return st;
    hipLaunchKernelGGL(sample_gather_kernel, dim3(gather_blocks(ctx, int64_t(n))), dim3(256), 0, ctx->stream, dc, dr, int(nr),
                       dt, int(nt), dp, int(np), dpts, int64_t(n), dout);
    BH_HIP(hipGetLastError());
    return sg.fetch(out, dout, n * kCoeffs * 8);
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
    const size_t need = rounded(n_coeffs * 8) + rounded(nr * 8) + rounded(nt * 8) + rounded(np * 8) +
                        rounded(n_points_flat * 8) + rounded(n_initial * 8) + rounded(n * kStokes * 8);
    Stage sg;
    if ((st = stage_open(ctx, need, sg))) return st;
    double *dc = sg.put(coeffs, n_coeffs, &st);
    double *dr = sg.put(r_grid, nr, &st);
    double *dt = sg.put(theta_grid, nt, &st);
    double *dp = sg.put(phi_grid, np, &st);
    double *dpts = sg.put(points, n_points_flat, &st);
    double *di = sg.put(initial, n_initial, &st);
    double *dout = sg.put<double>(nullptr, n * kStokes, &st);
    if (st) return st;
    hipLaunchKernelGGL(sample_step_kernel, dim3(blocks_for(int64_t(n), 256)), dim3(256), 0, ctx->stream, dc, dr,
                       int(nr), dt, int(nt), dp, int(np), dpts, int64_t(n), ds_cm, di, int(bcast), invalid_policy,
                       dout);
    BH_HIP(hipGetLastError());
    return sg.fetch(out, dout, n * kStokes * 8);
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
    Stage sg;
    if ((st = stage_open(ctx, rounded(n_points_flat * 8) + rounded(n), sg))) return st;
    double *dpts = sg.put(points, n_points_flat, &st);
    uint8_t *dout = sg.put<uint8_t>(nullptr, n, &st);
    if (st) return st;
    hipLaunchKernelGGL(valid_mask_kernel, dim3(blocks_for(int64_t(n), 256)), dim3(256), 0, ctx->stream, r_grid[0],
                       r_grid[nr - 1], theta_grid[0], theta_grid[nt - 1], dpts, int64_t(n), dout);
    BH_HIP(hipGetLastError());
    return sg.fetch(out, dout, n);
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
    const int64_t n_cells = int64_t(nr - 1) * (nt - 1) * np;
    BlockGeom g{};
    const bool partitioned = n >= kBucketMinPoints && n < (int64_t(1) << 32) && n_cells < (int64_t(1) << 31) &&
                             choose_blocks(nr, nt, np, n, g) && n >= kMinPointsPerBlock * g.nblocks;
    begin_timing(ctx);
    if (partitioned) {
        if ((st = sample_partitioned(ctx, g, d_coeffs, d_r, nr, d_theta, nt, d_phi, np, d_points, n, d_out)))
            return st;
    } else {
        hipLaunchKernelGGL(sample_gather_kernel, dim3(gather_blocks(ctx, n)), dim3(256), 0, ctx->stream, d_coeffs, d_r,
                           int(nr), d_theta, int(nt), d_phi, int(np), d_points, n, d_out);
        BH_HIP(hipGetLastError());
    }
    end_timing(ctx, 1);   // one timed call: the sort and the gather together
    return BH_OK;
}

}  // extern "C"
