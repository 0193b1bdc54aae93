// MLP surrogate: BF16 MFMA GEMMs with FP32 accumulation, forward / backward / Adam on gfx950.
//
// Network: x (n_in) -> [Linear width, SiLU] x depth -> Linear n_out.  Parameters are FP32
// masters updated by Adam, with two BF16 copies feeding v_mfma_f32_16x16x32_bf16: W as
// stored ([fan_in][fan_out], used by the input-gradient GEMM) and W^T ([fan_out][fan_in],
// used by the forward GEMM), so both GEMMs read operands with k contiguous.  Shapes are
// padded to multiples of 8 (n_in, n_out), so every global operand load is one 16-byte
// vector; padded rows and columns hold zeros and receive zero gradients, so they stay zero.
//
// Parameter vector layout (FP32, padded): for each layer l, W_l[fi_pad][fo_pad] row-major,
// then b_l[fo_pad].

#include "bh_internal.h"

#include <cmath>
#include <cstring>

namespace bh {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;     // output tile edge per workgroup
constexpr int kMaxLayersAdam = 20;   // at least depth + 1 for the largest depth bh_mlp_create accepts
// batch rows per split-K slice of a weight-gradient GEMM (fewer slices leave CUs idle), and at most this many
// slices (their partial results are summed by weight_reduce_kernel)
constexpr int kDwRows = 256;
constexpr int kDwMaxSplits = 32;
constexpr int kKStep = 32;    // K per MFMA pass

enum Epi { EPI_STORE = 0, EPI_BIAS_SILU = 1, EPI_BIAS = 2, EPI_DSILU = 3, EPI_MSE = 5 };

struct GemmArgs {
    const __bf16 *A;   // op(A): M x K, stored [M][K] (k contiguous)
    const __bf16 *B;   // op(B): K x N, stored [N][K] (k contiguous)
    int M, N, K;
    int lda, ldb;
    int n_store;       // columns n < n_store are written
    const float *bias; // EPI_BIAS*: N values
    float *Z;          // EPI_BIAS_SILU: pre-activation out (optional), EPI_DSILU: in; [M][ldz]
    int ldz;
    float *Cf;         // FP32 result (optional)
    __bf16 *Cb;        // BF16 result (optional)
    __bf16 *Cbt;       // the same BF16 values transposed, [n][m] with row stride ldct (optional)
    long long ldct;
    int ldc;
    long long z_stride;   // elements between the results of successive grid.z slices (split K) in Cf
    int c_trans;       // Cf stored [n][m] (ld = ldc) instead of [m][n]
    int k_split;       // K range of one grid.z slice (multiple of kKStep)
    float *colpart;    // EPI_DSILU, EPI_MSE: each row tile's column sums of the result, [blockIdx.y][N]
    // EPI_MSE: weighted squared error against t [M][n_real] with per-output weights w [n_real] and per-row
    // weights rw [M] (null: 1) on the first `masked` outputs; the result is dL/dy, and each workgroup's
    // share of the loss goes to loss_part[blockIdx.y gridDim.x + blockIdx.x]
    const float *t, *w, *rw;
    int n_real, masked;
    float *loss_part;
};

// SiLU and its derivative with the hardware reciprocal (1 ulp) instead of a full-precision divide.
__device__ __forceinline__ float silu(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// A 64 x kStageK (rows x k) operand tile, stored [row][k] with row stride ld, moves through registers into
// LDS as [row][k]: every thread loads kStageLoads 16-byte pieces (all in flight together), then stores
// them.  Zeros past the operand's rows and past the slice's K range.  Every operand is k-contiguous: the
// weight-gradient GEMMs read the transposed copies of their inputs that the producing kernels write.
constexpr int kStageK = 128;
constexpr int kLdsStage = kStageK + 8;
constexpr int kStageLoads = kStageK / kKStep;

__device__ __forceinline__ void load_stage(const __bf16 *src, int ld, int rows, int K, int r0, int k0, uint4 (&v)[kStageLoads],
                                           int tid) {
    const int r = tid >> 2;
#pragma unroll
    for (int s = 0; s < kStageLoads; ++s) {
        v[s] = uint4{0u, 0u, 0u, 0u};
        const int kk = s * kKStep + (tid & 3) * 8;
        if (r0 + r < rows && k0 + kk < K) {
            const __bf16 *q = src + size_t(r0 + r) * ld + k0 + kk;
            if (k0 + kk + 8 <= K) {
                v[s] = *reinterpret_cast<const uint4 *>(q);
            } else {   // the last, partial group of the K range: zeros past it
                unsigned short e[8];
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    e[i] = k0 + kk + i < K ? *reinterpret_cast<const unsigned short *>(q + i) : (unsigned short)0;
                v[s] = __builtin_bit_cast(uint4, e);
            }
        }
    }
}

__device__ __forceinline__ void store_stage(const uint4 (&v)[kStageLoads], __bf16 (*dst)[kLdsStage], int tid) {
#pragma unroll
    for (int s = 0; s < kStageLoads; ++s) *reinterpret_cast<uint4 *>(&dst[tid >> 2][s * kKStep + (tid & 3) * 8]) = v[s];
}

// C tile (64 x 64) = op(A) op(B) over this slice's K range, in stages of kStageK: the next stage's loads
// are in flight while the current stage's MFMAs run.  Four waves, each a 32 x 32 quadrant of 2 x 2 MFMA
// 16x16x32 tiles, the k-steps in order (so the sums do not depend on the staging).  Lane l of an MFMA
// holds A[l%16][(l/16)*8 .. +7] and B[(l/16)*8 .. +7][l%16] and receives D[(l/16)*4 + q][l%16], q = 0..3.
template <int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
    __shared__ __align__(16) __bf16 sA[kTile][kLdsStage];
    __shared__ __align__(16) __bf16 sB[kTile][kLdsStage];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int kbeg = blockIdx.z * g.k_split;
    const int kend = min(g.K, kbeg + g.k_split);
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int kb = (lane >> 4) * 8, rl = lane & 15;
    uint4 va[kStageLoads], vb[kStageLoads];
    load_stage(g.A, g.lda, g.M, kend, m0, kbeg, va, tid);
    load_stage(g.B, g.ldb, g.N, kend, n0, kbeg, vb, tid);
    for (int k0 = kbeg; k0 < kend; k0 += kStageK) {
        store_stage(va, sA, tid);
        store_stage(vb, sB, tid);
        __syncthreads();
        if (k0 + kStageK < kend) {
            load_stage(g.A, g.lda, g.M, kend, m0, k0 + kStageK, va, tid);
            load_stage(g.B, g.ldb, g.N, kend, n0, k0 + kStageK, vb, tid);
        }
#pragma unroll
        for (int ks = 0; ks < kStageLoads; ++ks) {
            if (k0 + ks * kKStep >= kend) break;
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8 *>(&sA[wm + i * 16 + rl][ks * kKStep + kb]);
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8 *>(&sB[wn + j * 16 + rl][ks * kKStep + kb]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    float csum[2] = {0.f, 0.f};   // EPI_DSILU / EPI_MSE: this lane's part of its two columns' sums
    float lsum = 0.f;             // EPI_MSE: this lane's part of the loss
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            // the lane's four rows m of column n, for the transposed copy (one 8-byte store when all are rows)
            union {
                __bf16 h[4];
                uint2 u;
            } tv;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + q;
                const int n = n0 + wn + j * 16 + rl;
                tv.h[q] = __bf16(0.f);
                if (m >= g.M || n >= g.n_store) continue;
                float v = acc[i][j][q];
                if constexpr (EPI == EPI_BIAS_SILU) {
                    v += g.bias[n];
                    if (g.Z) g.Z[size_t(m) * g.ldz + n] = v;
                    v = silu(v);
                } else if constexpr (EPI == EPI_BIAS) {
                    v += g.bias[n];
                } else if constexpr (EPI == EPI_DSILU) {
                    v *= dsilu(g.Z[size_t(m) * g.ldz + n]);
                    csum[j] += v;
                } else if constexpr (EPI == EPI_MSE) {
                    const float y = v + g.bias[n];
                    float gr = 0.f;
                    if (n < g.n_real) {
                        const float wo = (g.rw && n < g.masked) ? g.w[n] * g.rw[m] : g.w[n];
                        const float d = y - g.t[size_t(m) * g.n_real + n];
                        gr = wo == 0.f ? 0.f : 2.0f * wo * d / float(g.M);
                        lsum += wo == 0.f ? 0.f : wo * d * d / float(g.M);
                    }
                    v = gr;
                    csum[j] += v;
                }
                const size_t ci = g.c_trans ? size_t(n) * g.ldc + m : size_t(m) * g.ldc + n;
                if (g.Cf) g.Cf[size_t(blockIdx.z) * size_t(g.z_stride) + ci] = v;
                if (g.Cb) g.Cb[ci] = __bf16(v);
                tv.h[q] = __bf16(v);
            }
            // the transposed copy goes through LDS (the staging tile is free after the main loop's last
            // barrier): column n of the tile becomes one 128-byte row of Cbt
            if (g.Cbt) *reinterpret_cast<uint2 *>(&sA[wn + j * 16 + rl][wm + i * 16 + (lane >> 4) * 4]) = tv.u;
        }
    if (g.Cbt) {   // wave-uniform
        __syncthreads();
#pragma unroll
        for (int e = tid; e < kTile * 8; e += 256) {
            const int nl = e >> 3, ml = (e & 7) * 8;
            // rows of the tile past M hold zeros, and the copy's row stride reaches the tile's end
            if (n0 + nl < g.n_store)
                *reinterpret_cast<uint4 *>(g.Cbt + size_t(n0 + nl) * size_t(g.ldct) + m0 + ml) =
                    *reinterpret_cast<const uint4 *>(&sA[nl][ml]);
        }
    }
    if constexpr (EPI == EPI_DSILU || EPI == EPI_MSE) {
        // the four lanes of a column (lane / 16) hold its 4 x 2 x 4 values of this wave's 32 rows; the two
        // waves of a column range (wm = 0, 32) meet in LDS, and the tile's column sums are stored (no atomics,
        // so that the bias gradient, reduced over the tiles in a fixed order, is the same on every run)
        __shared__ float cpart[2][32];
        __shared__ float lpart[4];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            csum[j] += __shfl_xor(csum[j], 16);
            csum[j] += __shfl_xor(csum[j], 32);
        }
        if (wm != 0 && lane < 16) {
            cpart[wave & 1][rl] = csum[0];
            cpart[wave & 1][16 + rl] = csum[1];
        }
        if constexpr (EPI == EPI_MSE) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
            if (lane == 0) lpart[wave] = lsum;
        }
        __syncthreads();
        if (wm == 0 && lane < 16) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int n = n0 + wn + j * 16 + rl;
                if (n < g.n_store) g.colpart[size_t(blockIdx.y) * g.N + n] = csum[j] + cpart[wave & 1][j * 16 + rl];
            }
        }
        if constexpr (EPI == EPI_MSE) {
            if (tid == 0) g.loss_part[blockIdx.y * gridDim.x + blockIdx.x] = lpart[0] + lpart[1] + lpart[2] + lpart[3];
        }
    }
}

// Where each layer's bias gradient lies: b_off[l] in the parameter vector, fo[l] entries, and its row tiles'
// column sums at part_off[l] in the partial-sum buffer ([tiles][fo[l]]).
struct BiasMap {
    int layers, tiles;
    long long b_off[kMaxLayersAdam], part_off[kMaxLayersAdam];
    int fo[kMaxLayersAdam];
};

// Bias gradients: grad[b_off[l] + n] = the sum of the row tiles' column sums, always added in the same order.
// A workgroup takes 64 biases; its 16 waves sum every 16th tile (tiles w, w + 16, ...), and wave 0 adds the
// 16 partial sums in order.
__global__ __launch_bounds__(1024) void bias_reduce_kernel(const float *part, BiasMap B, float *grad) {
    __shared__ float acc[16][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int i = int(blockIdx.x) * 64 + lane, l = 0;
    while (l < B.layers && i >= B.fo[l]) i -= B.fo[l++];
    float s = 0.f;
    if (l < B.layers) {
        const float *p = part + B.part_off[l] + i;
#pragma unroll 4
        for (int t = w; t < B.tiles; t += 16) s += p[size_t(t) * B.fo[l]];
    }
    acc[w][lane] = s;
    __syncthreads();
    if (w == 0 && l < B.layers) {
        float tot = 0.f;
#pragma unroll
        for (int k = 0; k < 16; ++k) tot += acc[k][lane];
        grad[B.b_off[l] + i] = tot;
    }
}

// Where each layer's weight gradient lies: w_off[l] in the parameter vector, n[l] = fi * fo entries, and its
// split-K slices' partial results at part_off[l] ([splits][n[l]]).
struct WeightMap {
    int layers, splits;
    long long w_off[kMaxLayersAdam], part_off[kMaxLayersAdam], n[kMaxLayersAdam];
};

// Weight gradients: grad[w_off[l] + i] = the sum of the slices' partial results, added in slice order, so the
// gradient is the same on every run (split-K with atomics would add them in arrival order).
__global__ __launch_bounds__(256) void weight_reduce_kernel(const float *part, WeightMap W, float *grad) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    int l = 0;
    while (l < W.layers && i >= W.n[l]) i -= W.n[l++];
    if (l >= W.layers) return;
    const float *p = part + W.part_off[l] + i;
    float s = p[0];
    for (int z = 1; z < W.splits; ++z) s += p[size_t(z) * size_t(W.n[l])];
    grad[W.w_off[l] + i] = s;
}

// Sum of n floats into *out, by one workgroup (the loss from the output GEMM's per-workgroup parts).
__global__ __launch_bounds__(256) void sum_kernel(const float *x, int n, float *out) {
    __shared__ float part[4];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) s += x[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = part[0] + part[1] + part[2] + part[3];
}

// FP32 [rows][n] -> BF16 [rows][n_pad], zero padded
// BF16 inputs [rows][n_pad] (zero-padded columns), and when yt is set their transpose [n_pad][ldt].
__global__ void pack_input_kernel(const float *x, int rows, int n, int n_pad, __bf16 *y, __bf16 *yt, long long ldt) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(rows) * n_pad) return;
    const int c = int(i % n_pad);
    const size_t r = i / n_pad;
    const __bf16 v = __bf16(c < n ? x[r * n + c] : 0.0f);
    y[i] = v;
    if (yt) yt[size_t(c) * size_t(ldt) + r] = v;
}

__global__ void cast_bf16_kernel(const float *x, __bf16 *y, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = __bf16(x[i]);
}

// BF16 W^T of one layer: wt[n][k] = w[k][n]
__global__ void transpose_bf16_kernel(const float *w, int K, int N, __bf16 *wt) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(K) * N) return;
    const int k = int(i / N), n = int(i % N);
    wt[size_t(n) * K + k] = __bf16(w[i]);
}

// Where the weight matrices lie in the parameter vector: layer l's W [fi][fo] at w_off[l] (biases follow it).
struct LayerMap {
    int layers;
    long long w_off[kMaxLayersAdam];
    int fi[kMaxLayersAdam], fo[kMaxLayersAdam];
};

// Adam update of the FP32 masters, with the BF16 copy and, for a weight, its BF16 W^T entry.
__global__ void adam_kernel(float *w, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2, __bf16 *wb, __bf16 *wtb, LayerMap L) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    wb[i] = __bf16(wi);
    for (int l = 0; l < L.layers; ++l) {
        const long long e = (long long)i - L.w_off[l];
        if (e >= 0 && e < (long long)L.fi[l] * L.fo[l]) {
            const int k = int(e / L.fo[l]), c = int(e % L.fo[l]);
            wtb[L.w_off[l] + (long long)c * L.fi[l] + k] = __bf16(wi);
        }
    }
}

// ---- fused inference: all layers of a 128-row tile in one workgroup ------------------------------
// Width-256 networks.  Activations stay in LDS as BF16 [128][256] (each wave owns 32 rows, so a wave
// overwrites only rows it alone reads); each layer's W^T is staged in LDS 32 k-columns at a time
// (double buffered, one barrier per slice), and each wave computes its 32 x 256 block with 2 x 16
// MFMA 16x16x32 tiles accumulating in FP32.  The first layer (n_in <= 8 inputs) runs on the VALU with
// the same BF16 rounding of inputs and weights as the MFMA path.  The output layer (up to 1024
// columns) runs in passes of 256 columns and writes FP32 transposed ([n_out][ld]).
constexpr int kFRows = 128;             // rows per workgroup
constexpr int kFWidth = 256;            // hidden width
constexpr int kFActLd = kFWidth + 8;    // LDS row stride (bf16) of the activations
constexpr int kFK = 32;                 // k per MFMA step
constexpr int kFWaveCols = 64;          // output columns per wave and pass (4 waves x 64 = 256)
constexpr int kFRowTiles = kFRows / 16;
constexpr int kFColTiles = kFWaveCols / 16;
constexpr int kFMaxLayers = 10;

struct FusedArgs {
    const float *x;          // [rows][n_in]
    int rows, n_in, depth, n_out, n_out_pad;
    const float *w32;        // FP32 parameters (layer-0 weights, biases)
    const __bf16 *wtb;       // BF16 W^T per layer
    long long w_off[kFMaxLayers], b_off[kFMaxLayers];
    float *y;                // [n_out][ld], rows r0 .. of this launch at column offset 0
    _Float16 *yh;            // when set, the outputs go here as FP16 instead (same layout)
    long long ld;
    unsigned wt_bytes;       // bytes of the W^T buffer (buffer-resource bound)
};

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Byte offset of lane l's weight fragment row: W^T[col][(l/16)*8 ..] of the layer at wbase; columns
// past n_cols point beyond the buffer, so their loads return zeros.
__device__ __forceinline__ unsigned w_frag_offset(unsigned wbase, int col, int n_cols, int kq) {
    return col < n_cols ? wbase + unsigned(col) * unsigned(kFWidth * 2) + unsigned(kq * 2) : 0x80000000u;
}

// Weight (W^T) fragment of a 16x16x32 MFMA straight from global memory through a buffer resource:
// 16 contiguous bytes per lane at k-step kt (the weights of a layer, 128 KB, stay in L2 and are shared
// by every workgroup).
__device__ __forceinline__ bf16x8 load_w_frag(__amdgpu_buffer_rsrc_t wrs, unsigned voff, int kt) {
    return __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(wrs, int(voff), kt * kFK * 2, 0));
}

// Bias + SiLU of a hidden-layer tile computed as D = W^T act^T (lane l holds output columns
// cb .. cb+3, cb = c0 + (l/16)*4, of row l%16), written to the activations as one 8-byte store.
__device__ __forceinline__ void store_hidden(__bf16 (*act)[kFActLd], int row, int cb, f32x4 bias, f32x4 d) {
    union {
        __bf16 h[4];
        uint2 u;
    } p;
#pragma unroll
    for (int q = 0; q < 4; ++q) p.h[q] = __bf16(silu(d[q] + bias[q]));
    *reinterpret_cast<uint2 *>(&act[row][cb]) = p.u;
}

// The four biases of columns cb .. cb+3.
__device__ __forceinline__ f32x4 bias4(const float *bias, int cb) {
    return f32x4{bias[cb], bias[cb + 1], bias[cb + 2], bias[cb + 3]};
}

// acc[i][j] += the 16 x 16 tile (row tile i_lo + i, column tile j at column cb + 16 j) of act x W over
// the full k = 256; TRANS computes each tile as W^T act^T instead (lane l: columns (l/16)*4 + q of
// row l%16), else as act W (lane l: rows (l/16)*4 + q of column l%16).
template <int NI, bool TRANS>
__device__ __forceinline__ void gemm_tiles(const __bf16 (*act)[kFActLd], __amdgpu_buffer_rsrc_t wrs, unsigned wbase,
                                           int n_cols, int cb, int i_lo, f32x4 (&acc)[NI][kFColTiles], int rl, int kq) {
    unsigned voff[kFColTiles];
    bf16x8 w[kFColTiles];
#pragma unroll
    for (int j = 0; j < kFColTiles; ++j) {
        voff[j] = w_frag_offset(wbase, cb + j * 16 + rl, n_cols, kq);
        w[j] = load_w_frag(wrs, voff[j], 0);
    }
#pragma unroll
    for (int kt = 0; kt < kFWidth / kFK; ++kt) {
        bf16x8 wn[kFColTiles];
        if (kt + 1 < kFWidth / kFK) {
#pragma unroll
            for (int j = 0; j < kFColTiles; ++j) wn[j] = load_w_frag(wrs, voff[j], kt + 1);
        }
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&act[(i_lo + i) * 16 + rl][kt * kFK + kq]);
#pragma unroll
            for (int j = 0; j < kFColTiles; ++j)
                acc[i][j] = TRANS ? __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[j], a, acc[i][j], 0, 0, 0)
                                  : __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w[j], acc[i][j], 0, 0, 0);
        }
        if (kt + 1 < kFWidth / kFK) {
#pragma unroll
            for (int j = 0; j < kFColTiles; ++j) w[j] = wn[j];
        }
    }
}

// Output-layer tiles (act W form) + bias into y[col][row]: one 16-byte store of four rows per tile.
template <int NI>
__device__ __forceinline__ void store_output(const FusedArgs &F, const float *bias, int cb, int row0,
                                             const f32x4 (&acc)[NI][kFColTiles], int lane) {
#pragma unroll
    for (int j = 0; j < kFColTiles; ++j) {
        const int c = cb + j * 16 + (lane & 15);
        if (c >= F.n_out) continue;
        const float bv = bias[c];
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int row = row0 + i * 16 + (lane >> 4) * 4;
            if (F.yh) {   // FP16 outputs: four rows in one 8-byte store
                _Float16 *h = F.yh + size_t(c) * F.ld + row;
                if (row + 3 < F.rows && (F.ld & 3) == 0) {
                    typedef _Float16 h4 __attribute__((ext_vector_type(4)));
                    *reinterpret_cast<h4 *>(h) = h4{_Float16(acc[i][j][0] + bv), _Float16(acc[i][j][1] + bv),
                                                    _Float16(acc[i][j][2] + bv), _Float16(acc[i][j][3] + bv)};
                } else {
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (row + q < F.rows) h[q] = _Float16(acc[i][j][q] + bv);
                }
                continue;
            }
            float *o = F.y + size_t(c) * F.ld + row;
            if (row + 3 < F.rows && (F.ld & 3) == 0) {
                *reinterpret_cast<float4 *>(o) =
                    make_float4(acc[i][j][0] + bv, acc[i][j][1] + bv, acc[i][j][2] + bv, acc[i][j][3] + bv);
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (row + q < F.rows) o[q] = acc[i][j][q] + bv;
            }
        }
    }
}

// Fused transposed inference of the width-256 network.  A workgroup carries 128 rows through every
// layer with the activations resident in LDS (BF16).  Wave w computes all 128 rows x 64 output
// columns of a pass: 8 x 4 MFMA 16x16x32 tiles, activation fragments read from LDS, weight fragments
// read from global memory (L2), the next k-step's weight fragments in flight during the current
// MFMAs.  Hidden layers compute each tile as D = W^T act^T, so a lane ends with four consecutive
// columns of one row and writes them back with one 8-byte LDS store; two barriers per hidden layer
// (all reads of the layer's input done; outputs written).  The output layer computes D = act W, so a
// lane ends with four consecutive rows of one column: one 16-byte store into y[col][row].  Layer 0
// (n_in <= 8) is one zero-padded MFMA k-step.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void mlp_fused_kernel(FusedArgs F) {
    __shared__ __align__(16) __bf16 act[kFRows][kFActLd];
    __shared__ __align__(16) __bf16 xs[kFRows][8];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int r0 = blockIdx.x * kFRows;
    const int rl = lane & 15, kq = (lane >> 4) * 8;
    const __amdgpu_buffer_rsrc_t wrs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<__bf16 *>(F.wtb), (short)0, int(F.wt_bytes), int(kRsrcWord3));

    // layer 0: A = W0^T (k 0..7, zero above), B = x^T (BF16-rounded inputs)
    for (int i = tid; i < kFRows * 8; i += 256) {
        const int r = i >> 3, k = i & 7;
        xs[r][k] = __bf16((r0 + r < F.rows && k < F.n_in) ? F.x[size_t(r0 + r) * F.n_in + k] : 0.f);
    }
    __syncthreads();
    {
        const float *W0 = F.w32 + F.w_off[0];   // [8][256] FP32 master weights
        const float *b0 = F.w32 + F.b_off[0];
        const int c0 = wave * kFWaveCols;
        bf16x8 w[kFColTiles];
#pragma unroll
        for (int j = 0; j < kFColTiles; ++j) {
            w[j] = bf16x8{};
            if (lane < 16) {
#pragma unroll
                for (int k = 0; k < 8; ++k) w[j][k] = __bf16(k < F.n_in ? W0[k * kFWidth + c0 + j * 16 + rl] : 0.f);
            }
        }
        f32x4 bv[kFColTiles];
#pragma unroll
        for (int j = 0; j < kFColTiles; ++j) bv[j] = bias4(b0, c0 + j * 16 + (lane >> 4) * 4);
#pragma unroll
        for (int i = 0; i < kFRowTiles; ++i) {
            const bf16x8 xf = lane < 16 ? *reinterpret_cast<const bf16x8 *>(&xs[i * 16 + rl][0]) : bf16x8{};
#pragma unroll
            for (int j = 0; j < kFColTiles; ++j)
                store_hidden(act, i * 16 + rl, c0 + j * 16 + (lane >> 4) * 4, bv[j],
                             __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[j], xf, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0));
        }
    }
    __syncthreads();

    for (int l = 1; l < F.depth; ++l) {   // hidden layers
        const unsigned wbase = unsigned(F.w_off[l] * 2);   // W^T [256][256] of this layer, bytes
        const float *bias = F.w32 + F.b_off[l];
        const int c0 = wave * kFWaveCols;
        f32x4 acc[kFRowTiles][kFColTiles];
#pragma unroll
        for (int i = 0; i < kFRowTiles; ++i)
#pragma unroll
            for (int j = 0; j < kFColTiles; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
        gemm_tiles<kFRowTiles, true>(act, wrs, wbase, kFWidth, c0, 0, acc, rl, kq);
        f32x4 bv[kFColTiles];
#pragma unroll
        for (int j = 0; j < kFColTiles; ++j) bv[j] = bias4(bias, c0 + j * 16 + (lane >> 4) * 4);
        __syncthreads();   // every wave has read this layer's input activations
#pragma unroll
        for (int i = 0; i < kFRowTiles; ++i)
#pragma unroll
            for (int j = 0; j < kFColTiles; ++j) store_hidden(act, i * 16 + rl, c0 + j * 16 + (lane >> 4) * 4, bv[j], acc[i][j]);
        __syncthreads();   // the next layer reads all 256 columns
    }
    {   // output layer, passes of 256 columns; a last pass of <= 64 columns is split over the waves by rows
        const int n_cols = F.n_out_pad;
        const unsigned wbase = unsigned(F.w_off[F.depth] * 2);   // W^T [n_cols][256], bytes
        const float *bias = F.w32 + F.b_off[F.depth];
        for (int n0 = 0; n0 < n_cols; n0 += kFWidth) {
            if (n_cols - n0 > kFWaveCols) {
                const int c0 = n0 + wave * kFWaveCols;
                if (c0 < n_cols) {   // wave-uniform
                    f32x4 acc[kFRowTiles][kFColTiles];
#pragma unroll
                    for (int i = 0; i < kFRowTiles; ++i)
#pragma unroll
                        for (int j = 0; j < kFColTiles; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
                    gemm_tiles<kFRowTiles, false>(act, wrs, wbase, n_cols, c0, 0, acc, rl, kq);
                    store_output<kFRowTiles>(F, bias, c0, r0, acc, lane);
                }
            } else {
                constexpr int kSplit = kFRowTiles / 4;
                f32x4 acc[kSplit][kFColTiles];
#pragma unroll
                for (int i = 0; i < kSplit; ++i)
#pragma unroll
                    for (int j = 0; j < kFColTiles; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
                gemm_tiles<kSplit, false>(act, wrs, wbase, n_cols, n0, wave * kSplit, acc, rl, kq);
                store_output<kSplit>(F, bias, n0, r0 + wave * kSplit * 16, acc, lane);
            }
        }
    }
}

}  // namespace bh

using namespace bh;

struct bh_mlp {
    bh_ctx *ctx = nullptr;
    int n_in = 0, n_in_pad = 0, width = 0, depth = 0, n_out = 0, n_out_pad = 0;
    std::vector<int> fi, fo;                   // padded fan-in / fan-out per linear layer
    std::vector<size_t> w_off, b_off;          // offsets into the parameter vector
    size_t n_params = 0;
    float *w32 = nullptr;                      // FP32 master parameters
    __bf16 *wb = nullptr;                      // BF16 copy, same layout
    __bf16 *wtb = nullptr;                     // BF16 W^T per layer (at w_off)
    float *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    long long adam_t = 0;
    int cap = 0;                               // rows the activation buffers hold
    bool cap_train = false;                    // training buffers present
    __bf16 *xb = nullptr;                      // [cap][n_in_pad]
    std::vector<__bf16 *> act;                 // per hidden layer [cap][width]
    std::vector<float *> zpre;                 // per hidden layer pre-activation [cap][width]
    __bf16 *dyb = nullptr;                     // [cap][n_out_pad] loss gradient of the outputs
    __bf16 *dab = nullptr, *dab2 = nullptr;    // [cap][widest] pre-activation gradients, alternating
    // training: the same BF16 inputs, activations and gradients transposed ([columns][ldt]), written by the
    // kernels that produce them, so the weight-gradient GEMMs read both operands k-contiguous
    long long ldt = 0;
    __bf16 *xbT = nullptr, *dybT = nullptr, *dabT = nullptr, *dab2T = nullptr;
    std::vector<__bf16 *> actT;
    float *loss = nullptr;
    float *loss_part = nullptr;                // per-workgroup loss of the output GEMM (one per 64 x 64 tile)
    int loss_parts = 0;
    float *bias_part = nullptr;                // per row tile column sums of each layer's gradient [tiles][fo]
    float *dw_part = nullptr;                  // split-K partial weight gradients, per layer [splits][fi * fo]
    int dw_splits = 0;                         // slices the buffer holds
    int n_bias = 0;                            // biases in all (sum of fo)
    std::vector<int> fo_prefix;                // biases of the layers before l
    bool fused = true;                         // inference through mlp_fused_kernel when the shape allows
};

namespace bh {

template <int EPI>
static int run_gemm(bh_ctx *ctx, const GemmArgs &g0, int splits = 1) {
    GemmArgs g = g0;
    if (g.n_store == 0) g.n_store = g.N;
    g.k_split = ((g.K + splits - 1) / splits + kKStep - 1) / kKStep * kKStep;
    const int z = (g.K + g.k_split - 1) / g.k_split;
    dim3 grid((g.N + kTile - 1) / kTile, (g.M + kTile - 1) / kTile, z);
    hipLaunchKernelGGL((gemm_kernel<EPI>), grid, dim3(256), 0, ctx->stream, g);
    BH_HIP(hipGetLastError());
    return BH_OK;
}

static unsigned blocks_for(size_t n) { return unsigned((n + 255) / 256); }

// Split-K slices of the weight gradients for a batch of `rows` (monotonic in rows).
static int dw_splits_for(int rows) {
    if (rows < 1024) return 1;
    return rows / kDwRows < kDwMaxSplits ? rows / kDwRows : kDwMaxSplits;
}

static size_t weight_count(const bh_mlp *m) {
    size_t n = 0;
    for (int l = 0; l <= m->depth; ++l) n += size_t(m->fi[l]) * m->fo[l];
    return n;
}

static void free_acts(bh_mlp *m) {
    if (m->xb) (void)hipFree(m->xb);
    for (auto p : m->act) (void)hipFree(p);
    for (auto p : m->zpre)
        if (p) (void)hipFree(p);
    for (void *p : {(void *)m->dyb, (void *)m->dab, (void *)m->dab2, (void *)m->loss_part, (void *)m->bias_part,
                    (void *)m->dw_part, (void *)m->xbT, (void *)m->dybT, (void *)m->dabT, (void *)m->dab2T})
        if (p) (void)hipFree(p);
    for (auto p : m->actT)
        if (p) (void)hipFree(p);
    m->actT.clear();
    m->xbT = m->dybT = m->dabT = m->dab2T = nullptr;
    m->ldt = 0;
    m->loss_part = m->bias_part = m->dw_part = nullptr;
    m->dw_splits = 0;
    m->loss_parts = 0;
    m->xb = nullptr;
    m->act.clear();
    m->zpre.clear();
    m->dyb = m->dab = m->dab2 = nullptr;
    m->cap = 0;
    m->cap_train = false;
}

static int ensure_cap(bh_mlp *m, int rows, bool train) {
    if (rows <= m->cap && (!train || m->cap_train)) return BH_OK;
    if (rows < m->cap) rows = m->cap;
    train = train || m->cap_train;
    (void)hipStreamSynchronize(m->ctx->stream);
    free_acts(m);
    const size_t R = size_t(rows);
    const int widest = m->width > m->n_out_pad ? m->width : m->n_out_pad;
    bool ok = hipMalloc(&m->xb, R * m->n_in_pad * 2) == hipSuccess;
    for (int l = 0; l < m->depth && ok; ++l) {
        __bf16 *a = nullptr;
        float *z = nullptr;
        ok = hipMalloc(&a, R * m->width * 2) == hipSuccess;
        if (ok && train) ok = hipMalloc(&z, R * m->width * 4) == hipSuccess;
        m->act.push_back(a);
        m->zpre.push_back(z);
    }
    const int parts = int(((R + kTile - 1) / kTile) * ((m->n_out_pad + kTile - 1) / kTile));
    if (ok && train)
        ok = hipMalloc(&m->dyb, R * m->n_out_pad * 2) == hipSuccess && hipMalloc(&m->dab, R * widest * 2) == hipSuccess &&
             hipMalloc(&m->dab2, R * widest * 2) == hipSuccess &&
             hipMalloc(&m->loss_part, size_t(parts) * 4) == hipSuccess &&
             hipMalloc(&m->bias_part, ((R + kTile - 1) / kTile) * size_t(m->n_bias) * 4) == hipSuccess &&
             hipMalloc(&m->dw_part, size_t(dw_splits_for(rows)) * weight_count(m) * 4) == hipSuccess;
    if (ok && train) {
        const size_t ldt = (R + 63) / 64 * 64;
        ok = hipMalloc(&m->xbT, ldt * m->n_in_pad * 2) == hipSuccess &&
             hipMalloc(&m->dybT, ldt * m->n_out_pad * 2) == hipSuccess &&
             hipMalloc(&m->dabT, ldt * widest * 2) == hipSuccess && hipMalloc(&m->dab2T, ldt * widest * 2) == hipSuccess;
        for (int l = 0; l < m->depth && ok; ++l) {
            __bf16 *a = nullptr;
            ok = hipMalloc(&a, ldt * m->width * 2) == hipSuccess;
            m->actT.push_back(a);
        }
        m->ldt = (long long)ldt;
    }
    if (ok && train) {
        m->loss_parts = parts;
        m->dw_splits = dw_splits_for(rows);
    }
    if (!ok) {
        free_acts(m);
        return set_error(BH_ERR_MEMORY, "MLP activation allocation for %d rows failed", rows);
    }
    m->cap = rows;
    m->cap_train = train;
    return BH_OK;
}

static void refresh_transposed(bh_mlp *m) {
    for (int l = 0; l <= m->depth; ++l)
        hipLaunchKernelGGL(transpose_bf16_kernel, dim3(blocks_for(size_t(m->fi[l]) * m->fo[l])), dim3(256), 0,
                           m->ctx->stream, m->w32 + m->w_off[l], m->fi[l], m->fo[l], m->wtb + m->w_off[l]);
}

// BF16 copies (W and W^T) of the FP32 masters.
static int refresh_bf16(bh_mlp *m) {
    hipLaunchKernelGGL(cast_bf16_kernel, dim3(blocks_for(m->n_params)), dim3(256), 0, m->ctx->stream, m->w32, m->wb,
                       m->n_params);
    refresh_transposed(m);
    BH_HIP(hipGetLastError());
    return BH_OK;
}

// Forward of `rows` rows of x (FP32 [rows][n_in]): the n_out real columns into y_out ([rows][ld_out], or
// [n_out][ld_out] when trans).  train: the hidden layers only, keeping the pre-activations; the output
// layer then runs in backward(), whose epilogue turns its result into the loss gradient.
static int forward(bh_mlp *m, const float *d_x, int rows, bool train, float *y_out, int trans, int ld_out) {
    bh_ctx *ctx = m->ctx;
    hipLaunchKernelGGL(pack_input_kernel, dim3(blocks_for(size_t(rows) * m->n_in_pad)), dim3(256), 0, ctx->stream, d_x,
                       rows, m->n_in, m->n_in_pad, m->xb, train ? m->xbT : nullptr, m->ldt);
    const __bf16 *in = m->xb;
    int st;
    for (int l = 0; l <= (train ? m->depth - 1 : m->depth); ++l) {
        GemmArgs g{};
        g.A = in;
        g.lda = m->fi[l];
        g.B = m->wtb + m->w_off[l];
        g.ldb = m->fi[l];
        g.M = rows;
        g.K = m->fi[l];
        g.N = m->fo[l];
        g.bias = m->w32 + m->b_off[l];
        if (l < m->depth) {
            g.Z = train ? m->zpre[l] : nullptr;
            g.ldz = g.N;
            g.Cb = m->act[l];
            g.ldc = g.N;
            g.Cbt = train ? m->actT[l] : nullptr;
            g.ldct = m->ldt;
            st = run_gemm<EPI_BIAS_SILU>(ctx, g);
            in = m->act[l];
        } else {
            g.Cf = y_out;
            g.n_store = m->n_out;
            g.c_trans = trans;
            g.ldc = ld_out;
            st = run_gemm<EPI_BIAS>(ctx, g);
        }
        if (st) return st;
    }
    return BH_OK;
}

// Start of layer l's row-tile column sums in m->bias_part ([cap tiles][fo[l]] per layer, layers in order).
static size_t bias_part_base(const bh_mlp *m, int l) {
    return size_t((m->cap + kTile - 1) / kTile) * size_t(m->fo_prefix[l]);
}

// The batch loss of the last backward() into m->loss (the sum of the output GEMM's per-workgroup parts).
static int loss_total(bh_mlp *m, int rows) {
    const int parts = ((rows + kTile - 1) / kTile) * ((m->n_out_pad + kTile - 1) / kTile);
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, m->ctx->stream, m->loss_part, parts, m->loss);
    BH_HIP(hipGetLastError());
    return BH_OK;
}

// Output layer and backward of the weighted MSE loss at the hidden activations of the last training
// forward; gradients into m->grad, the loss in parts into m->loss_part (loss_total() adds them up).  Bias
// gradients are the column sums of each
// layer's pre-activation gradient, formed in the epilogue of the GEMM that produces it.
static int backward(bh_mlp *m, int rows, const float *d_t, const float *d_w, const float *d_rw, int masked) {
    bh_ctx *ctx = m->ctx;
    int st;
    {   // y = act W + b, then dL/dy (BF16), its column sums (output bias gradient) and the loss
        const int l = m->depth;
        GemmArgs g{};
        g.A = l == 0 ? m->xb : m->act[l - 1];
        g.lda = m->fi[l];
        g.B = m->wtb + m->w_off[l];
        g.ldb = m->fi[l];
        g.M = rows;
        g.K = m->fi[l];
        g.N = m->fo[l];
        g.bias = m->w32 + m->b_off[l];
        g.Cb = m->dyb;
        g.ldc = g.N;
        g.Cbt = m->dybT;
        g.ldct = m->ldt;
        g.colpart = m->bias_part + bias_part_base(m, l);
        g.t = d_t;
        g.w = d_w;
        g.rw = d_rw;
        g.n_real = m->n_out;
        g.masked = masked;
        g.loss_part = m->loss_part;
        if ((st = run_gemm<EPI_MSE>(ctx, g))) return st;
    }
    const __bf16 *dcur = m->dyb;   // gradient w.r.t. this layer's pre-activation, BF16
    const __bf16 *dcurT = m->dybT;  // the same, transposed
    // weight gradients in split-K slices of the batch rows (as run_gemm cuts them), summed afterwards
    const int splits = dw_splits_for(rows);
    const int k_split = ((rows + splits - 1) / splits + kKStep - 1) / kKStep * kKStep;
    WeightMap W{};
    W.layers = m->depth + 1;
    W.splits = (rows + k_split - 1) / k_split;
    if (W.splits > m->dw_splits) return set_error(BH_ERR_STATE, "weight-gradient slices exceed their buffer");
    long long part_off = 0;
    for (int l = 0; l <= m->depth; ++l) {
        W.w_off[l] = (long long)m->w_off[l];
        W.n[l] = (long long)m->fi[l] * m->fo[l];
        W.part_off[l] = part_off;
        part_off += W.n[l] * m->dw_splits;
    }
    for (int l = m->depth; l >= 0; --l) {
        const int K = m->fi[l], N = m->fo[l];
        // dW [K][N] = x^T dZ, split over rows, from the transposed copies ([K][rows] and [N][rows])
        GemmArgs gw{};
        gw.A = l == 0 ? m->xbT : m->actT[l - 1];
        gw.lda = int(m->ldt);
        gw.B = dcurT;
        gw.ldb = int(m->ldt);
        gw.M = K;
        gw.N = N;
        gw.K = rows;
        gw.Cf = m->dw_part + W.part_off[l];
        gw.ldc = N;
        gw.z_stride = W.n[l];
        if ((st = run_gemm<EPI_STORE>(ctx, gw, splits))) return st;
        if (l == 0) break;
        // dZ_prev = (dZ W^T) * silu'(z_prev)
        __bf16 *dnext = (dcur == m->dab) ? m->dab2 : m->dab;
        __bf16 *dnextT = (dcur == m->dab) ? m->dab2T : m->dabT;
        GemmArgs gd{};
        gd.A = dcur;
        gd.lda = N;
        gd.B = m->wb + m->w_off[l];   // W [K][N]: op(B)[n_out][k_in] with n_out contiguous
        gd.ldb = N;
        gd.M = rows;
        gd.N = K;
        gd.K = N;
        gd.Z = m->zpre[l - 1];
        gd.ldz = K;
        gd.Cb = dnext;
        gd.ldc = K;
        gd.Cbt = dnextT;
        gd.ldct = m->ldt;
        gd.colpart = m->bias_part + bias_part_base(m, l - 1);
        if ((st = run_gemm<EPI_DSILU>(ctx, gd))) return st;
        dcur = dnext;
        dcurT = dnextT;
    }
    hipLaunchKernelGGL(weight_reduce_kernel, dim3(blocks_for(weight_count(m))), dim3(256), 0, ctx->stream, m->dw_part, W,
                       m->grad);
    BiasMap B{};
    B.layers = m->depth + 1;
    B.tiles = (rows + kTile - 1) / kTile;
    for (int l = 0; l <= m->depth; ++l) {
        B.b_off[l] = (long long)m->b_off[l];
        B.part_off[l] = (long long)bias_part_base(m, l);
        B.fo[l] = m->fo[l];
    }
    hipLaunchKernelGGL(bias_reduce_kernel, dim3(unsigned((m->n_bias + 63) / 64)), dim3(1024), 0, ctx->stream, m->bias_part,
                       B, m->grad);
    BH_HIP(hipGetLastError());
    return BH_OK;
}

__global__ void columns_affine_kernel(float *d, size_t n, int cols, const float *mul, const float *add) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = int(i % cols);
    d[i] = d[i] * mul[c] + add[c];
}

int mlp_in_cols(const bh_mlp *m) { return m->n_in; }
bh_ctx *mlp_context(const bh_mlp *m) { return m->ctx; }
int mlp_out_cols(const bh_mlp *m) { return m->n_out; }

static bool fused_ok(const bh_mlp *m) {
    return m->fused && m->width == kFWidth && m->n_in_pad == 8 && m->depth + 1 <= kFMaxLayers &&
           m->n_out_pad <= 4 * kFWidth && m->n_params * 2 < (size_t(1) << 31);
}

// Fused inference of `rows` rows into y [n_out][ld] (transposed), FP32, or FP16 into yh when it is set.
static int forward_fused(bh_mlp *m, const float *d_x, int64_t rows, float *d_y, long long ld,
                         _Float16 *d_yh = nullptr) {
    FusedArgs F{};
    F.x = d_x;
    F.rows = int(rows);
    F.n_in = m->n_in;
    F.depth = m->depth;
    F.n_out = m->n_out;
    F.n_out_pad = m->n_out_pad;
    F.w32 = m->w32;
    F.wtb = m->wtb;
    F.wt_bytes = unsigned(m->n_params * 2);
    for (int l = 0; l <= m->depth; ++l) {
        F.w_off[l] = (long long)m->w_off[l];
        F.b_off[l] = (long long)m->b_off[l];
    }
    F.y = d_y;
    F.yh = d_yh;
    F.ld = ld;
    hipLaunchKernelGGL(mlp_fused_kernel, dim3(unsigned((rows + kFRows - 1) / kFRows)), dim3(256), 0, m->ctx->stream, F);
    BH_HIP(hipGetLastError());
    return BH_OK;
}

int mlp_forward_enqueue(bh_mlp *m, const float *d_x, int64_t rows, float *d_y, int transposed) {
    if (transposed && fused_ok(m)) return rows > 0 ? forward_fused(m, d_x, rows, d_y, rows) : BH_OK;
    const int chunk = 1 << 18;
    int st;
    if (rows > 0 && (st = ensure_cap(m, int(rows < chunk ? rows : chunk), false))) return st;
    for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
        const int n = int(rows - r0 < chunk ? rows - r0 : chunk);
        float *dst = transposed ? d_y + r0 : d_y + r0 * m->n_out;
        if ((st = forward(m, d_x + r0 * m->n_in, n, false, dst, transposed, transposed ? int(rows) : m->n_out)))
            return st;
    }
    BH_HIP(hipGetLastError());
    return BH_OK;
}

bool mlp_half_outputs(const bh_mlp *m) { return fused_ok(m); }

int mlp_forward_enqueue_f16(bh_mlp *m, const float *d_x, int64_t rows, _Float16 *d_y, long long ld) {
    if (!fused_ok(m)) return set_error(BH_ERR_VALUE, "FP16 network outputs need the fused width-256 inference");
    if (rows <= 0) return BH_OK;
    if (rows > (int64_t(1) << 31) - 1) return set_error(BH_ERR_VALUE, "too many network rows");
    return forward_fused(m, d_x, rows, nullptr, ld, d_y);
}

static int mlp_check(bh_mlp *m) {
    if (!m) return set_error(BH_ERR_VALUE, "NULL MLP handle");
    return ctx_check(m->ctx);
}

}  // namespace bh

extern "C" {

int bh_mlp_create(bh_ctx *ctx, int32_t n_in, int32_t width, int32_t depth, int32_t n_out, bh_mlp **out) {
    clear_error();
    int st = ctx_check(ctx);
    if (st) return st;
    if (!out) return set_error(BH_ERR_VALUE, "NULL output handle");
    if (n_in < 1 || n_in > 1024 || width < 8 || width > 4096 || width % 8 || depth < 1 || depth > 16 || n_out < 1 ||
        n_out > 8192)
        return set_error(BH_ERR_VALUE, "MLP shape out of range (width must be a multiple of 8)");
    bh_mlp *m = new bh_mlp();
    m->ctx = ctx;
    m->n_in = n_in;
    m->n_in_pad = (n_in + 7) / 8 * 8;
    m->width = width;
    m->depth = depth;
    m->n_out = n_out;
    m->n_out_pad = (n_out + 7) / 8 * 8;
    size_t off = 0;
    for (int l = 0; l <= depth; ++l) {
        const int a = l == 0 ? m->n_in_pad : width, b = l == depth ? m->n_out_pad : width;
        m->fi.push_back(a);
        m->fo.push_back(b);
        m->w_off.push_back(off);
        off += size_t(a) * b;
        m->b_off.push_back(off);
        off += b;
        m->fo_prefix.push_back(m->n_bias);
        m->n_bias += b;
    }
    m->n_params = off;
    bool ok = hipMalloc(&m->w32, off * 4) == hipSuccess && hipMalloc(&m->wb, off * 2) == hipSuccess &&
              hipMalloc(&m->wtb, off * 2) == hipSuccess && hipMalloc(&m->grad, off * 4) == hipSuccess &&
              hipMalloc(&m->adam_m, off * 4) == hipSuccess && hipMalloc(&m->adam_v, off * 4) == hipSuccess &&
              hipMalloc(&m->loss, 4) == hipSuccess;
    ok = ok && hipMemsetAsync(m->w32, 0, off * 4, ctx->stream) == hipSuccess &&
         hipMemsetAsync(m->adam_m, 0, off * 4, ctx->stream) == hipSuccess &&
         hipMemsetAsync(m->adam_v, 0, off * 4, ctx->stream) == hipSuccess && refresh_bf16(m) == BH_OK &&
         hipStreamSynchronize(ctx->stream) == hipSuccess;
    if (!ok) {
        bh_mlp_destroy(m);
        return set_error(BH_ERR_MEMORY, "MLP parameter allocation failed");
    }
    *out = m;
    return BH_OK;
}

int bh_mlp_destroy(bh_mlp *m) {
    if (!m) return BH_OK;
    if (m->ctx) {
        (void)hipSetDevice(m->ctx->device);
        (void)hipStreamSynchronize(m->ctx->stream);
    }
    free_acts(m);
    for (void *p : {(void *)m->w32, (void *)m->wb, (void *)m->wtb, (void *)m->grad, (void *)m->adam_m,
                    (void *)m->adam_v, (void *)m->loss})
        if (p) (void)hipFree(p);
    delete m;
    return BH_OK;
}

int bh_mlp_shape(bh_mlp *m, int32_t *n_in_pad, int32_t *n_out_pad, int64_t *n_params) {
    if (!m) return set_error(BH_ERR_VALUE, "NULL MLP handle");
    if (n_in_pad) *n_in_pad = m->n_in_pad;
    if (n_out_pad) *n_out_pad = m->n_out_pad;
    if (n_params) *n_params = int64_t(m->n_params);
    return BH_OK;
}

int bh_mlp_set_params(bh_mlp *m, const float *host) {
    clear_error();
    int st = mlp_check(m);
    if (st) return st;
    if (!host) return set_error(BH_ERR_VALUE, "NULL parameter buffer");
    BH_HIP(hipMemcpyAsync(m->w32, host, m->n_params * 4, hipMemcpyHostToDevice, m->ctx->stream));
    if ((st = refresh_bf16(m))) return st;
    BH_HIP(hipMemsetAsync(m->adam_m, 0, m->n_params * 4, m->ctx->stream));
    BH_HIP(hipMemsetAsync(m->adam_v, 0, m->n_params * 4, m->ctx->stream));
    m->adam_t = 0;
    BH_HIP(hipStreamSynchronize(m->ctx->stream));
    return BH_OK;
}

int bh_mlp_get_params(bh_mlp *m, float *host) {
    clear_error();
    int st = mlp_check(m);
    if (st) return st;
    if (!host) return set_error(BH_ERR_VALUE, "NULL parameter buffer");
    BH_HIP(hipMemcpyAsync(host, m->w32, m->n_params * 4, hipMemcpyDeviceToHost, m->ctx->stream));
    BH_HIP(hipStreamSynchronize(m->ctx->stream));
    return BH_OK;
}

int bh_mlp_get_adam(bh_mlp *m, float *host_m, float *host_v, int64_t *steps) {
    clear_error();
    int st = mlp_check(m);
    if (st) return st;
    if (!host_m || !host_v || !steps) return set_error(BH_ERR_VALUE, "NULL Adam state buffer");
    BH_HIP(hipMemcpyAsync(host_m, m->adam_m, m->n_params * 4, hipMemcpyDeviceToHost, m->ctx->stream));
    BH_HIP(hipMemcpyAsync(host_v, m->adam_v, m->n_params * 4, hipMemcpyDeviceToHost, m->ctx->stream));
    BH_HIP(hipStreamSynchronize(m->ctx->stream));
    *steps = m->adam_t;
    return BH_OK;
}

int bh_mlp_set_adam(bh_mlp *m, const float *host_m, const float *host_v, int64_t steps) {
    clear_error();
    int st = mlp_check(m);
    if (st) return st;
    if (!host_m || !host_v) return set_error(BH_ERR_VALUE, "NULL Adam state buffer");
    if (steps < 0) return set_error(BH_ERR_VALUE, "the Adam step count must be non-negative");
    BH_HIP(hipMemcpyAsync(m->adam_m, host_m, m->n_params * 4, hipMemcpyHostToDevice, m->ctx->stream));
    BH_HIP(hipMemcpyAsync(m->adam_v, host_v, m->n_params * 4, hipMemcpyHostToDevice, m->ctx->stream));
    BH_HIP(hipStreamSynchronize(m->ctx->stream));
    m->adam_t = steps;
    return BH_OK;
}

int bh_mlp_forward(bh_mlp *m, const float *d_x, int64_t rows, float *d_y, int32_t transposed) {
    clear_error();
    int st = mlp_check(m);
    if (st) return st;
    if (!d_x || !d_y || rows < 0 || rows > (int64_t(1) << 30)) return set_error(BH_ERR_VALUE, "bad MLP forward arguments");
    if (rows > 0 && (st = ensure_cap(m, int(rows < (1 << 18) ? rows : (1 << 18)), false))) return st;
    begin_timing(m->ctx);
    if ((st = mlp_forward_enqueue(m, d_x, rows, d_y, transposed))) return st;
    end_timing(m->ctx, rows > 0 ? 1 : 0);
    return BH_OK;
}

int bh_mlp_set_fused(bh_mlp *m, int32_t fused) {
    if (!m) return set_error(BH_ERR_VALUE, "NULL MLP handle");
    m->fused = fused != 0;
    return BH_OK;
}

int bh_columns_affine(bh_ctx *ctx, float *d, int64_t rows, int32_t cols, const float *mul, const float *add) {
    clear_error();
    int st = ctx_check(ctx);
    if (st) return st;
    if (!d || !mul || !add || rows < 0 || cols < 1) return set_error(BH_ERR_VALUE, "bad affine arguments");
    float *dm = nullptr;
    if (hipMalloc(&dm, size_t(cols) * 8) != hipSuccess) return set_error(BH_ERR_MEMORY, "affine table allocation failed");
    hipError_t e = hipMemcpyAsync(dm, mul, size_t(cols) * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dm + cols, add, size_t(cols) * 4, hipMemcpyHostToDevice, ctx->stream);
    const size_t n = size_t(rows) * cols;
    if (e == hipSuccess && n)
        hipLaunchKernelGGL(columns_affine_kernel, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, d, n, cols, dm, dm + cols);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dm);
    if (e != hipSuccess) return set_error(BH_ERR_DEVICE, "affine failed: %s", hipGetErrorString(e));
    return BH_OK;
}

// Gradient of the weighted MSE loss (targets d_t [rows][n_out], per-output weights d_w [n_out]) at the
// current parameters, without an update.  grad_host: n_params values (padded layout) or null.
int bh_mlp_gradient(bh_mlp *m, const float *d_x, const float *d_t, const float *d_w, const float *d_row_w,
                    int32_t masked_cols, int32_t rows, float *grad_host, float *loss_host) {
    clear_error();
    int st = mlp_check(m);
    if (st) return st;
    if (!d_x || !d_t || !d_w || rows < 1) return set_error(BH_ERR_VALUE, "bad MLP gradient arguments");
    if ((st = ensure_cap(m, rows, true))) return st;
    if ((st = forward(m, d_x, rows, true, nullptr, 0, 0))) return st;
    if (masked_cols < 0 || masked_cols > m->n_out) return set_error(BH_ERR_VALUE, "masked_cols must lie in [0, n_out]");
    if ((st = backward(m, rows, d_t, d_w, d_row_w, masked_cols))) return st;
    if (grad_host) BH_HIP(hipMemcpyAsync(grad_host, m->grad, m->n_params * 4, hipMemcpyDeviceToHost, m->ctx->stream));
    if (loss_host) {
        if ((st = loss_total(m, rows))) return st;
        BH_HIP(hipMemcpyAsync(loss_host, m->loss, 4, hipMemcpyDeviceToHost, m->ctx->stream));
    }
    BH_HIP(hipStreamSynchronize(m->ctx->stream));
    return BH_OK;
}

// One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on a batch: forward, weighted-MSE backward,
// update of the FP32 masters and both BF16 copies.  loss_host (optional) gets the batch loss.
int bh_mlp_train_step(bh_mlp *m, const float *d_x, const float *d_t, const float *d_w, const float *d_row_w,
                      int32_t masked_cols, int32_t rows, float lr, float *loss_host) {
    clear_error();
    int st = mlp_check(m);
    if (st) return st;
    if (!d_x || !d_t || !d_w || rows < 1 || !(lr > 0.f)) return set_error(BH_ERR_VALUE, "bad MLP training arguments");
    if ((st = ensure_cap(m, rows, true))) return st;
    begin_timing(m->ctx);
    if ((st = forward(m, d_x, rows, true, nullptr, 0, 0))) return st;
    if (masked_cols < 0 || masked_cols > m->n_out) return set_error(BH_ERR_VALUE, "masked_cols must lie in [0, n_out]");
    if ((st = backward(m, rows, d_t, d_w, d_row_w, masked_cols))) return st;
    m->adam_t += 1;
    const double b1 = 0.9, b2 = 0.999;
    const float c1 = float(1.0 - std::pow(b1, double(m->adam_t))), c2 = float(1.0 - std::pow(b2, double(m->adam_t)));
    LayerMap L{};
    L.layers = m->depth + 1;
    for (int l = 0; l <= m->depth; ++l) {
        L.w_off[l] = (long long)m->w_off[l];
        L.fi[l] = m->fi[l];
        L.fo[l] = m->fo[l];
    }
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(m->n_params)), dim3(256), 0, m->ctx->stream, m->w32, m->grad,
                       m->adam_m, m->adam_v, m->n_params, lr, float(b1), float(b2), 1e-8f, c1, c2, m->wb, m->wtb, L);
    BH_HIP(hipGetLastError());
    end_timing(m->ctx, 1);
    if (loss_host) {
        if ((st = loss_total(m, rows))) return st;
        BH_HIP(hipMemcpyAsync(loss_host, m->loss, 4, hipMemcpyDeviceToHost, m->ctx->stream));
        BH_HIP(hipStreamSynchronize(m->ctx->stream));
    }
    return BH_OK;
}

}  // extern "C"
