// Internal declarations shared by the bhgrrt HIP sources (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bhgrrt.h"

namespace bh {

// dword 3 of a raw buffer resource on gfx950 (raw access, no format conversion); loads past
// num_records return zero
constexpr unsigned kRsrcWord3 = 0x00020000u;
constexpr int kCoeffs = 11;
constexpr int kStokes = 4;
constexpr double kTwoPi = 2.0 * 3.14159265358979323846;
constexpr double kEpsTheta = 1.0e-7;   // kerr.py:27

// Thread-local last-error text (bh_last_error).
int set_error(int code, const char *fmt, ...);
void clear_error();

#define BH_HIP(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::bh::set_error(BH_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

struct GridHint {      // index guess u = scale * f(x) + offset, f = log2 or identity
    float scale;
    float offset;
    int log;
    float first, last;   // g[0] and g[n - 1], kept in registers for the range tests
};

struct BrickDev {
    bool live = false;
    int storage = BH_STORE_F32;
    int layout = BH_LAYOUT_SOA;
    int nr = 0, nt = 0, np = 0;
    size_t plane = 0;          // nr * nt * np elements
    size_t bytes = 0;          // device bytes of the coefficient array
    void *coeffs = nullptr;    // SoA [11][nr][nt][np], or corner cells [nr-1][nt-1][np][8][11]
    float *grids = nullptr;    // r | theta | phi, float32
    std::vector<float> hgrids; // host copy of the grids
    GridHint hr{}, ht{}, hp{};
};

constexpr int kMaxQueues = 8;
constexpr int kCounterWords = 64;
constexpr int kWorkOffset = 16;   // two uint64 work counters after the queue counters

struct Tuning {
    int block_threads = 256;
    int blocks_per_cu = 0;
    int steps_per_sync = 48;   // measured best for the 1024^2 renders (32..64 within noise; 8 is 8 % slower)
    int refill_min = 64;
    int queues = 2;            // measured: 2 queues beat 8 (one per XCD) by 3-5 % on single renders
    int order = 1;        // 1: tiles nearest the critical curve (longest rays) first
    int coop_steps = 32;  // lone steps before a lone ray asks an idle wave for help (0: never)
};

}  // namespace bh

struct bh_ctx {
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    int timed_launches = 0;
    unsigned int *d_counters = nullptr;   // ray-queue counters, kCounterWords words
    const unsigned int *last_counters = nullptr;   // counter block of the last render launch (null: d_counters)
    std::vector<bh::BrickDev> bricks;
    bh::Tuning tuning[3];
    int *d_order = nullptr;             // tile schedule of the last scene (device)
    size_t order_cap = 0;
    std::vector<double> order_key;      // scene parameters the schedule was built for
    unsigned *sur_count = nullptr;      // fallback counter of the last surrogate render (in its work buffer)
    void *d_synchro = nullptr;          // coefficient-precompute tables (grow-only)
    void *d_sampler_work = nullptr;     // large-N sampler: cell counts, point keys, ranks, order (grow-only)
    size_t sampler_work_cap = 0;
    size_t synchro_cap = 0;
    std::vector<double> h_synchro;      // host copy the table upload reads from
};

struct bh_mlp;

namespace bh {
int ctx_check(bh_ctx *ctx);
// Records ev_start on the context stream; end_timing records ev_stop.
void begin_timing(bh_ctx *ctx);
void end_timing(bh_ctx *ctx, int launches);
// Fill n words of device memory with v on the context stream (graph-safe, no memset node).
int fill_u32(bh_ctx *ctx, unsigned int *dst, unsigned int v, int n);
__global__ void fill_u32_kernel(unsigned int *dst, unsigned int v, int n);
// Process-wide context on the current device used by the host-buffer entry points.
bh_ctx *default_ctx(int *status);
std::mutex &default_ctx_mutex();
// Grow-only device staging buffer owned by the default context.
int staging(size_t bytes, void **ptr);
// Grow-only pinned, coherent host buffer that kernels address directly (host and device addresses).
int host_staging(size_t bytes, void **host, void **dev);
// MLP inference enqueued on the MLP's context stream, without timing events:
// d_x [rows][n_in] -> d_y [rows][n_out] or [n_out][rows] (transposed).
int mlp_forward_enqueue(bh_mlp *m, const float *d_x, int64_t rows, float *d_y, int transposed);
// Fused transposed inference with FP16 outputs y [n_out][ld] (the width-256 networks of the surrogate).
int mlp_forward_enqueue_f16(bh_mlp *m, const float *d_x, int64_t rows, _Float16 *d_y, long long ld);
// Whether mlp_forward_enqueue_f16 serves this network.
bool mlp_half_outputs(const bh_mlp *m);
int mlp_in_cols(const bh_mlp *m);
bh_ctx *mlp_context(const bh_mlp *m);
int mlp_out_cols(const bh_mlp *m);
}  // namespace bh
