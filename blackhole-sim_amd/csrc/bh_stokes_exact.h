// Exact constant-coefficient Stokes step (polarized_transfer.py:51-70), host and device code.
#pragma once

#include <limits>

#include "bh_internal.h"

namespace bh {

template <typename R>
__host__ __device__ inline R abs_hd(R x) { return x < R(0) ? -x : x; }

// The step's arithmetic is written with explicit fused multiply-adds and built with contraction off, so it
// rounds the same wherever it is inlined: the kernels call it from several code paths (a ray's own lane,
// a lone ray's wave, the consumer of a split lone ray) that must give the same image.
template <typename R>
__host__ __device__ inline R fma_hd(R a, R b, R c) {
    if constexpr (sizeof(R) == 4) return __builtin_fmaf(a, b, c);
    else return __builtin_fma(a, b, c);
}
// a0 b0 + a1 b1 + a2 b2 + a3 b3 + c, fused left to right
template <typename R>
__host__ __device__ inline R dot4_hd(R a0, R b0, R a1, R b1, R a2, R b2, R a3, R b3, R c) {
    return fma_hd(a3, b3, fma_hd(a2, b2, fma_hd(a1, b1, fma_hd(a0, b0, c))));
}

// Exact constant-coefficient step of dS/ds = j - K S over ds (polarized_transfer.py:51-70, which takes
// scipy's expm of the augmented matrix [[-K, j], [0, 0]] ds and applies it to [S; 1]).  With
// A = -K h and b = j h for h = ds / 2^p (p halvings bring the infinity norm of A to <= 1/2),
// exp([[A, b], [0, 0]]) = [[I + A F, F b], [0, 1]] where F = sum_k A^k / (k+1)!, truncated at order
// TAYLOR (FP64 14, FP32 8: remainder below the rounding unit at norm 1/2) and evaluated by Horner.
// Up to kExactSubsteps halvings the 2^p substeps S <- S + F (A S + b) are applied as matrix-vector
// products (a 4x4 matrix and two 4-vectors live); beyond that (huge ds: optically thick cells,
// pole-clamped rays) the block exponential (P, w) is formed once and squared p times,
// (P, w) -> (P P, P w + w).  A non-finite alpha ds (an infinite ds or coefficient, or NaN) gives NaN at
// once, as scipy's expm does for the reference (polarized_transfer.py:66); a NaN ds passes the ds <= 0
// test like the reference's and ends there too.  Only the kernels built for the exact rule contain this
// code; it is host and device code so that the CPU tests can run it.
constexpr int kExactSubsteps = 3;

template <typename R>
__host__ __device__ inline void exact_operator(const R c[kCoeffs], R ds, R A[4][4], R b[4], int &p) {
#pragma clang fp contract(off)
    const R ai = c[4], aq = c[5], au = c[6], av = c[7], rv = c[8], rq = c[9], ru = c[10];
    const R K[4][4] = {{ai, aq, au, av}, {aq, ai, rv, -ru}, {au, -rv, ai, rq}, {av, ru, -rq, ai}};
    R norm = R(0);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const R row = abs_hd(K[i][0]) + abs_hd(K[i][1]) + abs_hd(K[i][2]) + abs_hd(K[i][3]);
        norm = row > norm ? row : norm;
    }
    norm *= ds;
    p = 0;
    if (!(norm <= std::numeric_limits<R>::max())) {   // infinite or NaN: scipy's expm returns NaN there
        p = -1;
        return;
    }
    R h = ds;
    while (norm > R(0.5)) {   // at most ~1100 halvings: the finite norm reaches 1/2 before h underflows
        norm *= R(0.5);
        h *= R(0.5);
        ++p;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) A[i][j] = -K[i][j] * h;
        b[i] = c[i] * h;
    }
}

template <typename R>
__host__ __device__ inline void stokes_exact_squaring(R s[4], const R c[kCoeffs], R ds) {
#pragma clang fp contract(off)
    constexpr int TAYLOR = sizeof(R) == 8 ? 14 : 8;
    R A[4][4], b[4];
    int p;
    exact_operator(c, ds, A, b, p);
    if (p < 0) {
        for (int i = 0; i < 4; ++i) s[i] = std::numeric_limits<R>::quiet_NaN();
        return;
    }
    R F[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) F[i][j] = R(i == j);
    for (int k = TAYLOR + 1; k >= 2; --k) {   // F = I + A F / k
        const R inv = R(1) / R(k);
        R G[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                G[i][j] = fma_hd(dot4_hd(A[i][0], F[0][j], A[i][1], F[1][j], A[i][2], F[2][j], A[i][3], F[3][j], R(0)), inv,
                                 R(i == j));
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) F[i][j] = G[i][j];
    }
    R P[4][4], w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            P[i][j] = dot4_hd(A[i][0], F[0][j], A[i][1], F[1][j], A[i][2], F[2][j], A[i][3], F[3][j], R(i == j));
        w[i] = dot4_hd(F[i][0], b[0], F[i][1], b[1], F[i][2], b[2], F[i][3], b[3], R(0));
    }
    for (int q = 0; q < p; ++q) {   // (P, w) <- (P P, P w + w)
        R P2[4][4], w2[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                P2[i][j] = dot4_hd(P[i][0], P[0][j], P[i][1], P[1][j], P[i][2], P[2][j], P[i][3], P[3][j], R(0));
            w2[i] = dot4_hd(P[i][0], w[0], P[i][1], w[1], P[i][2], w[2], P[i][3], w[3], w[i]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int j = 0; j < 4; ++j) P[i][j] = P2[i][j];
            w[i] = w2[i];
        }
    }
    R out[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) out[i] = dot4_hd(P[i][0], s[0], P[i][1], s[1], P[i][2], s[2], P[i][3], s[3], w[i]);
#pragma unroll
    for (int i = 0; i < 4; ++i) s[i] = out[i];
}

template <typename R>
__host__ __device__ inline void stokes_exact(R s[4], const R c[kCoeffs], R ds) {
#pragma clang fp contract(off)
    if (ds <= R(0)) return;   // the reference returns S unchanged (a NaN ds falls through)
    constexpr int TAYLOR = sizeof(R) == 8 ? 14 : 8;
    R A[4][4], b[4];
    int p;
    exact_operator(c, ds, A, b, p);
    if (p < 0) {   // a non-finite alpha ds: NaN, as the reference's expm gives
        for (int i = 0; i < 4; ++i) s[i] = std::numeric_limits<R>::quiet_NaN();
        return;
    }
    if (p > kExactSubsteps) {
        stokes_exact_squaring(s, c, ds);
        return;
    }
    for (int n = 1 << p; n > 0; --n) {
        R v[4], t[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            v[i] = dot4_hd(A[i][0], s[0], A[i][1], s[1], A[i][2], s[2], A[i][3], s[3], b[i]);
            t[i] = v[i];
        }
#pragma unroll
        for (int k = TAYLOR + 1; k >= 2; --k) {   // t = v + A t / k
            const R inv = R(1) / R(k);
            R u[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                u[i] = fma_hd(dot4_hd(A[i][0], t[0], A[i][1], t[1], A[i][2], t[2], A[i][3], t[3], R(0)), inv, v[i]);
#pragma unroll
            for (int i = 0; i < 4; ++i) t[i] = u[i];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) s[i] += t[i];
    }
}

}  // namespace bh
