// cov_dist.h -- the uint8 axis-DP Hamming distance transform (kernel c).
//
// dist[w] = min over codewords c of d_H(w, c).  Start from D = 0 on codewords
// and 255 elsewhere; for each axis j, every fiber (the q words that differ only
// in digit j) becomes D[x] <- min(D[x], 1 + min over the fiber).  After all n
// axes D is the exact minimum distance (the R-fold dilation of
// cov/verify_independent.py:128-141 reads the same information as
// #{D <= r}).  Each pass streams the N-byte array once in and once out.
//
// Two pass kernels:
//   k_axis_inner  the last `a` axes (stride < q^a) of a contiguous tile of
//                 q^a words, all in LDS: one read and one write of the tile
//                 for a axes
//   k_axis_outer  one axis j with stride s = q^(n-1-j) >= 16: sixteen
//                 consecutive fibers per thread through 16-byte rows
// plus a byte-wise fallback k_axis_any for any stride.
#pragma once
#include "cov_common.h"

namespace cov {

constexpr int DT_TILE = 32768;   // bytes of LDS per inner-axes tile

__global__ void __launch_bounds__(WG) k_dist_fill(uint8_t* D, int64_t N)
{
    const int64_t n16 = N / 16;
    uint4 ff = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n16; i += (int64_t)gridDim.x * WG)
        reinterpret_cast<uint4*>(D)[i] = ff;
    for (int64_t w = n16 * 16 + (int64_t)blockIdx.x * WG + threadIdx.x; w < N; w += (int64_t)gridDim.x * WG)
        D[w] = 0xFF;
}

__global__ void __launch_bounds__(WG) k_dist_seed(const Cell C, uint8_t* D)
{
    for (int i = blockIdx.x * WG + threadIdx.x; i < C.M; i += gridDim.x * WG) D[C.cidx[i]] = 0;
}

// the last `a` axes of tiles of T = q^a contiguous words, in LDS
__global__ void __launch_bounds__(WG) k_axis_inner(uint8_t* D, int64_t N, int q, int a, int64_t T)
{
    __shared__ uint8_t t8[DT_TILE];
    const int64_t per = DT_TILE / T;              // whole tiles per workgroup
    const int64_t span = per * T;
    for (int64_t b0 = (int64_t)blockIdx.x * span; b0 < N; b0 += (int64_t)gridDim.x * span) {
        const int64_t len = (N - b0 < span) ? (N - b0) : span;
        for (int64_t x = threadIdx.x; x < len; x += WG) t8[x] = D[b0 + x];
        __syncthreads();
        int64_t s = 1;
        for (int ax = 0; ax < a; ax++) {
            // fibers of stride s inside each tile: len / q of them
            const int64_t nf = len / q;
            for (int64_t f = threadIdx.x; f < nf; f += WG) {
                const int64_t hi = f / s, lo = f - hi * s;
                const int64_t base = hi * s * q + lo;
                int m = 255;
                for (int v = 0; v < q; v++) { const int x = t8[base + v * s]; m = x < m ? x : m; }
                const int m1 = m < 255 ? m + 1 : 255;
                for (int v = 0; v < q; v++) {
                    const int x = t8[base + v * s];
                    if (m1 < x) t8[base + v * s] = (uint8_t)m1;
                }
            }
            __syncthreads();
            s *= q;
        }
        for (int64_t x = threadIdx.x; x < len; x += WG) D[b0 + x] = t8[x];
        __syncthreads();
    }
}

// one axis of stride s (any), one fiber per thread, byte accesses
__global__ void __launch_bounds__(WG) k_axis_any(uint8_t* D, int64_t N, int q, int64_t s)
{
    const int64_t nf = N / q;
    for (int64_t f = (int64_t)blockIdx.x * WG + threadIdx.x; f < nf; f += (int64_t)gridDim.x * WG) {
        const int64_t hi = f / s, lo = f - hi * s;
        const int64_t base = hi * s * q + lo;
        int m = 255;
        for (int v = 0; v < q; v++) { const int x = D[base + v * s]; m = x < m ? x : m; }
        const int m1 = m < 255 ? m + 1 : 255;
        for (int v = 0; v < q; v++) {
            const int x = D[base + v * s];
            if (m1 < x) D[base + v * s] = (uint8_t)m1;
        }
    }
}

__device__ __forceinline__ uint32_t bytes_min(uint32_t a, uint32_t b)
{
    // per-byte unsigned minimum
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t x = (a >> (8 * k)) & 0xFF, y = (b >> (8 * k)) & 0xFF;
        r |= (x < y ? x : y) << (8 * k);
    }
    return r;
}
__device__ __forceinline__ uint32_t bytes_inc_sat(uint32_t a)
{
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t x = (a >> (8 * k)) & 0xFF;
        r |= (x < 255 ? x + 1 : 255) << (8 * k);
    }
    return r;
}

// one axis of stride s, 16 consecutive fibers per thread when the row of 16
// bytes stays inside one stride segment (s % 16 == 0 not required: rows are
// taken at lo multiples of 16 and the ragged tail of each segment is handled
// byte-wise by the same thread)
__global__ void __launch_bounds__(WG) k_axis_outer(uint8_t* D, int64_t N, int q, int64_t s)
{
    const int64_t rows_per_seg = (s + 15) / 16;       // 16-byte rows per stride segment
    const int64_t nseg = N / (s * q);                  // independent fiber groups
    const int64_t total = nseg * rows_per_seg;
    for (int64_t r = (int64_t)blockIdx.x * WG + threadIdx.x; r < total; r += (int64_t)gridDim.x * WG) {
        const int64_t g = r / rows_per_seg, k = r - g * rows_per_seg;
        const int64_t lo = k * 16;
        const int64_t base = g * s * q + lo;
        const int64_t nb = (s - lo < 16) ? (s - lo) : 16;
        if (nb == 16 && ((base & 3) == 0) && ((s & 3) == 0)) {
            uint4 m = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
            for (int v = 0; v < q; v++) {
                const uint32_t* p = reinterpret_cast<const uint32_t*>(D + base + v * s);
                m.x = bytes_min(m.x, p[0]); m.y = bytes_min(m.y, p[1]);
                m.z = bytes_min(m.z, p[2]); m.w = bytes_min(m.w, p[3]);
            }
            m.x = bytes_inc_sat(m.x); m.y = bytes_inc_sat(m.y);
            m.z = bytes_inc_sat(m.z); m.w = bytes_inc_sat(m.w);
            for (int v = 0; v < q; v++) {
                uint32_t* p = reinterpret_cast<uint32_t*>(D + base + v * s);
                p[0] = bytes_min(m.x, p[0]); p[1] = bytes_min(m.y, p[1]);
                p[2] = bytes_min(m.z, p[2]); p[3] = bytes_min(m.w, p[3]);
            }
        } else {
            for (int64_t b = 0; b < nb; b++) {
                int mm = 255;
                for (int v = 0; v < q; v++) { const int x = D[base + b + v * s]; mm = x < mm ? x : mm; }
                const int m1 = mm < 255 ? mm + 1 : 255;
                for (int v = 0; v < q; v++) {
                    const int x = D[base + b + v * s];
                    if (m1 < x) D[base + b + v * s] = (uint8_t)m1;
                }
            }
        }
    }
}

// histogram of D over 0..n (n+1 bins, values above n land in bin n+1)
__global__ void __launch_bounds__(WG) k_dist_hist(const uint8_t* __restrict__ D, int64_t N, int n,
                                                  unsigned long long* hist)
{
    __shared__ unsigned int h[MAXN + 2];
    if (threadIdx.x < MAXN + 2) h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t n16 = N / 16;
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n16; i += (int64_t)gridDim.x * WG) {
        const uint4 v = reinterpret_cast<const uint4*>(D)[i];
        const uint32_t e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int b = 0; b < 4; b++) {
                int x = (e[k] >> (8 * b)) & 0xFF;
                atomicAdd(&h[x <= n ? x : n + 1], 1u);
            }
    }
    for (int64_t w = n16 * 16 + (int64_t)blockIdx.x * WG + threadIdx.x; w < N; w += (int64_t)gridDim.x * WG) {
        int x = D[w];
        atomicAdd(&h[x <= n ? x : n + 1], 1u);
    }
    __syncthreads();
    if (threadIdx.x < n + 2 && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

}  // namespace cov
