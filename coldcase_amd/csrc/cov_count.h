// cov_count.h -- the distance-count transform over Z_q^n and the maps built
// on it (gain and loss of the LNS engine).
//
// For a multiset S of words, N_d(x) = #{u in S : d_H(x, u) = d} for every
// word x and d = 0..R.  Layer d holds A_d; the DP runs axis by axis:
// processing axis j along a fiber of q words (stride q^j) splits the words u
// of S on whether u_j equals the fiber position a:
//     A_d[a] <- A_d[a] + T_{d-1} - A_{d-1}[a],   T_d = sum over the fiber of A_d
// (A_{-1} = 0), and after all n axes A_d = N_d exactly (the DP of
// cov/gemm2/dct.cu:1-11).  The order of the axes does not matter: every
// axis step is the same linear map on its own coordinate.
//
// Layout: (R+1) int32 layers of N words, layer-major, MSB-first word index.
// Passes over HBM, each a grid of LDS tiles that hold all R+1 layers of
// their words and apply several axes in LDS:
//   the inner pass   q^a consecutive words (the a fastest axes), reading
//                    only the source (the code's multiplicities, or cnt ==
//                    target) and writing every layer;
//   outer passes     q^b rows (axes j .. j+b-1, rows q^j apart) of W
//                    consecutive words, every layer read and written once;
//   the last pass    writes either every layer or only sum_d A_d into layer
//                    0 (the map: the gain when S = [cnt == 0], the loss when
//                    S = [cnt == 1]).
// Tiles never overlap, so each pass works in place.
#pragma once
#include "cov_common.h"

namespace cov {

constexpr int CT_SRC_LAYER = 0;     // layer 0 already holds the multiplicities
constexpr int CT_SRC_CNT = 1;       // layer 0 starts as [cnt == target]
constexpr int CT_OUT_LAYERS = 0;
constexpr int CT_OUT_SUM = 1;
constexpr int CT_NT = 512;          // threads per tile workgroup
constexpr int CT_LOAD_BATCH = 8;    // loads issued per thread before the LDS stores

struct CtPass {
    int64_t N;
    int q, R1;            // alphabet, layers (R + 1)
    int k;                // axes applied in the tile
    int wsh;              // W = 1 << wsh columns per tile
    int64_t rows;         // q^k
    int64_t rs;           // global distance between rows (q^j)
    int64_t span;         // q^(j+k): global words of one row group
    int64_t ncb;          // column blocks per row group: ceil(q^j / W)
    int64_t ntiles;       // (N / span) * ncb
    int first, last;      // the pass reads the source / writes the output
    int src, out, target;
};

// One workgroup per tile (grid-stride over tiles), CT_NT threads, dynamic
// LDS of R1 * rows * W int32.
__global__ void __launch_bounds__(CT_NT) k_ct_pass(CtPass p, int32_t* __restrict__ A,
                                                  const uint16_t* __restrict__ cnt)
{
    extern __shared__ int32_t L[];
    const int W = 1 << p.wsh;
    const int64_t T = p.rows << p.wsh;
    const int64_t colspan = p.rs;                       // q^j words per row segment (1 for the inner pass)
    for (int64_t tile = blockIdx.x; tile < p.ntiles; tile += gridDim.x) {
        const int64_t hi = tile / p.ncb;
        const int64_t lo0 = (tile - hi * p.ncb) << p.wsh;
        const int64_t base = hi * p.span + lo0;
        const int64_t wcols = colspan - lo0 < W ? colspan - lo0 : W;
        // load: every layer (or the source into layer 0 and zeros above),
        // CT_LOAD_BATCH loads in flight per thread
        for (int64_t e0 = threadIdx.x; e0 < T; e0 += (int64_t)CT_NT * CT_LOAD_BATCH) {
            for (int d = 0; d < (p.first ? 1 : p.R1); d++) {
                int32_t v[CT_LOAD_BATCH];
#pragma unroll
                for (int b = 0; b < CT_LOAD_BATCH; b++) {
                    const int64_t e = e0 + (int64_t)b * CT_NT;
                    v[b] = 0;
                    if (e < T) {
                        const int64_t r = e >> p.wsh, c = e & (W - 1);
                        if (c < wcols) {
                            const int64_t x = base + r * p.rs + c;
                            if (p.first && p.src == CT_SRC_CNT) v[b] = cnt[x] == (uint16_t)p.target;
                            else v[b] = A[(int64_t)d * p.N + x];
                        }
                    }
                }
#pragma unroll
                for (int b = 0; b < CT_LOAD_BATCH; b++) {
                    const int64_t e = e0 + (int64_t)b * CT_NT;
                    if (e < T) L[d * T + e] = v[b];
                }
            }
            if (p.first)
                for (int d = 1; d < p.R1; d++)
#pragma unroll
                    for (int b = 0; b < CT_LOAD_BATCH; b++) {
                        const int64_t e = e0 + (int64_t)b * CT_NT;
                        if (e < T) L[d * T + e] = 0;
                    }
        }
        __syncthreads();
        // the tile's axes, one after the other
        int64_t qi = 1;                                  // q^i: row distance of axis i
        for (int i = 0; i < p.k; i++) {
            const int64_t nfib = T / p.q;
            for (int64_t f = threadIdx.x; f < nfib; f += CT_NT) {
                const int64_t c = f & (W - 1);
                const int64_t fr = f >> p.wsh;           // fiber row index with digit i removed
                const int64_t rl = fr % qi, rh = fr / qi;
                const int64_t e0 = ((rh * p.q) * qi + rl) * W + c;
                const int64_t es = qi * W;
                int32_t t[MAXR + 1];
                for (int d = 0; d < p.R1 - 1; d++) {
                    int32_t s = 0;
                    for (int a = 0; a < p.q; a++) s += L[d * T + e0 + a * es];
                    t[d] = s;
                }
                for (int a = 0; a < p.q; a++) {
                    const int64_t e = e0 + a * es;
                    for (int d = p.R1 - 1; d >= 1; d--) L[d * T + e] += t[d - 1] - L[(d - 1) * T + e];
                }
            }
            qi *= p.q;
            __syncthreads();
        }
        // store
        for (int64_t e = threadIdx.x; e < T; e += CT_NT) {
            const int64_t r = e >> p.wsh, c = e & (W - 1);
            if (c >= wcols) continue;
            const int64_t x = base + r * p.rs + c;
            if (p.last && p.out == CT_OUT_SUM) {
                int32_t s = 0;
                for (int d = 0; d < p.R1; d++) s += L[d * T + e];
                A[x] = s;
            } else {
                for (int d = 0; d < p.R1; d++) A[(int64_t)d * p.N + x] = L[d * T + e];
            }
        }
        __syncthreads();
    }
}

// layer 0 := the multiplicity of every word in the code (one atomic per codeword)
__global__ void __launch_bounds__(WG) k_ct_scatter(const Cell C, int32_t* __restrict__ A0)
{
    for (int i = blockIdx.x * WG + threadIdx.x; i < C.M; i += gridDim.x * WG) atomicAdd(&A0[C.cidx[i]], 1);
}

// ------------------------------------------------------------- the map
// The map is layer 0 after a sum transform.

// max over the map (non-negative values): one atomicMax per workgroup
__global__ void __launch_bounds__(WG) k_map_max(const Cell C, const int32_t* __restrict__ A0, int64_t* out)
{
    __shared__ int32_t red[WAVES];
    int32_t m = 0;
    for (int64_t x = (int64_t)blockIdx.x * WG + threadIdx.x; x < C.N; x += (int64_t)gridDim.x * WG)
        m = max(m, A0[x]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < WAVES; w++) m = max(m, red[w]);
        atomicMax(reinterpret_cast<unsigned long long*>(out), (unsigned long long)m);
    }
}

// histogram of the positive values: bin floor((v - 1) * nbins / vmax), the
// last bin taking everything above (bins in LDS, one global add per bin and
// workgroup)
__global__ void __launch_bounds__(WG) k_map_hist(const Cell C, const int32_t* __restrict__ A0, int nbins,
                                                 int32_t vmax, unsigned long long* hist)
{
    extern __shared__ uint32_t hb[];
    for (int b = threadIdx.x; b < nbins; b += WG) hb[b] = 0;
    __syncthreads();
    for (int64_t x = (int64_t)blockIdx.x * WG + threadIdx.x; x < C.N; x += (int64_t)gridDim.x * WG) {
        const int32_t v = A0[x];
        if (v > 0) {
            int64_t b = ((int64_t)(v - 1) * nbins) / vmax;
            if (b >= nbins) b = nbins - 1;
            atomicAdd(&hb[b], 1u);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += WG)
        if (hb[b]) atomicAdd(&hist[b], (unsigned long long)hb[b]);
}

__global__ void __launch_bounds__(WG) k_map_read(const int32_t* __restrict__ A0, const int64_t* __restrict__ idx,
                                                 int k, int32_t* out)
{
    for (int i = blockIdx.x * WG + threadIdx.x; i < k; i += gridDim.x * WG) out[i] = A0[idx[i]];
}

// selects the words whose map value is at least thr (the ordered gather's
// selector)
struct SelMapGe {
    const int32_t* map;
    int32_t thr;
    __device__ __forceinline__ bool operator()(int64_t w) const { return map[w] >= thr; }
};

// per centre word: #{x in its radius-R ball : cnt[x] == target}, one
// workgroup per centre (the exact gain (target 0) or loss (target 1) of a
// single codeword at that word, cov/gemm2/dct.cu:200-221)
__global__ void __launch_bounds__(WG) k_ball_gather(const Cell C, const int64_t* __restrict__ centers, int k,
                                                    int target, int32_t* out)
{
    __shared__ WgRed red;
    for (int i = blockIdx.x; i < k; i += gridDim.x) {
        const int64_t w = centers[i];
        const uint64_t pk = pack_index(C, w);
        int64_t acc = 0;
        for (int r = 0; r <= C.R; r++) {
            const int64_t ns = C.bsub_off[r + 1] - C.bsub_off[r];
            int64_t P = 1;
            for (int s = 0; s < r; s++) P *= C.q1;
            const int64_t tot = ns * P;
            for (int64_t t = threadIdx.x; t < tot; t += WG) {
                const int64_t si = t / P;
                const uint32_t sub = C.bsub[C.bsub_off[r] + si];
                const int64_t x = pattern_word(C, w, pk, sub, r, (uint32_t)(t - si * P), nullptr);
                acc += cnt_get(C, x) == (uint16_t)target;
            }
        }
        const int64_t s = wg_sum(acc, red);
        if (threadIdx.x == 0) out[i] = (int32_t)s;
        __syncthreads();
    }
}

}  // namespace cov
