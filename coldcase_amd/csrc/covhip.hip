// covhip.hip -- host side of libcovhip.so: cell handles, tables, launches and
// the C ABI declared in include/covhip.h.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "covhip.h"
#include "cov_common.h"
#include "cov_cover.h"
#include "cov_dist.h"
#include "cov_batch.h"
#include "cov_step.h"
#include "cov_count.h"
#include "cov_plan.h"

using namespace cov;

// The allocation-free gather's scratch (cov_gather_uncovered_noalloc): NA_BLK
// block counts, the total on its own 256-byte line, NA_CAP output indices.
constexpr int64_t NA_BLK = 16384;
constexpr int64_t NA_CAP = 65536;
constexpr int64_t NA_WORDS = NA_BLK + 32 + NA_CAP;
static __device__ __align__(256) int64_t scratch[NA_WORDS];

struct cov_cell {
    Cell h{};                 // host copy of the device descriptor
    Cell* d = nullptr;        // device descriptor
    RecountPlan plan{};
    hipStream_t stream = nullptr;
    std::string err;
    int eval_groups = 0, commit_groups = 0;
    int64_t mem_bytes = 0;
    std::vector<void*> allocs;
    int64_t* d_small = nullptr;           // 64 int64 of device scratch
    // graph of GS steps
    hipGraphExec_t gexec = nullptr;
    int g_steps = 0, g_eg = 0, g_cg = 0;
    // timing
    bool timing = false;
    double ms[7] = {0, 0, 0, 0, 0, 0, 0};
    int64_t launches[7] = {0, 0, 0, 0, 0, 0, 0};
    std::vector<hipEvent_t> evs;          // event pool
    struct Pend { int kind; hipEvent_t a, b; };
    std::vector<Pend> pend;               // recorded, not yet read
    int64_t* trace = nullptr;             // device trace buffer
    int64_t tcap = 0;
    // distance transform
    uint32_t* cwb = nullptr;              // codeword bitmap, all zero between calls
    int64_t* dtstat = nullptr;            // the transform's statistics copies (DT_COPIES x DT_CSTRIDE)
    std::vector<int64_t> dtcopies;        // their host image
    int dt_lds = 0;                       // LDS bytes per transform tile (0: default)
    int dt_threads = 0;                   // threads per transform workgroup (0: default)
    bool dist_valid = false;              // dist holds the transform of the current code
    bool cnt_valid = true;                // cnt, planes and counts match the code
    bool th_valid = false;                // the coverer XORs and leaving counts match cnt
    int kick = 2;                         // the search's kicks per stagnation (as in the control block)
    int leave_mode = COV_LEAVE_AUTO;      // how the steps find the leaving side (cov_set_leave_mode)
    int th_on = 0;                        // the control block's th_on as last set from here
    double clk_hz = 1e8;                  // device wall clock rate (hipDeviceAttributeWallClockRate)
    int ncu = 256;                        // compute units (hipDeviceAttributeMultiprocessorCount)
    // count transform: (R+1) int32 layers of N words (allocated on first
    // use); layer 0 holds the map after a sum transform
    int32_t* ct = nullptr;
    // remove-and-repair descent: centres, their unique counts and the chosen
    // word (2 * cap + 1 int64), grown on demand
    int64_t* rr = nullptr;
    int64_t rr_cap = 0;
    int ct_lds = 0;                       // LDS bytes per transform tile (0: the default)
    int map_kind = -1;                    // source of the map in layer 0, -1: none
    // recorded on a batch's stream after each batched launch that includes
    // the cell: the cell's own stream waits for it before touching the cell,
    // and its memory is not freed before it
    hipEvent_t used = nullptr;
};

// Many cells stepped by the same launches (cov_batch.h).
struct cov_batch {
    std::vector<cov_cell*> cells;
    hipStream_t stream = nullptr;
    Cell* d_cells = nullptr;              // the cells' descriptors, in device memory
    int32_t* d_eoff = nullptr;            // evaluator workgroup ranges, ncell + 1
    int32_t* d_coff = nullptr;            // commit workgroup ranges, ncell + 1
    int eval_total = 0, commit_total = 0;
    hipGraphExec_t gexec = nullptr;
    std::string err;
    std::vector<const void*> accw;        // each cell's wide counts when the batch was made
};

namespace {

constexpr int GRAPH_STEPS = COV_GRAPH_STEPS;

#define HCK(h, x)                                                              \
    do {                                                                       \
        hipError_t e_ = (x);                                                   \
        if (e_ != hipSuccess) {                                                \
            if (h) (h)->err = std::string(#x) + ": " + hipGetErrorString(e_);  \
            return -(100 + (int)e_);                                           \
        }                                                                      \
    } while (0)

// device bytes of each array, in allocation order
struct Sizes {
    int64_t cnt, plane, blk, code, cpk, cidx, tabu, fsub, psub, bsub, ssub, ul_idx, ul_pk,
        shuf, acc, slot_cw, slot_cpk, rbuf, ctl, cell, small, arrive, dirty, jump, xor32, th, cst, cbs, cps, csr, dth;
};

Sizes sizes_of(const Shape& s)
{
    Sizes z{};
    z.cnt = ((s.N + 1) / 2 + 16) * 4;
    z.plane = (s.nzw + 16) * 4;
    z.blk = (s.nblk + 16) * 4;
    z.code = (int64_t)s.M * s.n + 64;
    z.cpk = (int64_t)s.M * 8;
    z.cidx = (int64_t)s.M * 8;
    z.tabu = (int64_t)s.M * s.n * s.q * 4;
    z.fsub = (int64_t)s.F * 4 + 4;
    z.psub = (int64_t)s.n * s.E * 4 + 4;
    z.bsub = ((int64_t)1 << s.n) * 4 + 4;
    z.ssub = ((int64_t)1 << s.n) * 4 + 4;
    z.ul_idx = s.ulcap * 8;
    z.ul_pk = s.ulcap * 8;
    z.shuf = (int64_t)s.M * s.n * 4 + 4;
    z.acc = (int64_t)s.M * ACCW * 4;
    z.slot_cw = (int64_t)s.M * 4;
    z.slot_cpk = (int64_t)s.M * 8;
    z.rbuf = (int64_t)RB * 8;
    z.ctl = sizeof(Ctl);
    z.cell = sizeof(Cell);
    z.small = 64 * 8;
    z.arrive = ARR_G * ARR_STRIDE * 4;
    z.dirty = DIRTY_CAP * 4;
    z.jump = RNG_JUMP_WORDS * 8;
    z.xor32 = ((s.N + 1) / 2 + 16) * 4;
    z.th = (int64_t)s.M * THW * 4;
    z.cst = SEL_WAVE * 16;
    z.cbs = SEL_WAVE * 8;
    z.cps = SEL_WAVE * 8;
    z.csr = (SEL_WAVE + 1) * 8;
    z.dth = (int64_t)s.M * THW * 4;
    return z;
}

int64_t total_bytes(const Shape& s)
{
    const Sizes z = sizes_of(s);
    // plus the distance array and codeword bitmap the transform allocates on first use
    return z.cnt + 3 * z.plane + z.blk + 2 * z.code + z.cpk + z.cidx + z.tabu + z.fsub + z.psub +
           z.bsub + z.ssub + z.ul_idx + z.ul_pk + z.shuf + z.acc + z.slot_cw + z.slot_cpk + z.rbuf +
           z.ctl + z.cell + z.small + z.arrive + z.dirty + z.jump + z.xor32 + z.th + z.cst + z.cbs + z.cps + z.csr + z.dth + (s.N + 64) +
           ((s.N >> 5) + 2) * 4;
}

template <typename T>
int dalloc(cov_cell* h, T** p, int64_t bytes)
{
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, (size_t)bytes);
    if (e != hipSuccess) {
        h->err = std::string("hipMalloc: ") + hipGetErrorString(e);
        return COV_ENOMEM;
    }
    // zeroed on the cell's own stream: a fill on the null stream would not be
    // ordered before the cell's later work on its non-blocking stream
    e = hipMemsetAsync(v, 0, (size_t)bytes, h->stream);
    if (e != hipSuccess) {
        (void)hipFree(v);
        h->err = std::string("hipMemsetAsync: ") + hipGetErrorString(e);
        return -(100 + (int)e);
    }
    h->allocs.push_back(v);
    h->mem_bytes += bytes;
    *p = (T*)v;
    return COV_OK;
}

// host -> device on the cell's stream, complete on return: callers pass stack
// values and vectors that die with their scope, and pageable sources are not
// guaranteed to be staged by the time an asynchronous copy returns (set-up
// paths only; the step loop uploads nothing)
template <typename T>
int upload(cov_cell* h, T* dst, const void* src, int64_t bytes)
{
    HCK(h, hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

#define CTL_PTR(h, field) (reinterpret_cast<char*>((h)->h.ctl) + offsetof(Ctl, field))

template <typename T>
int set_ctl(cov_cell* h, size_t off, T v)
{
    HCK(h, hipMemcpyAsync(reinterpret_cast<char*>(h->h.ctl) + off, &v, sizeof(T), hipMemcpyHostToDevice,
                          h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

int read_ctl(cov_cell* h, Ctl* out)
{
    if (h->used) HCK(h, hipStreamWaitEvent(h->stream, h->used, 0));
    HCK(h, hipMemcpyAsync(out, h->h.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

int push_cell(cov_cell* h)
{
    HCK(h, hipMemcpyAsync(h->d, &h->h, sizeof(Cell), hipMemcpyHostToDevice, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

int grid_for(int64_t items, int per_wg, int lo, int hi)
{
    int64_t g = (items + per_wg - 1) / per_wg;
    if (g < lo) g = lo;
    if (g > hi) g = hi;
    return (int)g;
}

void splitmix_seed(uint64_t seed, uint64_t s[4])
{
    for (int i = 0; i < 4; i++) {
        seed += 0x9E3779B97F4A7C15ULL;
        uint64_t z = seed;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        s[i] = z ^ (z >> 31);
    }
}

#define RC(x)                          \
    do {                               \
        int rc_ = (x);                 \
        if (rc_ != COV_OK) return rc_; \
    } while (0)

// Per-launch device time: an event before and after the launch on the cell's
// stream, read back in batches (no synchronisation between launches).
int timing_collect(cov_cell* h)
{
    if (h->pend.empty()) return COV_OK;
    HCK(h, hipEventSynchronize(h->pend.back().b));
    for (const auto& p : h->pend) {
        float ms = 0;
        HCK(h, hipEventElapsedTime(&ms, p.a, p.b));
        h->ms[p.kind] += ms;
        h->launches[p.kind] += 1;
        h->evs.push_back(p.a);
        h->evs.push_back(p.b);
    }
    h->pend.clear();
    return COV_OK;
}

int pool_event(cov_cell* h, hipEvent_t* e)
{
    if (h->evs.empty()) {
        HCK(h, hipEventCreate(e));
    } else {
        *e = h->evs.back();
        h->evs.pop_back();
    }
    return COV_OK;
}

// Launch on the cell's stream.  With timing on, the launch records a start
// and a stop event around the kernel's own execution (hipExtLaunchKernelGGL),
// so a kind accumulates kernel durations as a kernel trace reports them,
// without the gaps of separately recorded events.
template <typename K, typename... A>
int launch_t(cov_cell* h, int kind, K kern, dim3 g, dim3 b, uint32_t lds, A... args)
{
    if (!h->timing) {
        hipLaunchKernelGGL(kern, g, b, lds, h->stream, args...);
        HCK(h, hipGetLastError());
        return COV_OK;
    }
    if (h->pend.size() >= 4096) RC(timing_collect(h));
    hipEvent_t a = nullptr, e = nullptr;
    RC(pool_event(h, &a));
    RC(pool_event(h, &e));
    hipExtLaunchKernelGGL(kern, g, b, lds, h->stream, a, e, 0, args...);
    HCK(h, hipGetLastError());
    h->pend.push_back({kind, a, e});
    return COV_OK;
}

int do_recount(cov_cell* h)
{
    h->cnt_valid = true;
    h->th_valid = false;
    const int g = (int)std::min<int64_t>(h->plan.ntile, 8192);
    // the kernel's dynamic LDS limit is one value for every cell: set it to
    // this cell's tile before its launch
    HCK(h, hipFuncSetAttribute((const void*)k_recount, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)(h->plan.T * 4)));
    RC(launch_t(h, 3, k_recount, dim3(g), dim3(WG), (uint32_t)(h->plan.T * 4), h->h, h->plan));
    // (fills, not copies from this frame: the call returns before they run)
    HCK(h, hipMemsetAsync(CTL_PTR(h, unc), 0, 8, h->stream));
    HCK(h, hipMemsetAsync(CTL_PTR(h, ul_n), 0, 8, h->stream));
    HCK(h, hipMemsetAsync(CTL_PTR(h, ul_valid), 0, 4, h->stream));
    hipLaunchKernelGGL(k_derive, dim3(grid_for(h->h.nzw, WG, 1, 4096)), dim3(WG), 0, h->stream, h->h);
    HCK(h, hipGetLastError());
    return COV_OK;
}

int do_derive(cov_cell* h)
{
    HCK(h, hipMemsetAsync(CTL_PTR(h, unc), 0, 8, h->stream));
    hipLaunchKernelGGL(k_derive, dim3(grid_for(h->h.nzw, WG, 1, 4096)), dim3(WG), 0, h->stream, h->h);
    HCK(h, hipGetLastError());
    return COV_OK;
}

int list_rebuild(cov_cell* h)
{
    HCK(h, hipMemsetAsync(CTL_PTR(h, ul_n), 0, 8, h->stream));
    HCK(h, hipMemsetAsync(CTL_PTR(h, ul_ovf), 0, 4, h->stream));
    hipLaunchKernelGGL(k_list_rebuild, dim3(grid_for(h->h.nzw, WG, 1, 4096)), dim3(WG), 0, h->stream, h->h);
    HCK(h, hipGetLastError());
    Ctl c;
    RC(read_ctl(h, &c));
    const int32_t valid = c.ul_ovf ? 0 : 1;
    RC(set_ctl<int32_t>(h, offsetof(Ctl, ul_valid), valid));
    RC(set_ctl<int32_t>(h, offsetof(Ctl, ul_ovf), 0));
    if (!valid) RC(set_ctl<int64_t>(h, offsetof(Ctl, ul_n), 0));
    return COV_OK;
}

int launch_step(cov_cell* h)
{
    RC(launch_t(h, 0, k_eval, dim3(h->eval_groups), dim3(WG), 0, h->h));
    RC(launch_t(h, 1, k_commit, dim3(h->commit_groups), dim3(WG), 0, h->h));
    RC(launch_t(h, 5, k_post, dim3(1), dim3(WG), 0, h->h));
    return COV_OK;
}

int ensure_graph(cov_cell* h)
{
    if (h->gexec && h->g_eg == h->eval_groups && h->g_cg == h->commit_groups) return COV_OK;
    if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
    hipGraph_t g;
    HCK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
    for (int s = 0; s < GRAPH_STEPS; s++) {
        hipLaunchKernelGGL(k_eval, dim3(h->eval_groups), dim3(WG), 0, h->stream, h->h);
        hipLaunchKernelGGL(k_commit, dim3(h->commit_groups), dim3(WG), 0, h->stream, h->h);
        hipLaunchKernelGGL(k_post, dim3(1), dim3(WG), 0, h->stream, h->h);
    }
    HCK(h, hipStreamEndCapture(h->stream, &g));
    hipError_t e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HCK(h, e);
    h->g_steps = GRAPH_STEPS;
    h->g_eg = h->eval_groups;
    h->g_cg = h->commit_groups;
    return COV_OK;
}

// first min(cap, total) selected words in ascending order; returns the total
template <class Sel>
int64_t ordered_gather(cov_cell* h, Sel sel, int64_t* out, int64_t cap)
{
    const Cell& C = h->h;
    int64_t *dz = nullptr, *dout = nullptr;
    HCK(h, hipMallocAsync((void**)&dz, (size_t)C.nblk * 8 + 8, h->stream));
    hipLaunchKernelGGL(k_block_count<Sel>, dim3(grid_for(C.nblk, 1, 1, 8192)), dim3(WG), 0, h->stream, h->h, sel,
                       dz);
    HCK(h, hipGetLastError());
    std::vector<int64_t> bz((size_t)C.nblk);
    HCK(h, hipMemcpyAsync(bz.data(), dz, (size_t)C.nblk * 8, hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    int64_t total = 0;
    for (auto& v : bz) { const int64_t t = v; v = total; total += t; }
    const int64_t m = std::min(cap, total);
    if (m > 0) {
        HCK(h, hipMemcpyAsync(dz, bz.data(), (size_t)C.nblk * 8, hipMemcpyHostToDevice, h->stream));
        HCK(h, hipMallocAsync((void**)&dout, (size_t)m * 8, h->stream));
        hipLaunchKernelGGL(k_block_emit<Sel>, dim3(grid_for(C.nblk, 1, 1, 8192)), dim3(WG), 0, h->stream, h->h, sel,
                           dz, dout, m);
        HCK(h, hipGetLastError());
        HCK(h, hipMemcpyAsync(out, dout, (size_t)m * 8, hipMemcpyDeviceToHost, h->stream));
        HCK(h, hipFreeAsync(dout, h->stream));
    }
    HCK(h, hipFreeAsync(dz, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    return total;
}

// ------------------------------------------- the gather without allocation
// cov_gather_uncovered_noalloc keeps its block counts, their total and its
// output staging in one file-scope device array instead of allocating them
// per call: NA_BLK block counts (words up to NA_BLK * 4096 = 2^26), the
// total, then NA_CAP indices.  Each device holds its own copy for the life
// of the process; host code reaches it only through the address that
// hipGetSymbolAddress gives on that device (na_scratch_ptr).
// (NA_BLK, NA_CAP and the array are declared above the anonymous namespace.)

// exclusive prefix sum of v[0..nb) in place and its total, one workgroup
__global__ void __launch_bounds__(WG) k_scan_excl(int64_t* v, int64_t nb, int64_t* total)
{
    __shared__ WgScan sc;
    const int64_t per = (nb + WG - 1) / WG;
    const int64_t b0 = (int64_t)threadIdx.x * per;
    int64_t s = 0;
    for (int64_t k = 0; k < per; k++)
        if (b0 + k < nb) s += v[b0 + k];
    int64_t tot;
    int64_t run = wg_incl_sum(s, sc, &tot) - s;
    for (int64_t k = 0; k < per; k++)
        if (b0 + k < nb) {
            const int64_t t = v[b0 + k];
            v[b0 + k] = run;
            run += t;
        }
    if (threadIdx.x == 0) *total = tot;
}

constexpr int NA_DEVICES = 64;

// this device's address of `scratch`, resolved once per device (with that
// device current, as the calling thread's is); nullptr on failure
int64_t* na_scratch_ptr(cov_cell* h)
{
    static std::mutex mu;
    static int64_t* addr[NA_DEVICES] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= NA_DEVICES) {
        h->err = "no device for the scratch array";
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            h->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return nullptr;
        }
        addr[dev] = (int64_t*)p;
    }
    return addr[dev];
}

// ------------------------------------------------------- count transform
constexpr int CT_LDS_DEFAULT = 65536;

// The passes of a count transform of R + 1 layers: one inner pass over the a
// fastest axes (q^a words per tile), then outer passes of b axes each on
// tiles of q^b rows by W words, as many axes per pass as the LDS budget
// holds with W >= 64 (or the whole row segment when it is shorter).
int ct_plan(int q, int n, int R, int64_t N, int lds, std::vector<CtPass>& out)
{
    out.clear();
    const int R1 = R + 1;
    const int64_t wb = 4LL * R1;
    auto pw = [&](int e) { int64_t v = 1; for (int i = 0; i < e; i++) v *= q; return v; };
    int a = 0;
    while (a < n && pw(a + 1) * wb <= lds) a++;
    if (a == 0) return COV_ELIMIT;
    CtPass p{};
    p.N = N; p.q = q; p.R1 = R1;
    p.k = a; p.wsh = 0; p.rows = pw(a); p.rs = 1; p.span = pw(a); p.ncb = 1; p.ntiles = N / pw(a);
    out.push_back(p);
    int j = a;
    while (j < n) {
        const int64_t colspan = pw(j);
        int wmin = 0;
        while ((2LL << wmin) <= std::min<int64_t>(64, colspan)) wmin++;
        int b = n - j;
        while (b > 1 && pw(b) * (1LL << wmin) * wb > lds) b--;
        if (pw(b) * (1LL << wmin) * wb > lds) {
            while (wmin > 0 && pw(b) * (1LL << wmin) * wb > lds) wmin--;
            if (pw(b) * (1LL << wmin) * wb > lds) return COV_ELIMIT;
        }
        int wsh = wmin;
        while (wsh < 8 && (2LL << wsh) <= colspan && pw(b) * (2LL << wsh) * wb <= lds) wsh++;
        CtPass o{};
        o.N = N; o.q = q; o.R1 = R1; o.k = b; o.wsh = wsh; o.rows = pw(b); o.rs = colspan; o.span = pw(j + b);
        o.ncb = (colspan + (1LL << wsh) - 1) >> wsh;
        o.ntiles = (N / o.span) * o.ncb;
        out.push_back(o);
        j += b;
    }
    out.front().first = 1;
    out.back().last = 1;
    return COV_OK;
}

int run_count_transform(cov_cell* h, int source, int out_mode)
{
    const Cell& C = h->h;
    if (!h->ct) {
        HCK(h, hipMalloc((void**)&h->ct, (size_t)(C.R + 1) * (size_t)C.N * 4));
    }
    std::vector<CtPass> passes;
    const int lds = h->ct_lds ? h->ct_lds : CT_LDS_DEFAULT;
    if (ct_plan(C.q, C.n, C.R, C.N, lds, passes) != COV_OK) {
        h->err = "count transform: no tile plan for this shape and LDS budget";
        return COV_ELIMIT;
    }
    if (source == 0) {
        HCK(h, hipMemsetAsync(h->ct, 0, (size_t)C.N * 4, h->stream));
        hipLaunchKernelGGL(k_ct_scatter, dim3(grid_for(C.M, WG, 1, 256)), dim3(WG), 0, h->stream, h->h, h->ct);
        HCK(h, hipGetLastError());
    }
    for (CtPass p : passes) {
        p.src = source == 0 ? CT_SRC_LAYER : CT_SRC_CNT;
        p.target = source == 2 ? 1 : 0;
        p.out = out_mode;
        const size_t lb = (size_t)p.R1 * (size_t)(p.rows << p.wsh) * 4;
        HCK(h, hipFuncSetAttribute((const void*)k_ct_pass, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb));
        const int g = (int)std::min<int64_t>(p.ntiles, 16384);
        RC(launch_t(h, 6, k_ct_pass, dim3(g), dim3(CT_NT), (uint32_t)lb, p, h->ct,
                    (const uint16_t*)reinterpret_cast<const uint16_t*>(C.cnt32)));
    }
    h->map_kind = out_mode == CT_OUT_SUM ? source : -1;
    return COV_OK;
}

}  // namespace

// =============================================================== the ABI
extern "C" {

int cov_version(void) { return 1; }

const char* cov_strerror(int code)
{
    switch (code) {
    case COV_OK: return "ok";
    case COV_EPARAM: return "bad parameters";
    case COV_ESPACE: return "q^n too large";
    case COV_ENOMEM: return "device allocation failed";
    case COV_ESTATE: return "call not valid in this state";
    case COV_ELIMIT: return "input beyond a stated capacity";
    default: return code <= -100 ? hipGetErrorString((hipError_t)(-code - 100)) : "unknown error";
    }
}

int64_t cov_mem_bytes(int q, int n, int R, int M)
{
    Shape s;
    const int rc = shape_of(q, n, R, M, &s);
    if (rc != COV_OK) return rc;
    return total_bytes(s);
}

const char* cov_last_error(const cov_cell_t* h) { return h ? h->err.c_str() : "null handle"; }

int cov_create(int q, int n, int R, int M, cov_cell_t** out)
{
    if (!out) return COV_EPARAM;
    *out = nullptr;
    Shape s;
    RC(shape_of(q, n, R, M, &s));
    cov_cell* h = new cov_cell();
    auto fail = [&](int rc) {
        for (void* p : h->allocs) (void)hipFree(p);
        if (h->stream) (void)hipStreamDestroy(h->stream);
        if (h->used) (void)hipEventDestroy(h->used);
        delete h;
        return rc;
    };
    {
        const hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
        if (e != hipSuccess) { h->stream = nullptr; return fail(-(100 + (int)e)); }
    }
    {
        const hipError_t e = hipEventCreateWithFlags(&h->used, hipEventDisableTiming);
        if (e != hipSuccess) { h->used = nullptr; return fail(-(100 + (int)e)); }
    }
    {
        int dev = 0, khz = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
            h->clk_hz = (double)khz * 1e3;
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
            h->ncu = ncu;
    }
    Cell& C = h->h;
    SubsetTables tabs;
    plan_cell(s, C, tabs);
    const Sizes z = sizes_of(s);
    int rc = COV_OK;
#define AL(ptr, bytes) if ((rc = dalloc(h, &(ptr), (bytes))) != COV_OK) return fail(rc)
    AL(C.cnt32, z.cnt);
    AL(C.zb, z.plane);
    AL(C.ob, z.plane);
    AL(C.inl, z.plane);
    AL(C.blk, z.blk);
    AL(C.code, z.code);
    AL(C.best, z.code);
    AL(C.cpk, z.cpk);
    AL(C.cidx, z.cidx);
    AL(C.tabu, z.tabu);
    AL(C.fsub, z.fsub);
    AL(C.psub, z.psub);
    AL(C.bsub, z.bsub);
    AL(h->plan.ssub, z.ssub);
    AL(C.ul_idx, z.ul_idx);
    AL(C.ul_pk, z.ul_pk);
    AL(C.shuf, z.shuf);
    AL(C.acc, z.acc);
    AL(C.slot_cw, z.slot_cw);
    AL(C.slot_cpk, z.slot_cpk);
    AL(C.rbuf, z.rbuf);
    AL(C.ctl, z.ctl);
    AL(h->d, z.cell);
    AL(h->d_small, z.small);
    AL(C.arrive, z.arrive);
    AL(C.dirty, z.dirty);
    AL(C.jump, z.jump);
    AL(C.xor32, z.xor32);
    AL(C.th, z.th);
    AL(C.cst, z.cst);
    AL(C.cbs, z.cbs);
    AL(C.cps, z.cps);
    AL(C.csr, z.csr);
    AL(C.dth, z.dth);
#undef AL
    C.scratch64 = (decltype(C.scratch64))h->d_small;

    // recount plan: tiles of q^(n-t) words with a common t-digit prefix
    RecountPlan& pl = h->plan;
    pl.t = 0;
    pl.T = s.N;
    while (pl.T > TILE_MAX) { pl.t++; pl.T /= q; }
    pl.ntile = s.N / pl.T;
    std::vector<int> suf;
    for (int j = pl.t; j < n; j++) suf.push_back(j);
    std::vector<uint32_t> ss;
    int64_t sacc = 0;
    for (int r = 0; r <= MAXR + 1; r++) {
        pl.ssub_off[r] = (int64_t)ss.size();
        pl.spat[r] = sacc;
        int64_t pr = 1;
        for (int i = 0; i < r; i++) pr *= (q - 1);
        pl.spr[r] = pr;
        pl.sinv[r] = 1.0f / (float)pr;
        if (r <= R && r <= (int)suf.size()) {
            std::vector<uint32_t> t = subsets(suf, r);
            ss.insert(ss.end(), t.begin(), t.end());
            sacc += (int64_t)t.size() * pr;
        }
    }
    {
        // the rng jump matrices, computed once per process
        static const std::vector<uint64_t> jt = [] {
            std::vector<uint64_t> v(RNG_JUMP_WORDS);
            rng_jump_tables(v.data());
            return v;
        }();
        if ((rc = upload(h, C.jump, jt.data(), (int64_t)jt.size() * 8)) != COV_OK) return fail(rc);
    }
    if ((rc = upload(h, C.fsub, tabs.fs.data(), (int64_t)tabs.fs.size() * 4)) != COV_OK) return fail(rc);
    if (!tabs.ps.empty() && (rc = upload(h, C.psub, tabs.ps.data(), (int64_t)tabs.ps.size() * 4)) != COV_OK)
        return fail(rc);
    if ((rc = upload(h, C.bsub, tabs.bs.data(), (int64_t)tabs.bs.size() * 4)) != COV_OK) return fail(rc);
    if ((rc = upload(h, pl.ssub, ss.data(), (int64_t)ss.size() * 4)) != COV_OK) return fail(rc);

    Ctl c{};
    c.stall = 5000;
    c.kick = 2;
    c.cand = 100000;
    c.budget = 1024;
    c.tenure = 0;
    c.it_stop = 0;
    c.target = -1;
    c.ttt_it = -1;
    if ((rc = upload(h, C.ctl, &c, sizeof(Ctl))) != COV_OK) return fail(rc);
    if ((rc = push_cell(h)) != COV_OK) return fail(rc);
    if (pl.T * 4 > 65536) {
        h->err = "recount tile does not fit in LDS";
        return fail(COV_ELIMIT);
    }
    // default launch shape: about 32 candidate codewords per step
    const int64_t eval_waves = 32 * ((int64_t)s.F + (int64_t)(R + 1) * s.E) *
                               std::max<int64_t>(1, s.P / 64);
    h->eval_groups = grid_for(eval_waves / 4, 4, 16, 1024);
    h->commit_groups = grid_for((int64_t)s.E * s.P, WG, 1, 2048);
    // the uploads above read this frame's tables
    {
        const hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return fail(-(100 + (int)e));
    }
    *out = h;
    return COV_OK;
}

int cov_destroy(cov_cell_t* h)
{
    if (!h) return COV_EPARAM;
    if (h->used) (void)hipEventSynchronize(h->used);    // batched launches that include the cell
    (void)hipStreamSynchronize(h->stream);
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    for (void* p : h->allocs) (void)hipFree(p);
    if (h->trace) (void)hipFree(h->trace);
    if (h->h.dist) (void)hipFree(h->h.dist);
    if (h->cwb) (void)hipFree(h->cwb);
    if (h->dtstat) (void)hipFree(h->dtstat);
    if (h->ct) (void)hipFree(h->ct);
    if (h->rr) (void)hipFree(h->rr);
    for (const auto& p : h->pend) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (hipEvent_t e : h->evs) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(h->stream);
    if (h->used) (void)hipEventDestroy(h->used);
    delete h;
    return COV_OK;
}

int cov_sync(cov_cell_t* h)
{
    if (!h) return COV_EPARAM;
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

int cov_set_code_flags(cov_cell_t* h, const uint8_t* digits, int flags)
{
    if (h) h->th_valid = false;
    if (h) h->dist_valid = false;
    if (!h || !digits || (flags & ~COV_NO_RECOUNT)) return COV_EPARAM;
    const Cell& C = h->h;
    for (int64_t k = 0; k < (int64_t)C.M * C.n; k++)
        if (digits[k] >= C.q) { h->err = "digit out of range"; return COV_EPARAM; }
    RC(upload(h, C.code, digits, (int64_t)C.M * C.n));
    hipLaunchKernelGGL(k_pack_code, dim3(grid_for(C.M, WG, 1, 256)), dim3(WG), 0, h->stream, h->h);
    HCK(h, hipGetLastError());
    if (flags & COV_NO_RECOUNT) {
        h->cnt_valid = false;
        HCK(h, hipStreamSynchronize(h->stream));
        return COV_OK;
    }
    RC(do_recount(h));
    HCK(h, hipMemcpyAsync(C.best, C.code, (size_t)C.M * C.n, hipMemcpyDeviceToDevice, h->stream));
    Ctl c;
    RC(read_ctl(h, &c));
    RC(set_ctl<int64_t>(h, offsetof(Ctl, best_unc), c.unc));
    RC(set_ctl<int32_t>(h, offsetof(Ctl, ndirty), 0));
    return COV_OK;
}

int cov_set_code(cov_cell_t* h, const uint8_t* digits)
{
    return cov_set_code_flags(h, digits, 0);
}

int cov_get_code(cov_cell_t* h, uint8_t* digits)
{
    if (!h || !digits) return COV_EPARAM;
    HCK(h, hipMemcpyAsync(digits, h->h.code, (size_t)h->h.M * h->h.n, hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

int cov_get_best(cov_cell_t* h, uint8_t* digits, int64_t* best_uncovered)
{
    if (!h) return COV_EPARAM;
    if (digits) HCK(h, hipMemcpyAsync(digits, h->h.best, (size_t)h->h.M * h->h.n, hipMemcpyDeviceToHost, h->stream));
    Ctl c;
    RC(read_ctl(h, &c));
    if (best_uncovered) *best_uncovered = c.best_unc;
    return COV_OK;
}

int cov_recount(cov_cell_t* h)
{
    if (!h) return COV_EPARAM;
    RC(do_recount(h));
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

int cov_ball_update(cov_cell_t* h, const int64_t* centers, int k, int delta)
{
    if (h) h->th_valid = false;
    if (h && !h->cnt_valid) { h->err = "coverage counts not built (cov_recount)"; return COV_ESTATE; }
    if (!h || (k > 0 && !centers) || (delta != 1 && delta != -1) || k < 0) return COV_EPARAM;
    if (k == 0) return COV_OK;
    const Cell& C = h->h;
    for (int i = 0; i < k; i++)
        if (centers[i] < 0 || centers[i] >= C.N) { h->err = "centre out of range"; return COV_EPARAM; }
    if (k == 1) {
        int64_t w = centers[0];
        uint64_t pk = 0;
        for (int j = C.n - 1; j >= 0; j--) { pk |= (uint64_t)(w % C.q) << (4 * j); w /= C.q; }
        hipLaunchKernelGGL(k_ball_one, dim3(grid_for(C.V, WG, 1, 1024)), dim3(WG), 0, h->stream, h->h,
                           centers[0], pk, delta);
        HCK(h, hipGetLastError());
    } else {
        int64_t* dc = nullptr;
        HCK(h, hipMallocAsync((void**)&dc, (size_t)k * 8, h->stream));
        HCK(h, hipMemcpyAsync(dc, centers, (size_t)k * 8, hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(k_ball_multi, dim3(grid_for((int64_t)k * C.V, WG, 1, 4096)), dim3(WG), 0,
                           h->stream, h->h, dc, k, delta);
        HCK(h, hipGetLastError());
        HCK(h, hipFreeAsync(dc, h->stream));
        RC(do_derive(h));
        RC(set_ctl<int32_t>(h, offsetof(Ctl, ul_valid), 0));
    }
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

int64_t cov_count_uncovered(cov_cell_t* h)
{
    if (h && !h->cnt_valid) { h->err = "coverage counts not built (cov_recount)"; return COV_ESTATE; }
    if (!h) return COV_EPARAM;
    HCK(h, hipMemsetAsync(h->d_small, 0, 8, h->stream));
    hipLaunchKernelGGL(k_count_zero, dim3(grid_for(h->h.N / 8, WG, 1, 2048)), dim3(WG), 0, h->stream, h->h,
                       h->d_small);
    HCK(h, hipGetLastError());
    int64_t r = 0;
    HCK(h, hipMemcpyAsync(&r, h->d_small, 8, hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    return r;
}

int64_t cov_gather_uncovered(cov_cell_t* h, int64_t* out, int64_t cap)
{
    if (h && !h->cnt_valid) { h->err = "coverage counts not built (cov_recount)"; return COV_ESTATE; }
    if (!h || cap < 0 || (cap > 0 && !out)) return COV_EPARAM;
    return ordered_gather(h, SelZero{reinterpret_cast<const uint16_t*>(h->h.cnt32)}, out, cap);
}

int64_t cov_gather_uncovered_noalloc(cov_cell_t* h, int64_t* out, int64_t cap)
{
    if (h && !h->cnt_valid) { h->err = "coverage counts not built (cov_recount)"; return COV_ESTATE; }
    if (!h || cap < 0 || (cap > 0 && !out)) return COV_EPARAM;
    const Cell& C = h->h;
    if (C.nblk > NA_BLK) { h->err = "more than 2^26 words for the allocation-free gather"; return COV_ELIMIT; }
    if (cap > NA_CAP) { h->err = "cap above 65536 for the allocation-free gather"; return COV_ELIMIT; }
    int64_t* sp = na_scratch_ptr(h);
    if (!sp) return COV_ESTATE;
    int64_t* bz = sp;                   // block counts, then their offsets
    int64_t* tot = sp + NA_BLK;         // total selected
    int64_t* stage = sp + NA_BLK + 32;  // the first min(cap, total) indices
    if (h->used) HCK(h, hipStreamWaitEvent(h->stream, h->used, 0));
    const SelZero sel{reinterpret_cast<const uint16_t*>(C.cnt32)};
    const int g = grid_for(C.nblk, 1, 1, 8192);
    hipLaunchKernelGGL(k_block_count<SelZero>, dim3(g), dim3(WG), 0, h->stream, h->h, sel, bz);
    HCK(h, hipGetLastError());
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(WG), 0, h->stream, bz, C.nblk, tot);
    HCK(h, hipGetLastError());
    if (cap > 0) {
        hipLaunchKernelGGL(k_block_emit<SelZero>, dim3(g), dim3(WG), 0, h->stream, h->h, sel, bz, stage, cap);
        HCK(h, hipGetLastError());
    }
    int64_t total = 0;
    HCK(h, hipMemcpyAsync(&total, tot, 8, hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    const int64_t m = std::min(cap, total);
    if (m > 0) {
        HCK(h, hipMemcpyAsync(out, stage, (size_t)m * 8, hipMemcpyDeviceToHost, h->stream));
        HCK(h, hipStreamSynchronize(h->stream));
    }
    return total;
}

int64_t cov_gather_far(cov_cell_t* h, int r, int64_t* out, int64_t cap)
{
    if (!h || r < 0 || cap < 0 || (cap > 0 && !out)) return COV_EPARAM;
    if (!h->h.dist || !h->dist_valid) {
        h->err = "no distance transform of the current code";
        return COV_ESTATE;
    }
    return ordered_gather(h, SelFar{h->h.dist, r}, out, cap);
}

int64_t cov_first_uncovered_from(cov_cell_t* h, int64_t start)
{
    if (h && !h->cnt_valid) { h->err = "coverage counts not built (cov_recount)"; return COV_ESTATE; }
    if (!h || start < 0 || start >= h->h.N) return COV_EPARAM;
    hipLaunchKernelGGL(k_first_from, dim3(1), dim3(WG), 0, h->stream, h->h, start, h->d_small);
    HCK(h, hipGetLastError());
    int64_t r = 0;
    HCK(h, hipMemcpyAsync(&r, h->d_small, 8, hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    return r;
}

int cov_read_cnt(cov_cell_t* h, int64_t off, int64_t len, uint16_t* out)
{
    if (h && !h->cnt_valid) { h->err = "coverage counts not built (cov_recount)"; return COV_ESTATE; }
    if (!h || off < 0 || len < 0 || off + len > h->h.N || (len && !out)) return COV_EPARAM;
    const uint16_t* src = reinterpret_cast<const uint16_t*>(h->h.cnt32) + off;
    HCK(h, hipMemcpyAsync(out, src, (size_t)len * 2, hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

int cov_eval_moves(cov_cell_t* h, const uint8_t* u, int k, const int32_t* cw, int64_t* dvec)
{
    if (h && !h->cnt_valid) { h->err = "coverage counts not built (cov_recount)"; return COV_ESTATE; }
    if (!h || !u || k < 0 || (k > 0 && (!cw || !dvec))) return COV_EPARAM;
    const Cell& C = h->h;
    if (k > C.M) return COV_ELIMIT;
    if (k == 0) return COV_OK;
    uint64_t upk = 0;
    for (int j = 0; j < C.n; j++) {
        if (u[j] >= C.q) return COV_EPARAM;
        upk |= (uint64_t)u[j] << (4 * j);
    }
    for (int t = 0; t < k; t++)
        if (cw[t] < 0 || cw[t] >= C.M) return COV_EPARAM;
    RC(upload(h, C.slot_cw, cw, (int64_t)k * 4));
    HCK(h, hipMemcpyAsync(CTL_PTR(h, upk), &upk, 8, hipMemcpyHostToDevice, h->stream));
    HCK(h, hipMemsetAsync(C.acc, 0, (size_t)k * ACCW * 4, h->stream));
    const int64_t units = (int64_t)k * (C.F + (int64_t)C.n * C.E);
    hipLaunchKernelGGL(k_eval_probe, dim3(grid_for(units, 4 * 8, 1, 2048)), dim3(WG), 0, h->stream, h->h, k);
    HCK(h, hipGetLastError());
    int64_t* dd = nullptr;
    HCK(h, hipMallocAsync((void**)&dd, (size_t)k * C.n * 8, h->stream));
    hipLaunchKernelGGL(k_eval_finish, dim3(grid_for((int64_t)k * C.n, WG, 1, 256)), dim3(WG), 0, h->stream, h->h,
                       k, dd);
    HCK(h, hipGetLastError());
    HCK(h, hipMemcpyAsync(dvec, dd, (size_t)k * C.n * 8, hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipFreeAsync(dd, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

int cov_eval_codewords(cov_cell_t* h, int k, const int32_t* cw, int64_t* dvec)
{
    if (h && !h->cnt_valid) { h->err = "coverage counts not built (cov_recount)"; return COV_ESTATE; }
    if (!h || k < 0 || (k > 0 && (!cw || !dvec))) return COV_EPARAM;
    const Cell& C = h->h;
    if (k > C.M) return COV_ELIMIT;
    if (k == 0) return COV_OK;
    for (int t = 0; t < k; t++)
        if (cw[t] < 0 || cw[t] >= C.M) return COV_EPARAM;
    if (h->used) HCK(h, hipStreamWaitEvent(h->stream, h->used, 0));
    RC(upload(h, C.slot_cw, cw, (int64_t)k * 4));
    HCK(h, hipMemsetAsync(C.acc, 0, (size_t)k * ACCW * 4, h->stream));
    int32_t* negw = nullptr;
    int64_t* dd = nullptr;
    HCK(h, hipMallocAsync((void**)&negw, (size_t)k * WIDE_W * 4, h->stream));
    HCK(h, hipMemsetAsync(negw, 0, (size_t)k * WIDE_W * 4, h->stream));
    const int64_t units = (int64_t)k * (C.F + (int64_t)C.n * C.q1 * C.E);
    hipLaunchKernelGGL(k_eval_probe_all, dim3(grid_for(units, 4 * 8, 1, 4096)), dim3(WG), 0, h->stream, h->h, k,
                       negw);
    HCK(h, hipGetLastError());
    HCK(h, hipMallocAsync((void**)&dd, (size_t)k * C.n * C.q * 8, h->stream));
    hipLaunchKernelGGL(k_eval_finish_all, dim3(grid_for((int64_t)k * C.n * C.q, WG, 1, 1024)), dim3(WG), 0,
                       h->stream, h->h, k, negw, dd);
    HCK(h, hipGetLastError());
    HCK(h, hipMemcpyAsync(dvec, dd, (size_t)k * C.n * C.q * 8, hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipFreeAsync(dd, h->stream));
    HCK(h, hipFreeAsync(negw, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

int cov_commit(cov_cell_t* h, int i, int p, int v, int64_t* dunc)
{
    if (h) h->th_valid = false;
    if (h && !h->cnt_valid) { h->err = "coverage counts not built (cov_recount)"; return COV_ESTATE; }
    if (h) h->dist_valid = false;
    if (!h) return COV_EPARAM;
    const Cell& C = h->h;
    if (i < 0 || i >= C.M || p < 0 || p >= C.n || v < 0 || v >= C.q) return COV_EPARAM;
    hipLaunchKernelGGL(k_set_move, dim3(1), dim3(64), 0, h->stream, h->h, i, p, v);
    HCK(h, hipGetLastError());
    hipLaunchKernelGGL(k_commit_api, dim3(h->commit_groups), dim3(WG), 0, h->stream, h->h);
    HCK(h, hipGetLastError());
    Ctl c;
    RC(read_ctl(h, &c));
    if (dunc) *dunc = c.dsel;
    return COV_OK;
}

// all passes of the transform on the cell's stream; the last pass leaves in
// hh either the full histogram (full) or {#{dist <= rcut}, max dist}
static int run_transform(cov_cell_t* h, int full, int rcut, int64_t hh[DT_HB])
{
    Cell& C = h->h;
    if (!C.dist) {
        void* v = nullptr;
        HCK(h, hipMalloc(&v, (size_t)C.N + 64));
        C.dist = (decltype(C.dist))v;
        RC(push_cell(h));
    }
    if (!h->cwb) {
        const size_t wb = ((size_t)(C.N >> 5) + 2) * 4;
        HCK(h, hipMalloc((void**)&h->cwb, wb));
        HCK(h, hipMemsetAsync(h->cwb, 0, wb, h->stream));
    }
    DtPass pass[MAXN + 1];
    const int np = dt_plan(C.q, C.n, C.N, h->dt_lds ? h->dt_lds : DT_LDS_DEFAULT, pass, true, true, h->ncu);
    if (np <= 0) {
        h->err = "distance transform: no tile plan for this shape and LDS budget";
        return COV_ELIMIT;
    }
    pass[np - 1].full = full;
    pass[np - 1].rcut = rcut;
    const size_t sbytes = (size_t)DT_COPIES * DT_CSTRIDE * 8;
    if (!h->dtstat) HCK(h, hipMalloc((void**)&h->dtstat, sbytes));
    unsigned long long* dh = reinterpret_cast<unsigned long long*>(h->dtstat);
    HCK(h, hipMemsetAsync(dh, 0, sbytes, h->stream));
    const int gm = grid_for(C.M, WG, 1, 64);
    hipLaunchKernelGGL(k_cwb_mark, dim3(gm), dim3(WG), 0, h->stream, h->h, h->cwb, 1);
    HCK(h, hipGetLastError());
    for (int i = 0; i < np; i++) {
        const DtPass& p = pass[i];
        const int g = (int)std::min<int64_t>(p.ntiles, 8192);
        const int nt = h->dt_threads ? h->dt_threads : DT_THREADS_DEFAULT;
        int lrc = COV_OK;
        if (p.seed) {
            const bool ok = dt_seed_dispatch(p.q, [&](auto qb) {
                using QB = decltype(qb);
                lrc = launch_t(h, 2, (k_dt_seed<QB::q>), dim3(g), dim3(DT_COL_NT), (uint32_t)p.lds, p, C.dist,
                               h->cwb, dh);
            });
            if (!ok) {
                h->err = "distance transform: no seed kernel for this q";
                return COV_ELIMIT;
            }
        } else if (p.tcol) {
            const bool ok = dt_tcol_dispatch(p.q, p.b1, p.nax - p.b1, [&](auto qb) {
                using QB = decltype(qb);
                const hipError_t e = hipFuncSetAttribute((const void*)k_dt_tcol<QB::q, QB::b1, QB::b2>,
                                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds);
                if (e != hipSuccess) {
                    lrc = -(100 + (int)e);
                    return;
                }
                lrc = launch_t(h, 2, (k_dt_tcol<QB::q, QB::b1, QB::b2>), dim3(p.grid), dim3(p.nt), (uint32_t)p.lds, p,
                               C.dist, dh);
            });
            if (!ok) {
                h->err = "distance transform: no two-stage column kernel for this (q, axes)";
                return COV_ELIMIT;
            }
        } else if (p.col) {
            const unsigned gc = (p.nthr + DT_COL_NT - 1) / DT_COL_NT;
            const bool ok = dt_col_dispatch(p.q, p.nax, [&](auto qb) {
                using QB = decltype(qb);
                lrc = launch_t(h, 2, (k_dt_col<QB::q, QB::b>), dim3(gc), dim3(DT_COL_NT), 0, p, C.dist, dh);
            });
            if (!ok) {
                h->err = "distance transform: no column kernel for this (q, axes)";
                return COV_ELIMIT;
            }
        } else if (nt == 1024)
            lrc = launch_t(h, 2, k_dt_pass<1024>, dim3(g), dim3(1024), (uint32_t)dt_lds_bytes(p), p, C.dist, h->cwb, dh);
        else if (nt == 512)
            lrc = launch_t(h, 2, k_dt_pass<512>, dim3(g), dim3(512), (uint32_t)dt_lds_bytes(p), p, C.dist, h->cwb, dh);
        else
            lrc = launch_t(h, 2, k_dt_pass<256>, dim3(g), dim3(256), (uint32_t)dt_lds_bytes(p), p, C.dist, h->cwb, dh);
        RC(lrc);
        if (i == 0) {
            hipLaunchKernelGGL(k_cwb_mark, dim3(gm), dim3(WG), 0, h->stream, h->h, h->cwb, 0);
            HCK(h, hipGetLastError());
        }
    }
    h->dtcopies.resize((size_t)DT_COPIES * DT_CSTRIDE);
    HCK(h, hipMemcpyAsync(h->dtcopies.data(), dh, sbytes, hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    dt_fold_copies(h->dtcopies.data(), full != 0, hh);
    return COV_OK;
}

int cov_distance_transform(cov_cell_t* h, uint8_t* dist, int64_t* hist, int* radius)
{
    if (!h) return COV_EPARAM;
    const Cell& C = h->h;
    int64_t hh[DT_HB];
    RC(run_transform(h, 1, 0, hh));
    if (dist) HCK(h, hipMemcpyAsync(dist, C.dist, (size_t)C.N, hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    int rad = 0;
    for (int d = 0; d <= C.n + 1; d++)
        if (hh[d]) rad = d;
    if (hist)
        for (int d = 0; d <= C.n; d++) hist[d] = hh[d];
    if (radius) *radius = rad;
    h->dist_valid = true;
    return COV_OK;
}

int cov_distance_stats(cov_cell_t* h, int r, int64_t* within, int* radius)
{
    if (!h || r < 0) return COV_EPARAM;
    const Cell& C = h->h;
    int64_t hh[DT_HB];
    RC(run_transform(h, 0, r > 127 ? 127 : r, hh));
    HCK(h, hipStreamSynchronize(h->stream));
    if (within) *within = hh[0];
    if (radius) *radius = hh[1] > C.n ? C.n + 1 : (int)hh[1];
    h->dist_valid = true;
    return COV_OK;
}

int cov_count_transform(cov_cell_t* h, int source, int32_t* layers, int32_t* map)
{
    if (!h || source < 0 || source > 2) return COV_EPARAM;
    if (source > 0 && !h->cnt_valid) { h->err = "coverage counts not built (cov_recount)"; return COV_ESTATE; }
    const Cell& C = h->h;
    if ((C.R + 1) * C.N > ((int64_t)1 << 34)) { h->err = "count transform layers above 64 GB"; return COV_ELIMIT; }
    RC(run_count_transform(h, source, layers ? CT_OUT_LAYERS : CT_OUT_SUM));
    if (layers)
        HCK(h, hipMemcpyAsync(layers, h->ct, (size_t)(C.R + 1) * (size_t)C.N * 4, hipMemcpyDeviceToHost, h->stream));
    else if (map)
        HCK(h, hipMemcpyAsync(map, h->ct, (size_t)C.N * 4, hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

int cov_set_count_tile(cov_cell_t* h, int lds_bytes)
{
    if (!h || lds_bytes < 0 || lds_bytes > 160 * 1024) return COV_EPARAM;
    std::vector<CtPass> ps;
    if (lds_bytes && ct_plan(h->h.q, h->h.n, h->h.R, h->h.N, lds_bytes, ps) != COV_OK) return COV_EPARAM;
    h->ct_lds = lds_bytes;
    return COV_OK;
}

int cov_map_max(cov_cell_t* h, int32_t* max)
{
    if (!h || !max) return COV_EPARAM;
    if (h->map_kind < 0) { h->err = "no map (cov_count_transform)"; return COV_ESTATE; }
    HCK(h, hipMemsetAsync(h->d_small, 0, 8, h->stream));
    hipLaunchKernelGGL(k_map_max, dim3(grid_for(h->h.N, WG * 8, 1, 4096)), dim3(WG), 0, h->stream, h->h, h->ct,
                       h->d_small);
    HCK(h, hipGetLastError());
    int64_t r = 0;
    HCK(h, hipMemcpyAsync(&r, h->d_small, 8, hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    *max = (int32_t)r;
    return COV_OK;
}

int cov_map_hist(cov_cell_t* h, int nbins, int32_t vmax, int64_t* hist)
{
    if (!h || !hist || nbins < 1 || nbins > 8192 || vmax < 1) return COV_EPARAM;
    if (h->map_kind < 0) { h->err = "no map (cov_count_transform)"; return COV_ESTATE; }
    unsigned long long* dh = nullptr;
    HCK(h, hipMallocAsync((void**)&dh, (size_t)nbins * 8, h->stream));
    HCK(h, hipMemsetAsync(dh, 0, (size_t)nbins * 8, h->stream));
    hipLaunchKernelGGL(k_map_hist, dim3(grid_for(h->h.N, WG * 16, 1, 2048)), dim3(WG), (size_t)nbins * 4, h->stream,
                       h->h, h->ct, nbins, vmax, dh);
    HCK(h, hipGetLastError());
    HCK(h, hipMemcpyAsync(hist, dh, (size_t)nbins * 8, hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipFreeAsync(dh, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

int64_t cov_map_extract(cov_cell_t* h, int32_t thr, int64_t* idx, int32_t* val, int64_t cap)
{
    if (!h || cap < 0 || (cap > 0 && (!idx || !val))) return COV_EPARAM;
    if (h->map_kind < 0) { h->err = "no map (cov_count_transform)"; return COV_ESTATE; }
    const int64_t tot = ordered_gather(h, SelMapGe{h->ct, thr}, idx, cap);
    if (tot < 0) return tot;
    const int64_t m = std::min(cap, tot);
    if (m > 0) RC(cov_map_read(h, idx, (int)m, val));
    return tot;
}

int cov_map_read(cov_cell_t* h, const int64_t* idx, int k, int32_t* out)
{
    if (!h || k < 0 || (k > 0 && (!idx || !out))) return COV_EPARAM;
    if (h->map_kind < 0) { h->err = "no map (cov_count_transform)"; return COV_ESTATE; }
    if (k == 0) return COV_OK;
    for (int i = 0; i < k; i++)
        if (idx[i] < 0 || idx[i] >= h->h.N) { h->err = "word index out of range"; return COV_EPARAM; }
    int64_t* di = nullptr;
    int32_t* dv = nullptr;
    HCK(h, hipMallocAsync((void**)&di, (size_t)k * 8, h->stream));
    HCK(h, hipMallocAsync((void**)&dv, (size_t)k * 4, h->stream));
    HCK(h, hipMemcpyAsync(di, idx, (size_t)k * 8, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_map_read, dim3(grid_for(k, WG, 1, 1024)), dim3(WG), 0, h->stream, h->ct, di, k, dv);
    HCK(h, hipGetLastError());
    HCK(h, hipMemcpyAsync(out, dv, (size_t)k * 4, hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipFreeAsync(di, h->stream));
    HCK(h, hipFreeAsync(dv, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

int cov_ball_gather(cov_cell_t* h, const int64_t* centers, int k, int target, int32_t* out)
{
    if (h && !h->cnt_valid) { h->err = "coverage counts not built (cov_recount)"; return COV_ESTATE; }
    if (!h || k < 0 || target < 0 || target > 65535 || (k > 0 && (!centers || !out))) return COV_EPARAM;
    if (k == 0) return COV_OK;
    for (int i = 0; i < k; i++)
        if (centers[i] < 0 || centers[i] >= h->h.N) { h->err = "centre out of range"; return COV_EPARAM; }
    int64_t* dc = nullptr;
    int32_t* dv = nullptr;
    HCK(h, hipMallocAsync((void**)&dc, (size_t)k * 8, h->stream));
    HCK(h, hipMallocAsync((void**)&dv, (size_t)k * 4, h->stream));
    HCK(h, hipMemcpyAsync(dc, centers, (size_t)k * 8, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_ball_gather, dim3(std::min(k, 4096)), dim3(WG), 0, h->stream, h->h, dc, k, target, dv);
    HCK(h, hipGetLastError());
    HCK(h, hipMemcpyAsync(out, dv, (size_t)k * 4, hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipFreeAsync(dc, h->stream));
    HCK(h, hipFreeAsync(dv, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

int cov_load_words(cov_cell_t* h, const int64_t* words, int m)
{
    if (!h || m < 0 || (m > 0 && !words)) return COV_EPARAM;
    const Cell& C = h->h;
    for (int i = 0; i < m; i++)
        if (words[i] < 0 || words[i] >= C.N) { h->err = "word index out of range"; return COV_EPARAM; }
    h->th_valid = false;
    h->dist_valid = false;
    if (h->used) HCK(h, hipStreamWaitEvent(h->stream, h->used, 0));
    HCK(h, hipMemsetAsync(C.cnt32, 0, (size_t)((C.N + 1) / 2) * 4, h->stream));
    if (m > 0) {
        int64_t* dc = nullptr;
        HCK(h, hipMallocAsync((void**)&dc, (size_t)m * 8, h->stream));
        HCK(h, hipMemcpyAsync(dc, words, (size_t)m * 8, hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(k_ball_multi, dim3(grid_for((int64_t)m * C.V, WG, 1, 4096)), dim3(WG), 0, h->stream,
                           h->h, dc, m, 1);
        HCK(h, hipGetLastError());
        HCK(h, hipFreeAsync(dc, h->stream));
    }
    h->cnt_valid = true;
    RC(do_derive(h));
    RC(set_ctl<int32_t>(h, offsetof(Ctl, ul_valid), 0));
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

int cov_search_params(cov_cell_t* h, int tenure, int64_t stall, int kick, int cand)
{
    if (!h || tenure < 0 || stall < 0 || kick < 0 || cand < 1) return COV_EPARAM;
    h->th_valid = false;
    h->kick = kick;
    RC(set_ctl<int32_t>(h, offsetof(Ctl, tenure), tenure));
    RC(set_ctl<int64_t>(h, offsetof(Ctl, stall), stall));
    RC(set_ctl<int32_t>(h, offsetof(Ctl, kick), kick));
    RC(set_ctl<int32_t>(h, offsetof(Ctl, cand), cand));
    return COV_OK;
}

int cov_search_wide(cov_cell_t* h, int on)
{
    if (!h || (on != 0 && on != 1)) return COV_EPARAM;
    if (on && !h->h.accw) {
        RC(dalloc(h, &h->h.accw, (int64_t)h->h.M * WIDE_W * 4));
        RC(push_cell(h));
    }
    RC(set_ctl<int32_t>(h, offsetof(Ctl, wide), on));
    return COV_OK;
}

int cov_search_init(cov_cell_t* h, uint64_t seed, const uint8_t* seed_code, int mseed, int greedy)
{
    if (h) h->th_valid = false;
    if (h) h->dist_valid = false;
    if (!h || mseed < 0 || (mseed > 0 && !seed_code)) return COV_EPARAM;
    Cell& C = h->h;
    const int n = C.n, M = C.M;
    for (int64_t k = 0; k < (int64_t)mseed * n; k++)
        if (seed_code[k] >= C.q) { h->err = "digit out of range"; return COV_EPARAM; }
    // fresh control block, keeping the search knobs
    Ctl old;
    RC(read_ctl(h, &old));
    Ctl c{};
    c.tenure = old.tenure; c.stall = old.stall; c.kick = old.kick; c.cand = old.cand; c.budget = old.budget;
    c.wide = old.wide;
    c.target = -1;
    c.ttt_it = -1;
    splitmix_seed(seed, c.rst);
    RC(upload(h, C.ctl, &c, sizeof(Ctl)));
    h->th_on = 0;
    HCK(h, hipMemsetAsync(C.tabu, 0, (size_t)M * n * C.q * 4, h->stream));
    auto index_of = [&](const uint8_t* w) {
        int64_t x = 0;
        for (int j = 0; j < n; j++) x += (int64_t)w[j] * C.pw[j];
        return x;
    };
    auto pack_of = [&](const uint8_t* w) {
        uint64_t pk = 0;
        for (int j = 0; j < n; j++) pk |= (uint64_t)w[j] << (4 * j);
        return pk;
    };
    if (mseed > M) {
        // remove-and-repair descent (:1156-1189): counts of all seed words, then
        // repeatedly drop the first word with the fewest uniquely covered
        // words.  The choices stay on the device (three launches a removal,
        // no wait); the dropped words are read back once at the end.
        // (k_rr_pick orders by count << 32 | position: counts below 2^31)
        if (C.V >= ((int64_t)1 << 31)) { h->err = "remove-and-repair: ball of 2^31 words or more"; return COV_ELIMIT; }
        std::vector<uint8_t> sc(seed_code, seed_code + (size_t)mseed * n);
        std::vector<int64_t> si((size_t)mseed);
        for (int i = 0; i < mseed; i++) si[i] = index_of(&sc[(size_t)i * n]);
        if (h->rr_cap < mseed) {
            if (h->rr) (void)hipFree(h->rr);
            h->rr = nullptr;
            h->rr_cap = 0;
            HCK(h, hipMalloc((void**)&h->rr, (2 * (size_t)mseed + 1) * 8));
            h->rr_cap = mseed;
        }
        int64_t* dc = h->rr;
        int64_t* du = h->rr + mseed;
        int64_t* dsel = h->rr + 2 * (int64_t)mseed;
        HCK(h, hipMemsetAsync(C.cnt32, 0, (size_t)((C.N + 1) / 2) * 4, h->stream));
        RC(upload(h, dc, si.data(), (int64_t)mseed * 8));
        hipLaunchKernelGGL(k_ball_multi, dim3(grid_for((int64_t)mseed * C.V, WG, 1, 4096)), dim3(WG), 0, h->stream,
                           h->h, dc, mseed, 1);
        HCK(h, hipGetLastError());
        RC(do_derive(h));
        for (int left = mseed; left > M; left--) {
            hipLaunchKernelGGL(k_unique_cover, dim3(std::min(mseed, 4096)), dim3(WG), 0, h->stream, h->h, dc, mseed,
                               du);
            hipLaunchKernelGGL(k_rr_pick, dim3(1), dim3(WG), 0, h->stream, dc, du, mseed, dsel);
            hipLaunchKernelGGL(k_ball_drop, dim3(grid_for(C.V, WG, 1, 1024)), dim3(WG), 0, h->stream, h->h, dsel);
            HCK(h, hipGetLastError());
        }
        HCK(h, hipMemcpyAsync(si.data(), dc, (size_t)mseed * 8, hipMemcpyDeviceToHost, h->stream));
        HCK(h, hipStreamSynchronize(h->stream));
        // keep the listed words in their order (:1185-1188)
        int kept = 0;
        for (int i = 0; i < mseed; i++)
            if (si[i] >= 0) {
                if (kept != i) std::memmove(&sc[(size_t)kept * n], &sc[(size_t)i * n], (size_t)n);
                kept++;
            }
        if (kept != M) { h->err = "remove-and-repair kept the wrong number of words"; return COV_ESTATE; }
        RC(upload(h, C.code, sc.data(), (int64_t)M * n));
        hipLaunchKernelGGL(k_pack_code, dim3(grid_for(M, WG, 1, 256)), dim3(WG), 0, h->stream, h->h);
        HCK(h, hipGetLastError());
        RC(do_recount(h));
    } else {
        const int have = mseed;
        if (have > 0) RC(upload(h, C.code, seed_code, (int64_t)have * n));
        if (have < M && greedy) {
            // greedy placement on uncovered words (:1197-1234)
            HCK(h, hipMemsetAsync(C.cnt32, 0, (size_t)((C.N + 1) / 2) * 4, h->stream));
            hipLaunchKernelGGL(k_pack_code, dim3(grid_for(M, WG, 1, 256)), dim3(WG), 0, h->stream, h->h);
            HCK(h, hipGetLastError());
            RC(do_derive(h));
            for (int i = 0; i < have; i++) {
                hipLaunchKernelGGL(k_ball_cw, dim3(grid_for(C.V, WG, 1, 1024)), dim3(WG), 0, h->stream, h->h, i, 1);
                HCK(h, hipGetLastError());
            }
            for (int i = have; i < M; i++) {
                hipLaunchKernelGGL(k_greedy_pick, dim3(1), dim3(WG), 0, h->stream, h->h, i);
                hipLaunchKernelGGL(k_ball_cw, dim3(grid_for(C.V, WG, 1, 1024)), dim3(WG), 0, h->stream, h->h, i, 1);
                HCK(h, hipGetLastError());
            }
        } else {
            hipLaunchKernelGGL(k_init_random, dim3(1), dim3(RNG_SPLIT * 64), 0, h->stream, h->h, have);
            hipLaunchKernelGGL(k_pack_code, dim3(grid_for(M, WG, 1, 256)), dim3(WG), 0, h->stream, h->h);
            HCK(h, hipGetLastError());
            RC(do_recount(h));
        }
    }
    HCK(h, hipMemcpyAsync(C.best, C.code, (size_t)M * n, hipMemcpyDeviceToDevice, h->stream));
    Ctl now;
    RC(read_ctl(h, &now));
    RC(set_ctl<int64_t>(h, offsetof(Ctl, best_unc), now.unc));
    RC(set_ctl<int32_t>(h, offsetof(Ctl, ndirty), 0));
    if (now.unc <= list_keep(C)) RC(list_rebuild(h));
    else RC(set_ctl<int32_t>(h, offsetof(Ctl, ul_valid), 0));
    return COV_OK;
}

// Whether the steps keep the leaving counts (th) or walk the candidates'
// spheres.  Kept, every word whose count crosses 1 costs a few atomics on its
// coverer's row, so it pays where such words are rare: at an average coverage
// M*V/N of 4 or more (e^-4 (4 + 8) = 22% of the words have count 1 or 2, and
// far fewer at the usual 6 to 7).  More kicks per stagnation than the stale
// list holds also mean walking.
bool th_wanted(const cov_cell* h)
{
    if (h->kick + 1 > STALE_MAX || h->leave_mode == COV_LEAVE_WALK) return false;
    if (h->leave_mode == COV_LEAVE_KEPT) return true;
    const Cell& C = h->h;
    return (double)C.M * (double)C.V >= 4.0 * (double)C.N;
}

// before a search launch on stream st: (re)build the coverer XOR and the
// leaving counts unless they are up to date, or switch the steps to walking
int th_build(cov_cell* h, hipStream_t st)
{
    const int on = th_wanted(h) ? 1 : 0;
    if (on && h->th_valid) return COV_OK;
    if (!on && h->th_on == 0) return COV_OK;
    const Cell& C = h->h;
    if (on) {
        HCK(h, hipMemsetAsync(C.xor32, 0, (size_t)((C.N + 1) / 2) * 4, st));
        hipLaunchKernelGGL(k_xor_build, dim3(grid_for((int64_t)C.M * C.V, WG, 1, 8192)), dim3(WG), 0, st, h->h);
        HCK(h, hipGetLastError());
        HCK(h, hipMemsetAsync(C.th, 0, (size_t)C.M * THW * 4, st));
        HCK(h, hipMemsetAsync(C.dth, 0, (size_t)C.M * THW * 4, st));
        hipLaunchKernelGGL(k_th_all, dim3(grid_for((int64_t)C.M * C.F, WAVES, 1, 8192)), dim3(WG), 0, st, h->h);
        HCK(h, hipGetLastError());
    }
    hipLaunchKernelGGL(k_th_setup, dim3(1), dim3(64), 0, st, h->h, on);
    HCK(h, hipGetLastError());
    h->th_valid = on != 0;
    h->th_on = on;
    return COV_OK;
}

int cov_search_launch(cov_cell_t* h, int64_t nsteps, int mode)
{
    if (h && !h->cnt_valid) { h->err = "coverage counts not built (cov_recount)"; return COV_ESTATE; }
    if (h) h->dist_valid = false;
    if (!h || nsteps < 0 || (mode != COV_MODE_PLAIN && mode != COV_MODE_GRAPH)) return COV_EPARAM;
    if (nsteps == 0) return COV_OK;
    RC(th_build(h, h->stream));
    hipLaunchKernelGGL(k_begin_run, dim3(1), dim3(WG), 0, h->stream, h->h, nsteps);
    HCK(h, hipGetLastError());
    int64_t left = nsteps;
    if (mode == COV_MODE_GRAPH && !h->timing) {
        RC(ensure_graph(h));
        while (left >= h->g_steps) {
            HCK(h, hipGraphLaunch(h->gexec, h->stream));
            left -= h->g_steps;
        }
    }
    for (; left > 0; left--) RC(launch_step(h));
    return COV_OK;
}

int64_t cov_search_run(cov_cell_t* h, int64_t nsteps, int mode, int64_t* trace)
{
    if (!h || nsteps < 0) return COV_EPARAM;
    // the trace buffer and its capacity live in the control block, so captured
    // graphs stay valid whether or not a run records a trace
    if (trace && h->tcap < nsteps) {
        HCK(h, hipStreamSynchronize(h->stream));    // a launched run may still write the old buffer
        if (h->trace) (void)hipFree(h->trace);
        h->trace = nullptr;
        h->tcap = 0;
        HCK(h, hipMalloc((void**)&h->trace, (size_t)nsteps * TRW * 8));
        h->tcap = nsteps;
    }
    int64_t* tp = trace ? h->trace : nullptr;
    const int64_t tc = trace ? h->tcap : 0;
    HCK(h, hipMemcpyAsync(CTL_PTR(h, trace), &tp, 8, hipMemcpyHostToDevice, h->stream));
    HCK(h, hipMemcpyAsync(CTL_PTR(h, tcap), &tc, 8, hipMemcpyHostToDevice, h->stream));
    RC(cov_search_launch(h, nsteps, mode));
    HCK(h, hipStreamSynchronize(h->stream));
    Ctl c;
    RC(read_ctl(h, &c));
    if (c.fault == 1) { h->err = "inconsistent counters: no candidate move"; return COV_ESTATE; }
    if (c.fault == 2) { h->err = "candidate sample larger than the rng ring allows"; return COV_ELIMIT; }
    const int64_t ran = c.tpos;
    if (trace && ran > 0) {
        HCK(h, hipMemcpyAsync(trace, h->trace, (size_t)std::min(ran, nsteps) * TRW * 8, hipMemcpyDeviceToHost,
                              h->stream));
        HCK(h, hipStreamSynchronize(h->stream));
    }
    return ran;
}

int cov_search_state(cov_cell_t* h, cov_state_t* st)
{
    if (!h || !st) return COV_EPARAM;
    Ctl c;
    RC(read_ctl(h, &c));
    st->it = c.it;
    st->uncovered = c.unc;
    st->best_uncovered = c.best_unc;
    st->kicks = c.nkicks;
    st->since_improve = c.since;
    st->draws = c.ruse;
    st->list_n = c.ul_n;
    st->done = c.done;
    st->list_valid = c.ul_valid;
    st->fault = c.fault;
    st->reserved = 0;
    return COV_OK;
}

int cov_set_launch(cov_cell_t* h, int eval_groups, int commit_groups)
{
    if (!h || eval_groups < 1 || commit_groups < 1 || eval_groups > 65535 || commit_groups > 65535)
        return COV_EPARAM;
    h->eval_groups = eval_groups;
    h->commit_groups = commit_groups;
    return COV_OK;
}

int cov_set_leave_mode(cov_cell_t* h, int mode)
{
    if (!h || mode < COV_LEAVE_AUTO || mode > COV_LEAVE_KEPT) return COV_EPARAM;
    h->leave_mode = mode;
    h->th_valid = false;
    return COV_OK;
}

int cov_get_launch(const cov_cell_t* h, int* eval_groups, int* commit_groups)
{
    if (!h) return COV_EPARAM;
    if (eval_groups) *eval_groups = h->eval_groups;
    if (commit_groups) *commit_groups = h->commit_groups;
    return COV_OK;
}

int cov_set_transform_tile(cov_cell_t* h, int lds_bytes, int threads)
{
    if (!h || lds_bytes < 0 || (threads != 0 && threads != 256 && threads != 512 && threads != 1024))
        return COV_EPARAM;
    h->dt_threads = threads;
    DtPass pass[MAXN + 1];
    const Cell& C = h->h;
    if (lds_bytes && dt_plan(C.q, C.n, C.N, lds_bytes, pass, true, true, h->ncu) <= 0) return COV_EPARAM;
    h->dt_lds = lds_bytes;
    return COV_OK;
}

int cov_timing(cov_cell_t* h, int on)
{
    if (!h) return COV_EPARAM;
    RC(timing_collect(h));
    h->timing = on != 0;
    for (int k = 0; k < 7; k++) { h->ms[k] = 0; h->launches[k] = 0; }
    return COV_OK;
}

int cov_timing_read(cov_cell_t* h, double* ms, int64_t* launches, int nkinds)
{
    if (!h || nkinds < 0 || nkinds > 7) return COV_EPARAM;
    RC(timing_collect(h));
    for (int k = 0; k < nkinds; k++) {
        if (ms) ms[k] = h->ms[k];
        if (launches) launches[k] = h->launches[k];
    }
    return COV_OK;
}

int cov_search_phases(cov_cell_t* h, double* us, int n)
{
    if (!h || n < 0 || n > 12 || (n && !us)) return COV_EPARAM;
    Ctl c;
    RC(read_ctl(h, &c));
    for (int k = 0; k < n; k++) us[k] = (double)c.tph[k] * 1e6 / h->clk_hz;
    return COV_OK;
}

int cov_search_limits(cov_cell_t* h, int64_t target, double seconds)
{
    if (!h || !(seconds == seconds)) return COV_EPARAM;
    if (target < -1) target = -1;
    int64_t ticks = 0;
    if (seconds > 0) {
        const double t = seconds * h->clk_hz;
        ticks = t >= 9.0e18 ? (int64_t)9e18 : std::max<int64_t>(1, (int64_t)t);
    } else if (seconds < 0) {
        ticks = 1;                        // already past: no further step begins
    }
    if (h->used) HCK(h, hipStreamWaitEvent(h->stream, h->used, 0));
    hipLaunchKernelGGL(k_set_limits, dim3(1), dim3(64), 0, h->stream, h->h, target, ticks);
    HCK(h, hipGetLastError());
    HCK(h, hipStreamSynchronize(h->stream));
    return COV_OK;
}

int cov_search_ttt(cov_cell_t* h, int64_t* ttt_it, double* age_s)
{
    if (!h) return COV_EPARAM;
    Ctl c;
    RC(read_ctl(h, &c));
    hipLaunchKernelGGL(k_clock, dim3(1), dim3(64), 0, h->stream, h->d_small);
    HCK(h, hipGetLastError());
    int64_t now = 0;
    HCK(h, hipMemcpyAsync(&now, h->d_small, 8, hipMemcpyDeviceToHost, h->stream));
    HCK(h, hipStreamSynchronize(h->stream));
    if (ttt_it) *ttt_it = c.ttt_it;
    if (age_s) *age_s = c.ttt_it >= 0 ? (double)(now - c.t_ttt) / h->clk_hz : -1.0;
    return COV_OK;
}

int cov_query(cov_cell_t* h)
{
    if (!h) return COV_EPARAM;
    const hipError_t e = hipStreamQuery(h->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    HCK(h, e);
    return 0;
}

int cov_search_bytes(cov_cell_t* h, int64_t* eval_bytes, int64_t* commit_bytes, int64_t* move_bytes,
                     int64_t* move_steps)
{
    if (!h) return COV_EPARAM;
    Ctl c;
    RC(read_ctl(h, &c));
    if (eval_bytes) *eval_bytes = c.eval_bytes;
    if (commit_bytes) *commit_bytes = c.commit_bytes;
    if (move_bytes) *move_bytes = c.move_bytes;
    if (move_steps) *move_steps = c.move_steps;
    return COV_OK;
}

// ============================================================ batches
int cov_batch_create(cov_cell_t* const* cells, int k, cov_batch_t** out)
{
    if (!cells || k < 1 || !out) return COV_EPARAM;
    *out = nullptr;
    for (int i = 0; i < k; i++) {
        if (!cells[i] || !cells[i]->cnt_valid) return COV_EPARAM;
        for (int j = 0; j < i; j++)
            if (cells[j] == cells[i]) return COV_EPARAM;
    }
    cov_batch* b = new cov_batch();
    auto fail = [&](int rc) {
        cov_batch_destroy(b);
        return rc;
    };
    b->cells.assign(cells, cells + k);
    for (cov_cell* c : b->cells) {
        const hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return fail(-(100 + (int)e));
    }
    {
        const hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
        if (e != hipSuccess) { b->stream = nullptr; return fail(-(100 + (int)e)); }
    }
    std::vector<Cell> tab(k);
    std::vector<int32_t> eo(k + 1, 0), co(k + 1, 0);
    for (int i = 0; i < k; i++) {
        tab[i] = b->cells[i]->h;
        b->accw.push_back((const void*)b->cells[i]->h.accw);
        eo[i + 1] = eo[i] + b->cells[i]->eval_groups;
        co[i + 1] = co[i] + b->cells[i]->commit_groups;
    }
    if (eo[k] > (1 << 24) || co[k] > (1 << 24)) return fail(COV_ELIMIT);
    b->eval_total = eo[k];
    b->commit_total = co[k];
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, sizeof(Cell) * k);
    if (e != hipSuccess) return fail(COV_ENOMEM);
    b->d_cells = (Cell*)v;
    e = hipMalloc(&v, 4 * (k + 1));
    if (e != hipSuccess) return fail(COV_ENOMEM);
    b->d_eoff = (int32_t*)v;
    e = hipMalloc(&v, 4 * (k + 1));
    if (e != hipSuccess) return fail(COV_ENOMEM);
    b->d_coff = (int32_t*)v;
    if ((e = hipMemcpyAsync(b->d_cells, tab.data(), sizeof(Cell) * k, hipMemcpyHostToDevice, b->stream)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(b->d_eoff, eo.data(), 4 * (k + 1), hipMemcpyHostToDevice, b->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(b->d_coff, co.data(), 4 * (k + 1), hipMemcpyHostToDevice, b->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(b->stream)) != hipSuccess)
        return fail(-(100 + (int)e));
    *out = b;
    return COV_OK;
}

int cov_batch_destroy(cov_batch_t* b)
{
    if (!b) return COV_EPARAM;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->d_cells) (void)hipFree(b->d_cells);
    if (b->d_eoff) (void)hipFree(b->d_eoff);
    if (b->d_coff) (void)hipFree(b->d_coff);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return COV_OK;
}

const char* cov_batch_last_error(const cov_batch_t* b)
{
    return b ? b->err.c_str() : "null batch";
}

int cov_batch_size(const cov_batch_t* b)
{
    return b ? (int)b->cells.size() : COV_EPARAM;
}

int cov_batch_launch(cov_batch_t* b, int64_t nsteps, int mode)
{
    if (!b || nsteps < 0 || (mode != COV_MODE_PLAIN && mode != COV_MODE_GRAPH)) return COV_EPARAM;
    if (nsteps == 0) return COV_OK;
    const int k = (int)b->cells.size();
    for (int i = 0; i < k; i++)
        if ((const void*)b->cells[i]->h.accw != b->accw[i]) {
            b->err = "cell " + std::to_string(i) + " switched to --wide after the batch was made";
            return COV_ESTATE;
        }
    for (cov_cell* c : b->cells) {
        c->dist_valid = false;
        const int rc = th_build(c, b->stream);
        if (rc != COV_OK) { b->err = c->err; return rc; }
    }
    // (k_begin_run_b also turns the cells' trace records off)
    hipLaunchKernelGGL(k_begin_run_b, dim3(k), dim3(WG), 0, b->stream, b->d_cells, nsteps);
    HCK(b, hipGetLastError());
    int64_t left = nsteps;
    if (mode == COV_MODE_GRAPH) {
        if (!b->gexec) {
            hipGraph_t g;
            HCK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
            for (int st = 0; st < GRAPH_STEPS; st++) {
                hipLaunchKernelGGL(k_eval_b, dim3(b->eval_total), dim3(WG), 0, b->stream, b->d_cells, b->d_eoff, k);
                hipLaunchKernelGGL(k_commit_b, dim3(b->commit_total), dim3(WG), 0, b->stream, b->d_cells, b->d_coff,
                                   k);
                hipLaunchKernelGGL(k_post_b, dim3(k), dim3(WG), 0, b->stream, b->d_cells);
            }
            HCK(b, hipStreamEndCapture(b->stream, &g));
            const hipError_t e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HCK(b, e);
        }
        while (left >= GRAPH_STEPS) {
            HCK(b, hipGraphLaunch(b->gexec, b->stream));
            left -= GRAPH_STEPS;
        }
    }
    for (; left > 0; left--) {
        hipLaunchKernelGGL(k_eval_b, dim3(b->eval_total), dim3(WG), 0, b->stream, b->d_cells, b->d_eoff, k);
        hipLaunchKernelGGL(k_commit_b, dim3(b->commit_total), dim3(WG), 0, b->stream, b->d_cells, b->d_coff, k);
        hipLaunchKernelGGL(k_post_b, dim3(k), dim3(WG), 0, b->stream, b->d_cells);
        HCK(b, hipGetLastError());
    }
    for (cov_cell* c : b->cells) HCK(b, hipEventRecord(c->used, b->stream));
    return COV_OK;
}

int cov_batch_sync(cov_batch_t* b)
{
    if (!b) return COV_EPARAM;
    HCK(b, hipStreamSynchronize(b->stream));
    for (size_t i = 0; i < b->cells.size(); i++) {
        Ctl c;
        const int rc = read_ctl(b->cells[i], &c);
        if (rc != COV_OK) { b->err = b->cells[i]->err; return rc; }
        if (c.fault == 1) { b->err = "cell " + std::to_string(i) + ": inconsistent counters"; return COV_ESTATE; }
        if (c.fault == 2) { b->err = "cell " + std::to_string(i) + ": candidate sample too large"; return COV_ELIMIT; }
    }
    return COV_OK;
}

int cov_batch_query(cov_batch_t* b)
{
    if (!b) return COV_EPARAM;
    const hipError_t e = hipStreamQuery(b->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    b->err = std::string("hipStreamQuery: ") + hipGetErrorString(e);
    return -(100 + (int)e);
}

int cov_batch_run(cov_batch_t* b, int64_t nsteps, int mode)
{
    RC(cov_batch_launch(b, nsteps, mode));
    return cov_batch_sync(b);
}

}  // extern "C"
