// cov_batch.h -- many cells advanced by the same launches.
//
// A batched step is the three launches of the single-cell step (evaluator,
// selection and commit, post) over the workgroups of every cell at once: the grid is cut into
// consecutive ranges, cell i owning workgroups [off[i], off[i+1]), and each
// workgroup runs the single-cell body for its cell with its index inside the
// range.  Every cell keeps its own control block, counters, rng ring and tabu
// table, so a cell's trajectory is the one it has when run alone; the cells'
// narrow single-workgroup phases run side by side on different CUs.
#pragma once
#include "cov_step.h"

namespace cov {

// the cell whose range holds workgroup b (off is ascending, off[0] = 0)
__device__ inline int batch_cell(const int32_t* __restrict__ off, int ncell, int b)
{
    int lo = 0, hi = ncell - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(WG, 4) k_eval_b(const Cell* __restrict__ cells, const int32_t* __restrict__ off,
                                                  int ncell)
{
    const int i = batch_cell(off, ncell, blockIdx.x);
    eval_body(cells[i], blockIdx.x - off[i], off[i + 1] - off[i]);
}

__global__ void __launch_bounds__(WG) k_commit_b(const Cell* __restrict__ cells, const int32_t* __restrict__ off,
                                                 int ncell)
{
    const int i = batch_cell(off, ncell, blockIdx.x);
    commit_body(cells[i], blockIdx.x - off[i], off[i + 1] - off[i]);
}

// one workgroup per cell
__global__ void __launch_bounds__(WG) k_post_b(const Cell* __restrict__ cells)
{
    post_body(cells[blockIdx.x]);
}

// one workgroup per cell
__global__ void __launch_bounds__(WG) k_begin_run_b(const Cell* __restrict__ cells, int64_t nsteps)
{
    begin_run_body(cells[blockIdx.x], nsteps, true);
}

}  // namespace cov
