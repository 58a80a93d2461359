// batch.hip -- the batched form: thousands of independent, uneven sorts
// (coupled-cluster 4D/6D tiles) executed by ONE persistent launch.
//
// Host side (spg_batch_commit): every task gets its own Plan (tile box
// limited to kBatchTile elements so one kernel instantiation serves all
// tasks), the tasks' tiles are concatenated into one global tile sequence,
// and the sequence is cut into one contiguous, element-balanced range per
// workgroup.  Device side (k_batch): a workgroup walks its range; it
// recomputes its per-thread slot offsets only when it crosses into the next
// task, so the index setup is paid a few times per workgroup, not per tile.
// Results are identical to per-call launches: the same plan index math and
// the same per-element update, and no reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "spg_internal.hpp"

namespace spg {

__device__ inline double update_rt(int mode, double x, double y, double f, double g)
{
    if (mode == MODE_SORT) return update<MODE_SORT>(x, y, f, g);
    if (mode == MODE_ACC) return update<MODE_ACC>(x, y, f, g);
    return update<MODE_AXPBY>(x, y, f, g);
}

__global__ __launch_bounds__(kBatchBlock) void k_batch(const BatchTask* __restrict__ tasks,
                                                       const uint32_t* __restrict__ prefix,
                                                       const uint32_t* __restrict__ wg_start,
                                                       const uint32_t* __restrict__ wg_task,
                                                       uint32_t half)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    uint32_t g = wg_start[blockIdx.x];
    const uint32_t gend = wg_start[blockIdx.x + 1];
    uint32_t ti = wg_task[blockIdx.x];
    uint32_t buf = 0;
    while (g < gend) {
        const BatchTask& T = tasks[ti];
        const Plan& P = T.P;
        const uint32_t t0 = prefix[ti];
        const uint32_t tend = min(gend, prefix[ti + 1]);
        const double f = P.alpha, gb = P.beta;
        const int mode = P.mode;
        const double* __restrict__ in = T.in;
        double* __restrict__ out = T.out;
        if (P.kind == KIND_TILED) {
            uint32_t r_off[kBatchSlots], r_lds[kBatchSlots], r_mk[kBatchSlots];
            uint32_t w_off[kBatchSlots], w_lds[kBatchSlots], w_mk[kBatchSlots];
#pragma unroll
            for (int k = 0; k < kBatchSlots; ++k) {
                const int e = tid + k * kBatchBlock;
                read_slot(P, e, r_off[k], r_lds[k], r_mk[k]);
                write_slot(P, e, w_off[k], w_lds[k], w_mk[k]);
            }
            for (; g < tend; ++g) {
                uint32_t bin, bout, lim;
                tile_origin(P, g - t0, bin, bout, lim);
                double* L = lds + buf;
                double v[kBatchSlots];
                // unconditional loads (invalid slots read the tile origin)
#pragma unroll
                for (int k = 0; k < kBatchSlots; ++k)
                    v[k] = ld<kBatchNT>(in + bin + (slot_ok(r_mk[k], lim) ? r_off[k] : 0u));
#pragma unroll
                for (int k = 0; k < kBatchSlots; ++k)
                    if (slot_ok(r_mk[k], lim)) L[r_lds[k]] = v[k];
                __syncthreads();
                double x[kBatchSlots], y[kBatchSlots];
                bool ok[kBatchSlots];
#pragma unroll
                for (int k = 0; k < kBatchSlots; ++k) {
                    ok[k] = slot_ok(w_mk[k], lim);
                    y[k] = 0.0;
                    x[k] = L[ok[k] ? w_lds[k] : 0u];
                }
                if (mode != MODE_SORT) {
#pragma unroll
                    for (int k = 0; k < kBatchSlots; ++k)
                        y[k] = ld<kBatchNT>(out + bout + (ok[k] ? w_off[k] : 0u));
                }
#pragma unroll
                for (int k = 0; k < kBatchSlots; ++k)
                    if (ok[k]) st<kBatchNT>(out + bout + w_off[k], update_rt(mode, x[k], y[k], f, gb));
                buf = buf ? 0u : half;
            }
        } else {
            // rows plan: tile = (row, chunk of T.chunk contiguous elements)
            const uint32_t C = (uint32_t)P.C;
            for (; g < tend; ++g) {
                const uint32_t local = g - t0;
                const uint32_t row = local / T.nch, ch = local - row * T.nch;
                uint32_t bi, bo;
                row_origin(P, row, bi, bo);
                double x[kBatchSlots], y[kBatchSlots];
                bool ok[kBatchSlots];
#pragma unroll
                for (int k = 0; k < kBatchSlots; ++k) {
                    const uint32_t col = ch * T.chunk + tid + k * kBatchBlock;
                    ok[k] = col < C && (uint32_t)(tid + k * kBatchBlock) < T.chunk;
                    x[k] = 0.0;
                    y[k] = 0.0;
                    if (ok[k]) x[k] = ld<kBatchNT>(in + bi + col);
                }
                if (mode != MODE_SORT) {
#pragma unroll
                    for (int k = 0; k < kBatchSlots; ++k)
                        if (ok[k]) y[k] = ld<kBatchNT>(out + bo + ch * T.chunk + tid + k * kBatchBlock);
                }
#pragma unroll
                for (int k = 0; k < kBatchSlots; ++k)
                    if (ok[k])
                        st<kBatchNT>(out + bo + ch * T.chunk + tid + k * kBatchBlock,
                                 update_rt(mode, x[k], y[k], f, gb));
            }
        }
        ++ti;
    }
}


spg_tile_params batch_tile_params()
{
    spg_tile_params tp;
    spg_tile_params_default(&tp);
    tp.max_tile = kBatchTile;
    tp.block = kBatchBlock;
    tp.flags = 0;
    return tp;
}

int make_batch_task(const Plan& P, const double* in, double* out, BatchTask* t)
{
    std::memset(t, 0, sizeof *t);
    t->P = P;
    t->in = in;
    t->out = out;
    if (P.kind == KIND_TILED) {
        if (P.V > kBatchTile) return SPG_EINVAL;
        t->ntiles = (uint32_t)P.ntiles;
    } else {
        const uint32_t C = (uint32_t)P.C;
        t->chunk = C < (uint32_t)kBatchTile ? C : (uint32_t)kBatchTile;
        t->nch = (C + t->chunk - 1) / t->chunk;
        t->ntiles = (uint32_t)P.ntiles * t->nch;
    }
    return SPG_OK;
}

int batch_grid_capacity(uint32_t half)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    const int lds = (int)(2 * half * sizeof(double));
    if (lds > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_batch),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(k_batch),
                                                     kBatchBlock, (size_t)lds) != hipSuccess || occ <= 0)
        occ = 1;
    return cus * occ;
}

int schedule_batch(const BatchTask* tasks, size_t n, int64_t G_max, std::vector<uint32_t>& prefix,
                   std::vector<uint32_t>& wg_start, std::vector<uint32_t>& wg_task, uint32_t* half_out,
                   uint32_t* total, int* G_out)
{
    prefix.assign(n + 1, 0);
    std::vector<double> wprefix(n + 1, 0.0);
    uint64_t tot = 0;
    uint32_t half = 1;
    for (size_t i = 0; i < n; ++i) {
        const BatchTask& t = tasks[i];
        tot += t.ntiles;
        if (tot > 0x7fffffffULL) return SPG_EINVAL_DIM;
        prefix[i + 1] = (uint32_t)tot;
        wprefix[i + 1] = wprefix[i] + (double)t.P.nelem * (t.P.mode == MODE_SORT ? 2.0 : 3.0);
        if (t.P.kind == KIND_TILED) half = std::max<uint32_t>(half, (uint32_t)(t.P.Tx * t.P.lds_row));
    }
    int64_t G = G_max > 0 ? G_max : batch_grid_capacity(half);
    if (G > (int64_t)tot) G = (int64_t)tot;
    if (G < 1) G = 1;
    // element-balanced contiguous ranges: workgroup w starts at the tile
    // where the cumulative traffic reaches w/G of the total
    wg_start.assign(G + 1, 0);
    wg_task.assign(G + 1, 0);
    const double W = wprefix[n];
    size_t ti = 0;
    for (int64_t w = 0; w <= G; ++w) {
        if (w == G || n == 0) {
            wg_start[w] = (uint32_t)tot;
            wg_task[w] = (uint32_t)n;
            continue;
        }
        const double target = W * (double)w / (double)G;
        while (ti + 1 < n && wprefix[ti + 1] <= target) ++ti;
        const BatchTask& t = tasks[ti];
        const double per_tile = (wprefix[ti + 1] - wprefix[ti]) / std::max<uint32_t>(t.ntiles, 1);
        uint32_t local = per_tile > 0 ? (uint32_t)((target - wprefix[ti]) / per_tile) : 0;
        if (local >= t.ntiles) local = t.ntiles ? t.ntiles - 1 : 0;
        uint32_t gs = prefix[ti] + local;
        if (w > 0 && gs < wg_start[w - 1]) gs = wg_start[w - 1];
        wg_start[w] = gs;
        // the task containing gs (an empty range leaves it unused)
        size_t tj = ti;
        while (tj < n && prefix[tj + 1] <= gs) ++tj;
        wg_task[w] = (uint32_t)tj;
    }
    *half_out = half;
    *total = (uint32_t)tot;
    *G_out = (int)G;
    return SPG_OK;
}

hipError_t launch_batch_kernel(const BatchTask* d_tasks, const uint32_t* d_prefix,
                               const uint32_t* d_wg_start, const uint32_t* d_wg_task, int G,
                               uint32_t half, hipStream_t st)
{
    const size_t lds = 2 * (size_t)half * sizeof(double);
    hipLaunchKernelGGL(k_batch, dim3((unsigned)G), dim3(kBatchBlock), lds, st, d_tasks, d_prefix,
                       d_wg_start, d_wg_task, half);
    return hipGetLastError();
}

}  // namespace spg

using namespace spg;

struct spg_batch {
    std::vector<BatchTask> tasks;
    std::vector<std::pair<uintptr_t, uintptr_t>> outs, ins;  // byte ranges
    int64_t bytes = 0;       // algorithmic bytes of one launch
    int64_t elems = 0;
    bool committed = false;
    int dev = -1;
    int grid = 0;
    uint32_t half = 0;       // elements of one LDS buffer
    BatchTask* d_tasks = nullptr;
    uint32_t* d_prefix = nullptr;
    uint32_t* d_wg_start = nullptr;
    uint32_t* d_wg_task = nullptr;
    uint32_t total_tiles = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
};

namespace {

int bfail(int rc, const char* msg)
{
    set_error(rc, msg);
    return rc;
}

void free_device(spg_batch* b)
{
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->d_tasks) (void)hipFree(b->d_tasks);
    if (b->d_prefix) (void)hipFree(b->d_prefix);
    if (b->d_wg_start) (void)hipFree(b->d_wg_start);
    if (b->d_wg_task) (void)hipFree(b->d_wg_task);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    b->d_tasks = nullptr;
    b->d_prefix = b->d_wg_start = b->d_wg_task = nullptr;
}

bool any_overlap(std::vector<std::pair<uintptr_t, uintptr_t>> v)
{
    std::sort(v.begin(), v.end());
    for (size_t i = 1; i < v.size(); ++i)
        if (v[i].first < v[i - 1].second) return true;
    return false;
}

}  // namespace

extern "C" {

int spg_batch_create(spg_batch** out)
{
    if (!out) return bfail(SPG_EINVAL, "null argument");
    *out = new spg_batch();
    return SPG_OK;
}

int spg_batch_add(spg_batch* b, const double* unsorted, double* sorted, int rank, const int* dims,
                  const int* perm, double factor, double beta)
{
    if (!b) return bfail(SPG_EINVAL, "null batch");
    if (b->committed) return bfail(SPG_ESTATE, "batch already committed");
    const spg_tile_params tp = batch_tile_params();
    Plan P;
    int rc = get_plan(rank, dims, perm, factor, beta, &tp, &P, false);
    if (rc != SPG_OK) return bfail(rc, spg_strerror(rc));
    if (P.nelem == 0) return SPG_OK;
    if (!unsorted || !sorted) return bfail(SPG_EINVAL, "null array");
    BatchTask t;
    if (make_batch_task(P, unsorted, sorted, &t) != SPG_OK)
        return bfail(SPG_EINVAL, "tile too large for the batched kernel");
    const uintptr_t len = (uintptr_t)t.P.nelem * sizeof(double);
    b->outs.push_back({(uintptr_t)sorted, (uintptr_t)sorted + len});
    b->ins.push_back({(uintptr_t)unsorted, (uintptr_t)unsorted + len});
    b->tasks.push_back(t);
    b->elems += t.P.nelem;
    b->bytes += t.P.nelem * (int64_t)sizeof(double) * (t.P.mode == MODE_SORT ? 2 : 3);
    return SPG_OK;
}

int spg_batch_commit(spg_batch* b, void* stream)
{
    if (!b) return bfail(SPG_EINVAL, "null batch");
    if (b->committed) return bfail(SPG_ESTATE, "batch already committed");
    // outputs must be disjoint from each other and from every input
    if (any_overlap(b->outs)) return bfail(SPG_EALIAS, "two outputs overlap");
    {
        std::vector<std::pair<uintptr_t, uintptr_t>> all = b->outs;
        std::sort(all.begin(), all.end());
        for (auto& r : b->ins) {
            auto it = std::upper_bound(all.begin(), all.end(), std::make_pair(r.first, ~(uintptr_t)0));
            if (it != all.end() && it->first < r.second) return bfail(SPG_EALIAS, "an output overlaps an input");
            if (it != all.begin() && std::prev(it)->second > r.first)
                return bfail(SPG_EALIAS, "an output overlaps an input");
        }
    }
    const size_t n = b->tasks.size();
    std::vector<uint32_t> prefix, wg_start, wg_task;
    uint32_t half = 1, tot = 0;
    int G = 1;
    // grid: the resident capacity for the batch's largest LDS tile
    int rc = schedule_batch(b->tasks.data(), n, 0, prefix, wg_start, wg_task, &half, &tot, &G);
    if (rc != SPG_OK) return bfail(rc, "too many tiles in one batch");
    b->total_tiles = tot;
    b->half = half;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return bfail(SPG_EHIP, "no device");
    b->dev = dev;
    b->grid = (int)G;
    free_device(b);
    hipStream_t st = (hipStream_t)stream;
    if (hipMalloc(&b->d_tasks, sizeof(BatchTask) * std::max<size_t>(n, 1)) != hipSuccess ||
        hipMalloc(&b->d_prefix, sizeof(uint32_t) * (n + 1)) != hipSuccess ||
        hipMalloc(&b->d_wg_start, sizeof(uint32_t) * (G + 1)) != hipSuccess ||
        hipMalloc(&b->d_wg_task, sizeof(uint32_t) * (G + 1)) != hipSuccess) {
        (void)hipGetLastError();
        free_device(b);
        return bfail(SPG_ENOMEM, "work-list allocation failed");
    }
    if (n && hipMemcpyAsync(b->d_tasks, b->tasks.data(), sizeof(BatchTask) * n, hipMemcpyHostToDevice, st) != hipSuccess)
        return bfail(SPG_EHIP, "work-list upload failed");
    if (hipMemcpyAsync(b->d_prefix, prefix.data(), sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(b->d_wg_start, wg_start.data(), sizeof(uint32_t) * (G + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(b->d_wg_task, wg_task.data(), sizeof(uint32_t) * (G + 1), hipMemcpyHostToDevice, st) != hipSuccess)
        return bfail(SPG_EHIP, "work-list upload failed");
    if (hipStreamSynchronize(st) != hipSuccess) return bfail(SPG_EHIP, "stream synchronize failed");
    b->committed = true;
    return SPG_OK;
}

int spg_batch_launch(spg_batch* b, void* stream)
{
    if (!b) return bfail(SPG_EINVAL, "null batch");
    if (!b->committed) return bfail(SPG_ESTATE, "batch not committed");
    if (b->total_tiles == 0) return SPG_OK;
    const hipError_t e = launch_batch_kernel(b->d_tasks, b->d_prefix, b->d_wg_start, b->d_wg_task,
                                             b->grid, b->half, (hipStream_t)stream);
    if (e != hipSuccess) return bfail(SPG_EHIP, hipGetErrorString(e));
    return SPG_OK;
}

int spg_batch_graph_launch(spg_batch* b, void* stream)
{
    if (!b) return bfail(SPG_EINVAL, "null batch");
    if (!b->committed) return bfail(SPG_ESTATE, "batch not committed");
    hipStream_t st = (hipStream_t)stream;
    if (!b->graph_exec) {
        // capture needs a non-null stream
        if (!st) return bfail(SPG_EINVAL, "graph capture needs a non-null stream");
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return bfail(SPG_EHIP, "begin capture failed");
        const int rc = spg_batch_launch(b, stream);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (rc != SPG_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess || !g) return bfail(SPG_EHIP, "end capture failed");
        b->graph = g;
        if (hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0) != hipSuccess)
            return bfail(SPG_EHIP, "graph instantiate failed");
    }
    if (hipGraphLaunch(b->graph_exec, st) != hipSuccess) return bfail(SPG_EHIP, "graph launch failed");
    return SPG_OK;
}

int spg_batch_destroy(spg_batch* b)
{
    if (!b) return SPG_OK;
    free_device(b);
    delete b;
    return SPG_OK;
}

int64_t spg_batch_bytes(const spg_batch* b) { return b ? b->bytes : 0; }

int spg_batch_info(const spg_batch* b, char* buf, int buflen)
{
    if (!b || !buf) return bfail(SPG_EINVAL, "null argument");
    char tmp[256];
    std::snprintf(tmp, sizeof tmp,
                  "{\"tasks\":%zu,\"tiles\":%u,\"grid\":%d,\"lds_bytes\":%zu,\"bytes\":%lld,"
                  "\"elems\":%lld,\"committed\":%d}",
                  b->tasks.size(), b->total_tiles, b->grid, 2 * (size_t)b->half * sizeof(double),
                  (long long)b->bytes, (long long)b->elems, b->committed ? 1 : 0);
    if (buflen <= (int)std::strlen(tmp)) return bfail(SPG_EINVAL, "buffer too small");
    std::strcpy(buf, tmp);
    return SPG_OK;
}

}  // extern "C"
