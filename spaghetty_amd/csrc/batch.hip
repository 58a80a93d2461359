// batch.hip -- the batched form: thousands of independent, uneven sorts
// (coupled-cluster 4D/6D tiles) executed by ONE persistent launch.
//
// Host side (spg_batch_commit): every task gets its own Plan (tile box
// limited to kBatchTile elements, or twice that for tasks that move element
// pairs when SPG_BATCH_WIDE is set, so one kernel instantiation serves all
// tasks), and each task's tiles are cut into chunks of up to kChunkTiles
// consecutive tiles, listed in task order.  Device side (k_batch): the
// persistent workgroups take the chunks round-robin and recompute their
// per-thread slot offsets only when a chunk belongs to another task than
// the previous one.  Results are identical to per-call launches: the same
// plan index math and the same per-element update, and no reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "spg_internal.hpp"
#include "spg_tile.hpp"
#include "spg_vtile.hpp"

namespace spg {

__device__ inline double update_rt(int mode, double x, double y, double f, double g)
{
    if (mode == MODE_SORT) return update<MODE_SORT>(x, y, f, g);
    if (mode == MODE_ACC) return update<MODE_ACC>(x, y, f, g);
    return update<MODE_AXPBY>(x, y, f, g);
}

// Waves per SIMD a k_batch instantiation is compiled for.  The lean kernel
// (no 8-pair path, no classic path) takes 108 VGPRs unpipelined, so four
// 256-thread workgroups fit a CU; with both paths it takes 166 (three).
__host__ __device__ constexpr int batch_waves(bool pipe, int paths)
{
    return paths == 0 ? (pipe ? 3 : 4) : (pipe ? 2 : 3);
}

// Workgroup w takes chunks w, w+G, w+2G, ... of the chunk list, so at any
// moment the grid works on neighbouring chunks (DRAM and L2 locality, as in
// the per-call walk); a chunk is a run of consecutive tiles of one task, so
// the per-thread slot offsets are recomputed at most once per chunk.
// PATHS: the task classes this instantiation can run besides single
// elements, 4-pair tiles and rows (kPathWide: 8-pair tiles; kPathClassic:
// tiles whose slot offsets exceed a buffer resource's range).  Every path
// costs registers for the whole kernel, so the host launches the leanest
// instantiation that covers a batch.
template <bool NT, bool PIPE, int PATHS>
__global__ __launch_bounds__(kBatchBlock) __attribute__((amdgpu_waves_per_eu(batch_waves(PIPE, PATHS)))) void k_batch(const BatchTask* __restrict__ tasks,
                                                       const BatchChunk* __restrict__ chunks,
                                                       uint32_t nchunks, uint32_t half)
{
    constexpr bool kWide = (PATHS & kPathWide) != 0;
    constexpr bool kClassic = (PATHS & kPathClassic) != 0;
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    uint32_t buf = 0;
    uint32_t last = 0xffffffffu;
    // branch-free tasks: 4 or 8 pairs, or 8 single slots; the second mask
    // words serve pairs over odd extents (4 pairs)
    VSlots<kBatchSlots, true> S;
    for (uint32_t j = blockIdx.x; j < nchunks; j += gridDim.x) {
        const BatchChunk c = chunks[j];
        const BatchTask& T = tasks[c.task];
        const Plan& P = T.P;
        const double* __restrict__ in = T.in;
        double* __restrict__ out = T.out;
        const uint32_t g = c.first, gend = c.first + c.count;
        if (P.kind == KIND_TILED && P.tk == 1) {
            constexpr int K2 = kBatchSlots / 2;
            const bool pairs = P.vec == 2;
            const bool wide = kWide && pairs && P.slots > K2;  // 8 pairs: tiles up to 2 * kBatchTile
            const bool odd = pairs && !wide && P.pair8;       // pairs over odd extents
            if (c.task != last) {
                if (wide)
                    vslots_init<2, kBatchSlots>(P, tid, kBatchBlock, S);
                else if (odd)
                    vslots_init<2, K2, VSlots<kBatchSlots, true>, true>(P, tid, kBatchBlock, S);
                else if (pairs)
                    vslots_init<2, K2>(P, tid, kBatchBlock, S);
                else
                    vslots_init<1, kBatchSlots>(P, tid, kBatchBlock, S);
                last = c.task;
            }
            if (odd) {
                if (P.mode == MODE_SORT)
                    buf = vtile_run<2, K2, MODE_SORT, NT, NT, PIPE, true>(P, in, out, g, gend, 1u, lds, buf, half, S);
                else if (P.mode == MODE_ACC)
                    buf = vtile_run<2, K2, MODE_ACC, NT, NT, PIPE, true>(P, in, out, g, gend, 1u, lds, buf, half, S);
                else
                    buf = vtile_run<2, K2, MODE_AXPBY, NT, NT, PIPE, true>(P, in, out, g, gend, 1u, lds, buf, half,
                                                                        S);
            } else if (wide) {
                if (P.mode == MODE_SORT)
                    buf = vtile_run<2, kBatchSlots, MODE_SORT, NT, NT, PIPE>(P, in, out, g, gend, 1u, lds, buf,
                                                                            half, S);
                else if (P.mode == MODE_ACC)
                    buf = vtile_run<2, kBatchSlots, MODE_ACC, NT, NT, PIPE>(P, in, out, g, gend, 1u, lds, buf,
                                                                           half, S);
                else
                    buf = vtile_run<2, kBatchSlots, MODE_AXPBY, NT, NT, PIPE>(P, in, out, g, gend, 1u, lds,
                                                                             buf, half, S);
            } else if (pairs) {
                if (P.mode == MODE_SORT)
                    buf = vtile_run<2, K2, MODE_SORT, NT, NT, PIPE>(P, in, out, g, gend, 1u, lds, buf, half, S);
                else if (P.mode == MODE_ACC)
                    buf = vtile_run<2, K2, MODE_ACC, NT, NT, PIPE>(P, in, out, g, gend, 1u, lds, buf, half, S);
                else
                    buf = vtile_run<2, K2, MODE_AXPBY, NT, NT, PIPE>(P, in, out, g, gend, 1u, lds, buf, half, S);
            } else {
                if (P.mode == MODE_SORT)
                    buf = vtile_run<1, kBatchSlots, MODE_SORT, NT, NT, PIPE>(P, in, out, g, gend, 1u, lds, buf,
                                                                            half, S);
                else if (P.mode == MODE_ACC)
                    buf = vtile_run<1, kBatchSlots, MODE_ACC, NT, NT, PIPE>(P, in, out, g, gend, 1u, lds, buf,
                                                                           half, S);
                else
                    buf = vtile_run<1, kBatchSlots, MODE_AXPBY, NT, NT, PIPE>(P, in, out, g, gend, 1u, lds,
                                                                             buf, half, S);
            }
        } else if (P.kind == KIND_TILED) {
            if constexpr (!kClassic) continue;  // never scheduled on this instantiation
            // classic tasks (slot offsets beyond a buffer resource's range)
            // recompute their slots every chunk: keeping them across chunks
            // would hold a second slot set in registers for the whole kernel
            uint32_t r_off[kBatchSlots], r_lds[kBatchSlots], r_mk[kBatchSlots];
            uint32_t w_off[kBatchSlots], w_lds[kBatchSlots], w_mk[kBatchSlots];
#pragma unroll
            for (int k = 0; k < kBatchSlots; ++k) {
                const int e = tid + k * kBatchBlock;
                read_slot(P, e, r_off[k], r_lds[k], r_mk[k]);
                write_slot(P, e, w_off[k], w_lds[k], w_mk[k]);
            }
            last = 0xffffffffu;
            if (P.mode == MODE_SORT)
                buf = tile_run<kBatchSlots, MODE_SORT, (NT ? 1 : 0), PIPE>(P, in, out, g, gend, 0u, 1u, lds, buf,
                                                                 half, r_off, r_lds, r_mk, w_off,
                                                                 w_lds, w_mk);
            else if (P.mode == MODE_ACC)
                buf = tile_run<kBatchSlots, MODE_ACC, (NT ? 1 : 0), PIPE>(P, in, out, g, gend, 0u, 1u, lds, buf,
                                                                half, r_off, r_lds, r_mk, w_off,
                                                                w_lds, w_mk);
            else
                buf = tile_run<kBatchSlots, MODE_AXPBY, (NT ? 1 : 0), PIPE>(P, in, out, g, gend, 0u, 1u, lds,
                                                                  buf, half, r_off, r_lds, r_mk,
                                                                  w_off, w_lds, w_mk);
        } else {
            // rows plan: tile = (row, chunk of T.chunk contiguous elements)
            const uint32_t C = (uint32_t)P.C;
            const double f = P.alpha, gb = P.beta;
            const int mode = P.mode;
            for (uint32_t local = g; local < gend; ++local) {
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
                    if (ok[k]) x[k] = ld<NT>(in + bi + col);
                }
                if (mode != MODE_SORT) {
#pragma unroll
                    for (int k = 0; k < kBatchSlots; ++k)
                        if (ok[k]) y[k] = ld<NT>(out + bo + ch * T.chunk + tid + k * kBatchBlock);
                }
#pragma unroll
                for (int k = 0; k < kBatchSlots; ++k)
                    if (ok[k])
                        st<NT>(out + bo + ch * T.chunk + tid + k * kBatchBlock,
                               update_rt(mode, x[k], y[k], f, gb));
            }
        }
    }
}

typedef void (*BatchFn)(const BatchTask*, const BatchChunk*, uint32_t, uint32_t);

template <int PATHS>
BatchFn batch_fn_p(int flags)
{
    const bool nt = flags & SPG_TILE_NT, pipe = flags & SPG_TILE_PIPE;
    if (nt) return pipe ? k_batch<true, true, PATHS> : k_batch<true, false, PATHS>;
    return pipe ? k_batch<false, true, PATHS> : k_batch<false, false, PATHS>;
}

BatchFn batch_fn(int flags, int paths)
{
    return paths ? batch_fn_p<kPathWide | kPathClassic>(flags) : batch_fn_p<0>(flags);
}

int task_paths(const BatchTask& t)
{
    if (t.P.kind != KIND_TILED) return 0;
    if (t.P.tk != 1) return kPathClassic;
    return (t.P.vec == 2 && t.P.slots > kBatchSlots / 2) ? kPathWide : 0;
}

spg_tile_params batch_tile_params()
{
    spg_tile_params tp;
    spg_tile_params_default(&tp);
    tp.max_tile = kBatchTile;
    tp.block = kBatchBlock;
    tp.flags = 0;
    // one block per tile where the two fastest axes swap places: tiles of
    // several blocks ran the 2,837-tile 4D coupled-cluster mix 0.8% slower
    // (the batch fills the device with many tasks' tiles anyway)
    tp.tw = 1;
    return tp;
}

int batch_plan(int rank, const int* dims, const int* perm, double factor, double beta, Plan* out)
{
    spg_tile_params tp = batch_tile_params();
    int rc = get_plan(rank, dims, perm, factor, beta, &tp, out, false);
    if (rc != SPG_OK || out->kind != KIND_ROWS || out->rank < 2 || out->C >= kBatchRowsMin) return rc;
    // a rows tile of a short row leaves most of the workgroup idle: plan
    // the task as tiles over several rows instead (the 2,837-tile 4D
    // coupled-cluster mix ran 1.3% faster so)
    Plan T;
    tp.kind = KIND_TILED;
    if (get_plan(rank, dims, perm, factor, beta, &tp, &T, false) == SPG_OK) *out = T;
    return SPG_OK;
}

int make_batch_task(const Plan& P, const double* in, double* out, BatchTask* t)
{
    std::memset(t, 0, sizeof *t);
    t->P = P;
    t->in = in;
    t->out = out;
    if (P.kind == KIND_TILED) {
        // pair tasks of the branch-free path may take 8 pairs per thread
        const bool pairs = P.tk == 1 && P.vec == 2;
        if (P.V > (pairs ? 2 * kBatchTile : kBatchTile) || P.block != kBatchBlock) return SPG_EINVAL;
        // pairs over odd extents run with 4 pairs per thread (the 8-pair
        // path does not split partial tiles); their pairs are 8-byte
        // aligned whatever the arrays' alignment
        if (pairs && P.pair8 && P.V > kBatchTile) return SPG_EINVAL;
        if (pairs && !P.pair8 && ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15u)) {
            // as in a per-call launch, arrays that are only 8-byte aligned
            // move single elements (a tile of more than kBatchTile elements
            // cannot)
            if (P.V > kBatchTile) return SPG_EINVAL;
            t->P.vec = 1;
            t->P.slots = (P.V + kBatchBlock - 1) / kBatchBlock;
        }
        t->ntiles = (uint32_t)P.ntiles;
    } else {
        const uint32_t C = (uint32_t)P.C;
        t->chunk = C < (uint32_t)kBatchTile ? C : (uint32_t)kBatchTile;
        t->nch = (C + t->chunk - 1) / t->chunk;
        t->ntiles = (uint32_t)P.ntiles * t->nch;
    }
    return SPG_OK;
}

int batch_grid_capacity(uint32_t half, int flags, int paths)
{
    const BatchFn fn = batch_fn(flags, paths);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    const int lds = (int)(((flags & SPG_TILE_LDS1) ? 1 : 2) * half * sizeof(double));
    // one occupancy query per device, instantiation and LDS size (the
    // serving engine sizes every group's grid)
    static std::mutex mu;
    static std::map<std::tuple<int, BatchFn, int>, int> cache;
    std::lock_guard<std::mutex> lk(mu);
    const auto key = std::make_tuple(dev, fn, lds);
    const auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    if (lds > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(fn),
                                                     kBatchBlock, (size_t)lds) != hipSuccess || occ <= 0)
        occ = 1;
    // more resident workgroups do not pay: on the 2,837-tile 4D mix the lean
    // instantiation ran 3.92-3.95 TB/s with four per CU (14.1 waves resident)
    // against 3.98-4.03 with three
    occ = std::min(occ, kBatchMaxPerCU);
    cache[key] = cus * occ;
    return cus * occ;
}

int schedule_batch(const BatchTask* tasks, size_t n, int64_t G_max, int flags,
                   std::vector<BatchChunk>& chunks, uint32_t* half_out, uint32_t* total, int* G_out,
                   int* paths_out)
{
    chunks.clear();
    uint64_t tot = 0;
    uint32_t half = 1;
    int paths = 0;
    for (size_t i = 0; i < n; ++i) {
        const BatchTask& t = tasks[i];
        paths |= task_paths(t);
        tot += t.ntiles;
        if (tot > 0x7fffffffULL) return SPG_EINVAL_DIM;
        if (t.P.kind == KIND_TILED)
            half = std::max<uint32_t>(half, t.P.tk == 1 ? (uint32_t)t.P.half : (uint32_t)t.P.lds_tile);
        for (uint32_t f = 0; f < t.ntiles; f += kChunkTiles) {
            BatchChunk c;
            c.task = (uint32_t)i;
            c.first = f;
            c.count = std::min<uint32_t>(kChunkTiles, t.ntiles - f);
            c.pad_ = 0;
            chunks.push_back(c);
        }
    }
    if (paths) paths = kPathWide | kPathClassic;  // the two instantiations: lean and full
    int64_t G = batch_grid_capacity(half, flags, paths);
    if (G_max > 0 && G > G_max) G = G_max;
    if (G > (int64_t)chunks.size()) G = (int64_t)chunks.size();
    if (G < 1) G = 1;
    *half_out = half;
    *total = (uint32_t)tot;
    *G_out = (int)G;
    *paths_out = paths;
    return SPG_OK;
}

hipError_t launch_batch_kernel(const BatchTask* d_tasks, const BatchChunk* d_chunks,
                               uint32_t nchunks, int G, uint32_t half, int flags, int paths,
                               hipStream_t st)
{
    // SPG_TILE_LDS1: one LDS buffer per workgroup (the kernel sees half 0)
    const bool one = flags & SPG_TILE_LDS1;
    const size_t lds = (one ? 1 : 2) * (size_t)half * sizeof(double);
    const BatchFn fn = batch_fn(flags, paths);
    if (lds > 65536) {
        // more than the default dynamic LDS limit: raise it for this kernel
        // (once per kernel and size; the serving loop launches without the
        // occupancy query that otherwise sets it)
        static std::mutex mu;
        static size_t raised[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        std::lock_guard<std::mutex> lk(mu);
        const int slot = (flags & 3) | (paths ? 4 : 0);
        if (raised[slot] < lds) {
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
            raised[slot] = lds;
        }
    }
    hipLaunchKernelGGL(fn, dim3((unsigned)G), dim3(kBatchBlock), lds, st, d_tasks, d_chunks, nchunks,
                       one ? 0u : half);
    return hipGetLastError();
}

}  // namespace spg

using namespace spg;

struct spg_batch {
    std::vector<BatchTask> tasks;  // every task, planned for k_batch, in add order
    std::vector<std::pair<uintptr_t, uintptr_t>> outs, ins;  // byte ranges, per task
    std::vector<uint32_t> wave_first{0};  // first task of each wave
    int64_t bytes = 0;       // algorithmic bytes of one launch
    int64_t elems = 0;
    bool committed = false;
    int dev = -1;
    // after commit: per wave, one k_batch launch over the tasks it runs (a
    // slice of `run` and of the chunk list), then the wave's big tasks, each
    // launched on its own per-call plan
    struct Wave {
        uint32_t task0 = 0, chunk0 = 0, nchunks = 0, tiles = 0;
        int grid = 0;
        uint32_t half = 0;   // elements of one LDS buffer
        int paths = 0;       // k_batch instantiation (task classes it must run)
        uint32_t big0 = 0, nbig = 0;
    };
    std::vector<Wave> waves;
    std::vector<BatchTask> run;  // the tasks k_batch runs, wave after wave
    struct Big {
        Plan P;
        const double* in;
        double* out;
    };
    std::vector<Big> bigs;
    BatchTask* d_tasks = nullptr;
    BatchChunk* d_chunks = nullptr;
    uint32_t nchunks = 0;
    uint32_t total_tiles = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int flags = kBatchDefaultFlags;
    struct Args {
        int rank;
        int dims[MAXR], perm[MAXR];
        double factor, beta;
    };
    std::vector<Args> args;  // per task, to re-plan at commit (SPG_BATCH_WIDE, big tasks)
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
    if (b->d_chunks) (void)hipFree(b->d_chunks);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    b->d_tasks = nullptr;
    b->d_chunks = nullptr;
}

bool any_overlap(std::vector<std::pair<uintptr_t, uintptr_t>> v)
{
    std::sort(v.begin(), v.end());
    for (size_t i = 1; i < v.size(); ++i)
        if (v[i].first < v[i - 1].second) return true;
    return false;
}

}  // namespace

namespace {

// ---- the batched form without device allocation (spg_batch_run_noalloc) ----
// The work list lives in this file-scope device array: every GPU the library
// runs on holds one copy until the process ends.  Host code never uses the
// array's name as an address: at a device's first call the array's address
// on that device is taken with hipGetSymbolAddress and kept per device.
constexpr size_t kNoallocBytes = SPG_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
static_assert(sizeof(BatchTask) % alignof(BatchTask) == 0 && sizeof(BatchChunk) == 16, "work-list records");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

namespace {

std::mutex g_scratch_mu;                       // guards the table and the staging copy
void* g_scratch_dev[kMaxDevices] = {nullptr};  // device address of scratch, per device
unsigned char g_scratch_host[kNoallocBytes];   // host image of the work list

// Device address of scratch on the current device (taken once per device).
// Called with g_scratch_mu held.
int scratch_address(int dev, unsigned char** out)
{
    if (dev < 0 || dev >= kMaxDevices) return SPG_EINVAL;
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return SPG_EHIP;
        }
        g_scratch_dev[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_dev[dev]);
    return SPG_OK;
}

// Host half of a commit: checks, wave by wave, that outputs are disjoint
// from each other and from every input; sets tasks of at least
// kBatchBigElems elements aside to run on their own per-call plans (unless
// SPG_BATCH_INLINE); re-plans pair tasks for SPG_BATCH_WIDE; and builds each
// wave's chunk list, grid and kernel instantiation.  Allocates no device
// memory; `chunks` receives every wave's chunks, wave after wave.
int batch_prepare(spg_batch* b, std::vector<BatchChunk>& chunks)
{
    const size_t n = b->tasks.size();
    std::vector<uint32_t> bounds = b->wave_first;
    bounds.push_back((uint32_t)n);
    b->waves.clear();
    b->run.clear();
    b->bigs.clear();
    chunks.clear();
    b->total_tiles = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return bfail(SPG_EHIP, "no device");
    b->dev = dev;
    spg_tile_params wide = batch_tile_params();
    wide.max_tile = 2 * kBatchTile;
    for (size_t w = 0; w + 1 < bounds.size(); ++w) {
        const uint32_t t0 = bounds[w], t1 = bounds[w + 1];
        if (t1 <= t0) continue;
        // outputs must be disjoint from each other and from every input
        std::vector<std::pair<uintptr_t, uintptr_t>> outs(b->outs.begin() + t0, b->outs.begin() + t1);
        if (any_overlap(outs)) return bfail(SPG_EALIAS, "two outputs overlap");
        std::sort(outs.begin(), outs.end());
        for (uint32_t i = t0; i < t1; ++i) {
            const auto& r = b->ins[i];
            auto it = std::upper_bound(outs.begin(), outs.end(), std::make_pair(r.first, ~(uintptr_t)0));
            if (it != outs.end() && it->first < r.second) return bfail(SPG_EALIAS, "an output overlaps an input");
            if (it != outs.begin() && std::prev(it)->second > r.first)
                return bfail(SPG_EALIAS, "an output overlaps an input");
        }
        spg_batch::Wave wv;
        wv.task0 = (uint32_t)b->run.size();
        wv.big0 = (uint32_t)b->bigs.size();
        for (uint32_t i = t0; i < t1; ++i) {
            BatchTask t = b->tasks[i];
            const spg_batch::Args& a = b->args[i];
            if (!(b->flags & SPG_BATCH_INLINE) && t.P.nelem >= kBatchBigElems) {
                spg_batch::Big g;
                const int rc = plan_for_call(t.in, t.out, a.rank, a.dims, a.perm, a.factor, a.beta,
                                             nullptr, &g.P, true);
                if (rc != SPG_OK) return bfail(rc, "planning a big task failed");
                g.in = t.in;
                g.out = t.out;
                b->bigs.push_back(g);
                continue;
            }
            if ((b->flags & SPG_BATCH_WIDE) && t.P.kind == KIND_TILED && t.P.tk == 1 && t.P.vec == 2) {
                // tasks that run as pairs take tiles of up to 2 * kBatchTile elements
                Plan P;
                BatchTask wt;
                if (get_plan(a.rank, a.dims, a.perm, a.factor, a.beta, &wide, &P, false) == SPG_OK &&
                    P.kind == KIND_TILED && P.tk == 1 && P.vec == 2 && make_batch_task(P, t.in, t.out, &wt) == SPG_OK)
                    t = wt;
            }
            b->run.push_back(t);
        }
        wv.nbig = (uint32_t)b->bigs.size() - wv.big0;
        const uint32_t nrun = (uint32_t)b->run.size() - wv.task0;
        if (nrun) {
            std::vector<BatchChunk> wc;
            uint32_t half = 1, tot = 0;
            int G = 1, paths = 0;
            // grid: the resident capacity for the wave's largest LDS tile
            const int rc = schedule_batch(b->run.data() + wv.task0, nrun, 0, b->flags, wc, &half, &tot, &G,
                                          &paths);
            if (rc != SPG_OK) return bfail(rc, "too many tiles in one wave");
            wv.chunk0 = (uint32_t)chunks.size();
            wv.nchunks = (uint32_t)wc.size();
            wv.tiles = tot;
            wv.grid = G;
            wv.half = half;
            wv.paths = paths;
            chunks.insert(chunks.end(), wc.begin(), wc.end());
            b->total_tiles += tot;
        }
        b->waves.push_back(wv);
    }
    b->nchunks = (uint32_t)chunks.size();
    return SPG_OK;
}

// Enqueues every wave: its k_batch launch over the work list at (d_tasks,
// d_chunks), then its big tasks.
int batch_enqueue(const spg_batch* b, const BatchTask* d_tasks, const BatchChunk* d_chunks, hipStream_t st)
{
    for (const spg_batch::Wave& w : b->waves) {
        if (w.tiles) {
            const hipError_t e = launch_batch_kernel(d_tasks + w.task0, d_chunks + w.chunk0, w.nchunks, w.grid,
                                                     w.half, b->flags, w.paths, st);
            if (e != hipSuccess) return bfail(SPG_EHIP, hipGetErrorString(e));
        }
        for (uint32_t k = 0; k < w.nbig; ++k) {
            const spg_batch::Big& g = b->bigs[w.big0 + k];
            const int rc = launch_plan(g.P, g.in, g.out, st);
            if (rc != SPG_OK) return rc;
        }
    }
    return SPG_OK;
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
    Plan P;
    int rc = batch_plan(rank, dims, perm, factor, beta, &P);
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
    spg_batch::Args a;
    std::memset(&a, 0, sizeof a);
    a.rank = rank;
    for (int q = 0; q < rank; ++q) {
        a.dims[q] = dims[q];
        a.perm[q] = perm[q];
    }
    a.factor = factor;
    a.beta = beta;
    b->args.push_back(a);
    b->elems += t.P.nelem;
    b->bytes += t.P.nelem * (int64_t)sizeof(double) * (t.P.mode == MODE_SORT ? 2 : 3);
    return SPG_OK;
}

int spg_batch_next_wave(spg_batch* b)
{
    if (!b) return bfail(SPG_EINVAL, "null batch");
    if (b->committed) return bfail(SPG_ESTATE, "batch already committed");
    if (b->wave_first.back() < (uint32_t)b->tasks.size()) b->wave_first.push_back((uint32_t)b->tasks.size());
    return SPG_OK;
}

int spg_batch_set_flags(spg_batch* b, int flags)
{
    if (!b) return bfail(SPG_EINVAL, "null batch");
    if (b->committed) return bfail(SPG_ESTATE, "batch already committed");
    if (flags < 0 || (flags & ~kBatchFlagMask)) return bfail(SPG_EINVAL, "invalid flags");
    b->flags = flags;
    return SPG_OK;
}


int spg_batch_commit(spg_batch* b, void* stream)
{
    if (!b) return bfail(SPG_EINVAL, "null batch");
    if (b->committed) return bfail(SPG_ESTATE, "batch already committed");
    std::vector<BatchChunk> chunks;
    const int rc = batch_prepare(b, chunks);
    if (rc != SPG_OK) return rc;
    const size_t n = b->run.size();
    free_device(b);
    hipStream_t st = (hipStream_t)stream;
    if (hipMalloc(&b->d_tasks, sizeof(BatchTask) * std::max<size_t>(n, 1)) != hipSuccess ||
        hipMalloc(&b->d_chunks, sizeof(BatchChunk) * std::max<size_t>(chunks.size(), 1)) != hipSuccess) {
        (void)hipGetLastError();
        free_device(b);
        return bfail(SPG_ENOMEM, "work-list allocation failed");
    }
    if (n && hipMemcpyAsync(b->d_tasks, b->run.data(), sizeof(BatchTask) * n, hipMemcpyHostToDevice, st) !=
                 hipSuccess)
        return bfail(SPG_EHIP, "work-list upload failed");
    if (!chunks.empty() && hipMemcpyAsync(b->d_chunks, chunks.data(), sizeof(BatchChunk) * chunks.size(),
                                          hipMemcpyHostToDevice, st) != hipSuccess)
        return bfail(SPG_EHIP, "work-list upload failed");
    if (hipStreamSynchronize(st) != hipSuccess) return bfail(SPG_EHIP, "stream synchronize failed");
    b->committed = true;
    return SPG_OK;
}

int spg_batch_launch(spg_batch* b, void* stream)
{
    if (!b) return bfail(SPG_EINVAL, "null batch");
    if (!b->committed) return bfail(SPG_ESTATE, "batch not committed");
    return batch_enqueue(b, b->d_tasks, b->d_chunks, (hipStream_t)stream);
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

int spg_batch_run_noalloc(int ntasks, const double* const* unsorted, double* const* sorted,
                          const int* ranks, const int* dims, const int* perms, const double* factors,
                          const double* betas, int flags, void* stream)
{
    if (ntasks < 0 || (ntasks > 0 && (!unsorted || !sorted || !ranks || !dims || !perms || !factors || !betas)))
        return bfail(SPG_EINVAL, "null argument");
    if (ntasks == 0) return SPG_OK;
    if (flags < 0 || (flags & ~kBatchFlagMask)) return bfail(SPG_EINVAL, "invalid flags");
    // the same validation and planning as spg_batch_add / spg_batch_commit,
    // on a batch description that lives only on the host
    spg_batch b;
    b.flags = flags;
    for (int i = 0; i < ntasks; ++i) {
        const int rc = spg_batch_add(&b, unsorted[i], sorted[i], ranks[i], dims + (size_t)i * SPG_MAX_RANK,
                                     perms + (size_t)i * SPG_MAX_RANK, factors[i], betas[i]);
        if (rc != SPG_OK) return rc;
    }
    std::vector<BatchChunk> chunks;
    int rc = batch_prepare(&b, chunks);
    if (rc != SPG_OK) return rc;
    const size_t n = b.run.size();
    const size_t o_ch = sizeof(BatchTask) * n;
    const size_t need = o_ch + sizeof(BatchChunk) * chunks.size();
    if (need > kNoallocBytes) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "work list of %zu bytes exceeds the %zu-byte no-allocation scratch",
                      need, kNoallocBytes);
        return bfail(SPG_EINVAL, msg);
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    unsigned char* d = nullptr;
    rc = scratch_address(b.dev, &d);
    if (rc != SPG_OK) return bfail(rc, "cannot take the address of the scratch array");
    hipStream_t st = (hipStream_t)stream;
    if (need) {
        std::memcpy(g_scratch_host, b.run.data(), o_ch);
        std::memcpy(g_scratch_host + o_ch, chunks.data(), sizeof(BatchChunk) * chunks.size());
        // a copy from pageable host memory: the runtime has taken the bytes
        // when the call returns, so the next call may refill the host image
        if (hipMemcpyAsync(d, g_scratch_host, need, hipMemcpyHostToDevice, st) != hipSuccess)
            return bfail(SPG_EHIP, "work-list upload failed");
    }
    return batch_enqueue(&b, reinterpret_cast<const BatchTask*>(d), reinterpret_cast<const BatchChunk*>(d + o_ch),
                         st);
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
    int grid = 0;
    size_t lds = 0;
    for (const spg_batch::Wave& w : b->waves) {
        grid = std::max(grid, w.grid);
        lds = std::max(lds, ((b->flags & SPG_TILE_LDS1) ? 1 : 2) * (size_t)w.half * sizeof(double));
    }
    char tmp[320];
    std::snprintf(tmp, sizeof tmp,
                  "{\"tasks\":%zu,\"tiles\":%u,\"chunks\":%u,\"grid\":%d,\"lds_bytes\":%zu,"
                  "\"bytes\":%lld,\"elems\":%lld,\"committed\":%d,\"waves\":%zu,\"big_tasks\":%zu}",
                  b->tasks.size(), b->total_tiles, b->nchunks, grid, lds, (long long)b->bytes,
                  (long long)b->elems, b->committed ? 1 : 0, b->committed ? b->waves.size() : b->wave_first.size(),
                  b->bigs.size());
    if (buflen <= (int)std::strlen(tmp)) return bfail(SPG_EINVAL, "buffer too small");
    std::strcpy(buf, tmp);
    return SPG_OK;
}

}  // extern "C"
