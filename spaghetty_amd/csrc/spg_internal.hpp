// spg_internal.hpp -- host-side declarations shared by the library's sources.
#pragma once
#include <string>
#include <vector>

#include "spaghetty_amd.h"
#include "spg_plan.hpp"

namespace spg {

// Problem after validation, extent-1 removal and fusion (fastest first).
struct Canon {
    int r = 0;
    int n[MAXR] = {0};
    int out_axis[MAXR] = {0};  // input axis at output position q
    int64_t nelem = 0;
};

// Validates (rank, dims, perm) in NWChem order and canonicalises it.
int canonicalize(int rank, const int* dims, const int* perm, Canon* c);

// Mode for a given five-case g (beta).
inline int mode_of(double beta) { return beta == 0.0 ? MODE_SORT : (beta == 1.0 ? MODE_ACC : MODE_AXPBY); }

// Builds the plan for a canonical problem.  `tp` may force any subset of
// the tile parameters (fields < 0 / 0 mean "choose").  Returns a status.
int build_plan(const Canon& c, int mode, const spg_tile_params* tp, Plan* out);

// Heuristic cost of a plan (byte-equivalents of HBM traffic + overheads).
double plan_cost(const Plan& P);

// Key of a canonical problem in the dispatch table.
std::string dispatch_key(const Canon& c, int mode);

// Tuned parameters for a key, if the dispatch table has them.
bool dispatch_lookup(const std::string& key, spg_tile_params* out, bool reused);

// Human/JSON description of a plan.
std::string describe(const Plan& P);

// Executes a plan on the host with exactly the kernels' index math; used
// only to check plans (coverage and values) without a GPU.
int emulate(const Plan& P, const double* in, double* out, std::vector<int>* hits);

// Thread-local last error.
void set_error(int code, const char* msg);

// Plan for a call (cached); `resolve` fixes the launch geometry for the
// current device.
int get_plan(int rank, const int* dims, const int* perm, double alpha, double beta,
             const spg_tile_params* tp, Plan* out, bool resolve);

// The plan a per-call sort of these arrays runs (spg_sort_with's planning:
// tuned or default tile, single elements for arrays that are only 8-byte
// aligned), with its argument checks; with `resolve` the launch geometry is
// fixed for the current device, else launch_plan fixes it at launch.
int plan_for_call(const double* unsorted, double* sorted, int rank, const int* dims, const int* perm,
                  double factor, double beta, const spg_tile_params* params, Plan* out, bool resolve);
// The same plan for arrays whose 16-byte alignment is known but whose
// addresses are not (no argument checks).
int plan_for_alignment(bool aligned16, int rank, const int* dims, const int* perm, double factor, double beta,
                       const spg_tile_params* params, Plan* out, bool resolve);
// Enqueues a planned sort (kernels.hip).
int launch_plan(const Plan& P, const double* in, double* out, hipStream_t st);

// ---- batched walk (batch.hip), shared by the batched and serving forms ----
constexpr int kBatchBlock = 256;
constexpr int kBatchSlots = 8;
constexpr int kBatchTile = kBatchBlock * kBatchSlots;  // 2048 elements
// Access flags of the batched walk.  Tiles are small transposes whose edges
// write partial cache lines that the L2 merges, so access is ordinary
// (temporal), and unpipelined: the kernel then takes 166 VGPRs, three
// workgroups of 256 threads fit a CU (two LDS buffers of up to 2,048
// elements each), and on the 2,837-tile 4D coupled-cluster mix it moves
// 4.05 TB/s, against 3.96 with 4,096-element pair tiles (their LDS fits
// only two workgroups per CU), 3.87 pipelined (203 VGPRs: two workgroups)
// and 4.03-4.05 with one LDS buffer.  A batch's default.
constexpr int kBatchDefaultFlags = 0;
// The serving form's flags: its groups run on several streams at once, and
// pipelining measured best there (saturated 2,000-request mix: 3.58-3.60
// TB/s, against 3.53 unpipelined).
constexpr int kServeFlags = SPG_TILE_PIPE;
// Flags a batch accepts.
constexpr int kBatchFlagMask = SPG_TILE_NT | SPG_TILE_PIPE | SPG_TILE_LDS1 | SPG_BATCH_WIDE | SPG_BATCH_INLINE;
// Tasks of at least this many elements fill the device on their own: a
// batch (and a serving group) launches each of them on its own per-call
// plan (tuned tiles of up to 12,288 elements, 768 threads), after the
// persistent launch over the smaller ones.
constexpr int64_t kBatchBigElems = 1 << 22;
// Resident workgroups per CU a batched launch is sized for.
constexpr int kBatchMaxPerCU = 3;


constexpr uint32_t kChunkTiles = 8;  // consecutive tiles of one task per work item

struct BatchChunk {
    uint32_t task, first, count, pad_;
};

struct BatchTask {
    Plan P;
    const double* in;
    double* out;
    uint32_t ntiles;  // tiles of this task in the batch walk
    uint32_t chunk;   // rows plans: elements of a row per tile
    uint32_t nch;     // rows plans: tiles per row
    uint32_t pad_;
};


// Tile parameters every batched task is planned with.
spg_tile_params batch_tile_params();
// Rows plans of batched tasks need at least this many contiguous elements
// per row; shorter shared rows are planned as tiles.
constexpr int kBatchRowsMin = 256;
// Plan of a batched task (batch_tile_params; short rows planned as tiles).
int batch_plan(int rank, const int* dims, const int* perm, double factor, double beta, Plan* out);
// Fills a batch task from a plan; SPG_EINVAL if the plan's tile is too big.
int make_batch_task(const Plan& P, const double* in, double* out, BatchTask* t);
// Task classes beyond single elements, 4-pair tiles and rows; a batch runs
// on the k_batch instantiation that covers its tasks (0: the lean one).
constexpr int kPathWide = 1;     // tiles of 8 pairs per thread (SPG_BATCH_WIDE)
constexpr int kPathClassic = 2;  // slot offsets beyond a buffer resource's range
// Workgroups one batched launch can keep resident for LDS buffers of `half`
// elements on the current device.
int batch_grid_capacity(uint32_t half, int flags, int paths);
// Cuts the tiles of `n` tasks into chunks of up to kChunkTiles consecutive
// tiles of one task, in task order, picks the kernel instantiation (paths)
// and sizes the grid: the resident capacity for the tasks' largest tile,
// capped at G_max when G_max > 0.
int schedule_batch(const BatchTask* tasks, size_t n, int64_t G_max, int flags,
                   std::vector<BatchChunk>& chunks, uint32_t* half, uint32_t* total, int* G_out,
                   int* paths);
// Enqueues the persistent batched kernel.
hipError_t launch_batch_kernel(const BatchTask* d_tasks, const BatchChunk* d_chunks,
                               uint32_t nchunks, int G, uint32_t half, int flags, int paths,
                               hipStream_t st);

}  // namespace spg
