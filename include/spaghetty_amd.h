/*
 * spaghetty_amd.h -- C ABI of libspaghetty_amd.so, the MI355X (gfx950) tensor
 * sort library that drops in for spaghetty's generated tce_sort_4 library.
 *
 * Array convention (NWChem TCE, trunk/old_sort.f:33-95): dims are listed
 * slowest -> fastest (the LAST dim is stride-1; Fortran column-major
 * U(d,c,b,a) for dims a,b,c,d), the permutation is 1-based, and the input
 * element at 1-based coordinates id() lands at
 *     ib = id(p_R) + jd(p_R)*(id(p_{R-1})-1 + jd(p_{R-1})*(...))
 * of the output, whose extents are therefore (jd(p_1), ..., jd(p_R)).
 *     sort:             s(ib) = factor*u(ia)
 *     sort-accumulate:  s(ib) = s(ib) + factor*u(ia)   (two roundings)
 *     five-case form:   s(ib) = g*s(ib) + factor*u(ia) (g = beta)
 * Element counts are limited to 2^31-1 (32-bit index arithmetic, as the
 * reference's default INTEGER).
 */
#ifndef SPAGHETTY_AMD_H
#define SPAGHETTY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define SPG_OK           0
#define SPG_EINVAL_PERM  1  /* permutation is not a permutation of 1..rank  */
#define SPG_EINVAL_DIM   2  /* negative dim, rank out of range, > 2^31-1 elements */
#define SPG_EALIAS       3  /* input and output overlap                     */
#define SPG_ENOMEM       4  /* device allocation failed                     */
#define SPG_EHIP         5  /* HIP runtime error                            */
#define SPG_EINVAL       6  /* other invalid argument                       */
#define SPG_EBUDGET      7  /* request larger than the serving HBM budget   */
#define SPG_ESTATE       8  /* call not valid in the object's current state */

#define SPG_MAX_RANK 8

/* ---- NWChem drop-in entries (host pointers, synchronous) -----------------
 * Replace the generated dispatcher
 *     subroutine tce_sort_4(u,s,a,b,c,d,i,j,k,l,factor)
 * written by trunk/makelib.py:64-95 (and the hand-written
 * trunk/tce_sort4_switch.F:1-62); tce_sortacc_4 / tce_sort_6 /
 * tce_sortacc_6 keep NWChem's signatures (branches/spaghetty3/python/
 * tce_sort_hirata.F:62,143; branches/spaghetty2/tce_sortacc_hirata.F:49,120)
 * with the semantics of trunk/old_sort.f:33-95, 97-234, 512-574, 576-713.
 * Fortran passes everything by reference; the three spellings cover
 * gfortran (-funderscoring), -fno-underscoring and f2c-style manglings, as
 * trunk/tester_cutil.c:33-76 does for its helpers.  Like the reference
 * (makelib.py:90-91), an invalid permutation prints 'something is wrong...'
 * and leaves `sorted` untouched; spg_last_error() reports the cause. */
void tce_sort_4_(const double *unsorted, double *sorted, const int *a, const int *b,
                 const int *c, const int *d, const int *i, const int *j, const int *k,
                 const int *l, const double *factor);
void tce_sort_4(const double *unsorted, double *sorted, const int *a, const int *b,
                const int *c, const int *d, const int *i, const int *j, const int *k,
                const int *l, const double *factor);
void tce_sort_4__(const double *unsorted, double *sorted, const int *a, const int *b,
                  const int *c, const int *d, const int *i, const int *j, const int *k,
                  const int *l, const double *factor);
void tce_sortacc_4_(const double *unsorted, double *sorted, const int *a, const int *b,
                    const int *c, const int *d, const int *i, const int *j, const int *k,
                    const int *l, const double *factor);
void tce_sortacc_4(const double *unsorted, double *sorted, const int *a, const int *b,
                   const int *c, const int *d, const int *i, const int *j, const int *k,
                   const int *l, const double *factor);
void tce_sortacc_4__(const double *unsorted, double *sorted, const int *a, const int *b,
                     const int *c, const int *d, const int *i, const int *j, const int *k,
                     const int *l, const double *factor);
void tce_sort_6_(const double *unsorted, double *sorted, const int *a, const int *b,
                 const int *c, const int *d, const int *e, const int *f, const int *i,
                 const int *j, const int *k, const int *l, const int *m, const int *n,
                 const double *factor);
void tce_sort_6(const double *unsorted, double *sorted, const int *a, const int *b,
                const int *c, const int *d, const int *e, const int *f, const int *i,
                const int *j, const int *k, const int *l, const int *m, const int *n,
                const double *factor);
void tce_sort_6__(const double *unsorted, double *sorted, const int *a, const int *b,
                  const int *c, const int *d, const int *e, const int *f, const int *i,
                  const int *j, const int *k, const int *l, const int *m, const int *n,
                  const double *factor);
void tce_sortacc_6_(const double *unsorted, double *sorted, const int *a, const int *b,
                    const int *c, const int *d, const int *e, const int *f, const int *i,
                    const int *j, const int *k, const int *l, const int *m, const int *n,
                    const double *factor);
void tce_sortacc_6(const double *unsorted, double *sorted, const int *a, const int *b,
                   const int *c, const int *d, const int *e, const int *f, const int *i,
                   const int *j, const int *k, const int *l, const int *m, const int *n,
                   const double *factor);
void tce_sortacc_6__(const double *unsorted, double *sorted, const int *a, const int *b,
                     const int *c, const int *d, const int *e, const int *f, const int *i,
                     const int *j, const int *k, const int *l, const int *m, const int *n,
                     const double *factor);

/* Ranks 2 and 8 (NWChem tce_sort_2 / tce_sort_8 and their sortacc forms,
 * branches/spaghetty3/python/tce_sort_hirata.F:22-61, 303-559), with the
 * semantics of trunk/old_sort.f:10-31, 236-478, 489-510, 715-957; same
 * conventions as above. */
void tce_sort_2_(const double *unsorted, double *sorted, const int *a, const int *b,
                 const int *i, const int *j, const double *factor);
void tce_sort_2(const double *unsorted, double *sorted, const int *a, const int *b,
                const int *i, const int *j, const double *factor);
void tce_sort_2__(const double *unsorted, double *sorted, const int *a, const int *b,
                  const int *i, const int *j, const double *factor);
void tce_sortacc_2_(const double *unsorted, double *sorted, const int *a, const int *b,
                    const int *i, const int *j, const double *factor);
void tce_sortacc_2(const double *unsorted, double *sorted, const int *a, const int *b,
                   const int *i, const int *j, const double *factor);
void tce_sortacc_2__(const double *unsorted, double *sorted, const int *a, const int *b,
                     const int *i, const int *j, const double *factor);
void tce_sort_8_(const double *unsorted, double *sorted, const int *a, const int *b,
                 const int *c, const int *d, const int *e, const int *f, const int *g,
                 const int *h, const int *i, const int *j, const int *k, const int *l,
                 const int *m, const int *n, const int *o, const int *p, const double *factor);
void tce_sort_8(const double *unsorted, double *sorted, const int *a, const int *b,
                const int *c, const int *d, const int *e, const int *f, const int *g,
                const int *h, const int *i, const int *j, const int *k, const int *l,
                const int *m, const int *n, const int *o, const int *p, const double *factor);
void tce_sort_8__(const double *unsorted, double *sorted, const int *a, const int *b,
                  const int *c, const int *d, const int *e, const int *f, const int *g,
                  const int *h, const int *i, const int *j, const int *k, const int *l,
                  const int *m, const int *n, const int *o, const int *p, const double *factor);
void tce_sortacc_8_(const double *unsorted, double *sorted, const int *a, const int *b,
                    const int *c, const int *d, const int *e, const int *f, const int *g,
                    const int *h, const int *i, const int *j, const int *k, const int *l,
                    const int *m, const int *n, const int *o, const int *p, const double *factor);
void tce_sortacc_8(const double *unsorted, double *sorted, const int *a, const int *b,
                   const int *c, const int *d, const int *e, const int *f, const int *g,
                   const int *h, const int *i, const int *j, const int *k, const int *l,
                   const int *m, const int *n, const int *o, const int *p, const double *factor);
void tce_sortacc_8__(const double *unsorted, double *sorted, const int *a, const int *b,
                     const int *c, const int *d, const int *e, const int *f, const int *g,
                     const int *h, const int *i, const int *j, const int *k, const int *l,
                     const int *m, const int *n, const int *o, const int *p, const double *factor);

/* ---- status-returning entries ---------------------------------------------
 * Any rank 0..8.  The generated kernel family trans_<P>_loop_<L>_omp_any_f
 * (trunk/generator.py:107-159), its accumulate form (generator.py:143-144)
 * and the five-case (f,g) routines of trunk/generator2.py:133-443 and
 * trunk/generator6.py:137-220 (g = beta; beta 0 sorts, beta 1 accumulates)
 * all map here.  spg_sort_host takes host pointers and returns when `sorted`
 * holds the result; spg_sort takes device pointers and is ordered on
 * `stream` (a hipStream_t; NULL = the null stream). */
int spg_sort_host(const double *unsorted, double *sorted, int rank, const int *dims,
                  const int *perm, double factor, double beta);
int spg_sort(const double *unsorted, double *sorted, int rank, const int *dims,
             const int *perm, double factor, double beta, void *stream);

/* ---- the contraction around the sort (TTDT), on the device ---------------
 * The pattern NWChem's TCE wraps around tce_sort_4 for a two-tensor
 * contraction, e.g. i0(p3,p4,h1,h2) = Sum(p5,p6) v(p3,p4,p5,p6) *
 * t(p5,p6,h1,h2) in branches/old/python/macaroni/ccsd_t2_8_skeleton.F:
 * 162-183: A (dims adims, NWChem order) is sorted with aperm into a K x M
 * column-major matrix (its K fastest elements contiguous), B with bperm
 * into K x N, C' = alpha * A'^T B' (one DGEMM 'T','N', M x N column-major),
 * and C' -- read with dims cdims (product N*M, M fastest) -- is sorted
 * with cperm into C: C = beta*C + sorted(C') (beta 0 sorts, 1 accumulates
 * as tce_sortacc_4).  Ordered on `stream`; `work` holds M*K + K*N + M*N
 * doubles (NULL: a per-device buffer the library keeps; calls on different
 * streams that use it run one after the other on the device).  Device
 * pointers. */
int spg_ttdt(const double *a, int ra, const int *adims, const int *aperm, const double *b, int rb,
             const int *bdims, const int *bperm, double *c, int rc, const int *cdims,
             const int *cperm, int m, int n, int k, double alpha, double beta, double *work,
             void *stream);

/* ---- tile parameters (the autotuner's search space) ----------------------
 * kind: -1 choose, 0 LDS-tiled, 1 rows (shared contiguous axis, no LDS).
 * nx/ny: number of axes in the input-contiguous (X) / output-contiguous (Y)
 * group; tx/ty: chunk of the group's slowest axis; block: threads per
 * workgroup; grid_per_cu: persistent workgroups per CU (0 = occupancy);
 * xcd_remap: 1 keeps neighbouring tiles on one XCD; flags (-1 = choose):
 * SPG_TILE_NT uses nontemporal loads and stores (best for long contiguous
 * streams), SPG_TILE_NT_LOAD nontemporal input loads with ordinary output
 * access, SPG_TILE_PIPE overlaps the next tile's loads with the current
 * tile's stores.  LDS-tiled plans run the branch-free buffer-access kernel
 * with two elements (16 bytes) per access where the extents allow it;
 * SPG_TILE_SCALAR keeps it at one element, SPG_TILE_LDS1 gives it one LDS
 * buffer instead of two, SPG_TILE_CLASSIC selects the kernel with per-slot
 * branches instead.  Negative or zero fields mean
 * "choose". */
#define SPG_TILE_NT      1
#define SPG_TILE_PIPE    2
#define SPG_TILE_NT_LOAD 4
#define SPG_TILE_CLASSIC 8   /* per-slot branches and 8-byte accesses (k_tiled) */
#define SPG_TILE_SCALAR  16  /* k_vtile with one element per slot even where two fit */
#define SPG_TILE_LDS1    32  /* k_vtile with one LDS buffer (two barriers per tile) */
#define SPG_TILE_NOSPLIT 128 /* k_vtile pairs: LDS rows in tile order (no even/odd split) */
typedef struct spg_tile_params {
    int32_t kind;
    int32_t nx, tx, ny, ty;
    int32_t block;
    int32_t grid_per_cu;
    int32_t xcd_remap;
    int32_t max_tile;  /* largest tile in elements (0 = library default) */
    int32_t flags;
    int32_t walk_x;    /* tile walk: group width (0 = no grouping) */
    int32_t walk_mode; /* tile walk: digit order, 0..3 (see plan.cpp set_walk) */
    int32_t walk_chunk;/* tiles a workgroup takes per turn, consecutive in the walk (0/1 = one) */
    int32_t tw;        /* tiles of several blocks, where the two fastest axes of the input are the
                        * two fastest of the output, swapped (each tile of those two axes is then one
                        * contiguous block on both sides): a tile also spans tw values of the input's
                        * next axis (tw > 1) or -tw values of the output's next axis (tw < -1;
                        * the same axis where one is next on both sides); 0 = the planner's
                        * choice, 1 = one block per tile */
} spg_tile_params;

void spg_tile_params_default(spg_tile_params *p);

int spg_sort_with(const double *unsorted, double *sorted, int rank, const int *dims,
                  const int *perm, double factor, double beta,
                  const spg_tile_params *params, void *stream);

/* ---- dispatch table (replaces the loop-order table of
 * trunk/makelib.py:39-62) -------------------------------------------------
 * One JSON object per line: {"key": "...", "kind": k, "nx": .., "tx": ..,
 * "ny": .., "ty": .., "block": .., "grid_per_cu": .., "xcd_remap": .., ...}
 * with the fields of spg_tile_params; a missing "tw" reads as 1. 
 * spg_dispatch_key writes the key of a problem into buf. */
int spg_dispatch_load(const char *path);
int spg_dispatch_set(const char *key, const spg_tile_params *params);
int spg_dispatch_clear(void);

/* Plans for outputs written cold.  The main table is tuned for a caller that
 * sorts into the same output arrays again soon (NWChem's loops reuse their
 * scratch arrays): an output of up to ~256 MB is then still largely in the
 * device's last-level cache when it is written again, and the fastest plan
 * differs from the one for outputs written cold (a caller streaming through
 * many distinct arrays; the 54 (T) permutations at 16^6, sort and
 * accumulate: 64.4 instead of 69.8 us per call on average cold, 63.6
 * instead of 57.9 us reused).  spg_dispatch_load_streamed reads
 * such plans into a second table (the file dispatch_gfx950_streamed.jsonl
 * beside the library is read on first use, or the one SPG_DISPATCH_TABLE_
 * STREAMED names); spg_set_output_reuse(0) makes later calls consult it
 * first, spg_set_output_reuse(1) (the default) not; it returns the previous
 * setting and applies to the whole process.  SPG_OUTPUTS=streamed in the
 * environment starts the process with 0.  No counterpart in the reference
 * (its CPU kernels have no such cache to plan for). */
int spg_dispatch_load_streamed(const char *path);
int spg_set_output_reuse(int reused);
int spg_dispatch_key(int rank, const int *dims, const int *perm, double beta, char *buf,
                     int buflen);

/* ---- plan inspection (host only, no GPU needed) --------------------------
 * spg_plan_describe writes a JSON description of the plan the library
 * would run.  spg_plan_check runs that plan's index walk on the host over
 * host arrays (every output element must be produced exactly once); it
 * exists to test plans without a GPU and is not a compute path. */
int spg_plan_describe(int rank, const int *dims, const int *perm, double beta,
                      const spg_tile_params *params, char *buf, int buflen);
int spg_plan_check(const double *unsorted, double *sorted, int rank, const int *dims,
                   const int *perm, double factor, double beta,
                   const spg_tile_params *params);
/* Checks the index math's multiply-and-shift division against the hardware
 * division on edge cases and `trials` random operands (host only). */
int spg_selftest_fastdiv(int64_t trials);

/* ---- batched form ----------------------------------------------------------
 * Many independent sorts (uneven 4D/6D tiles) in one persistent launch.
 * Add tasks (device pointers), commit (uploads the work list; allocates),
 * then launch, or capture once into a hipGraph and replay.  Tasks may be cut
 * into waves (spg_batch_next_wave: the tasks added after it form the next
 * wave), which run one after the other on the stream: a job larger than
 * device memory reuses the same buffers wave after wave.  Within a wave,
 * outputs must not overlap each other or any input (checked at commit).
 * Each wave is one persistent launch over its tasks below 2^22 elements,
 * then one launch per larger task on that task's per-call plan (such a task
 * fills the device on its own); SPG_BATCH_INLINE puts every task in the
 * persistent launch. */
typedef struct spg_batch spg_batch;
int spg_batch_create(spg_batch **out);
int spg_batch_add(spg_batch *b, const double *unsorted, double *sorted, int rank,
                  const int *dims, const int *perm, double factor, double beta);
int spg_batch_next_wave(spg_batch *b);
/* Access flags of the batch's kernel (SPG_TILE_NT, SPG_TILE_PIPE,
 * SPG_TILE_LDS1: one LDS buffer per workgroup, so more workgroups fit a CU,
 * SPG_BATCH_WIDE: tasks whose tiles move element pairs take tiles of up
 * to 4096 elements instead of 2048, SPG_BATCH_INLINE; before commit); the
 * default is 0. */
#define SPG_BATCH_WIDE 64
#define SPG_BATCH_INLINE 256
int spg_batch_set_flags(spg_batch *b, int flags);
int spg_batch_commit(spg_batch *b, void *stream);
int spg_batch_launch(spg_batch *b, void *stream);
int spg_batch_graph_launch(spg_batch *b, void *stream);
int spg_batch_destroy(spg_batch *b);
int64_t spg_batch_bytes(const spg_batch *b);
int spg_batch_info(const spg_batch *b, char *buf, int buflen);

/* The batched form in one call that allocates no device memory.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same tasks and results as spg_batch_create + spg_batch_add for
 * each task + spg_batch_set_flags(flags) + spg_batch_commit +
 * spg_batch_launch, ordered on `stream`; task i has rank ranks[i], dims
 * dims[i*SPG_MAX_RANK ...] and permutation perms[i*SPG_MAX_RANK ...].  The
 * work list (1,080 bytes per task plus 16 per chunk of up to 8 tiles) is
 * staged in a fixed device array of SPG_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends; a larger work list returns SPG_EINVAL.  All calls
 * on one GPU share that array: calls on one stream run in order, calls that
 * overlap on different streams of one GPU race. */
#define SPG_NOALLOC_SCRATCH_BYTES (16 << 20)
int spg_batch_run_noalloc(int ntasks, const double *const *unsorted, double *const *sorted,
                          const int *ranks, const int *dims, const int *perms,
                          const double *factors, const double *betas, int flags, void *stream);

/* ---- serving form ------------------------------------------------------------
 * Continuous batching.  Requests (device pointers; `sorted` NULL = the
 * server allocates the output inside its budget and discards it when the
 * request completes) carry an arrival time in microseconds from
 * spg_server_start (negative = on submission).  An engine thread admits
 * arrived requests in FIFO order while the input+output bytes of unfinished
 * requests fit `budget_bytes`, holds back requests whose data overlap an
 * earlier unfinished request's output (so accumulations into one tensor keep
 * arrival order), and launches each admitted group as one batched persistent
 * kernel on one of `slots` streams, so groups overlap on the device.
 * SPG_SERVER_FAKE replaces the device with a bandwidth model (no GPU calls)
 * for testing the scheduler.  Results: per-request arrival, launch and
 * completion times; spg_server_stats gives p50/p90/p99 latency and GB/s. */
#define SPG_SERVER_FAKE 1
typedef struct spg_server spg_server;
int spg_server_create(int64_t budget_bytes, int slots, int max_batch, int flags, spg_server **out);
int spg_server_set_fake_rate(spg_server *s, double gbs, double latency_us);
int spg_server_submit(spg_server *s, const double *unsorted, double *sorted, int rank,
                      const int *dims, const int *perm, double factor, double beta,
                      double arrival_us, int64_t *id);
int spg_server_start(spg_server *s);
int spg_server_wait(spg_server *s, double timeout_s);
int spg_server_results(spg_server *s, int64_t n, double *arrival_us, double *admit_us,
                       double *done_us, int32_t *batch);
int spg_server_stats(spg_server *s, char *buf, int buflen);
int spg_server_destroy(spg_server *s);

/* ---- misc ----------------------------------------------------------------- */
int spg_last_error(void);
const char *spg_last_error_message(void);
const char *spg_strerror(int code);
int spg_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SPAGHETTY_AMD_H */
