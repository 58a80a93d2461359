// kernels.hip -- gfx950 kernels for b(perm(idx)) (+)= f * a(idx) on FP64.
//
// k_tiled: the tile box of a Plan (spg_plan.hpp) is read from HBM in input
//   order -- consecutive lanes take consecutive input addresses -- staged in
//   a padded LDS tile, and written in output order, so both HBM streams are
//   coalesced runs.  Workgroups are persistent: each thread computes its
//   slot offsets once and then walks tiles t = w, w+G, ...; the tile origin
//   is wave-uniform (scalar) work.  Two LDS buffers alternate, so one
//   barrier per tile separates a tile's LDS stores from its loads and the
//   next tile's stores from this tile's loads.
// k_rows: rows that are contiguous on both sides (shared fastest axis, or
//   the whole array after fusion) are streamed with 16-byte accesses.
//
// The update rounds the product and the sum separately (no FMA), matching
// the reference's y = y + x*s (trunk/old_sort.f:535).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>

#include "spg_internal.hpp"
#include "spg_tile.hpp"
#include "spg_vtile.hpp"

namespace spg {

// PIPE: the next tile's HBM loads are issued right after the barrier, so
// they are in flight while this tile is read from LDS and written out.
template <int K, int MODE, int NTM, bool PIPE>
__global__ __launch_bounds__(K >= 16 ? 256 : 512) void k_tiled(const double* __restrict__ in,
                                                double* __restrict__ out, const Plan P)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    const int B = blockDim.x;
    uint32_t r_off[K], r_lds[K], r_mk[K], w_off[K], w_lds[K], w_mk[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int e = tid + k * B;
        read_slot(P, e, r_off[k], r_lds[k], r_mk[k]);
        write_slot(P, e, w_off[k], w_lds[k], w_mk[k]);
    }
    const uint32_t G = gridDim.x;
    const uint32_t ntiles = (uint32_t)P.ntiles;
    const uint32_t buf_elems = (uint32_t)P.lds_tile;
    const uint32_t t = logical_wg(blockIdx.x, G, P.xcd_remap);
    tile_run<K, MODE, NTM, PIPE>(P, in, out, t, ntiles, 0u, G, lds, 0u, buf_elems, r_off, r_lds, r_mk,
                                w_off, w_lds, w_mk);
}

// k_vtile: the same tile walk with branch-free buffer accesses and VEC
// elements per slot (spg_vtile.hpp).  NTM as for k_tiled.
// Up to max_block_v(MODE, PIPE) threads with 8 pairs each (tiles above
// 4096 elements), else up to 512.
// ODD: pairs over odd extents (Plan.pair8), whose partial tiles may hold a
// pair's first element without its second.
template <int VEC, int K, int MODE, int NTM, bool PIPE, bool ODD = false>
__global__ __launch_bounds__(VEC == 2 && K == 8 ? max_block_v(MODE, PIPE) : 512) void k_vtile(
    const double* __restrict__ in, double* __restrict__ out, const Plan P)
{
    extern __shared__ double lds[];
    VSlots<K, ODD> S;
    vslots_init<VEC, K, VSlots<K, ODD>, ODD>(P, threadIdx.x, blockDim.x, S);
    const uint32_t G = gridDim.x;
    const uint32_t t = logical_wg(blockIdx.x, G, P.xcd_remap);
    const uint32_t half = P.lds1 ? 0u : (uint32_t)P.half;
    const uint32_t ntiles = (uint32_t)P.ntiles;
    const uint32_t c = (uint32_t)P.chunk;
    if (c <= 1) {
        vtile_run<VEC, K, MODE, (NTM != 0), (NTM == 1), PIPE, ODD>(P, in, out, t, ntiles, G, lds, 0u, half, S);
        return;
    }
    // chunks of c consecutive tiles: workgroup t takes chunks t, t+G, ...,
    // so one CU's successive tiles share rows (and address translations)
    uint32_t buf = 0;
    for (uint32_t j = t; j < (ntiles + c - 1) / c; j += G) {
        const uint32_t g0 = j * c, g1 = g0 + c < ntiles ? g0 + c : ntiles;
        buf = vtile_run<VEC, K, MODE, (NTM != 0), (NTM == 1), PIPE, ODD>(P, in, out, g0, g1, 1u, lds, buf, half,
                                                                         S);
    }
}

template <int MODE, int VEC, int U, bool NT>
__global__ __launch_bounds__(1024) void k_rows(const double* __restrict__ in,
                                               double* __restrict__ out, const Plan P)
{
    const uint32_t B = blockDim.x;
    const uint32_t per = B * VEC * U;
    const uint32_t C = (uint32_t)P.C;
    const uint32_t nch = (C + per - 1) / per;
    const uint32_t total = (uint32_t)P.ntiles * nch;
    const double f = P.alpha, g = P.beta;
    const uint32_t G = gridDim.x;
    for (uint32_t w = blockIdx.x; w < total; w += G) {
        const uint32_t row = w / nch, ch = w - row * nch;
        uint32_t bi, bo;
        row_origin(P, row, bi, bo);
        const uint32_t c0 = ch * per;
        if (VEC == 2) {
            double2 xv[U], yv[U];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t col = c0 + (u * B + threadIdx.x) * 2;
                ok[u] = col < C;
                if (ok[u]) {
                    xv[u].x = ld<NT>(in + bi + col);
                    xv[u].y = ld<NT>(in + bi + col + 1);
                }
            }
            if (MODE != MODE_SORT) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t col = c0 + (u * B + threadIdx.x) * 2;
                    if (ok[u]) {
                        yv[u].x = ld<NT>(out + bo + col);
                        yv[u].y = ld<NT>(out + bo + col + 1);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t col = c0 + (u * B + threadIdx.x) * 2;
                if (ok[u]) {
                    st<NT>(out + bo + col, update<MODE>(xv[u].x, MODE != MODE_SORT ? yv[u].x : 0.0, f, g));
                    st<NT>(out + bo + col + 1,
                           update<MODE>(xv[u].y, MODE != MODE_SORT ? yv[u].y : 0.0, f, g));
                }
            }
        } else {
            double xv[U], yv[U];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t col = c0 + u * B + threadIdx.x;
                ok[u] = col < C;
                xv[u] = 0.0;
                yv[u] = 0.0;
                if (ok[u]) xv[u] = ld<NT>(in + bi + col);
            }
            if (MODE != MODE_SORT) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t col = c0 + u * B + threadIdx.x;
                    if (ok[u]) yv[u] = ld<NT>(out + bo + col);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t col = c0 + u * B + threadIdx.x;
                if (ok[u]) st<NT>(out + bo + col, update<MODE>(xv[u], yv[u], f, g));
            }
        }
    }
}

namespace {

typedef void (*KernelFn)(const double*, double*, const Plan);

template <int MODE, int NTM, bool PIPE>
KernelFn tiled_fn(int K)
{
    switch (K) {
    case 1: return k_tiled<1, MODE, NTM, PIPE>;
    case 2: return k_tiled<2, MODE, NTM, PIPE>;
    case 4: return k_tiled<4, MODE, NTM, PIPE>;
    case 8: return k_tiled<8, MODE, NTM, PIPE>;
    default: return k_tiled<16, MODE, NTM, PIPE>;
    }
}

template <int NTM, bool PIPE>
KernelFn pick_tiled_v(int K, int mode)
{
    if (mode == MODE_SORT) return tiled_fn<MODE_SORT, NTM, PIPE>(K);
    if (mode == MODE_ACC) return tiled_fn<MODE_ACC, NTM, PIPE>(K);
    return tiled_fn<MODE_AXPBY, NTM, PIPE>(K);
}

template <bool PIPE>
KernelFn pick_tiled_p(int K, int mode, int nt)
{
    if (nt == 1) return pick_tiled_v<1, PIPE>(K, mode);
    if (nt == 2) return pick_tiled_v<2, PIPE>(K, mode);
    return pick_tiled_v<0, PIPE>(K, mode);
}

KernelFn pick_tiled(int slots, int mode, int nt, int pipe)
{
    int K = 1;
    while (K < slots) K <<= 1;
    return pipe ? pick_tiled_p<true>(K, mode, nt) : pick_tiled_p<false>(K, mode, nt);
}

template <int VEC, int MODE, int NTM, bool PIPE, bool ODD>
KernelFn vtile_fn(int K)
{
    switch (K) {
    case 1: return k_vtile<VEC, 1, MODE, NTM, PIPE, ODD>;
    case 2: return k_vtile<VEC, 2, MODE, NTM, PIPE, ODD>;
    case 4: return k_vtile<VEC, 4, MODE, NTM, PIPE, ODD>;
    case 8: return k_vtile<VEC, 8, MODE, NTM, PIPE, ODD>;
    default:
        if constexpr (VEC == 1) return k_vtile<1, 16, MODE, NTM, PIPE>;
        else return k_vtile<2, 8, MODE, NTM, PIPE, ODD>;  // unreachable: pairs use <= 8 slots
    }
}

template <int VEC, int NTM, bool PIPE, bool ODD>
KernelFn pick_vtile_m(int K, int mode)
{
    if (mode == MODE_SORT) return vtile_fn<VEC, MODE_SORT, NTM, PIPE, ODD>(K);
    if (mode == MODE_ACC) return vtile_fn<VEC, MODE_ACC, NTM, PIPE, ODD>(K);
    return vtile_fn<VEC, MODE_AXPBY, NTM, PIPE, ODD>(K);
}

template <int VEC, bool PIPE, bool ODD>
KernelFn pick_vtile_n(int K, int mode, int nt)
{
    if (nt == 1) return pick_vtile_m<VEC, 1, PIPE, ODD>(K, mode);
    if (nt == 2) return pick_vtile_m<VEC, 2, PIPE, ODD>(K, mode);
    return pick_vtile_m<VEC, 0, PIPE, ODD>(K, mode);
}

KernelFn pick_vtile(int vec, int slots, int mode, int nt, int pipe, int pair8)
{
    int K = 1;
    while (K < slots) K <<= 1;
    if (vec == 2 && pair8)
        return pipe ? pick_vtile_n<2, true, true>(K, mode, nt) : pick_vtile_n<2, false, true>(K, mode, nt);
    if (vec == 2) return pipe ? pick_vtile_n<2, true, false>(K, mode, nt) : pick_vtile_n<2, false, false>(K, mode, nt);
    return pipe ? pick_vtile_n<1, true, false>(K, mode, nt) : pick_vtile_n<1, false, false>(K, mode, nt);
}


constexpr int kRowsUnroll = 4;

template <int VEC, bool NT>
KernelFn rows_fn(int mode)
{
    if (mode == MODE_SORT) return k_rows<MODE_SORT, VEC, kRowsUnroll, NT>;
    if (mode == MODE_ACC) return k_rows<MODE_ACC, VEC, kRowsUnroll, NT>;
    return k_rows<MODE_AXPBY, VEC, kRowsUnroll, NT>;
}

KernelFn pick_rows(int vec, int mode, int nt)
{
    if (vec == 2) return nt ? rows_fn<2, true>(mode) : rows_fn<2, false>(mode);
    return nt ? rows_fn<1, true>(mode) : rows_fn<1, false>(mode);
}

// The kernel a plan runs.
KernelFn plan_kernel(const Plan& P)
{
    if (P.kind == KIND_ROWS) return pick_rows(P.vec, P.mode, P.nt);
    if (P.tk == 1) return pick_vtile(P.vec, P.slots, P.mode, P.nt, P.pipe, P.pair8);
    return pick_tiled(P.slots, P.mode, P.nt, P.pipe);
}

std::mutex g_occ_mu;
std::map<std::tuple<int, KernelFn, int, int>, int> g_occ;
std::map<int, int> g_cus;

int cu_count(int dev)
{
    std::lock_guard<std::mutex> lk(g_occ_mu);
    auto it = g_cus.find(dev);
    if (it != g_cus.end()) return it->second;
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
        n = 256;
    g_cus[dev] = n;
    return n;
}

int occupancy(int dev, KernelFn fn, int block, int lds)
{
    std::lock_guard<std::mutex> lk(g_occ_mu);
    auto key = std::make_tuple(dev, fn, block, lds);
    auto it = g_occ.find(key);
    if (it != g_occ.end()) return it->second;
    if (lds > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(fn), block,
                                                     (size_t)lds) != hipSuccess || nb <= 0)
        nb = 1;
    g_occ[key] = nb;
    return nb;
}

}  // namespace

// Resolves the launch geometry of a plan on the current device.
int resolve_launch(Plan* P)
{
    if (P->kind == KIND_EMPTY) return SPG_OK;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return SPG_EHIP;
    const int cus = cu_count(dev);
    KernelFn fn = plan_kernel(*P);
    int64_t work = P->ntiles;
    if (P->kind == KIND_TILED && P->tk == 1 && P->chunk > 1) work = (P->ntiles + P->chunk - 1) / P->chunk;
    if (P->kind == KIND_ROWS) {
        const int64_t per = (int64_t)P->block * P->vec * kRowsUnroll;
        work = P->ntiles * ((P->C + per - 1) / per);
    }
    int64_t G;
    if (P->grid < 0) {
        G = (int64_t)cus * (-P->grid);
    } else if (P->grid > 0) {
        return SPG_OK;  // already resolved
    } else {
        G = (int64_t)cus * occupancy(dev, fn, P->block, P->lds_bytes);
    }
    if (G > work) G = work;
    if (G > 8 && (G & 7)) G &= ~(int64_t)7;
    if (G < 1) G = 1;
    P->grid = (int32_t)G;
    return SPG_OK;
}

int launch_plan(const Plan& P0, const double* in, double* out, hipStream_t st)
{
    if (P0.kind == KIND_EMPTY || P0.nelem == 0) return SPG_OK;
    Plan P = P0;
    if (P.vec == 2 && !P.pair8 && ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15u)) {
        // pairs planned 16-byte aligned on arrays that are only 8-byte
        // aligned: single elements, which a tile above kMaxTileElems cannot
        // use (the caller re-plans).  pair8 plans take 8-byte aligned pairs
        // anyway, so the arrays' alignment does not matter to them.
        if (P.kind == KIND_TILED && (P.V > kMaxTileElems || P.block > 512)) {
            set_error(SPG_EINVAL, "tile too large for arrays that are not 16-byte aligned");
            return SPG_EINVAL;
        }
        P.vec = 1;
        if (P.kind == KIND_TILED) P.slots = (P.V + P.block - 1) / P.block;
    }
    if (P.grid <= 0) {
        const int rc = resolve_launch(&P);
        if (rc != SPG_OK) return rc;
    }
    KernelFn fn = plan_kernel(P);
    if (P.kind == KIND_TILED && P.lds_bytes > 65536) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        (void)occupancy(dev, fn, P.block, P.lds_bytes);  // sets the LDS attribute once
    }
    hipLaunchKernelGGL(fn, dim3((unsigned)P.grid), dim3((unsigned)P.block), (size_t)P.lds_bytes, st,
                       in, out, P);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(SPG_EHIP, hipGetErrorString(e));
        return SPG_EHIP;
    }
    return SPG_OK;
}

}  // namespace spg
