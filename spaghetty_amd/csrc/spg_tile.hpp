// spg_tile.hpp -- the per-tile device steps shared by the per-call and the
// batched kernels: HBM reads in input order, staging into LDS, LDS reads and
// HBM writes in output order (see spg_plan.hpp for the tile geometry).
#pragma once
#include <hip/hip_runtime.h>

#include "spg_plan.hpp"

namespace spg {

// One tile's HBM reads.  Loads are unconditional so they issue back to
// back: a slot outside the valid part of a partial tile (or beyond the
// tile) reads the tile origin, which is always a valid element, and its
// value is simply never staged.
// NTM: 0 ordinary access, 1 nontemporal loads and stores, 2 nontemporal
// input loads only (the output, whose tile edges are partial cache lines,
// stays cached).
template <int K, int NTM>
__device__ __attribute__((always_inline)) inline void tile_load(double (&v)[K], const double* __restrict__ in, uint32_t bin,
                                 uint32_t lim, const uint32_t (&off)[K], const uint32_t (&mk)[K])
{
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = ld<(NTM != 0)>(in + bin + (slot_ok(mk[k], lim) ? off[k] : 0u));
}

template <int K>
__device__ __attribute__((always_inline)) inline void tile_stage(const double (&v)[K], double* L, uint32_t lim,
                                  const uint32_t (&idx)[K], const uint32_t (&mk)[K])
{
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (slot_ok(mk[k], lim)) L[idx[k]] = v[k];
}

// One tile's LDS reads and HBM writes (output order); only the stores are
// predicated.
template <int K, int MODE, int NTM>
__device__ __attribute__((always_inline)) inline void tile_write(const double* L, double* __restrict__ out, uint32_t bout,
                                  uint32_t lim, const uint32_t (&off)[K], const uint32_t (&idx)[K],
                                  const uint32_t (&mk)[K], double f, double g)
{
    bool ok[K];
    double x[K], y[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        ok[k] = slot_ok(mk[k], lim);
        x[k] = L[ok[k] ? idx[k] : 0u];
    }
    if (MODE != MODE_SORT) {
#pragma unroll
        for (int k = 0; k < K; ++k) y[k] = ld<(NTM == 1)>(out + bout + (ok[k] ? off[k] : 0u));
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) y[k] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (ok[k]) st<(NTM == 1)>(out + bout + off[k], update<MODE>(x[k], y[k], f, g));
}

// A run of tiles [g, gend) of one plan, walked with stride `step`, by one
// workgroup whose per-thread slots are (r_*, w_*).  Two LDS buffers of
// `half` elements alternate; with half == 0 one buffer is reused, with a
// second barrier per tile.  Returns the LDS buffer offset to use next.
// PIPE overlaps each tile's stores with the next tile's loads.
template <int K, int MODE, int NTM, bool PIPE>
__device__ __attribute__((always_inline)) inline uint32_t tile_run(const Plan& P, const double* __restrict__ in,
                                    double* __restrict__ out, uint32_t g, uint32_t gend,
                                    uint32_t base, uint32_t step, double* lds, uint32_t buf,
                                    uint32_t half, const uint32_t (&r_off)[K],
                                    const uint32_t (&r_lds)[K], const uint32_t (&r_mk)[K],
                                    const uint32_t (&w_off)[K], const uint32_t (&w_lds)[K],
                                    const uint32_t (&w_mk)[K])
{
    const double f = P.alpha, gb = P.beta;
    if (g >= gend) return buf;
    if (PIPE) {
        double v[K];
        uint32_t bin, bout, lim;
        tile_origin(P, g - base, bin, bout, lim);
        tile_load<K, NTM>(v, in, bin, lim, r_off, r_mk);
        for (;;) {
            double* L = lds + buf;
            tile_stage<K>(v, L, lim, r_lds, r_mk);
            __syncthreads();
            const uint32_t gn = g + step;
            const bool more = gn < gend;
            uint32_t bin_n = 0, bout_n = 0, lim_n = 0;
            if (more) {
                tile_origin(P, gn - base, bin_n, bout_n, lim_n);
                tile_load<K, NTM>(v, in, bin_n, lim_n, r_off, r_mk);
            }
            tile_write<K, MODE, NTM>(L, out, bout, lim, w_off, w_lds, w_mk, f, gb);
            if (!half) __syncthreads();  // one buffer: the next tile is staged into it
            buf = buf ? 0u : half;
            if (!more) break;
            g = gn;
            bout = bout_n;
            lim = lim_n;
        }
    } else {
        for (; g < gend; g += step) {
            uint32_t bin, bout, lim;
            tile_origin(P, g - base, bin, bout, lim);
            double* L = lds + buf;
            double v[K];
            tile_load<K, NTM>(v, in, bin, lim, r_off, r_mk);
            tile_stage<K>(v, L, lim, r_lds, r_mk);
            __syncthreads();
            tile_write<K, MODE, NTM>(L, out, bout, lim, w_off, w_lds, w_mk, f, gb);
            if (!half) __syncthreads();
            buf = buf ? 0u : half;
        }
    }
    return buf;
}

}  // namespace spg
