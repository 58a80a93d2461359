// spg_vtile.hpp -- the branch-free tile steps of k_vtile: buffer-resource
// HBM accesses whose out-of-range offset masks a slot, and VEC consecutive
// elements per slot (VEC = 2: 16-byte loads and stores).
//
// Slot s of a tile covers the VEC elements e = VEC*s + {0..VEC-1} of the
// read order (c fastest, then xr, then yr) in the read phase and of the
// write order (c fastest, then yr, then xr) in the write phase.  A plan is
// eligible for VEC = 2 only when every such pair is two neighbouring
// addresses on its HBM side (vtile_vec2_ok in plan.cpp); its LDS partner is
// at +rstep (read phase) or +1 (write phase).  With even extents every pair
// starts 16-byte aligned and shares its partial-tile mask.  With odd extents
// (ODD, Plan.pair8) pairs may start at an odd element offset (the 16-byte
// access is then 8-byte aligned) and a partial tile may hold one element of
// a pair only: partial tiles of such plans move each element of a pair with
// an 8-byte access of its own, masked by the element's own mask word.
//
// Each tile rebases a buffer resource at its own origin, so slot offsets
// are 32-bit byte offsets relative to the tile; an offset of kOob lies
// beyond every resource's num_records, which makes the hardware return 0
// for a load and drop a store.  Slots beyond the tile (e >= V) carry kOob
// and stage into a dummy LDS word past the tile; slots of a partial tile
// outside its valid box get kOob for that tile only.
#pragma once
#include <hip/hip_runtime.h>

#include "spg_plan.hpp"

namespace spg {

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int32_t kRsrcWord3 = 0x00020000;    // raw buffer, 32-bit data format (gfx9 family)

__device__ __attribute__((always_inline)) inline __amdgpu_buffer_rsrc_t tile_rsrc(const double* base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), (short)0, kTileRecords,
                                             kRsrcWord3);
}

// Cache-policy operand of the buffer instructions: 2 sets the nontemporal bit.
template <bool NT>
constexpr int cpol() { return NT ? 2 : 0; }

template <int VEC, bool NT>
__device__ __attribute__((always_inline)) inline void bload(double (&v)[VEC], __amdgpu_buffer_rsrc_t r, uint32_t off)
{
    if constexpr (VEC == 2) {
        const u32x4 q = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, cpol<NT>());
        v[0] = __builtin_bit_cast(double, u32x2{q.x, q.y});
        v[1] = __builtin_bit_cast(double, u32x2{q.z, q.w});
    } else {
        v[0] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, cpol<NT>()));
    }
}

template <int VEC, bool NT>
__device__ __attribute__((always_inline)) inline void bstore(const double (&v)[VEC], __amdgpu_buffer_rsrc_t r, uint32_t off)
{
    if constexpr (VEC == 2) {
        const u32x2 a = __builtin_bit_cast(u32x2, v[0]);
        const u32x2 b = __builtin_bit_cast(u32x2, v[1]);
        __builtin_amdgcn_raw_buffer_store_b128(u32x4{a.x, a.y, b.x, b.y}, r, off, 0, cpol<NT>());
    } else {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v[0]), r, off, 0, cpol<NT>());
    }
}

// Per-thread slot tables of a plan: byte offsets relative to the tile
// origin (kOob beyond the tile), LDS element indices (the dummy word beyond
// the tile) and the packed (x top, y top) coordinates for partial tiles.
// With ODD the mask words of each pair's second element follow (r_mk2,
// w_mk2).
template <int K, bool ODD = false>
struct VSlots {
    uint32_t r_off[K], r_lds[K], r_mk[K];
    uint32_t w_off[K], w_lds[K], w_mk[K];
    uint32_t r_mk2[ODD ? K : 1], w_mk2[ODD ? K : 1];
};

// The helpers below take any slot table with at least K entries (SL), so
// one table can serve a pair kernel (K pairs) and a single-element kernel
// (2K slots) alike.
template <int VEC, int K, class SL, bool ODD = false>
__device__ __attribute__((always_inline)) inline void vslots_init(const Plan& P, int tid, int B, SL& S)
{
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int e = VEC * (tid + k * B);
        uint32_t off, lds, mk;
        read_slot(P, e, off, lds, mk);
        const bool in_tile = e < P.V;
        S.r_off[k] = in_tile ? off * 8u : kOob;
        S.r_lds[k] = in_tile ? lds : (uint32_t)P.vdummy;
        S.r_mk[k] = in_tile ? mk : 0u;
        write_slot(P, e, off, lds, mk);
        S.w_off[k] = in_tile ? off * 8u : kOob;
        S.w_lds[k] = in_tile ? lds : (uint32_t)P.vdummy;
        S.w_mk[k] = in_tile ? mk : 0u;
        if constexpr (ODD) {
            read_slot(P, e + 1, off, lds, mk);
            S.r_mk2[k] = in_tile ? mk : 0u;
            write_slot(P, e + 1, off, lds, mk);
            S.w_mk2[k] = in_tile ? mk : 0u;
        }
    }
}

// Offsets of one side of one tile.  eo: one per slot (or pair): a full
// tile uses the tables as they are; a partial tile sends the slots outside
// its valid box out of range.  Partial tiles of ODD plans (split) move the
// two elements of a pair separately: eo then holds the first element's
// offset and e1 the second's, each out of range where that element is.
template <int K, bool ODD>
struct VOff {
    uint32_t eo[K];
    uint32_t e1[ODD ? K : 1];
    bool split;  // uniform: a partial tile of an ODD plan
};

template <int K, bool ODD>
__device__ __attribute__((always_inline)) inline void voff(VOff<K, ODD>& o, const uint32_t* off, const uint32_t* mk,
                                                          const uint32_t* mk2, uint32_t lim, bool full)
{
    o.split = false;
    if (full) {
#pragma unroll
        for (int k = 0; k < K; ++k) o.eo[k] = off[k];
        return;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) o.eo[k] = slot_ok(mk[k], lim) ? off[k] : kOob;
    if constexpr (ODD) {
        o.split = true;
#pragma unroll
        for (int k = 0; k < K; ++k) o.e1[k] = slot_ok(mk2[k], lim) ? off[k] + 8u : kOob;
    }
}

template <int VEC, int K, bool NT, bool ODD>
__device__ __attribute__((always_inline)) inline void vload(double (&v)[K][VEC], __amdgpu_buffer_rsrc_t r,
                                                           const VOff<K, ODD>& o)
{
    if constexpr (ODD) {
        if (o.split) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                v[k][0] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, o.eo[k], 0, cpol<NT>()));
                v[k][1] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, o.e1[k], 0, cpol<NT>()));
            }
            return;
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) bload<VEC, NT>(v[k], r, o.eo[k]);
}

template <int VEC, int K, bool NT, bool ODD>
__device__ __attribute__((always_inline)) inline void vstore1(const double (&v)[VEC], __amdgpu_buffer_rsrc_t r,
                                                             const VOff<K, ODD>& o, int k)
{
    if constexpr (ODD) {
        if (o.split) {
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v[0]), r, o.eo[k], 0, cpol<NT>());
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v[1]), r, o.e1[k], 0, cpol<NT>());
            return;
        }
    }
    bstore<VEC, NT>(v, r, o.eo[k]);
}

template <int VEC, int K, bool NT, bool ODD, class SL>
__device__ __attribute__((always_inline)) inline void vtile_load(double (&v)[K][VEC], const double* in, uint32_t bin, uint32_t lim,
                                  bool full, const SL& S)
{
    const __amdgpu_buffer_rsrc_t r = tile_rsrc(in + bin);
    VOff<K, ODD> o;
    voff<K, ODD>(o, &S.r_off[0], &S.r_mk[0], &S.r_mk2[0], lim, full);
    vload<VEC, K, NT, ODD>(v, r, o);
}

template <int VEC, int K, class SL>
__device__ __attribute__((always_inline)) inline void vtile_stage(const double (&v)[K][VEC], double* L, uint32_t rstep, const SL& S)
{
#pragma unroll
    for (int k = 0; k < K; ++k) {
        L[S.r_lds[k]] = v[k][0];
        if constexpr (VEC == 2) L[S.r_lds[k] + rstep] = v[k][1];
    }
}

// single: one LDS buffer serves every tile, so all waves must have read it
// before any wave stages the next tile: a barrier after the LDS reads.
template <int VEC, int K, int MODE, bool NT, bool ODD, class SL>
__device__ __attribute__((always_inline)) inline void vtile_write(const double* L, double* out, uint32_t bout, uint32_t lim, bool full,
                                   const SL& S, double f, double g, bool single)
{
    const __amdgpu_buffer_rsrc_t r = tile_rsrc(out + bout);
    VOff<K, ODD> o;
    voff<K, ODD>(o, &S.w_off[0], &S.w_mk[0], &S.w_mk2[0], lim, full);
    double x[K][VEC], y[K][VEC];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        x[k][0] = L[S.w_lds[k]];
        if constexpr (VEC == 2) x[k][1] = L[S.w_lds[k] + 1];
    }
    if (single) __syncthreads();
    if constexpr (MODE != MODE_SORT) vload<VEC, K, NT, ODD>(y, r, o);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double ov[VEC];
#pragma unroll
        for (int j = 0; j < VEC; ++j)
            ov[j] = update<MODE>(x[k][j], MODE != MODE_SORT ? y[k][j] : 0.0, f, g);
        vstore1<VEC, K, NT, ODD>(ov, r, o, k);
    }
}

// The old output values of one tile (accumulate and five-case forms).
template <int VEC, int K, bool NT, bool ODD, class SL>
__device__ __attribute__((always_inline)) inline void vtile_load_out(double (&y)[K][VEC], const double* out, uint32_t bout, uint32_t lim,
                                      bool full, const SL& S)
{
    const __amdgpu_buffer_rsrc_t r = tile_rsrc(out + bout);
    VOff<K, ODD> o;
    voff<K, ODD>(o, &S.w_off[0], &S.w_mk[0], &S.w_mk2[0], lim, full);
    vload<VEC, K, NT, ODD>(y, r, o);
}

// One tile's LDS reads, update with the old output values y (already
// loaded) and stores.  single: a barrier after the LDS reads (one buffer).
template <int VEC, int K, int MODE, bool NT, bool ODD, class SL>
__device__ __attribute__((always_inline)) inline void vtile_write_with(const double* L, double* out, uint32_t bout, uint32_t lim, bool full,
                                        const SL& S, const double (&y)[K][VEC], double f, double g, bool single)
{
    const __amdgpu_buffer_rsrc_t r = tile_rsrc(out + bout);
    VOff<K, ODD> o;
    voff<K, ODD>(o, &S.w_off[0], &S.w_mk[0], &S.w_mk2[0], lim, full);
    double x[K][VEC];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        x[k][0] = L[S.w_lds[k]];
        if constexpr (VEC == 2) x[k][1] = L[S.w_lds[k] + 1];
    }
    if (single) __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double ov[VEC];
#pragma unroll
        for (int j = 0; j < VEC; ++j) ov[j] = update<MODE>(x[k][j], MODE != MODE_SORT ? y[k][j] : 0.0, f, g);
        vstore1<VEC, K, NT, ODD>(ov, r, o, k);
    }
}

// Whether a pipelined accumulate carries the next tile's old output values
// with its prefetch (a second set of K*VEC doubles); with 16 single slots
// that would spill, so those load the current tile's old values right
// after the barrier, ahead of the prefetch.
template <int VEC, int K, int MODE>
__host__ __device__ constexpr bool pipe_out_ahead()
{
    return MODE != MODE_SORT && K <= 8;
}

// State of the pipelined walk between tiles: the tile about to be staged
// (its input values already in v, and its old output values in y when they
// travel with the prefetch), its output origin and valid extents.
template <int VEC, int K>
struct VPipe {
    double v[K][VEC];
    double y[K][VEC];
    uint32_t g, bout, lim, buf;
};

// One pipelined step: stage tile g, then -- after the barrier and before
// tile g's stores -- issue the next tile's input loads (and, for an
// accumulate, its old output loads, or this tile's ahead of the prefetch),
// then read tile g from LDS and store it.  Vector memory operations complete
// in issue order for the wait counter on this chip family, so the next
// step's wait for the prefetched values never waits for this tile's stores,
// and no load waits behind the prefetch.  Returns whether another tile
// follows.
template <int VEC, int K, int MODE, bool NTL, bool NTS, bool ODD, class SL>
__device__ __attribute__((always_inline)) inline bool vtile_pipe_step(
    const Plan& P, const double* __restrict__ in, double* __restrict__ out, VPipe<VEC, K>& st, uint32_t gend,
    uint32_t step, double* lds, uint32_t half, uint32_t full_lim, const SL& S)
{
    constexpr bool ahead = pipe_out_ahead<VEC, K, MODE>();
    double* L = lds + st.buf;
    vtile_stage<VEC, K>(st.v, L, (uint32_t)P.rstep, S);
    __syncthreads();
    const uint32_t gn = st.g + step;
    const bool more = gn < gend;
    double y[K][VEC];
    if constexpr (ahead) {
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int j = 0; j < VEC; ++j) y[k][j] = st.y[k][j];
    } else if constexpr (MODE != MODE_SORT) {
        vtile_load_out<VEC, K, NTS, ODD>(y, out, st.bout, st.lim, st.lim == full_lim, S);
    }
    uint32_t bin_n = 0, bout_n = 0, lim_n = 0;
    if (more) {
        tile_origin(P, gn, bin_n, bout_n, lim_n);
        vtile_load<VEC, K, NTL, ODD>(st.v, in, bin_n, lim_n, lim_n == full_lim, S);
        if constexpr (ahead) vtile_load_out<VEC, K, NTS, ODD>(st.y, out, bout_n, lim_n, lim_n == full_lim, S);
    }
    vtile_write_with<VEC, K, MODE, NTS, ODD>(L, out, st.bout, st.lim, st.lim == full_lim, S, y, P.alpha,
                                             P.beta, half == 0);
    st.buf = st.buf ? 0u : half;
    st.g = gn;
    st.bout = bout_n;
    st.lim = lim_n;
    return more;
}

// Tiles [g, gend) of one plan with stride `step`, by one workgroup; LDS
// buffers of `half` elements alternate (one barrier per tile), or, with
// half == 0, one buffer is reused (two barriers per tile, half the LDS, so
// more workgroups fit on a CU).  PIPE issues the next tile's loads right
// after the barrier, so they are in flight while this tile is read from LDS
// and written; its first step is peeled off the loop, so the loop head is
// reached only after a prefetch followed by stores and the compiler's wait
// there covers the prefetch alone.  Returns the next buffer offset.
// NTL / NTS: nontemporal loads of the input / accesses of the output.
template <int VEC, int K, int MODE, bool NTL, bool NTS, bool PIPE, bool ODD = false, class SL>
__device__ __attribute__((always_inline)) inline uint32_t vtile_run(const Plan& P, const double* __restrict__ in,
                                     double* __restrict__ out, uint32_t g, uint32_t gend,
                                     uint32_t step, double* lds, uint32_t buf, uint32_t half,
                                     const SL& S)
{
    const double f = P.alpha, gb = P.beta;
    const uint32_t full_lim = (uint32_t)P.m_ch | ((uint32_t)P.ty << 16);
    const uint32_t rstep = (uint32_t)P.rstep;
    const bool single = half == 0;
    if (g >= gend) return buf;
    if (PIPE) {
        VPipe<VEC, K> st;
        uint32_t bin;
        st.g = g;
        st.buf = buf;
        tile_origin(P, g, bin, st.bout, st.lim);
        vtile_load<VEC, K, NTL, ODD>(st.v, in, bin, st.lim, st.lim == full_lim, S);
        if constexpr (pipe_out_ahead<VEC, K, MODE>())
            vtile_load_out<VEC, K, NTS, ODD>(st.y, out, st.bout, st.lim, st.lim == full_lim, S);
        // (with 16 single slots the duplicated step makes the compiler keep
        // a private copy of the plan: those run the loop without the peel)
        if (K >= 16 || vtile_pipe_step<VEC, K, MODE, NTL, NTS, ODD>(P, in, out, st, gend, step, lds, half, full_lim, S))
            while (vtile_pipe_step<VEC, K, MODE, NTL, NTS, ODD>(P, in, out, st, gend, step, lds, half, full_lim, S)) {
            }
        return st.buf;
    }
    for (; g < gend; g += step) {
        uint32_t bin, bout, lim;
        tile_origin(P, g, bin, bout, lim);
        const bool full = lim == full_lim;
        double* L = lds + buf;
        double v[K][VEC];
        vtile_load<VEC, K, NTL, ODD>(v, in, bin, lim, full, S);
        vtile_stage<VEC, K>(v, L, rstep, S);
        __syncthreads();
        vtile_write<VEC, K, MODE, NTS, ODD>(L, out, bout, lim, full, S, f, gb, single);
        buf = buf ? 0u : half;
    }
    return buf;
}

}  // namespace spg
