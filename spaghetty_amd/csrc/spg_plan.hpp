// spg_plan.hpp -- the execution plan of one tensor sort and the index math
// shared by the HIP kernels and the host-side plan checker.
//
// Problem (NWChem tce_sort_N / tce_sortacc_N, trunk/old_sort.f:33-95,
// 512-574): dims jd(1..R) listed slowest -> fastest, permutation (i,j,k,l..)
// 1-based; the input element at 1-based coordinates id() goes to
//   ib = id(p_R) + jd(p_R)*(id(p_{R-1})-1 + ...)      (old_sort.f:52)
// with value   s(ib) = f*u(ia)            (sort)
//              s(ib) = s(ib) + f*u(ia)    (sort-accumulate, two roundings)
//              s(ib) = g*s(ib) + f*u(ia)  (five-case form, generator2.py:396)
//
// Internally axes are numbered FASTEST FIRST.  After dropping extent-1 axes
// and fusing input axes that stay adjacent and in order in the output, the
// problem is a rank-r permutation with extents n[0..r-1] (n[0] contiguous in
// the input) and out_axis[q] = input axis at output position q (q = 0 is the
// output's contiguous axis).
//
// "Tiled" plan: one workgroup tile is the box
//     [shared axis C] x [X group: input axes s..s+nx-1] x [Y group: output
//                                             positions s..s+ny-1]
// where s = 1 and C = n[0] when the input's contiguous axis is also the
// output's (out_axis[0] == 0), else s = 0 and C = 1.  The X group is
// contiguous in the input (its fused index xr has input offset C*xr), the Y
// group contiguous in the output (output offset C*yr).  Every axis of a
// group is whole except the group's slowest ("top") axis, which is cut into
// chunks of tx (ty) so that a tile holds V = C*Tx*Ty elements,
// Tx = x_low*tx, Ty = y_low*ty.  The remaining axes (Z) are walked tile by
// tile.  A tile is read in input order (c fastest, then xr, then yr) and
// staged in LDS, then written in output order (c fastest, then yr, then xr):
// both HBM sides are contiguous runs.  Where the X and Y axes swap places
// (Plan.tw), a tile may span several values of one more axis W, slowest in
// both orders.
//
// "Rows" plan: when the shared contiguous axis is long, every row of C
// elements is contiguous on both sides; rows are streamed without LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spg {

constexpr int MAXR = 8;
constexpr int MAXG = MAXR + 2;  // tile grid digits: x chunk, y chunk, Z axes
constexpr int kMaxTileElems = 4096;    // largest tile (elements) of k_tiled or single elements
// k_vtile with 8 pairs per thread: up to kMaxBlockV threads (three waves per
// SIMD, at most 168 VGPRs each).  A pipelined accumulate also holds the next
// tile's loads and the output's old values (178 VGPRs) and stays within 512
// threads; the other forms fit (sort 129, accumulate 138).
constexpr int kMaxBlockV = 768;
constexpr int kMaxTileElemsV = 16 * kMaxBlockV;  // 12288: largest tile of any plan
constexpr int64_t kBigTileElems = 1 << 22;       // default plans above this may use such tiles
__host__ __device__ constexpr int max_block_v(int mode, bool pipe)
{
    return (mode == 0 || !pipe) ? kMaxBlockV : 512;
}
__host__ __device__ constexpr int max_tile_v(int mode, bool pipe) { return 16 * max_block_v(mode, pipe); }
constexpr int kMaxSlots = 16;          // largest elements per thread per tile
constexpr uint32_t kBadSlot = 0xffffffffu;
constexpr uint32_t kOob = 0x80000000u;        // k_vtile: byte offset past every tile resource
constexpr int32_t kTileRecords = 0x7fffff00;  // k_vtile: num_records of a tile resource (bytes)

enum Kind : int32_t { KIND_TILED = 0, KIND_ROWS = 1, KIND_EMPTY = 2 };

// Division by a divisor fixed at plan time, for dividends below 2^31:
// q = floor(n * m / 2^(32 + sh)) with l = ceil(log2 d), sh = l - 1 and
// m = ceil(2^(31 + l) / d) (Granlund & Montgomery, "Division by invariant
// integers using multiplication", PLDI 1994).  The error of n*m/2^(31+l)
// against n/d is below n / 2^(31+l) < 2^-l <= 1/d, so the floor is exact,
// and m < 2^32.  d = 1 is stored as m = 0 (q = n).  One 32-bit high
// multiply and a shift: a scalar instruction pair for wave-uniform values.
struct FastDiv {
    uint32_t m, sh;
};

inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f;
    if (d <= 1) {
        f.m = 0;
        f.sh = 0;
        return f;
    }
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    f.sh = l - 1;
    f.m = (uint32_t)(((1ull << (31 + l)) + d - 1) / d);
    return f;
}

__host__ __device__ __attribute__((always_inline)) inline uint32_t fdiv(uint32_t n, const FastDiv& f)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t hi = __umulhi(n, f.m);
#else
    const uint32_t hi = (uint32_t)(((uint64_t)n * f.m) >> 32);
#endif
    return f.m ? (hi >> f.sh) : n;
}

enum Mode : int32_t { MODE_SORT = 0, MODE_ACC = 1, MODE_AXPBY = 2 };

struct Plan {
    // problem
    int64_t nelem;
    double alpha;   // factor
    double beta;    // 0: sort, 1: sort-accumulate, else five-case g
    int32_t mode;   // Mode
    int32_t kind;   // Kind
    int32_t rank;   // rank after fusion
    int32_t ext[MAXR];       // fused extents, fastest first
    int32_t out_axis[MAXR];  // out_axis[q]: input axis at output position q

    // tiled / rows description
    int32_t C;                // shared contiguous extent (1 if none)
    int32_t nx, ny, nz;       // group sizes
    int32_t x_ext[MAXR], x_ostr[MAXR];   // X axes: extent, output stride
    int32_t y_ext[MAXR], y_istr[MAXR];   // Y axes: extent, input stride
    int32_t z_ext[MAXR], z_istr[MAXR], z_ostr[MAXR];
    int32_t x_low, y_low;     // products of the whole lower axes
    int32_t tx, ty;           // chunk of the top axis
    int32_t x_top, y_top;     // extent of the top axis (1 if group empty)
    int32_t x_nchunk, y_nchunk;
    int32_t x_top_istr, x_top_ostr, y_top_istr, y_top_ostr;
    int32_t Tx, Ty, V;        // tile extents, elements per tile
    int32_t lds_row;          // LDS row pitch (elements) of the [xr][yr*C+c] tile
    int64_t ntiles;           // tiled: tiles; rows: rows

    // launch
    int32_t block;            // threads per workgroup
    int32_t slots;            // slots per thread per tile (ceil(V/(vec*block)))
    int32_t grid;             // workgroups (persistent when < ntiles)
    int32_t xcd_remap;        // 1: give each XCD a contiguous run of tiles
    int32_t vec;              // elements per thread access (1 or 2): rows, k_vtile
    int32_t lds_bytes;        // dynamic LDS per workgroup
    int32_t nt;               // 1: nontemporal loads and stores, 2: nontemporal loads only
    int32_t pipe;             // 1: next tile's loads overlap this tile's writes
    int32_t tk;               // tiled: 0 k_tiled, 1 k_vtile (buffer accesses, vec per slot)
    int32_t rstep;            // k_vtile: LDS distance of the two elements of a read pair
    int32_t vdummy;           // k_vtile: LDS index of the dummy word past the tile
    int32_t half;             // k_vtile: LDS elements of one buffer (tile + dummy words)
    int32_t lds1;             // k_vtile: 1 = one LDS buffer (two barriers per tile)
    int32_t xsplit;           // k_vtile pairs without a shared axis: even x rows in the
                              // first half of the LDS tile, odd ones in the second
    int32_t pair8;            // k_vtile pairs over odd extents: pairs may start at odd
                              // element offsets (8-byte aligned 16-byte accesses), and
                              // a partial tile may hold one element of a pair only
    // Tiles over several blocks.  When the X and Y groups are the two
    // fastest non-shared axes of both sides, swapped (out_axis[s] = s+1,
    // out_axis[s+1] = s), a tile of the whole X and Y axes is one contiguous
    // block on both sides, however small.  Such a tile may also span tw
    // values of one more axis W (one LDS plane per value), so that small
    // blocks still make large tiles.  The partial-tile mask then bounds w
    // instead of the X top axis, which is whole.
    int32_t tw;               // W values per tile (1: no W axis)
    int32_t w_axis;           // canonical axis of W (-1: none)
    int32_t w_top, w_nchunk, w_istr, w_ostr;
    int32_t lds_plane;        // LDS elements of one W plane (Tx * lds_row)
    int32_t lds_tile;         // LDS elements of one tile (tw planes)
    int32_t m_top, m_ch;      // masked X-side coordinate: extent and chunk
                              // (x_top and tx, or w_top and tw)
    // walk order (tile_origin): the tile grid's digits (x chunk, y chunk, Z
    // axes) fastest first, one pair of them (gp, gp+1) optionally grouped
    int32_t ng;                         // digits
    int32_t g_ext[MAXG], g_istr[MAXG], g_ostr[MAXG];
    int32_t g_xc, g_yc;                 // digit holding the x / y chunk (-1: none)
    int32_t gp, gw;                     // grouped digit and group width (gw 0: plain)
    int32_t gwlast, gblk, gwg, glast;   // last group's width, blk = ext[gp]*ext[gp+1],
                                        // wg = gw*ext[gp+1], index of the last group
    int32_t walk_mode, wx;              // as requested (for describe)
    int32_t chunk;                      // k_vtile: consecutive tiles per workgroup turn

    // the divisors of the index math, prepared for multiplication
    FastDiv fd_C, fd_Tx, fd_Ty, fd_xlow, fd_ylow, fd_xnch, fd_ynch;
    FastDiv fd_g[MAXG], fd_gblk, fd_gwg, fd_gw, fd_gwlast;
    FastDiv fd_x[MAXR], fd_y[MAXR], fd_z[MAXR];
};

// ---------------------------------------------------------------------------
// per-slot index math (a slot is one element position inside a tile)
// ---------------------------------------------------------------------------

// LDS row of tile row xr.  With xsplit the read phase's pairs (xr, xr+1),
// taken by consecutive lanes, land in rows xr/2 and Tx/2 + xr/2: consecutive
// lanes then store one (odd) row pitch apart instead of two, which spreads
// them over all LDS banks, and a pair's two rows are a constant Tx/2 rows
// apart.
__host__ __device__ __attribute__((always_inline)) inline uint32_t lds_xrow(const Plan& P, uint32_t xr)
{
    return P.xsplit ? (xr & 1u) * ((uint32_t)P.Tx >> 1) + (xr >> 1) : xr;
}

// Read-phase slot e (input order: c, xr, yr, w): relative input offset,
// LDS index and the (x top or w, y top) coordinates used to mask partial
// tiles.
__host__ __device__ __attribute__((always_inline)) inline void read_slot(const Plan& P, int e, uint32_t& in_rel,
                                          uint32_t& lds, uint32_t& mk)
{
    if (e >= P.V) { in_rel = 0; lds = 0; mk = kBadSlot; return; }
    const uint32_t ue = (uint32_t)e;
    const uint32_t r = fdiv(ue, P.fd_C);
    const uint32_t c = ue - r * (uint32_t)P.C;
    const uint32_t yw = fdiv(r, P.fd_Tx);
    const uint32_t xr = r - yw * (uint32_t)P.Tx;
    const uint32_t w = fdiv(yw, P.fd_Ty);  // 0 without a W axis
    const uint32_t yr0 = yw - w * (uint32_t)P.Ty;
    const uint32_t xt = fdiv(xr, P.fd_xlow);
    uint32_t yr = yr0, off = 0, yt = 0;
    for (int j = 0; j < P.ny; ++j) {
        if (j == P.ny - 1) {
            yt = yr;
            off += yr * (uint32_t)P.y_istr[j];
        } else {
            const uint32_t q = fdiv(yr, P.fd_y[j]);
            off += (yr - q * (uint32_t)P.y_ext[j]) * (uint32_t)P.y_istr[j];
            yr = q;
        }
    }
    in_rel = c + (uint32_t)P.C * xr + off + w * (uint32_t)P.w_istr;
    lds = w * (uint32_t)P.lds_plane + lds_xrow(P, xr) * (uint32_t)P.lds_row + yr0 * (uint32_t)P.C + c;
    mk = (P.tw > 1 ? w : xt) | (yt << 16);
}

// Write-phase slot e (output order: c, yr, xr, w).
__host__ __device__ __attribute__((always_inline)) inline void write_slot(const Plan& P, int e, uint32_t& out_rel,
                                           uint32_t& lds, uint32_t& mk)
{
    if (e >= P.V) { out_rel = 0; lds = 0; mk = kBadSlot; return; }
    const uint32_t ue = (uint32_t)e;
    const uint32_t r = fdiv(ue, P.fd_C);
    const uint32_t c = ue - r * (uint32_t)P.C;
    const uint32_t xw = fdiv(r, P.fd_Ty);
    const uint32_t yr = r - xw * (uint32_t)P.Ty;
    const uint32_t w = fdiv(xw, P.fd_Tx);  // 0 without a W axis
    const uint32_t xr0 = xw - w * (uint32_t)P.Tx;
    const uint32_t yt = fdiv(yr, P.fd_ylow);
    uint32_t xr = xr0, off = 0, xt = 0;
    for (int j = 0; j < P.nx; ++j) {
        if (j == P.nx - 1) {
            xt = xr;
            off += xr * (uint32_t)P.x_ostr[j];
        } else {
            const uint32_t q = fdiv(xr, P.fd_x[j]);
            off += (xr - q * (uint32_t)P.x_ext[j]) * (uint32_t)P.x_ostr[j];
            xr = q;
        }
    }
    out_rel = c + (uint32_t)P.C * yr + off + w * (uint32_t)P.w_ostr;
    lds = w * (uint32_t)P.lds_plane + lds_xrow(P, xr0) * (uint32_t)P.lds_row + yr * (uint32_t)P.C + c;
    mk = (P.tw > 1 ? w : xt) | (yt << 16);
}


// Origin of tile t: input / output base offsets and the valid extents of
// the two masked coordinates (lim = rx | ry << 16: the X top axis, or W,
// and the Y top axis).  Plans keep every element offset
// below 2^31 (as the reference's default-INTEGER index arithmetic does), so
// the walk uses 32-bit unsigned arithmetic.
//
// Walk order.  The tile grid has digits (x chunk, y chunk, each Z axis)
// walked in the plan's digit order, fastest first.  Optionally the digit
// pair (gp, gp+1) is walked in groups of gw values of digit gp: digit gp
// fastest inside the group, then every value of digit gp+1, then the next
// group (the last group may be narrower, gwlast).  Tiles in flight at the
// same time then span about gw values of one digit and many of the other,
// so the runs they read and write together are long on both HBM sides;
// the group width trades DRAM page locality between the two sides.
__host__ __device__ __attribute__((always_inline)) inline void tile_origin(const Plan& P, uint32_t t, uint32_t& bin,
                                            uint32_t& bout, uint32_t& lim)
{
    uint32_t rest = t, bi = 0, bo = 0, xc = 0, yc = 0;
    bool skip = false;
    // unrolled over MAXG so every digit's divisor and strides are fixed
    // kernel argument offsets the compiler keeps in scalar registers
#pragma nounroll
    for (int i = 0; i < MAXG; ++i) {
        if (i >= P.ng) break;
        if (skip) {
            skip = false;
            continue;
        }
        uint32_t d0, d1 = 0;
        const bool grouped = P.gw != 0 && i == P.gp && i + 1 < MAXG;
        if (grouped) {
            const uint32_t blk = fdiv(rest, P.fd_gblk);
            const uint32_t r = rest - blk * (uint32_t)P.gblk;
            const uint32_t grp = fdiv(r, P.fd_gwg);
            const uint32_t w = r - grp * (uint32_t)P.gwg;
            const bool narrow = grp == (uint32_t)P.glast;
            const uint32_t a = narrow ? fdiv(w, P.fd_gwlast) : fdiv(w, P.fd_gw);
            const uint32_t width = narrow ? (uint32_t)P.gwlast : (uint32_t)P.gw;
            d0 = grp * (uint32_t)P.gw + (w - a * width);
            d1 = a;
            rest = blk;
        } else {
            const uint32_t q = fdiv(rest, P.fd_g[i]);
            d0 = rest - q * (uint32_t)P.g_ext[i];
            rest = q;
        }
        bi += d0 * (uint32_t)P.g_istr[i];
        bo += d0 * (uint32_t)P.g_ostr[i];
        if (i == P.g_xc) xc = d0;
        if (i == P.g_yc) yc = d0;
        if (grouped) {
            bi += d1 * (uint32_t)P.g_istr[i + 1];
            bo += d1 * (uint32_t)P.g_ostr[i + 1];
            if (i + 1 == P.g_xc) xc = d1;
            if (i + 1 == P.g_yc) yc = d1;
            skip = true;
        }
    }
    const int rxl = P.m_top - (int)xc * P.m_ch, ryl = P.y_top - (int)yc * P.ty;
    const int rx = rxl < P.m_ch ? rxl : P.m_ch;
    const int ry = ryl < P.ty ? ryl : P.ty;
    bin = bi;
    bout = bo;
    lim = (uint32_t)rx | ((uint32_t)ry << 16);
}

__host__ __device__ __attribute__((always_inline)) inline bool slot_ok(uint32_t mk, uint32_t lim)
{
    return (mk & 0xffffu) < (lim & 0xffffu) && (mk >> 16) < (lim >> 16);
}


// Row r of a rows plan: base offsets of its C contiguous elements.
__host__ __device__ __attribute__((always_inline)) inline void row_origin(const Plan& P, uint32_t r, uint32_t& bin,
                                           uint32_t& bout)
{
    uint32_t bi = 0, bo = 0;
    uint32_t rest = r;
    for (int j = 0; j < P.nz; ++j) {
        const uint32_t q = fdiv(rest, P.fd_z[j]);
        const uint32_t c = rest - q * (uint32_t)P.z_ext[j];
        bi += c * (uint32_t)P.z_istr[j];
        bo += c * (uint32_t)P.z_ostr[j];
        rest = q;
    }
    bin = bi;
    bout = bo;
}

// Logical tile-walker index of workgroup w: with xcd_remap the workgroups
// that the dispatcher places on one XCD (w % 8 == x) take a contiguous run
// of logical indices, so neighbouring tiles (which share cache lines at
// their edges) stay in one XCD's L2.
__host__ __device__ __attribute__((always_inline)) inline uint32_t logical_wg(uint32_t w, uint32_t G, int remap)
{
    if (!remap || (G & 7u)) return w;
    return (w & 7u) * (G >> 3) + (w >> 3);
}

// The element-wise update.  The reference rounds the product and the sum
// separately (y + (x*f) on a baseline x86-64 build without FMA), so
// contraction into a fused multiply-add is switched off here as well as on
// the command line.
// Global accesses.  Every input element is read once and every output
// element written once, so by default they bypass cache retention
// (nontemporal); the cached form is kept for the autotuner to compare.
template <bool NT>
__device__ __attribute__((always_inline)) inline double ld(const double* p)
{
    if (NT) return __builtin_nontemporal_load(p);
    return *p;
}
template <bool NT>
__device__ __attribute__((always_inline)) inline void st(double* p, double v)
{
    if (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

template <int MODE>
__host__ __device__ __attribute__((always_inline)) inline double update(double x, double y, double f, double g)
{
#pragma clang fp contract(off)
    if (MODE == MODE_SORT) return x * f;
    if (MODE == MODE_ACC) {
        const double p = x * f;
        return y + p;
    }
    const double p = x * f;
    const double q = y * g;
    return q + p;
}

}  // namespace spg
