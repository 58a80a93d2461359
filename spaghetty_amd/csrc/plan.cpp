// plan.cpp -- validation, canonicalisation and tile planning (host only).
//
// The reference picks, per permutation, the CPU loop order that ran fastest
// (trunk/makelib.py:39-62, found by brute force in trunk/generator.py:
// 162-313).  Here the same role is played by a plan: the tile box a
// workgroup stages through LDS, chosen by a cost model of HBM traffic and
// refined on the device by the autotuner (spaghetty_amd/autotune.py).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "spg_internal.hpp"

namespace spg {

int canonicalize(int rank, const int* dims, const int* perm, Canon* c)
{
    if (rank < 0 || rank > MAXR) return SPG_EINVAL_DIM;
    if (rank > 0 && (!dims || !perm)) return SPG_EINVAL;
    bool seen[MAXR] = {false};
    for (int q = 0; q < rank; ++q) {
        const int p = perm[q];
        if (p < 1 || p > rank || seen[p - 1]) return SPG_EINVAL_PERM;
        seen[p - 1] = true;
    }
    int64_t n = 1;
    for (int q = 0; q < rank; ++q) {
        if (dims[q] < 0) return SPG_EINVAL_DIM;
        n *= dims[q];
        if (n > 0x7fffffffLL) return SPG_EINVAL_DIM;
    }
    Canon k;
    k.nelem = n;
    // fastest-first numbering: input axis a <-> NWChem axis rank-a (1-based)
    int ext[MAXR], oax[MAXR];
    for (int a = 0; a < rank; ++a) ext[a] = dims[rank - 1 - a];
    for (int q = 0; q < rank; ++q) oax[q] = rank - perm[rank - 1 - q];
    if (n == 0) {
        k.r = 0;
        *c = k;
        return SPG_OK;
    }
    // drop extent-1 axes
    int keep_id[MAXR], r = 0;
    for (int a = 0; a < rank; ++a) keep_id[a] = ext[a] > 1 ? r++ : -1;
    int n2[MAXR], o2[MAXR], nq = 0;
    for (int a = 0; a < rank; ++a)
        if (keep_id[a] >= 0) n2[keep_id[a]] = ext[a];
    for (int q = 0; q < rank; ++q)
        if (keep_id[oax[q]] >= 0) o2[nq++] = keep_id[oax[q]];
    // fuse input axes a, a+1 that are adjacent and in order in the output
    bool merged = true;
    while (merged && r > 1) {
        merged = false;
        int pos[MAXR];
        for (int q = 0; q < r; ++q) pos[o2[q]] = q;
        for (int a = 0; a + 1 < r; ++a) {
            if (pos[a + 1] == pos[a] + 1) {
                n2[a] *= n2[a + 1];
                for (int b = a + 1; b + 1 < r; ++b) n2[b] = n2[b + 1];
                int w = 0;
                for (int q = 0; q < r; ++q) {
                    if (o2[q] == a + 1) continue;
                    o2[w++] = o2[q] > a + 1 ? o2[q] - 1 : o2[q];
                }
                --r;
                merged = true;
                break;
            }
        }
    }
    k.r = r;
    for (int a = 0; a < r; ++a) {
        k.n[a] = n2[a];
        k.out_axis[a] = o2[a];
    }
    *c = k;
    return SPG_OK;
}

std::string dispatch_key(const Canon& c, int mode)
{
    if (c.nelem == 0) return "empty";
    std::string s = "r" + std::to_string(c.r) + (mode == MODE_SORT ? ":s" : ":a");
    s += ":n";
    for (int a = 0; a < c.r; ++a) s += (a ? "x" : "") + std::to_string(c.n[a]);
    s += ":p";
    for (int a = 0; a < c.r; ++a) s += std::to_string(c.out_axis[a]);
    return s;
}

namespace {

struct Strides {
    int64_t in[MAXR];
    int64_t out[MAXR];
    int pos[MAXR];
};

Strides strides_of(const Canon& c)
{
    Strides s;
    int64_t m = 1;
    for (int a = 0; a < c.r; ++a) {
        s.in[a] = m;
        m *= c.n[a];
    }
    m = 1;
    for (int q = 0; q < c.r; ++q) {
        s.out[c.out_axis[q]] = m;
        s.pos[c.out_axis[q]] = q;
        m *= c.n[c.out_axis[q]];
    }
    return s;
}

void base_plan(const Canon& c, int mode, Plan* P)
{
    std::memset(P, 0, sizeof(Plan));
    P->nelem = c.nelem;
    P->mode = mode;
    P->rank = c.r;
    for (int a = 0; a < c.r; ++a) {
        P->ext[a] = c.n[a];
        P->out_axis[a] = c.out_axis[a];
    }
    P->C = 1;
    P->x_low = P->y_low = 1;
    P->tx = P->ty = 1;
    P->x_top = P->y_top = 1;
    P->x_nchunk = P->y_nchunk = 1;
    P->Tx = P->Ty = 1;
    P->V = 1;
    P->vec = 1;
    P->tw = 1;
    P->w_axis = -1;
    P->w_top = P->w_nchunk = 1;
    P->m_top = P->m_ch = 1;
}

// Rows plan: axis 0 is contiguous in input and output; rows are the rest.
void make_rows(const Canon& c, int mode, Plan* P)
{
    base_plan(c, mode, P);
    Strides st = strides_of(c);
    P->kind = KIND_ROWS;
    P->C = c.r ? c.n[0] : 1;
    int nz = 0;
    for (int a = 1; a < c.r; ++a) {
        P->z_ext[nz] = c.n[a];
        P->z_istr[nz] = (int32_t)st.in[a];
        P->z_ostr[nz] = (int32_t)st.out[a];
        ++nz;
    }
    P->nz = nz;
    P->ntiles = c.nelem / P->C;
}

// Tiled plan with the given groups and top chunks; false if invalid.
bool make_tiled(const Canon& c, int mode, int nx, int tx, int ny, int ty, Plan* P)
{
    base_plan(c, mode, P);
    Strides st = strides_of(c);
    const bool shared = c.r >= 1 && c.out_axis[0] == 0;
    const int s = shared ? 1 : 0;
    if (s + nx > c.r || s + ny > c.r) return false;
    if (!shared && (nx < 1 || ny < 1)) return false;
    bool used[MAXR] = {false};
    if (shared) used[0] = true;
    for (int j = 0; j < nx; ++j) used[s + j] = true;
    for (int j = 0; j < ny; ++j) {
        const int a = c.out_axis[s + j];
        if (used[a]) return false;
        used[a] = true;
    }
    P->kind = KIND_TILED;
    P->C = shared ? c.n[0] : 1;
    P->nx = nx;
    P->ny = ny;
    int64_t low = 1;
    for (int j = 0; j < nx; ++j) {
        const int a = s + j;
        P->x_ext[j] = c.n[a];
        P->x_ostr[j] = (int32_t)st.out[a];
        if (j < nx - 1) low *= c.n[a];
    }
    P->x_low = (int32_t)low;
    if (nx > 0) {
        P->x_top = c.n[s + nx - 1];
        if (tx < 1 || tx > P->x_top) return false;
        P->tx = tx;
        P->x_top_istr = (int32_t)st.in[s + nx - 1];
        P->x_top_ostr = (int32_t)st.out[s + nx - 1];
    }
    low = 1;
    for (int j = 0; j < ny; ++j) {
        const int a = c.out_axis[s + j];
        P->y_ext[j] = c.n[a];
        P->y_istr[j] = (int32_t)st.in[a];
        if (j < ny - 1) low *= c.n[a];
    }
    P->y_low = (int32_t)low;
    if (ny > 0) {
        const int a = c.out_axis[s + ny - 1];
        P->y_top = c.n[a];
        if (ty < 1 || ty > P->y_top) return false;
        P->ty = ty;
        P->y_top_istr = (int32_t)st.in[a];
        P->y_top_ostr = (int32_t)st.out[a];
    }
    P->x_nchunk = (P->x_top + P->tx - 1) / P->tx;
    P->y_nchunk = (P->y_top + P->ty - 1) / P->ty;
    const int64_t Tx = (int64_t)P->x_low * P->tx, Ty = (int64_t)P->y_low * P->ty;
    const int64_t V = (int64_t)P->C * Tx * Ty;
    if (V > kMaxTileElemsV || Tx > 0xfffe || Ty > 0xfffe) return false;
    if (P->tx > 0xfffe || P->ty > 0xfffe) return false;
    P->Tx = (int32_t)Tx;
    P->Ty = (int32_t)Ty;
    P->V = (int32_t)V;
    // Z: remaining axes, walked in input order
    int nz = 0;
    int64_t zprod = 1;
    for (int a = 0; a < c.r; ++a) {
        if (used[a]) continue;
        P->z_ext[nz] = c.n[a];
        P->z_istr[nz] = (int32_t)st.in[a];
        P->z_ostr[nz] = (int32_t)st.out[a];
        zprod *= c.n[a];
        ++nz;
    }
    P->nz = nz;
    P->ntiles = (int64_t)P->x_nchunk * P->y_nchunk * zprod;
    // LDS pitch: an odd number of doubles spreads the read phase's
    // column-wise stores (consecutive xr) over distinct banks.
    int row = P->Ty * P->C;
    if ((row & 1) == 0) row += 1;
    P->lds_row = row;
    P->lds_plane = P->Tx * row;
    P->lds_tile = P->lds_plane;
    P->m_top = P->x_top;
    P->m_ch = P->tx;
    return true;
}

// The two fastest non-shared axes swap places (X and Y one axis each, both
// whole in every tile): each tile is one block, contiguous on both sides.
bool swapped_pair(const Canon& c)
{
    const int s = c.r >= 1 && c.out_axis[0] == 0 ? 1 : 0;
    return c.r >= s + 3 && c.out_axis[s] == s + 1 && c.out_axis[s + 1] == s;
}

// Turns the one-block tile of a swapped pair (make_tiled with nx = ny = 1
// and whole chunks) into a tile of tw blocks along axis a, one of the
// remaining axes; false if that does not fit.
bool add_block_axis(const Canon& c, int a, int tw, Plan* P)
{
    const int s = P->C > 1 ? 1 : 0;
    if (!swapped_pair(c) || P->nx != 1 || P->ny != 1 || P->tx != P->x_top || P->ty != P->y_top) return false;
    if (a < s + 2 || a >= c.r || tw < 2 || tw > c.n[a] || c.n[a] > 0xfffe) return false;
    const int64_t V = (int64_t)P->V * tw;
    if (V > kMaxTileElemsV) return false;
    Strides st = strides_of(c);
    // Z: the axes past the block except W, in input order (as make_tiled)
    int nz = 0;
    int64_t zprod = 1;
    for (int b = s + 2; b < c.r; ++b) {
        if (b == a) continue;
        P->z_ext[nz] = c.n[b];
        P->z_istr[nz] = (int32_t)st.in[b];
        P->z_ostr[nz] = (int32_t)st.out[b];
        zprod *= c.n[b];
        ++nz;
    }
    P->nz = nz;
    P->tw = tw;
    P->w_axis = a;
    P->w_top = c.n[a];
    P->w_nchunk = (c.n[a] + tw - 1) / tw;
    P->w_istr = (int32_t)st.in[a];
    P->w_ostr = (int32_t)st.out[a];
    P->V = (int32_t)V;
    P->lds_tile = tw * P->lds_plane;
    P->m_top = P->w_top;
    P->m_ch = tw;
    P->ntiles = (int64_t)P->w_nchunk * zprod;
    return true;
}

// Magic-number forms of every divisor the index math uses.
void prepare_divisors(Plan* P)
{
    P->fd_C = make_fastdiv((uint32_t)P->C);
    P->fd_Tx = make_fastdiv((uint32_t)P->Tx);
    P->fd_Ty = make_fastdiv((uint32_t)P->Ty);
    P->fd_xlow = make_fastdiv((uint32_t)P->x_low);
    P->fd_ylow = make_fastdiv((uint32_t)P->y_low);
    P->fd_xnch = make_fastdiv((uint32_t)P->x_nchunk);
    P->fd_ynch = make_fastdiv((uint32_t)P->y_nchunk);
    for (int j = 0; j < MAXG; ++j) P->fd_g[j] = make_fastdiv((uint32_t)(j < P->ng ? P->g_ext[j] : 1));
    P->fd_gblk = make_fastdiv((uint32_t)P->gblk);
    P->fd_gwg = make_fastdiv((uint32_t)P->gwg);
    P->fd_gw = make_fastdiv((uint32_t)P->gw);
    P->fd_gwlast = make_fastdiv((uint32_t)P->gwlast);
    for (int j = 0; j < MAXR; ++j) {
        P->fd_x[j] = make_fastdiv((uint32_t)(j < P->nx ? P->x_ext[j] : 1));
        P->fd_y[j] = make_fastdiv((uint32_t)(j < P->ny ? P->y_ext[j] : 1));
        P->fd_z[j] = make_fastdiv((uint32_t)(j < P->nz ? P->z_ext[j] : 1));
    }
}

// k_vtile with two elements per slot: every pair must be two neighbouring
// addresses on its HBM side.  With a shared contiguous axis (C > 1) pairs
// run along c, so C even suffices (every other stride is a multiple of C,
// and pairs start at even offsets inside one partial-tile mask).  Without
// one, read pairs (xr, xr+1) run along the X group, whose fused index is
// the input offset, and write pairs (yr, yr+1) along the Y group, whose
// fused index is the output offset: Tx and Ty even keep every pair inside
// one tile row on both sides, whatever the extents.
bool vtile_vec2_ok(const Plan& P)
{
    if (P.C > 1) return P.C % 2 == 0;
    return P.Tx % 2 == 0 && P.Ty % 2 == 0;
}

// Pairs of a plan without a shared axis start at even offsets and keep
// their partial-tile mask when the input's and the output's contiguous
// extents are even (the valid extents x_top - xc*tx and y_top - yc*ty are
// then even, and every stride is even); otherwise pairs may start at odd
// offsets and a partial tile may cut one (Plan.pair8).
bool vtile_pair8(const Plan& P)
{
    return P.C == 1 && (P.ext[0] % 2 != 0 || P.ext[P.out_axis[0]] % 2 != 0);
}

// Largest tile-relative element offset of any slot; k_vtile addresses a
// tile through a buffer resource with 31-bit byte offsets.
int64_t max_slot_offset(const Plan& P)
{
    uint32_t mx = 0;
    for (int e = 0; e < P.V; ++e) {
        uint32_t off, lds, mk;
        read_slot(P, e, off, lds, mk);
        mx = std::max(mx, off);
        write_slot(P, e, off, lds, mk);
        mx = std::max(mx, off);
    }
    return (int64_t)mx;
}

// Walk order of a tiled plan (tile_origin): the grid digits and the grouped
// pair.  Modes (TileParams.walk_mode), with g = the group width:
//   0  x chunk, y chunk, Z axes in input order; g groups x chunks against y
//   1  x chunk, then the Z axis with the smallest input stride (the next
//      run of the input), y chunk, other Z axes; g groups that Z axis
//      against y chunks
//   2  y chunk, the Z axis with the smallest output stride, x chunk, other
//      Z axes; g groups that Z axis against x chunks
//   3  y chunk, x chunk, Z axes; g groups y chunks against x chunks
// A mode that needs a Z axis the plan lacks, or a group as wide as its
// digit, falls back to the plain order of that mode.  Digits of extent 1
// are dropped.  Returns false for an unknown mode.
bool set_walk(Plan* P, int mode, int g)
{
    P->ng = 0;
    P->g_xc = P->g_yc = -1;
    P->gp = P->gw = P->gwlast = P->gblk = P->gwg = 0;
    P->glast = -1;
    P->walk_mode = mode;
    P->wx = g;
    if (P->kind != KIND_TILED) return true;
    const bool valid = mode >= 0 && mode <= 3;
    if (!valid) mode = 0;
    struct D {
        int ext, istr, ostr, tag;  // tag: 0 x chunk, 1 y chunk, 2 Z axis
    };
    // the masked X-side digit: X top chunks, or W chunks (X is whole then)
    D xd = P->tw > 1 ? D{P->w_nchunk, P->tw * P->w_istr, P->tw * P->w_ostr, 0}
                     : D{P->x_nchunk, P->tx * P->x_top_istr, P->tx * P->x_top_ostr, 0};
    D yd{P->y_nchunk, P->ty * P->y_top_istr, P->ty * P->y_top_ostr, 1};
    std::vector<D> z;
    for (int j = 0; j < P->nz; ++j) z.push_back(D{P->z_ext[j], P->z_istr[j], P->z_ostr[j], 2});
    auto pick = [&z](bool by_in) {
        int b = -1;
        for (int j = 0; j < (int)z.size(); ++j)
            if (z[j].ext > 1 && (b < 0 || (by_in ? z[j].istr < z[b].istr : z[j].ostr < z[b].ostr))) b = j;
        return b;
    };
    std::vector<D> order;
    int gp = 0;
    if (mode == 0 || mode == 3) {
        order.push_back(mode == 0 ? xd : yd);
        order.push_back(mode == 0 ? yd : xd);
        for (auto& d : z) order.push_back(d);
    } else {
        const int b = pick(mode == 1);
        order.push_back(mode == 1 ? xd : yd);
        if (b >= 0) {
            order.push_back(z[b]);
            gp = 1;
        }
        order.push_back(mode == 1 ? yd : xd);
        for (int j = 0; j < (int)z.size(); ++j)
            if (j != b) order.push_back(z[j]);
    }
    // drop extent-1 digits, remembering where the grouped pair went
    std::vector<D> kept;
    int gp_kept = -1;
    for (int k = 0; k < (int)order.size(); ++k) {
        if (order[k].ext <= 1) continue;
        if (k == gp) gp_kept = (int)kept.size();
        kept.push_back(order[k]);
    }
    P->ng = (int)kept.size();
    for (int k = 0; k < P->ng; ++k) {
        P->g_ext[k] = kept[k].ext;
        P->g_istr[k] = kept[k].istr;
        P->g_ostr[k] = kept[k].ostr;
        if (kept[k].tag == 0) P->g_xc = k;
        if (kept[k].tag == 1) P->g_yc = k;
    }
    // the grouped pair: digit gp and the one after it, both still present
    const bool pair_ok = gp_kept >= 0 && gp_kept + 1 < P->ng;
    if (g > 0 && pair_ok && g < P->g_ext[gp_kept]) {
        P->gp = gp_kept;
        P->gw = g;
        const int groups = (P->g_ext[gp_kept] + g - 1) / g;
        P->glast = groups - 1;
        P->gwlast = P->g_ext[gp_kept] - (groups - 1) * g;
        P->gblk = P->g_ext[gp_kept] * P->g_ext[gp_kept + 1];
        P->gwg = g * P->g_ext[gp_kept + 1];
    }
    return valid;
}

// Launch parameters of a plan; false if the tile cannot run (a tile above
// kMaxTileElems needs k_vtile with pairs).
bool finish_launch(Plan* P, const spg_tile_params* tp)
{
    set_walk(P, tp ? tp->walk_mode : 0, tp ? tp->walk_x : 0);
    P->chunk = (tp && tp->walk_chunk > 1) ? tp->walk_chunk : 1;
    prepare_divisors(P);
    if (P->kind == KIND_TILED) {
        const int fl = (tp && tp->flags >= 0) ? tp->flags : 0;
        const bool big = P->V > kMaxTileElems;
        int block = (tp && tp->block > 0) ? tp->block : 0;
        if (big) {
            // pairs, at most 8 per thread: block >= V/16, up to kMaxBlockV
            // (512 for a pipelined accumulate)
            const int need = ((P->V + 16 * 64 - 1) / (16 * 64)) * 64;
            if (block < need) block = need;
            if (block > max_block_v(P->mode, (fl & SPG_TILE_PIPE) != 0)) return false;
        } else {
            if (!block) {
                block = P->V >= 1024 ? 256 : ((P->V + 63) / 64) * 64;
                if (block < 64) block = 64;
                if (block > 256) block = 256;
            }
            // at most kMaxSlots elements per thread; 16 slots only with <= 256
            // threads (the classic kernel's launch bounds), so tiles stay <= 4096
            if ((P->V + block - 1) / block > kMaxSlots)
                block = ((P->V + kMaxSlots * 64 - 1) / (kMaxSlots * 64)) * 64;
            if ((P->V + block - 1) / block > 8 && block > 256) block = 256;
            if (block > 512) block = 512;
        }
        P->block = block;
        P->slots = (P->V + block - 1) / block;
        P->lds_bytes = 2 * P->lds_tile * (int)sizeof(double);
        P->tk = 0;
        P->vec = 1;
        if (!(fl & SPG_TILE_CLASSIC) && (max_slot_offset(*P) + 2) * 8 < kTileRecords) {
            P->tk = 1;
            P->vec = (!(fl & SPG_TILE_SCALAR) && vtile_vec2_ok(*P)) ? 2 : 1;
            P->pair8 = P->vec == 2 && vtile_pair8(*P);
            P->rstep = P->C > 1 ? 1 : P->lds_row;
            P->slots = (P->V + P->vec * block - 1) / (P->vec * block);
            // the dummy word (and, for pairs, its partner at +rstep) lies
            // past the tile; buffers keep 16-byte alignment
            P->vdummy = P->lds_tile;
            int half = P->vdummy + P->rstep + 1;
            P->xsplit = 0;
            if (P->vec == 2 && P->C == 1 && P->lds_row > P->Ty && !(fl & SPG_TILE_NOSPLIT)) {
                // pairs (xr, xr+1) in rows xr/2 and Tx/2 + xr/2; the dummy
                // word and its partner are the pad words of rows 0 and
                // Tx/2, which the write phase never reads
                P->xsplit = 1;
                P->rstep = (P->Tx / 2) * P->lds_row;
                P->vdummy = P->Ty;
                half = P->lds_tile;
            }
            half = (half + 1) & ~1;
            P->half = half;
            P->lds1 = (fl & SPG_TILE_LDS1) ? 1 : 0;
            // two buffers of a big tile may not fit the CU's 160 KB
            if (2 * half * (int)sizeof(double) > 160 * 1024) P->lds1 = 1;
            P->lds_bytes = (P->lds1 ? 1 : 2) * half * (int)sizeof(double);
        }
        if (big && !(P->tk == 1 && P->vec == 2 && P->slots <= 8)) return false;
    } else if (P->kind == KIND_ROWS) {
        P->block = (tp && tp->block > 0) ? tp->block : 256;
        P->vec = (P->C % 2 == 0) ? 2 : 1;
        P->slots = 0;
        P->lds_bytes = 0;
    }
    P->grid = 0;  // resolved at launch from the device's CU count
    P->xcd_remap = (tp && tp->xcd_remap >= 0) ? tp->xcd_remap : 1;
    // Nontemporal access pays off on long contiguous streams (rows plans);
    // a transposing tile writes partial cache lines that the L2 should merge.
    if (tp && tp->flags >= 0) {
        P->nt = (tp->flags & SPG_TILE_NT) ? 1 : ((tp->flags & SPG_TILE_NT_LOAD) ? 2 : 0);
        P->pipe = (tp->flags & SPG_TILE_PIPE) ? 1 : 0;
    } else {
        P->nt = P->kind == KIND_ROWS ? 1 : 0;
        P->pipe = 0;
    }
    if (tp && tp->grid_per_cu > 0) P->grid = -tp->grid_per_cu;  // per-CU request
    return true;
}

// Cache lines (128 B) touched by a set of element offsets, sorted runs.
double lines_of(std::vector<uint32_t>& offs)
{
    if (offs.empty()) return 0.0;
    std::sort(offs.begin(), offs.end());
    double lines = 0.0;
    size_t i = 0;
    while (i < offs.size()) {
        size_t j = i + 1;
        while (j < offs.size() && offs[j] == offs[j - 1] + 1) ++j;
        // a run of L doubles at random alignment touches (8L + 120)/128
        // lines on average
        lines += (8.0 * (double)(j - i) + 120.0) / 128.0;
        i = j;
    }
    return lines;
}

// Fast estimate: one run per row on each side.
double quick_cost(const Plan& P)
{
    const double acc = P.mode == MODE_SORT ? 1.0 : 2.0;
    const double N = (double)P.nelem;
    if (P.kind == KIND_ROWS) {
        const double rows = (double)P.ntiles;
        const double lines = (8.0 * P.C + 120.0) / 128.0;
        return rows * lines * 128.0 * (1.0 + acc) + rows * 64.0;
    }
    const double lr = P.Ty * (8.0 * P.C * P.Tx + 120.0) / 128.0;
    const double lw = P.Tx * (8.0 * P.C * P.Ty + 120.0) / 128.0;
    const double per = N / P.V;
    double cost = per * (lr + lw * acc) * 128.0;
    cost += (double)P.ntiles * (384.0 + 1.0 * P.V);
    return cost;
}

}  // namespace

double plan_cost(const Plan& P)
{
    if (P.kind != KIND_TILED) return quick_cost(P);
    const double acc = P.mode == MODE_SORT ? 1.0 : 2.0;
    std::vector<uint32_t> ri, wo;
    ri.reserve(P.V);
    wo.reserve(P.V);
    for (int e = 0; e < P.V; ++e) {
        uint32_t off, lds, mk;
        read_slot(P, e, off, lds, mk);
        ri.push_back(off);
        write_slot(P, e, off, lds, mk);
        wo.push_back(off);
    }
    const double lr = lines_of(ri), lw = lines_of(wo);
    const double per = (double)P.nelem / P.V;
    double cost = per * (lr + lw * acc) * 128.0;
    cost += (double)P.ntiles * (384.0 + 1.0 * P.V);
    // too few tiles to fill 256 CUs: the tail is idle
    if (P.ntiles < 2048) cost *= 1.0 + 0.35 * (2048.0 - (double)P.ntiles) / 2048.0;
    return cost;
}

namespace {

std::vector<int> chunk_choices(int m, int64_t unit, int max_tile)
{
    std::vector<int> v;
    if (m <= 1) {
        v.push_back(1);
        return v;
    }
    for (int t = 1; t <= m; ++t) {
        if ((int64_t)t * unit > max_tile) break;
        bool take = m <= 48 || (m % t == 0) || (t % 8 == 0) || t == m;
        if (take) v.push_back(t);
    }
    if (v.empty()) v.push_back(1);
    return v;
}

}  // namespace

int build_plan(const Canon& c, int mode, const spg_tile_params* tp, Plan* out)
{
    Plan best;
    if (c.nelem == 0) {
        base_plan(c, mode, &best);
        best.kind = KIND_EMPTY;
        prepare_divisors(&best);
        *out = best;
        return SPG_OK;
    }
    if (tp && (tp->walk_mode < 0 || tp->walk_mode > 3 || tp->walk_x < 0 || tp->walk_chunk < 0 ||
               tp->walk_chunk > 4096))
        return SPG_EINVAL;
    const bool shared = c.r >= 1 && c.out_axis[0] == 0;
    // tiles above kMaxTileElems only for k_vtile with pairs: on request (the
    // autotuner's max_tile), or by default for problems of at least
    // kBigTileElems elements, where the longer runs pay (on small problems
    // fewer, larger tiles leave the grid under-filled)
    const bool pairs_ok = !(tp && tp->flags >= 0 && (tp->flags & (SPG_TILE_CLASSIC | SPG_TILE_SCALAR)));
    const bool pipe_req = tp && tp->flags >= 0 && (tp->flags & SPG_TILE_PIPE);
    const int max_tile = (tp && tp->max_tile > 0)
                             ? std::min(tp->max_tile, pairs_ok ? kMaxTileElemsV : kMaxTileElems)
                             : (pairs_ok && c.nelem >= kBigTileElems ? kMaxTileElemsV : kMaxTileElems);
    const int want_kind = tp ? tp->kind : -1;
    // tiles over several blocks: 0 the cost model's choice, 1 none, tw > 1
    // (tw < -1): tw (-tw) blocks along the input's (output's) next axis
    const int want_tw = tp ? tp->tw : 0;
    if (want_tw == -1 || (want_tw != 0 && want_tw != 1 && want_kind == KIND_ROWS)) return SPG_EINVAL;
    const bool forced_w = want_tw > 1 || want_tw < -1;
    if (forced_w && !swapped_pair(c)) return SPG_EINVAL;
    // rank <= 1 after fusion: a contiguous stream
    if (c.r <= 1 || want_kind == KIND_ROWS) {
        if (c.r > 1 && !shared) return SPG_EINVAL;  // rows need a shared contiguous axis
        make_rows(c, mode, &best);
        if (!finish_launch(&best, tp)) return SPG_EINVAL;
        *out = best;
        return SPG_OK;
    }
    struct Cand {
        Plan p;
        double q;
    };
    std::vector<Cand> cands;
    if (shared && want_kind != KIND_TILED) {
        Plan p;
        make_rows(c, mode, &p);
        prepare_divisors(&p);
        cands.push_back({p, quick_cost(p)});
    }
    const int s = shared ? 1 : 0;
    const int nx_lo = shared ? 0 : 1;
    if (forced_w) cands.clear();  // a forced W axis excludes the rows form
    for (int nx = nx_lo; s + nx <= c.r && !forced_w; ++nx) {
        if (tp && tp->nx > 0 && nx != tp->nx) continue;
        int64_t xlow = 1;
        for (int j = 0; j + 1 < nx; ++j) xlow *= c.n[s + j];
        if ((int64_t)(shared ? c.n[0] : 1) * xlow > max_tile) break;
        const int xtop = nx ? c.n[s + nx - 1] : 1;
        std::vector<int> txs = (tp && tp->tx > 0 && nx) ? std::vector<int>{tp->tx}
                                                       : chunk_choices(xtop, (shared ? c.n[0] : 1) * xlow, max_tile);
        for (int ny = nx_lo; s + ny <= c.r; ++ny) {
            if (tp && tp->ny > 0 && ny != tp->ny) continue;
            if (shared && nx == 0 && ny == 0) continue;
            int64_t ylow = 1;
            for (int j = 0; j + 1 < ny; ++j) ylow *= c.n[c.out_axis[s + j]];
            const int ytop = ny ? c.n[c.out_axis[s + ny - 1]] : 1;
            for (int tx : txs) {
                const int64_t unit = (int64_t)(shared ? c.n[0] : 1) * xlow * tx * ylow;
                if (unit > max_tile) continue;
                std::vector<int> tys = (tp && tp->ty > 0 && ny) ? std::vector<int>{tp->ty}
                                                               : chunk_choices(ytop, unit, max_tile);
                for (int ty : tys) {
                    Plan p;
                    if (!make_tiled(c, mode, nx, tx, ny, ty, &p)) continue;
                    prepare_divisors(&p);
                    if (p.V > max_tile) continue;
                    if (p.V > kMaxTileElems && (p.V > max_tile_v(mode, pipe_req) || !vtile_vec2_ok(p)))
                        continue;
                    cands.push_back({p, quick_cost(p)});
                }
            }
        }
    }
    // tiles of several blocks (Plan.tw), ranked by their exact line count
    if (want_tw != 1 && swapped_pair(c) && want_kind != KIND_ROWS && (!tp || tp->nx <= 0 || tp->nx == 1) &&
        (!tp || tp->ny <= 0 || tp->ny == 1) && (!tp || tp->tx <= 0 || tp->tx == c.n[s]) &&
        (!tp || tp->ty <= 0 || tp->ty == c.n[s + 1])) {
        const int64_t blk = (int64_t)(shared ? c.n[0] : 1) * c.n[s] * c.n[s + 1];
        const int axes[2] = {s + 2, c.out_axis[s + 2]};
        for (int side = 0; side < 2; ++side) {
            const int a = axes[side];
            // one axis is next on both sides: its chunks once (a forced tw < -1 names it too)
            if (side == 1 && a == axes[0] && want_tw >= -1) break;
            if ((want_tw > 1 && side != 0) || (want_tw < -1 && side != 1)) continue;
            const std::vector<int> tws = forced_w ? std::vector<int>{want_tw > 0 ? want_tw : -want_tw}
                                                  : chunk_choices(c.n[a], blk, max_tile);
            for (int tw : tws) {
                Plan p;
                if (!make_tiled(c, mode, 1, c.n[s], 1, c.n[s + 1], &p)) continue;
                if (!add_block_axis(c, a, tw, &p)) continue;
                prepare_divisors(&p);
                if (p.V > max_tile) continue;
                if (p.V > kMaxTileElems && (p.V > max_tile_v(mode, pipe_req) || !vtile_vec2_ok(p))) continue;
                cands.push_back({p, plan_cost(p)});
            }
        }
    }
    if (cands.empty()) return SPG_EINVAL;
    std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.q < b.q; });
    // the best two dozen by the fast estimate, re-ranked by the exact line
    // count, then the rest: the first candidate that can launch wins (a tile
    // above kMaxTileElems needs the 16-byte buffer path, which a tile whose
    // offsets leave a buffer resource's 2 GB range cannot take)
    const size_t top = std::min<size_t>(cands.size(), 24);
    for (size_t i = 0; i < top; ++i) cands[i].q = plan_cost(cands[i].p);
    std::stable_sort(cands.begin(), cands.begin() + top,
                     [](const Cand& a, const Cand& b) { return a.q < b.q; });
    for (const Cand& c : cands) {
        best = c.p;
        if (finish_launch(&best, tp)) {
            *out = best;
            return SPG_OK;
        }
    }
    return SPG_EINVAL;
}

std::string describe(const Plan& P)
{
    char buf[2048];
    std::string s = "{";
    auto arr = [](const int32_t* v, int n) {
        std::string t = "[";
        for (int i = 0; i < n; ++i) t += (i ? "," : "") + std::to_string(v[i]);
        return t + "]";
    };
    std::snprintf(buf, sizeof buf,
                  "\"kind\":%d,\"mode\":%d,\"nelem\":%lld,\"rank\":%d,\"C\":%d,\"nx\":%d,"
                  "\"ny\":%d,\"nz\":%d,\"tx\":%d,\"ty\":%d,\"Tx\":%d,\"Ty\":%d,\"V\":%d,"
                  "\"ntiles\":%lld,\"block\":%d,\"slots\":%d,\"lds_bytes\":%d,\"vec\":%d,"
                  "\"xcd_remap\":%d,\"nt\":%d,\"pipe\":%d,\"tk\":%d,\"lds1\":%d,\"wx\":%d,"
                  "\"walk_mode\":%d,\"gw\":%d,\"gp\":%d,\"ng\":%d,\"xsplit\":%d,\"chunk\":%d,\"pair8\":%d,"
                  "\"tw\":%d,\"w_axis\":%d,\"w_top\":%d,\"cost\":%.6g,",
                  P.kind, P.mode, (long long)P.nelem, P.rank, P.C, P.nx, P.ny, P.nz, P.tx, P.ty,
                  P.Tx, P.Ty, P.V, (long long)P.ntiles, P.block, P.slots, P.lds_bytes, P.vec,
                  P.xcd_remap, P.nt, P.pipe, P.tk, P.lds1, P.wx, P.walk_mode, P.gw, P.gp, P.ng, P.xsplit,
                  P.chunk, P.pair8, P.tw, P.w_axis, P.w_top,
                  P.kind == KIND_EMPTY ? 0.0 : plan_cost(P));
    s += buf;
    s += "\"ext\":" + arr(P.ext, P.rank) + ",\"out_axis\":" + arr(P.out_axis, P.rank);
    s += ",\"x_ext\":" + arr(P.x_ext, P.nx) + ",\"y_ext\":" + arr(P.y_ext, P.ny);
    s += ",\"z_ext\":" + arr(P.z_ext, P.nz) + "}";
    return s;
}

int emulate(const Plan& P, const double* in, double* out, std::vector<int>* hits)
{
    const double f = P.alpha, g = P.beta;
    auto upd = [&](double x, double y) {
        if (P.mode == MODE_SORT) return update<MODE_SORT>(x, y, f, g);
        if (P.mode == MODE_ACC) return update<MODE_ACC>(x, y, f, g);
        return update<MODE_AXPBY>(x, y, f, g);
    };
    if (hits) hits->assign((size_t)P.nelem, 0);
    if (P.kind == KIND_EMPTY) return SPG_OK;
    if (P.kind == KIND_ROWS) {
        for (int64_t r = 0; r < P.ntiles; ++r) {
            uint32_t bi, bo;
            row_origin(P, (uint32_t)r, bi, bo);
            for (int c = 0; c < P.C; ++c) {
                const uint32_t o = bo + c;
                if (o >= P.nelem || bi + c >= P.nelem) return SPG_EINVAL;
                out[o] = upd(in[bi + c], out[o]);
                if (hits) (*hits)[o]++;
            }
        }
        return SPG_OK;
    }
    if (P.tk == 1 && P.vec == 2) {
        // k_vtile pairs: element e+1 of a slot starting at even e must be the
        // next address on its HBM side, at +rstep (read) / +1 (write) in LDS.
        // Without pair8, every tile origin and pair is 16-byte aligned and
        // both elements of a pair are valid together for every partial-tile
        // extent (lim) the walk meets; with pair8, partial tiles move the
        // two elements separately, each under its own mask.
        std::vector<uint32_t> lims;
        for (int64_t t = 0; t < P.ntiles; ++t) {
            uint32_t bi, bo, lim;
            tile_origin(P, (uint32_t)t, bi, bo, lim);
            if (!P.pair8 && ((bi | bo) & 1u)) return SPG_EINVAL;
            if (std::find(lims.begin(), lims.end(), lim) == lims.end()) lims.push_back(lim);
        }
        if (P.V & 1) return SPG_EINVAL;
        auto same_mask = [&lims](uint32_t m0, uint32_t m1) {
            for (uint32_t lim : lims)
                if (slot_ok(m0, lim) != slot_ok(m1, lim)) return false;
            return true;
        };
        for (int e = 0; e + 1 < P.V; e += 2) {
            uint32_t o0, l0, m0, o1, l1, m1;
            read_slot(P, e, o0, l0, m0);
            read_slot(P, e + 1, o1, l1, m1);
            if (o1 != o0 + 1 || l1 != l0 + (uint32_t)P.rstep ||
                (!P.pair8 && ((o0 & 1u) || !same_mask(m0, m1))))
                return SPG_EINVAL;
            write_slot(P, e, o0, l0, m0);
            write_slot(P, e + 1, o1, l1, m1);
            if (o1 != o0 + 1 || l1 != l0 + 1 || (!P.pair8 && ((o0 & 1u) || !same_mask(m0, m1))))
                return SPG_EINVAL;
        }
    }
    if (P.tk == 1) {
        // the dummy word and its partner lie past the tile (or, with xsplit,
        // on pad words no slot reads), inside the buffer
        if (P.vdummy + P.rstep >= P.half) return SPG_EINVAL;
        if (P.xsplit) {
            const int row = P.lds_row, pad = P.Ty * P.C;
            if (row <= pad || P.vdummy % row != pad || (P.vdummy + P.rstep) % row != pad)
                return SPG_EINVAL;
        } else if (P.vdummy < P.lds_tile) {
            return SPG_EINVAL;
        }
        if (P.half < P.lds_tile) return SPG_EINVAL;
    }
    std::vector<double> lds((size_t)P.lds_tile, 0.0);
    std::vector<uint8_t> filled(lds.size(), 0);
    for (int64_t t = 0; t < P.ntiles; ++t) {
        uint32_t bi, bo, lim;
        tile_origin(P, (uint32_t)t, bi, bo, lim);
        std::fill(filled.begin(), filled.end(), 0);
        for (int e = 0; e < P.V; ++e) {
            uint32_t rel, l, mk;
            read_slot(P, e, rel, l, mk);
            if (!slot_ok(mk, lim)) continue;
            if ((int64_t)bi + rel >= P.nelem || l >= lds.size()) return SPG_EINVAL;
            lds[l] = in[bi + rel];
            filled[l] = 1;
        }
        for (int e = 0; e < P.V; ++e) {
            uint32_t rel, l, mk;
            write_slot(P, e, rel, l, mk);
            if (!slot_ok(mk, lim)) continue;
            if ((int64_t)bo + rel >= P.nelem || l >= lds.size() || !filled[l]) return SPG_EINVAL;
            out[bo + rel] = upd(lds[l], out[bo + rel]);
            if (hits) (*hits)[bo + rel]++;
        }
    }
    return SPG_OK;
}

}  // namespace spg
