// capi.cpp -- the C ABI of libspaghetty_amd.so (declared in
// include/spaghetty_amd.h): NWChem drop-in entries, status-returning
// entries, plan inspection and the dispatch table.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>

#include <fstream>
#include <atomic>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "spg_internal.hpp"

namespace spg {

namespace {

thread_local int t_err = SPG_OK;
thread_local std::string t_msg;

std::mutex g_plan_mu;
std::unordered_map<std::string, Plan> g_plans;
// bumped (under g_plan_mu) whenever the dispatch table changes: a plan built
// from the previous table is then not cached
uint64_t g_plan_gen = 0;

std::mutex g_disp_mu;
std::unordered_map<std::string, spg_tile_params> g_disp;
// Plans for outputs written cold (a caller streaming through many distinct
// output arrays), consulted first while g_reused is 0.  The main table's
// plans are tuned for outputs written again while much of them is still in
// the 256 MB last-level cache, as a loop reusing its arrays does; for
// outputs that fit in that cache the two regimes want different plans.
std::unordered_map<std::string, spg_tile_params> g_disp_streamed;
std::atomic<int> g_reused{1};

std::string params_str(const spg_tile_params& p)
{
    char b[160];
    std::snprintf(b, sizeof b, "k%d,%d,%d,%d,%d,b%d,g%d,x%d,m%d,f%d,w%d,%d,%d,t%d", p.kind, p.nx, p.tx,
                  p.ny, p.ty, p.block, p.grid_per_cu, p.xcd_remap, p.max_tile, p.flags, p.walk_x,
                  p.walk_mode, p.walk_chunk, p.tw);
    return b;
}

}  // namespace

void set_error(int code, const char* msg)
{
    t_err = code;
    t_msg = msg ? msg : "";
}

bool dispatch_lookup(const std::string& key, spg_tile_params* out, bool reused)
{
    std::lock_guard<std::mutex> lk(g_disp_mu);
    if (!reused) {
        auto it = g_disp_streamed.find(key);
        if (it != g_disp_streamed.end()) {
            *out = it->second;
            return true;
        }
    }
    auto it = g_disp.find(key);
    if (it == g_disp.end()) return false;
    *out = it->second;
    return true;
}

int resolve_launch(Plan* P);

// Parameters for a problem the dispatch table does not hold.  Below
// kUntunedSmallElems elements the cost model alone picks tiles too large to
// fill the device and ordinary stores; on 38 untuned 4D/6D problems of
// 0.33M-6.25M elements (24^4, 36^4, 50^4, 12^6; 5 or 4 permutations, sort and
// accumulate) nontemporal access with tiles of at most nelem/256 elements
// (256..4096, a power of two) ran 1.10-1.56x faster than the cost model's
// choice (geometric mean per size), none more than 1% slower.  Above it the rule
// did not hold (18^6: 0.84x, 80^4: 1.03x), so larger problems keep the cost
// model's plan.
constexpr int64_t kUntunedSmallElems = 1 << 23;

static bool untuned_params(const Canon& c, spg_tile_params* p)
{
    if (c.nelem > kUntunedSmallElems) return false;
    spg_tile_params_default(p);
    p->flags = SPG_TILE_NT;
    int cap = 256;
    while (cap < kMaxTileElems && (int64_t)cap * 2 * 256 <= c.nelem) cap *= 2;
    p->max_tile = cap;
    return true;
}

namespace {

// Cache key of a call: the raw arguments (before canonicalisation), the
// mode, any forced tile parameters and the device the launch geometry was
// resolved for (-1: not resolved).
struct RawKey {
    int32_t v[4 + 2 * MAXR + sizeof(spg_tile_params) / sizeof(int32_t)];
    bool operator==(const RawKey& o) const { return std::memcmp(v, o.v, sizeof v) == 0; }
};
struct RawKeyHash {
    size_t operator()(const RawKey& k) const
    {
        uint64_t h = 1469598103934665603ULL;  // FNV-1a over the words
        for (int32_t w : k.v) h = (h ^ (uint32_t)w) * 1099511628211ULL;
        return (size_t)h;
    }
};
std::unordered_map<RawKey, Plan, RawKeyHash> g_raw;

std::once_flag g_autoload;

// The dispatch table shipped beside the library (or named by
// SPG_DISPATCH_TABLE) is loaded on first use, so Fortran and C callers get
// the tuned plans too.
void autoload_dispatch()
{
    const char* regime = std::getenv("SPG_OUTPUTS");
    if (regime && std::strcmp(regime, "streamed") == 0) g_reused = 0;
    std::string dir;
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&autoload_dispatch), &info) && info.dli_fname) {
        const std::string lib = info.dli_fname;
        const size_t slash = lib.rfind('/');
        dir = slash == std::string::npos ? std::string(".") : lib.substr(0, slash);
    }
    auto table = [&dir](const char* var, const char* file) {
        const char* env = std::getenv(var);
        if (env && *env) return std::string(env);
        return dir.empty() ? std::string() : dir + "/" + file;
    };
    const std::string path = table("SPG_DISPATCH_TABLE", "dispatch_gfx950.jsonl");
    if (!path.empty() && path != "none" && std::ifstream(path)) spg_dispatch_load(path.c_str());
    const std::string streamed = table("SPG_DISPATCH_TABLE_STREAMED", "dispatch_gfx950_streamed.jsonl");
    if (!streamed.empty() && streamed != "none" && std::ifstream(streamed))
        spg_dispatch_load_streamed(streamed.c_str());
}

}  // namespace

// Plan for a problem: the dispatch table's tuned parameters when it has the
// problem's key, else the cost model's choice.  Plans are cached by their
// raw arguments; with `resolve` the launch geometry is fixed for the current
// device and cached as well.
int get_plan(int rank, const int* dims, const int* perm, double alpha, double beta,
             const spg_tile_params* tp, Plan* out, bool resolve)
{
    std::call_once(g_autoload, autoload_dispatch);
    const int mode = mode_of(beta);
    const bool reused = g_reused.load() != 0;
    RawKey rk;
    std::memset(&rk, 0, sizeof rk);
    bool keyed = rank >= 0 && rank <= MAXR && (rank == 0 || (dims && perm));
    int dev = -1;
    if (resolve && hipGetDevice(&dev) != hipSuccess) return SPG_EHIP;
    if (keyed) {
        rk.v[0] = rank;
        rk.v[1] = mode;
        rk.v[2] = dev;
        rk.v[3] = tp ? 1 : (reused ? 0 : 2);
        for (int q = 0; q < rank; ++q) {
            rk.v[4 + q] = dims[q];
            rk.v[4 + MAXR + q] = perm[q];
        }
        if (tp) std::memcpy(&rk.v[4 + 2 * MAXR], tp, sizeof(spg_tile_params));
        std::lock_guard<std::mutex> lk(g_plan_mu);
        auto it = g_raw.find(rk);
        if (it != g_raw.end()) {
            Plan P = it->second;
            P.alpha = alpha;
            P.beta = beta;
            *out = P;
            return SPG_OK;
        }
    }
    uint64_t gen;
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        gen = g_plan_gen;
    }
    Canon c;
    int rc = canonicalize(rank, dims, perm, &c);
    if (rc != SPG_OK) return rc;
    const std::string key = dispatch_key(c, mode);
    std::string ck = key + ":m" + std::to_string(mode);
    if (tp) ck += ":" + params_str(*tp);
    else if (!reused) ck += ":streamed";
    Plan P;
    bool have = false;
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        auto it = g_plans.find(ck);
        if (it != g_plans.end()) {
            P = it->second;
            have = true;
        }
    }
    if (!have) {
        spg_tile_params tuned;
        const spg_tile_params* use = tp;
        if (!tp && (dispatch_lookup(key, &tuned, reused) || untuned_params(c, &tuned))) use = &tuned;
        rc = build_plan(c, mode, use, &P);
        if (rc != SPG_OK && use == &tuned) rc = build_plan(c, mode, nullptr, &P);
        if (rc != SPG_OK) return rc;
        std::lock_guard<std::mutex> lk(g_plan_mu);
        if (gen == g_plan_gen) g_plans[ck] = P;
    }
    P.mode = mode;
    if (resolve && P.grid <= 0) {
        rc = resolve_launch(&P);
        if (rc != SPG_OK) return rc;
    }
    if (keyed) {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        if (gen == g_plan_gen) {
            if (g_raw.size() > 65536) g_raw.clear();  // bound the cache
            g_raw[rk] = P;
        }
    }
    P.alpha = alpha;
    P.beta = beta;
    *out = P;
    return SPG_OK;
}

namespace {

bool overlaps(const void* a, const void* b, int64_t n)
{
    if (n <= 0) return false;
    const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b, len = (uintptr_t)n * sizeof(double);
    return x < y + len && y < x + len;
}

// Host-pointer context: one stream and two growable device buffers per
// process (the drop-in entries are synchronous, like the reference's).
struct HostCtx {
    int dev = -1;
    hipStream_t stream = nullptr;
    double* din = nullptr;
    double* dout = nullptr;
    size_t cap = 0;
};
std::mutex g_host_mu;
HostCtx g_host;

int host_ctx_reserve(size_t bytes)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return SPG_EHIP;
    if (g_host.dev != dev) {
        // a different device: release the old device's stream and buffers
        // (with that device current) and start over
        if (g_host.dev >= 0 && hipSetDevice(g_host.dev) == hipSuccess) {
            if (g_host.stream) (void)hipStreamSynchronize(g_host.stream);
            if (g_host.din) (void)hipFree(g_host.din);
            if (g_host.dout) (void)hipFree(g_host.dout);
            if (g_host.stream) (void)hipStreamDestroy(g_host.stream);
        }
        if (hipSetDevice(dev) != hipSuccess) return SPG_EHIP;
        g_host = HostCtx();
        g_host.dev = dev;
    }
    if (!g_host.stream && hipStreamCreateWithFlags(&g_host.stream, hipStreamNonBlocking) != hipSuccess)
        return SPG_EHIP;
    if (bytes <= g_host.cap) return SPG_OK;
    if (g_host.din) (void)hipFree(g_host.din);
    if (g_host.dout) (void)hipFree(g_host.dout);
    g_host.din = g_host.dout = nullptr;
    g_host.cap = 0;
    size_t want = bytes < ((size_t)1 << 20) ? ((size_t)1 << 20) : bytes;
    if (hipMalloc(&g_host.din, want) != hipSuccess) {
        (void)hipGetLastError();
        return SPG_ENOMEM;
    }
    if (hipMalloc(&g_host.dout, want) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(g_host.din);
        g_host.din = nullptr;
        return SPG_ENOMEM;
    }
    g_host.cap = want;
    return SPG_OK;
}

int fail(int rc, const char* msg)
{
    set_error(rc, msg);
    return rc;
}

}  // namespace

}  // namespace spg

using namespace spg;

extern "C" {

int spg_version(void) { return 10000; }

int spg_last_error(void) { return t_err; }

const char* spg_last_error_message(void) { return t_msg.c_str(); }

const char* spg_strerror(int code)
{
    switch (code) {
    case SPG_OK: return "success";
    case SPG_EINVAL_PERM: return "invalid permutation";
    case SPG_EINVAL_DIM: return "invalid dimensions";
    case SPG_EALIAS: return "input and output overlap";
    case SPG_ENOMEM: return "device allocation failed";
    case SPG_EHIP: return "HIP runtime error";
    case SPG_EINVAL: return "invalid argument";
    case SPG_EBUDGET: return "request exceeds the HBM budget";
    case SPG_ESTATE: return "invalid state for this call";
    default: return "unknown error";
    }
}

void spg_tile_params_default(spg_tile_params* p)
{
    if (!p) return;
    p->kind = -1;
    p->nx = p->tx = p->ny = p->ty = 0;
    p->block = 0;
    p->grid_per_cu = 0;
    p->xcd_remap = -1;
    p->max_tile = 0;
    p->flags = -1;
    p->walk_x = 0;
    p->walk_mode = 0;
    p->walk_chunk = 0;
    p->tw = 0;
}

}  // extern "C"

namespace spg {

int plan_for_alignment(bool aligned16, int rank, const int* dims, const int* perm, double factor, double beta,
                       const spg_tile_params* params, Plan* out, bool resolve)
{
    Plan P;
    int rc = get_plan(rank, dims, perm, factor, beta, params, &P, resolve);
    if (rc != SPG_OK) return rc;
    if (P.nelem != 0 && !aligned16 && P.kind == KIND_TILED && P.V > kMaxTileElems && !P.pair8) {
        // a tile above kMaxTileElems needs 16-byte pairs, which plans over
        // even extents take 16-byte aligned: on arrays that are only 8-byte
        // aligned, plan again within the single-element limit (pair8 plans
        // take 8-byte aligned pairs anyway)
        spg_tile_params small;
        spg_tile_params_default(&small);
        small.max_tile = kMaxTileElems;
        small.flags = SPG_TILE_SCALAR;
        rc = get_plan(rank, dims, perm, factor, beta, &small, &P, resolve);
        if (rc != SPG_OK) return rc;
    }
    *out = P;
    return SPG_OK;
}

int plan_for_call(const double* unsorted, double* sorted, int rank, const int* dims, const int* perm,
                  double factor, double beta, const spg_tile_params* params, Plan* out, bool resolve)
{
    const bool aligned16 =
        ((reinterpret_cast<uintptr_t>(unsorted) | reinterpret_cast<uintptr_t>(sorted)) & 15u) == 0;
    Plan P;
    const int rc = plan_for_alignment(aligned16, rank, dims, perm, factor, beta, params, &P, resolve);
    if (rc != SPG_OK) return rc;
    if (P.nelem != 0) {
        if (!unsorted || !sorted) return SPG_EINVAL;
        if (overlaps(unsorted, sorted, P.nelem)) return SPG_EALIAS;
    }
    *out = P;
    return SPG_OK;
}

}  // namespace spg

extern "C" {

int spg_sort_with(const double* unsorted, double* sorted, int rank, const int* dims,
                  const int* perm, double factor, double beta, const spg_tile_params* params,
                  void* stream)
{
    set_error(SPG_OK, "");
    Plan P;
    int rc = plan_for_call(unsorted, sorted, rank, dims, perm, factor, beta, params, &P, true);
    if (rc == SPG_EALIAS) return fail(rc, "unsorted and sorted overlap");
    if (rc == SPG_EINVAL && (!unsorted || !sorted)) return fail(rc, "null array");
    if (rc != SPG_OK) return fail(rc, spg_strerror(rc));
    if (P.nelem == 0) return SPG_OK;
    rc = launch_plan(P, unsorted, sorted, (hipStream_t)stream);
    if (rc != SPG_OK) return fail(rc, t_msg.empty() ? spg_strerror(rc) : t_msg.c_str());
    return SPG_OK;
}

int spg_sort(const double* unsorted, double* sorted, int rank, const int* dims, const int* perm,
             double factor, double beta, void* stream)
{
    return spg_sort_with(unsorted, sorted, rank, dims, perm, factor, beta, nullptr, stream);
}

int spg_sort_host(const double* unsorted, double* sorted, int rank, const int* dims,
                  const int* perm, double factor, double beta)
{
    set_error(SPG_OK, "");
    Plan P;
    int rc = get_plan(rank, dims, perm, factor, beta, nullptr, &P, false);
    if (rc != SPG_OK) return fail(rc, spg_strerror(rc));
    if (P.nelem == 0) return SPG_OK;
    if (!unsorted || !sorted) return fail(SPG_EINVAL, "null array");
    if (overlaps(unsorted, sorted, P.nelem)) return fail(SPG_EALIAS, "unsorted and sorted overlap");
    const size_t bytes = (size_t)P.nelem * sizeof(double);
    std::lock_guard<std::mutex> lk(g_host_mu);
    rc = host_ctx_reserve(bytes);
    if (rc != SPG_OK) return fail(rc, spg_strerror(rc));
    hipStream_t s = g_host.stream;
    if (hipMemcpyAsync(g_host.din, unsorted, bytes, hipMemcpyHostToDevice, s) != hipSuccess)
        return fail(SPG_EHIP, "host-to-device copy failed");
    if (P.mode != MODE_SORT &&
        hipMemcpyAsync(g_host.dout, sorted, bytes, hipMemcpyHostToDevice, s) != hipSuccess)
        return fail(SPG_EHIP, "host-to-device copy failed");
    rc = launch_plan(P, g_host.din, g_host.dout, s);
    if (rc != SPG_OK) return fail(rc, t_msg.c_str());
    if (hipMemcpyAsync(sorted, g_host.dout, bytes, hipMemcpyDeviceToHost, s) != hipSuccess)
        return fail(SPG_EHIP, "device-to-host copy failed");
    if (hipStreamSynchronize(s) != hipSuccess) return fail(SPG_EHIP, "stream synchronize failed");
    return SPG_OK;
}

// ---- dispatch table --------------------------------------------------------

static bool json_int(const std::string& line, const char* field, int32_t* v)
{
    const std::string pat = std::string("\"") + field + "\"";
    size_t p = line.find(pat);
    if (p == std::string::npos) return false;
    p = line.find(':', p + pat.size());
    if (p == std::string::npos) return false;
    *v = (int32_t)std::strtol(line.c_str() + p + 1, nullptr, 10);
    return true;
}

}  // extern "C"

namespace spg {
namespace {

// Reads a table file (one JSON object per line) into `into`; returns the
// records read or an error status.
int load_table(const char* path, std::unordered_map<std::string, spg_tile_params>& into)
{
    set_error(SPG_OK, "");
    if (!path) return fail(SPG_EINVAL, "null path");
    std::ifstream f(path);
    if (!f) return fail(SPG_EINVAL, "cannot open dispatch table");
    std::string line;
    int n = 0;
    std::unordered_map<std::string, spg_tile_params> add;
    while (std::getline(f, line)) {
        size_t k = line.find("\"key\"");
        if (k == std::string::npos) continue;
        size_t q1 = line.find('"', line.find(':', k) + 1);
        size_t q2 = q1 == std::string::npos ? q1 : line.find('"', q1 + 1);
        if (q2 == std::string::npos) continue;
        const std::string key = line.substr(q1 + 1, q2 - q1 - 1);
        spg_tile_params p;
        spg_tile_params_default(&p);
        p.flags = 0;
        json_int(line, "kind", &p.kind);
        json_int(line, "nx", &p.nx);
        json_int(line, "tx", &p.tx);
        json_int(line, "ny", &p.ny);
        json_int(line, "ty", &p.ty);
        json_int(line, "block", &p.block);
        json_int(line, "grid_per_cu", &p.grid_per_cu);
        json_int(line, "xcd_remap", &p.xcd_remap);
        json_int(line, "max_tile", &p.max_tile);
        json_int(line, "flags", &p.flags);
        json_int(line, "walk_x", &p.walk_x);
        json_int(line, "walk_mode", &p.walk_mode);
        json_int(line, "walk_chunk", &p.walk_chunk);
        // records tuned before tiles of several blocks existed keep one block
        p.tw = 1;
        json_int(line, "tw", &p.tw);
        add[key] = p;
        ++n;
    }
    {
        std::lock_guard<std::mutex> lk(g_disp_mu);
        for (auto& kv : add) into[kv.first] = kv.second;
    }
    std::lock_guard<std::mutex> lk(g_plan_mu);
    g_plans.clear();
    g_raw.clear();
    ++g_plan_gen;
    return n;
}

}  // namespace
}  // namespace spg

extern "C" {

int spg_dispatch_load(const char* path) { return load_table(path, g_disp); }

int spg_dispatch_load_streamed(const char* path) { return load_table(path, g_disp_streamed); }

int spg_set_output_reuse(int reused)
{
    const int prev = g_reused.exchange(reused ? 1 : 0);
    return prev;
}

int spg_dispatch_set(const char* key, const spg_tile_params* params)
{
    if (!key || !params) return fail(SPG_EINVAL, "null argument");
    {
        std::lock_guard<std::mutex> lk(g_disp_mu);
        g_disp[key] = *params;
    }
    std::lock_guard<std::mutex> lk(g_plan_mu);
    g_plans.clear();
    g_raw.clear();
    ++g_plan_gen;
    return SPG_OK;
}

int spg_dispatch_clear(void)
{
    {
        std::lock_guard<std::mutex> lk(g_disp_mu);
        g_disp.clear();
        g_disp_streamed.clear();
    }
    std::lock_guard<std::mutex> lk(g_plan_mu);
    g_plans.clear();
    g_raw.clear();
    ++g_plan_gen;
    return SPG_OK;
}

int spg_dispatch_key(int rank, const int* dims, const int* perm, double beta, char* buf,
                     int buflen)
{
    Canon c;
    int rc = canonicalize(rank, dims, perm, &c);
    if (rc != SPG_OK) return fail(rc, spg_strerror(rc));
    const std::string k = dispatch_key(c, mode_of(beta));
    if (!buf || buflen <= (int)k.size()) return fail(SPG_EINVAL, "buffer too small");
    std::memcpy(buf, k.c_str(), k.size() + 1);
    return SPG_OK;
}

// ---- plan inspection ---------------------------------------------------------

int spg_plan_describe(int rank, const int* dims, const int* perm, double beta,
                      const spg_tile_params* params, char* buf, int buflen)
{
    Plan P;
    int rc = get_plan(rank, dims, perm, 1.0, beta, params, &P, false);
    if (rc != SPG_OK) return fail(rc, spg_strerror(rc));
    const std::string s = describe(P);
    if (!buf || buflen <= (int)s.size()) return fail(SPG_EINVAL, "buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return SPG_OK;
}

int spg_plan_check(const double* unsorted, double* sorted, int rank, const int* dims,
                   const int* perm, double factor, double beta, const spg_tile_params* params)
{
    Plan P;
    int rc = get_plan(rank, dims, perm, factor, beta, params, &P, false);
    if (rc != SPG_OK) return fail(rc, spg_strerror(rc));
    std::vector<int> hits;
    rc = emulate(P, unsorted, sorted, &hits);
    if (rc != SPG_OK) return fail(rc, "plan walk left the arrays or read an unfilled LDS slot");
    for (size_t i = 0; i < hits.size(); ++i)
        if (hits[i] != 1) return fail(SPG_EINVAL, "plan does not produce every element exactly once");
    return SPG_OK;
}

int spg_selftest_fastdiv(int64_t trials)
{
    // every divisor shape the plans use, against the hardware division
    uint64_t x = 0x9E3779B97F4A7C15ULL;
    auto next = [&x]() {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return x;
    };
    const uint32_t edges_d[] = {1, 2, 3, 5, 7, 10, 16, 20, 30, 40, 100, 255, 256, 257, 1000, 10000,
                                65535, 65536, 1u << 20, (1u << 30) + 1, 0x7fffffffu};
    const uint32_t edges_n[] = {0, 1, 2, 3, 99, 100, 101, 65535, 65536, 0x7ffffffeu, 0x7fffffffu};
    for (uint32_t d : edges_d) {
        const FastDiv f = make_fastdiv(d);
        for (uint32_t n : edges_n)
            if (fdiv(n, f) != n / d) return fail(SPG_EINVAL, "fast division mismatch");
        for (uint32_t k = 0; k < 4096; ++k) {
            const uint64_t n = (uint64_t)k * d;
            for (int64_t dn = -1; dn <= 1; ++dn) {
                const int64_t v = (int64_t)n + dn;
                if (v < 0 || v > 0x7fffffff) continue;
                if (fdiv((uint32_t)v, f) != (uint32_t)v / d) return fail(SPG_EINVAL, "fast division mismatch");
            }
        }
    }
    for (int64_t t = 0; t < trials; ++t) {
        const uint32_t d = (uint32_t)(next() % 0x7fffffffu) >> (next() % 31) | 1u;
        const uint32_t dd = (t & 1) ? d : d + 1;
        const uint32_t n = (uint32_t)(next() & 0x7fffffffu);
        if (fdiv(n, make_fastdiv(dd)) != n / dd) return fail(SPG_EINVAL, "fast division mismatch");
    }
    return SPG_OK;
}

// ---- NWChem drop-in entries -------------------------------------------------

static void dropin(const char* name, int rank, const double* u, double* s, const int* const* d,
                   const int* const* p, const double* factor, double beta)
{
    int dims[MAXR], perm[MAXR];
    for (int q = 0; q < rank; ++q) {
        dims[q] = *d[q];
        perm[q] = *p[q];
    }
    const int rc = spg_sort_host(u, s, rank, dims, perm, *factor, beta);
    if (rc == SPG_EINVAL_PERM) {
        // trunk/makelib.py:90-91: unknown permutation -> message, no-op
        std::printf(" something is wrong...\n");
        std::fflush(stdout);
    } else if (rc != SPG_OK) {
        std::fprintf(stderr, "%s: %s (%s)\n", name, spg_strerror(rc), spg_last_error_message());
    }
}

#define SPG_SORT4_BODY(NAME, BETA)                                                        \
    {                                                                                     \
        const int* d[4] = {a, b, c, d_};                                                  \
        const int* p[4] = {i, j, k, l};                                                   \
        dropin(NAME, 4, unsorted, sorted, d, p, factor, BETA);                            \
    }
#define SPG_SORT6_BODY(NAME, BETA)                                                        \
    {                                                                                     \
        const int* d[6] = {a, b, c, d_, e, f};                                            \
        const int* p[6] = {i, j, k, l, m, n};                                             \
        dropin(NAME, 6, unsorted, sorted, d, p, factor, BETA);                            \
    }
#define SPG_SORT2_BODY(NAME, BETA)                                                        \
    {                                                                                     \
        const int* d[2] = {a, b};                                                         \
        const int* p[2] = {i, j};                                                         \
        dropin(NAME, 2, unsorted, sorted, d, p, factor, BETA);                            \
    }
#define SPG_SORT8_BODY(NAME, BETA)                                                        \
    {                                                                                     \
        const int* d[8] = {a, b, c, d_, e, f, g, h};                                      \
        const int* p[8] = {i, j, k, l, m, n, o, p_};                                      \
        dropin(NAME, 8, unsorted, sorted, d, p, factor, BETA);                            \
    }
#define SPG_ARGS2                                                                         \
    const double *unsorted, double *sorted, const int *a, const int *b, const int *i,     \
        const int *j, const double *factor
#define SPG_ARGS8                                                                         \
    const double *unsorted, double *sorted, const int *a, const int *b, const int *c,     \
        const int *d_, const int *e, const int *f, const int *g, const int *h,            \
        const int *i, const int *j, const int *k, const int *l, const int *m,             \
        const int *n, const int *o, const int *p_, const double *factor
#define SPG_ARGS4                                                                         \
    const double *unsorted, double *sorted, const int *a, const int *b, const int *c,     \
        const int *d_, const int *i, const int *j, const int *k, const int *l,            \
        const double *factor
#define SPG_ARGS6                                                                         \
    const double *unsorted, double *sorted, const int *a, const int *b, const int *c,     \
        const int *d_, const int *e, const int *f, const int *i, const int *j,            \
        const int *k, const int *l, const int *m, const int *n, const double *factor

void tce_sort_4_(SPG_ARGS4) SPG_SORT4_BODY("tce_sort_4", 0.0)
void tce_sort_4(SPG_ARGS4) SPG_SORT4_BODY("tce_sort_4", 0.0)
void tce_sort_4__(SPG_ARGS4) SPG_SORT4_BODY("tce_sort_4", 0.0)
void tce_sortacc_4_(SPG_ARGS4) SPG_SORT4_BODY("tce_sortacc_4", 1.0)
void tce_sortacc_4(SPG_ARGS4) SPG_SORT4_BODY("tce_sortacc_4", 1.0)
void tce_sortacc_4__(SPG_ARGS4) SPG_SORT4_BODY("tce_sortacc_4", 1.0)
void tce_sort_6_(SPG_ARGS6) SPG_SORT6_BODY("tce_sort_6", 0.0)
void tce_sort_6(SPG_ARGS6) SPG_SORT6_BODY("tce_sort_6", 0.0)
void tce_sort_6__(SPG_ARGS6) SPG_SORT6_BODY("tce_sort_6", 0.0)
void tce_sortacc_6_(SPG_ARGS6) SPG_SORT6_BODY("tce_sortacc_6", 1.0)
void tce_sortacc_6(SPG_ARGS6) SPG_SORT6_BODY("tce_sortacc_6", 1.0)
void tce_sortacc_6__(SPG_ARGS6) SPG_SORT6_BODY("tce_sortacc_6", 1.0)
void tce_sort_2_(SPG_ARGS2) SPG_SORT2_BODY("tce_sort_2", 0.0)
void tce_sort_2(SPG_ARGS2) SPG_SORT2_BODY("tce_sort_2", 0.0)
void tce_sort_2__(SPG_ARGS2) SPG_SORT2_BODY("tce_sort_2", 0.0)
void tce_sortacc_2_(SPG_ARGS2) SPG_SORT2_BODY("tce_sortacc_2", 1.0)
void tce_sortacc_2(SPG_ARGS2) SPG_SORT2_BODY("tce_sortacc_2", 1.0)
void tce_sortacc_2__(SPG_ARGS2) SPG_SORT2_BODY("tce_sortacc_2", 1.0)
void tce_sort_8_(SPG_ARGS8) SPG_SORT8_BODY("tce_sort_8", 0.0)
void tce_sort_8(SPG_ARGS8) SPG_SORT8_BODY("tce_sort_8", 0.0)
void tce_sort_8__(SPG_ARGS8) SPG_SORT8_BODY("tce_sort_8", 0.0)
void tce_sortacc_8_(SPG_ARGS8) SPG_SORT8_BODY("tce_sortacc_8", 1.0)
void tce_sortacc_8(SPG_ARGS8) SPG_SORT8_BODY("tce_sortacc_8", 1.0)
void tce_sortacc_8__(SPG_ARGS8) SPG_SORT8_BODY("tce_sortacc_8", 1.0)

}  // extern "C"
