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
bool dispatch_lookup(const std::string& key, spg_tile_params* out);

// Human/JSON description of a plan.
std::string describe(const Plan& P);

// Executes a plan on the host with exactly the kernels' index math; used
// only to check plans (coverage and values) without a GPU.
int emulate(const Plan& P, const double* in, double* out, std::vector<int>* hits);

// Thread-local last error.
void set_error(int code, const char* msg);

}  // namespace spg
