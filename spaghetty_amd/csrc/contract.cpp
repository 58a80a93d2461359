// contract.cpp -- the contraction NWChem's TCE wraps around tce_sort_4
// ("TTDT": transpose, transpose, DGEMM, transpose), on the device.
//
// branches/old/python/macaroni/ccsd_t2_8_skeleton.F:162-183 computes
//     i0(p3,p4,h1,h2) = Sum(p5,p6) v(p3,p4,p5,p6) * t(p5,p6,h1,h2)
// as: sort A into a (K x M) column-major matrix (tce_sort_4 ... 4,3,2,1),
// sort B into (K x N) (... 2,1,4,3), C' = A'^T B' with dgemm('T','N'), and
// sort C' (M x N, column-major) into the result (... 2,1,4,3).  Here the
// three sorts run on the library's kernels and the product on rocBLAS's
// DGEMM, all ordered on one stream: no host round trip between the steps.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <utility>

#include "spg_internal.hpp"

namespace spg {
namespace {

struct DevState {
    rocblas_handle blas = nullptr;
    double* work = nullptr;
    size_t cap = 0;            // doubles
    hipEvent_t done = nullptr;  // recorded after the last step that reads `work`
};

std::mutex g_ttdt_mu;
std::map<int, DevState> g_ttdt;

int tfail(int rc, const char* msg)
{
    set_error(rc, msg);
    return rc;
}

int64_t prod(int r, const int* d)
{
    int64_t p = 1;
    for (int q = 0; q < r; ++q) p *= d[q];
    return p;
}

}  // namespace
}  // namespace spg

using namespace spg;

extern "C" {

int spg_ttdt(const double* a, int ra, const int* adims, const int* aperm, const double* b, int rb,
             const int* bdims, const int* bperm, double* c, int rc, const int* cdims,
             const int* cperm, int m, int n, int k, double alpha, double beta, double* work,
             void* stream)
{
    set_error(SPG_OK, "");
    if (m < 0 || n < 0 || k < 0) return tfail(SPG_EINVAL_DIM, "negative matrix size");
    if (ra < 0 || ra > MAXR || rb < 0 || rb > MAXR || rc < 0 || rc > MAXR)
        return tfail(SPG_EINVAL_DIM, "rank out of range");
    for (int q = 0; q < ra; ++q)
        if (adims[q] < 0) return tfail(SPG_EINVAL_DIM, "negative dim");
    for (int q = 0; q < rb; ++q)
        if (bdims[q] < 0) return tfail(SPG_EINVAL_DIM, "negative dim");
    for (int q = 0; q < rc; ++q)
        if (cdims[q] < 0) return tfail(SPG_EINVAL_DIM, "negative dim");
    const int64_t M = m, N = n, K = k;
    if (prod(ra, adims) != M * K || prod(rb, bdims) != K * N || prod(rc, cdims) != M * N)
        return tfail(SPG_EINVAL_DIM, "dims do not multiply to the matrix sizes");
    if (M * K > 0x7fffffffLL || K * N > 0x7fffffffLL || M * N > 0x7fffffffLL)
        return tfail(SPG_EINVAL_DIM, "more than 2^31-1 elements");
    if (M * N == 0) return SPG_OK;
    if (!a || !b || !c) return tfail(SPG_EINVAL, "null array");
    hipStream_t st = (hipStream_t)stream;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return tfail(SPG_EHIP, "no device");
    const size_t need = (size_t)(M * K + K * N + M * N);
    std::lock_guard<std::mutex> lk(g_ttdt_mu);
    DevState& S = g_ttdt[dev];
    if (!S.blas && rocblas_create_handle(&S.blas) != rocblas_status_success) {
        S.blas = nullptr;
        return tfail(SPG_EHIP, "rocblas_create_handle failed");
    }
    double* w = work;
    if (!w) {
        if (S.cap < need) {
            // the buffer may still be read by earlier calls on this device
            if (S.work) {
                if (hipDeviceSynchronize() != hipSuccess) return tfail(SPG_EHIP, "synchronize failed");
                (void)hipFree(S.work);
            }
            S.work = nullptr;
            S.cap = 0;
            if (hipMalloc(&S.work, need * sizeof(double)) != hipSuccess) {
                (void)hipGetLastError();
                return tfail(SPG_ENOMEM, "work buffer allocation failed");
            }
            S.cap = need;
        }
        w = S.work;
        // calls on other streams share this buffer: wait until the previous
        // call's last step (the final sort, which reads C') has run
        if (!S.done && hipEventCreateWithFlags(&S.done, hipEventDisableTiming) != hipSuccess) {
            S.done = nullptr;
            return tfail(SPG_EHIP, "event creation failed");
        }
        if (hipStreamWaitEvent(st, S.done, 0) != hipSuccess) return tfail(SPG_EHIP, "stream wait failed");
    }
    double* as = w;
    double* bs = as + M * K;
    double* cs = bs + K * N;
    if (K > 0) {
        int r = spg_sort(a, as, ra, adims, aperm, 1.0, 0.0, stream);
        if (r != SPG_OK) return r;
        r = spg_sort(b, bs, rb, bdims, bperm, 1.0, 0.0, stream);
        if (r != SPG_OK) return r;
    }
    if (rocblas_set_stream(S.blas, st) != rocblas_status_success)
        return tfail(SPG_EHIP, "rocblas_set_stream failed");
    const double zero = 0.0;
    // C' (M x N, column-major) = alpha * A'^T (M x K) * B' (K x N)
    if (rocblas_dgemm(S.blas, rocblas_operation_transpose, rocblas_operation_none, m, n, k, &alpha, as,
                      k > 0 ? k : 1, bs, k > 0 ? k : 1, &zero, cs, m) != rocblas_status_success)
        return tfail(SPG_EHIP, "rocblas_dgemm failed");
    const int r = spg_sort(cs, c, rc, cdims, cperm, 1.0, beta, stream);
    if (r != SPG_OK) return r;
    if (w == S.work && hipEventRecord(S.done, st) != hipSuccess) return tfail(SPG_EHIP, "event record failed");
    return SPG_OK;
}

}  // extern "C"
