// Allocation-free residual + squared l2 norm (cdr_hip_residual_norm2d_f64_noalloc).
//
// The per-block partial sums that cdr_hip_residual_norm2d_f64 keeps in its caller's work buffer
// live here in a file-scope device array, `scratch`.  Each GPU holds its own copy of the array at
// its own address: on the first call on a device (with that device current) the address is
// looked up once with hipGetSymbolAddress and kept per device, and only that address is passed to
// the kernels.  Nothing is allocated and nothing needs clearing: the residual kernel writes every
// partial that the final reduction reads.  Concurrent calls on one GPU share the array and race.
#include "../../include/cedar_hip.h"

#include <hip/hip_runtime.h>

#include <mutex>

#include "bmg2_launch.hpp"

namespace {

static __device__ __align__(256) double scratch[CDR_NOALLOC_NORM_PARTIALS];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
double *scratch_dev[kMaxDevices] = {};

// device address of `scratch` on the current device (looked up once per device), or null
double *scratch_ptr(int *status) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        *status = CDR_ERR_NO_DEVICE;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            *status = CDR_ERR_HIP;
            return nullptr;
        }
        scratch_dev[dev] = static_cast<double *>(p);
    }
    *status = CDR_OK;
    return scratch_dev[dev];
}

}  // namespace

extern "C" int cdr_hip_residual_norm2d_f64_noalloc(const double *so, int ns, const double *qf,
                                                   const double *q, double *res, int ii, int jj,
                                                   double *sumsq_dev, void *stream) {
    if (!so || !qf || !q || !res || !sumsq_dev || ii < 3 || jj < 3 ||
        (long long)ii * jj >= (1LL << 31) || (ns != 3 && ns != 5))
        return CDR_ERR_ARG;
    if (cdr::k2::norm_partials(ii, jj) > (std::size_t)CDR_NOALLOC_NORM_PARTIALS)
        return CDR_ERR_TOO_LARGE;
    int st = CDR_OK;
    double *partials = scratch_ptr(&st);
    if (!partials) return st;
    cdr::k2::residual_norm<double>(so, ns, qf, q, res, ii, jj, partials, sumsq_dev,
                                   reinterpret_cast<hipStream_t>(stream), 1);
    return hipGetLastError() == hipSuccess ? CDR_OK : CDR_ERR_HIP;
}
