// y = A q for the 2D and 3D stencil operators: the kernel behind the handle API's cedar_matvec
// (the reference runs src/2d/ftn/mpi/BMG2_SymStd_UTILS_matvec.f90 and
// src/3d/ftn/mpi/BMG3_SymStd_UTILS_matvec.f90 after a halo exchange; on one GPU the ghosts are
// the boundary).  One thread per interior point, lanes along i; the terms are summed left to
// right in the order the Fortran writes them and the library is built without contraction, so
// every value rounds exactly as the reference's does.  Ghosts of y are not written.
#include "../../include/cedar_hip.h"

#include <hip/hip_runtime.h>

#include <cstddef>

namespace {

inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

// 2D planes {c, w, s, sw, nw}
__global__ __launch_bounds__(256) void matvec2_k(const double *__restrict__ so, int nstncl,
                                                 const double *__restrict__ q,
                                                 double *__restrict__ y, int ii, int jj) {
    const int i = 1 + blockIdx.x * 64 + (threadIdx.x % 64);
    const int j = 1 + blockIdx.y * 4 + (threadIdx.x / 64);
    if (i > ii - 2 || j > jj - 2) return;
    const std::size_t P = (std::size_t)ii * jj;
    const std::size_t p = (std::size_t)j * ii + i;
    const double *c = so, *w = so + P, *s = so + 2 * P;
    double v = c[p] * q[p] - w[p] * q[p - 1] - w[p + 1] * q[p + 1] - s[p] * q[p - ii] -
               s[p + ii] * q[p + ii];
    if (nstncl == 5) {
        const double *sw = so + 3 * P, *nw = so + 4 * P;
        v = v - sw[p] * q[p - ii - 1] - nw[p + 1] * q[p - ii + 1] - nw[p + ii] * q[p + ii - 1] -
            sw[p + ii + 1] * q[p + ii + 1];
    }
    y[p] = v;
}

// 3D planes {p, pw, ps, b, psw, pnw, bw, bnw, bn, bne, be, bse, bs, bsw}
__global__ __launch_bounds__(256) void matvec3_k(const double *__restrict__ so, int nstncl,
                                                 const double *__restrict__ q,
                                                 double *__restrict__ y, int ii, int jj, int kk) {
    const int i = 1 + blockIdx.x * 64 + (threadIdx.x % 64);
    const int j = 1 + blockIdx.y * 4 + (threadIdx.x / 64);
    const int k = 1 + blockIdx.z;
    if (i > ii - 2 || j > jj - 2 || k > kk - 2) return;
    const std::size_t P = (std::size_t)ii * jj * kk, sj = ii, sk = (std::size_t)ii * jj;
    const std::size_t p = k * sk + j * sj + i;
    auto A = [&](int d, std::size_t at) { return so[d * P + at]; };
    enum { kp, kpw, kps, kb, kpsw, kpnw, kbw, kbnw, kbn, kbne, kbe, kbse, kbs, kbsw };
    const std::size_t n = p + sj, e = p + 1, ne = p + sj + 1, up = p + sk;
    double v;
    if (nstncl == 14) {
        v = A(kp, p) * q[p] - A(kpw, p) * q[p - 1] - A(kpnw, n) * q[p + sj - 1] -
            A(kps, n) * q[p + sj] - A(kpsw, ne) * q[p + sj + 1] - A(kpw, e) * q[p + 1] -
            A(kpnw, e) * q[p - sj + 1] - A(kps, p) * q[p - sj] - A(kpsw, p) * q[p - sj - 1] -
            A(kb, p) * q[p - sk] - A(kbw, p) * q[p - sk - 1] - A(kbnw, n) * q[p - sk + sj - 1] -
            A(kbn, n) * q[p - sk + sj] - A(kbne, ne) * q[p - sk + sj + 1] -
            A(kbe, e) * q[p - sk + 1] - A(kbse, e) * q[p - sk - sj + 1] -
            A(kbs, p) * q[p - sk - sj] - A(kbsw, p) * q[p - sk - sj - 1] -
            A(kb, up) * q[p + sk] - A(kbe, up) * q[p + sk - 1] -
            A(kbse, up + sj) * q[p + sk + sj - 1] - A(kbs, up + sj) * q[p + sk + sj] -
            A(kbsw, up + sj + 1) * q[p + sk + sj + 1] - A(kbw, up + 1) * q[p + sk + 1] -
            A(kbnw, up + 1) * q[p + sk - sj + 1] - A(kbn, up) * q[p + sk - sj] -
            A(kbne, up) * q[p + sk - sj - 1];
    } else {
        v = A(kp, p) * q[p] - A(kpw, p) * q[p - 1] - A(kps, n) * q[p + sj] -
            A(kpw, e) * q[p + 1] - A(kps, p) * q[p - sj] - A(kb, p) * q[p - sk] -
            A(kb, up) * q[p + sk];
    }
    y[p] = v;
}

bool too_big(long long n) { return n >= (1LL << 31); }

}  // namespace

extern "C" {

int cdr_hip_matvec2d_f64(const double *so, int nstncl, const double *q, double *y, int ii, int jj,
                         void *stream) {
    if (!so || !q || !y || (nstncl != 3 && nstncl != 5) || ii < 3 || jj < 3 ||
        too_big((long long)ii * jj))
        return CDR_ERR_ARG;
    dim3 grid((unsigned)((ii - 2 + 63) / 64), (unsigned)((jj - 2 + 3) / 4));
    matvec2_k<<<grid, 256, 0, S(stream)>>>(so, nstncl, q, y, ii, jj);
    return hipGetLastError() == hipSuccess ? CDR_OK : CDR_ERR_HIP;
}

int cdr_hip_matvec3d_f64(const double *so, int nstncl, const double *q, double *y, int ii, int jj,
                         int kk, void *stream) {
    if (!so || !q || !y || (nstncl != 4 && nstncl != 14) || ii < 3 || jj < 3 || kk < 3 ||
        kk - 2 > 65535 || too_big((long long)ii * jj * kk))
        return CDR_ERR_ARG;
    dim3 grid((unsigned)((ii - 2 + 63) / 64), (unsigned)((jj - 2 + 3) / 4), (unsigned)(kk - 2));
    matvec3_k<<<grid, 256, 0, S(stream)>>>(so, nstncl, q, y, ii, jj, kk);
    return hipGetLastError() == hipSuccess ? CDR_OK : CDR_ERR_HIP;
}

}  // extern "C"
