// C ABI of the 3D kernels: argument checks, launch, status.  See include/cedar_hip.h.
#include "../../include/cedar_hip.h"

#include <hip/hip_runtime.h>

#include "bmg2_launch.hpp"
#include "bmg3_launch.hpp"

namespace {

using namespace cdr;

inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }
inline int launched() { return hipGetLastError() == hipSuccess ? CDR_OK : CDR_ERR_HIP; }
inline bool ext3_ok(int ii, int jj, int kk) {
    return ii >= 3 && jj >= 3 && kk >= 3 && (long long)ii * jj * kk < (1LL << 31);
}
inline bool sten3_ok(int ns) { return ns == 4 || ns == 14; }
inline bool dir_ok(int ud) { return ud == CDR_DOWN || ud == CDR_UP; }
inline bool coarse1(int f, int c) { return c - 2 == (f - 3) / 2 + 1; }
inline bool coarse3_ok(int iif, int jjf, int kkf, int iic, int jjc, int kkc) {
    return ext3_ok(iif, jjf, kkf) && ext3_ok(iic, jjc, kkc) && coarse1(iif, iic) &&
           coarse1(jjf, jjc) && coarse1(kkf, kkc);
}

template <typename T>
int residual3_t(const T *so, int ns, const T *qf, const T *q, T *res, int ii, int jj, int kk,
                void *st) {
    if (!so || !qf || !q || !res || !ext3_ok(ii, jj, kk) || !sten3_ok(ns)) return CDR_ERR_ARG;
    k3::residual<T>(so, ns, qf, q, res, ii, jj, kk, S(st));
    return launched();
}

template <typename T>
int gs3_t(const T *so, int ns, const T *qf, T *q, const T *sor, int ii, int jj, int kk, int ud,
          void *st) {
    if (!so || !qf || !q || !sor || !ext3_ok(ii, jj, kk) || !sten3_ok(ns) || !dir_ok(ud))
        return CDR_ERR_ARG;
    k3::relax_gs<T>(so, ns, qf, q, sor, ii, jj, kk, ud, S(st));
    return launched();
}

template <typename T>
int restrict3_t(const T *ci, const T *q, T *qc, int iif, int jjf, int kkf, int iic, int jjc,
                int kkc, void *st) {
    if (!ci || !q || !qc || !coarse3_ok(iif, jjf, kkf, iic, jjc, kkc)) return CDR_ERR_ARG;
    k3::restrict_<T>(ci, q, qc, iif, jjf, kkf, iic, jjc, kkc, S(st));
    return launched();
}

template <typename T>
int interp3_t(T *q, const T *qc, const T *res, const T *so, int ns, const T *ci, int iic, int jjc,
              int kkc, int iif, int jjf, int kkf, void *st) {
    if (!q || !qc || !res || !so || !ci || !sten3_ok(ns) ||
        !coarse3_ok(iif, jjf, kkf, iic, jjc, kkc))
        return CDR_ERR_ARG;
    k3::interp_add<T>(q, qc, res, so, ns, ci, iic, jjc, kkc, iif, jjf, kkf, S(st));
    return launched();
}

template <typename T>
int solve_cg3_t(T *q, const T *qf, const double *abd, int ldab, double *bbd, int ii, int jj,
                int kk, void *st) {
    if (!q || !qf || !abd || !bbd || !ext3_ok(ii, jj, kk) || ldab < k3::cg_bandwidth(ii, jj) + 1)
        return CDR_ERR_ARG;
    k3::solve_cg<T>(q, qf, abd, ldab, bbd, ii, jj, kk, S(st));
    return launched();
}

// a periodic boundary id (BMG_get_bc of a nonzero mask, or its indefinite negative)
inline bool jpn3_ok(int jpn) { return jpn != 0 && k3::per_mask3(jpn) != 0; }
// the periodic coarsening pairs n fine points with n/2 coarse ones: even periodic extents
inline bool even_per3(int jpn, int iif, int jjf, int kkf) {
    const int m = k3::per_mask3(jpn);
    return (!(m & 1) || (iif - 2) % 2 == 0) && (!(m & 2) || (jjf - 2) % 2 == 0) &&
           (!(m & 4) || (kkf - 2) % 2 == 0);
}
constexpr int kDenseMax3 = 2048;  // coarsest unknowns of the periodic dense factor

template <typename T>
int gs3_per_t(const T *so, int ns, const T *qf, T *q, const T *sor, int ii, int jj, int kk, int ud,
              int jpn, void *st) {
    if (!so || !qf || !q || !sor || !ext3_ok(ii, jj, kk) || !sten3_ok(ns) || !dir_ok(ud) ||
        !jpn3_ok(jpn))
        return CDR_ERR_ARG;
    k3::relax_gs_per<T>(so, ns, qf, q, sor, ii, jj, kk, ud, jpn, S(st));
    return launched();
}

template <typename T>
int restrict3_per_t(const T *ci, T *q, T *qc, int iif, int jjf, int kkf, int iic, int jjc, int kkc,
                    int jpn, void *st) {
    if (!ci || !q || !qc || !coarse3_ok(iif, jjf, kkf, iic, jjc, kkc) || !jpn3_ok(jpn))
        return CDR_ERR_ARG;
    k3::restrict_per<T>(ci, q, qc, iif, jjf, kkf, iic, jjc, kkc, jpn, S(st));
    return launched();
}

template <typename T>
int interp3_per_t(T *q, const T *qc, const T *res, const T *so, int ns, const T *ci, int iic,
                  int jjc, int kkc, int iif, int jjf, int kkf, int jpn, void *st) {
    if (!q || !qc || !res || !so || !ci || !sten3_ok(ns) ||
        !coarse3_ok(iif, jjf, kkf, iic, jjc, kkc) || !jpn3_ok(jpn))
        return CDR_ERR_ARG;
    k3::interp_add_per<T>(q, qc, res, so, ns, ci, iic, jjc, kkc, iif, jjf, kkf, jpn, S(st));
    return launched();
}

template <typename T>
int solve_cg3_per_t(T *q, const T *qf, const double *abd, double *bbd, int ii, int jj, int kk,
                    int jpn, void *st) {
    if (!q || !qf || !abd || !bbd || !ext3_ok(ii, jj, kk) || !jpn3_ok(jpn) ||
        (long long)(ii - 2) * (jj - 2) * (kk - 2) > kDenseMax3)
        return CDR_ERR_ARG;
    k3::solve_cg_per<T>(q, qf, abd, bbd, ii, jj, kk, jpn, S(st));
    return launched();
}

}  // namespace

extern "C" {

size_t cdr_hip_norm_work3d(int ii, int jj, int kk) {
    return ext3_ok(ii, jj, kk) ? k3::norm_partials(ii, jj, kk) : 0;
}

int cdr_hip_residual3d_f64(const double *so, int ns, const double *qf, const double *q,
                           double *res, int ii, int jj, int kk, void *st) {
    return residual3_t<double>(so, ns, qf, q, res, ii, jj, kk, st);
}
int cdr_hip_residual3d_f32(const float *so, int ns, const float *qf, const float *q, float *res,
                           int ii, int jj, int kk, void *st) {
    return residual3_t<float>(so, ns, qf, q, res, ii, jj, kk, st);
}

int cdr_hip_residual_norm3d_f64(const double *so, int ns, const double *qf, const double *q,
                                double *res, int ii, int jj, int kk, double *work,
                                double *sumsq_dev, void *st) {
    if (!so || !qf || !q || !res || !work || !sumsq_dev || !ext3_ok(ii, jj, kk) || !sten3_ok(ns))
        return CDR_ERR_ARG;
    k3::residual_norm(so, ns, qf, q, res, ii, jj, kk, work, sumsq_dev, S(st));
    return launched();
}

int cdr_hip_setup_recip3d_f64(const double *so, double *sor, int ii, int jj, int kk, int ns,
                              void *st) {
    if (!so || !sor || !ext3_ok(ii, jj, kk) || !sten3_ok(ns)) return CDR_ERR_ARG;
    k3::setup_recip(so, ns, sor, ii, jj, kk, S(st));
    return launched();
}

int cdr_hip_relax_gs3d_f64(const double *so, int ns, const double *qf, double *q,
                           const double *sor, int ii, int jj, int kk, int ud, void *st) {
    return gs3_t<double>(so, ns, qf, q, sor, ii, jj, kk, ud, st);
}
int cdr_hip_relax_gs3d_f32(const float *so, int ns, const float *qf, float *q, const float *sor,
                           int ii, int jj, int kk, int ud, void *st) {
    return gs3_t<float>(so, ns, qf, q, sor, ii, jj, kk, ud, st);
}

int cdr_hip_restrict3d_f64(const double *ci, const double *q, double *qc, int iif, int jjf,
                           int kkf, int iic, int jjc, int kkc, void *st) {
    return restrict3_t<double>(ci, q, qc, iif, jjf, kkf, iic, jjc, kkc, st);
}
int cdr_hip_restrict3d_f32(const float *ci, const float *q, float *qc, int iif, int jjf, int kkf,
                           int iic, int jjc, int kkc, void *st) {
    return restrict3_t<float>(ci, q, qc, iif, jjf, kkf, iic, jjc, kkc, st);
}

int cdr_hip_interp_add3d_f64(double *q, const double *qc, const double *res, const double *so,
                             int ns, const double *ci, int iic, int jjc, int kkc, int iif,
                             int jjf, int kkf, void *st) {
    return interp3_t<double>(q, qc, res, so, ns, ci, iic, jjc, kkc, iif, jjf, kkf, st);
}
int cdr_hip_interp_add3d_f32(float *q, const float *qc, const float *res, const float *so, int ns,
                             const float *ci, int iic, int jjc, int kkc, int iif, int jjf, int kkf,
                             void *st) {
    return interp3_t<float>(q, qc, res, so, ns, ci, iic, jjc, kkc, iif, jjf, kkf, st);
}

int cdr_hip_setup_interp3d_f64(const double *so, int ns, double *ci, int iif, int jjf, int kkf,
                               int iic, int jjc, int kkc, void *st) {
    if (!so || !ci || !sten3_ok(ns) || !coarse3_ok(iif, jjf, kkf, iic, jjc, kkc))
        return CDR_ERR_ARG;
    k3::setup_interp(so, ns, ci, iif, jjf, kkf, iic, jjc, kkc, S(st));
    return launched();
}

int cdr_hip_setup_galerkin3d_f64(const double *so, int ns, const double *ci, double *soc,
                                 int iif, int jjf, int kkf, int iic, int jjc, int kkc, void *st) {
    if (!so || !ci || !soc || !sten3_ok(ns) || !coarse3_ok(iif, jjf, kkf, iic, jjc, kkc))
        return CDR_ERR_ARG;
    k3::setup_galerkin(so, ns, ci, soc, iif, jjf, kkf, iic, jjc, kkc, S(st));
    return launched();
}

int cdr_hip_cg_bandwidth3d(int ii, int jj) { return ii >= 3 && jj >= 3 ? k3::cg_bandwidth(ii, jj) : 0; }

int cdr_hip_setup_cg3d_f64(const double *so, int ns, double *abd, int ldab, int ii, int jj,
                           int kk, int *info_dev, void *st) {
    if (!so || !abd || !info_dev || !ext3_ok(ii, jj, kk) || !sten3_ok(ns) ||
        ldab < k3::cg_bandwidth(ii, jj) + 1)
        return CDR_ERR_ARG;
    k2::fill<int>(info_dev, 0, 1, S(st));
    k3::setup_cg(so, ns, abd, ldab, ii, jj, kk, info_dev, S(st));
    if (hipGetLastError() != hipSuccess) return CDR_ERR_HIP;
    int info = 0;
    if (hipMemcpyAsync(&info, info_dev, sizeof(int), hipMemcpyDeviceToHost, S(st)) != hipSuccess)
        return CDR_ERR_HIP;
    if (hipStreamSynchronize(S(st)) != hipSuccess) return CDR_ERR_HIP;
    return info != 0 ? CDR_ERR_NOT_SPD : CDR_OK;
}

int cdr_hip_solve_cg3d_f64(double *q, const double *qf, const double *abd, int ldab, double *bbd,
                           int ii, int jj, int kk, void *st) {
    return solve_cg3_t<double>(q, qf, abd, ldab, bbd, ii, jj, kk, st);
}
int cdr_hip_solve_cg3d_f32(float *q, const float *qf, const double *abd, int ldab, double *bbd,
                           int ii, int jj, int kk, void *st) {
    return solve_cg3_t<float>(q, qf, abd, ldab, bbd, ii, jj, kk, st);
}


int cdr_hip_relax_gs3d_per_f64(const double *so, int ns, const double *qf, double *q,
                               const double *sor, int ii, int jj, int kk, int ud, int jpn,
                               void *st) {
    return gs3_per_t<double>(so, ns, qf, q, sor, ii, jj, kk, ud, jpn, st);
}
int cdr_hip_relax_gs3d_per_f32(const float *so, int ns, const float *qf, float *q,
                               const float *sor, int ii, int jj, int kk, int ud, int jpn,
                               void *st) {
    return gs3_per_t<float>(so, ns, qf, q, sor, ii, jj, kk, ud, jpn, st);
}

int cdr_hip_restrict3d_per_f64(const double *ci, double *q, double *qc, int iif, int jjf, int kkf,
                               int iic, int jjc, int kkc, int jpn, void *st) {
    return restrict3_per_t<double>(ci, q, qc, iif, jjf, kkf, iic, jjc, kkc, jpn, st);
}
int cdr_hip_restrict3d_per_f32(const float *ci, float *q, float *qc, int iif, int jjf, int kkf,
                               int iic, int jjc, int kkc, int jpn, void *st) {
    return restrict3_per_t<float>(ci, q, qc, iif, jjf, kkf, iic, jjc, kkc, jpn, st);
}

int cdr_hip_interp_add3d_per_f64(double *q, const double *qc, const double *res, const double *so,
                                 int ns, const double *ci, int iic, int jjc, int kkc, int iif,
                                 int jjf, int kkf, int jpn, void *st) {
    return interp3_per_t<double>(q, qc, res, so, ns, ci, iic, jjc, kkc, iif, jjf, kkf, jpn, st);
}
int cdr_hip_interp_add3d_per_f32(float *q, const float *qc, const float *res, const float *so,
                                 int ns, const float *ci, int iic, int jjc, int kkc, int iif,
                                 int jjf, int kkf, int jpn, void *st) {
    return interp3_per_t<float>(q, qc, res, so, ns, ci, iic, jjc, kkc, iif, jjf, kkf, jpn, st);
}

int cdr_hip_setup_interp3d_per_f64(const double *so, int ns, double *ci, int iif, int jjf,
                                   int kkf, int iic, int jjc, int kkc, int jpn, void *st) {
    if (!so || !ci || !sten3_ok(ns) || !coarse3_ok(iif, jjf, kkf, iic, jjc, kkc) || !jpn3_ok(jpn))
        return CDR_ERR_ARG;
    k3::setup_interp_per(so, ns, ci, iif, jjf, kkf, iic, jjc, kkc, jpn, S(st));
    return launched();
}

int cdr_hip_setup_galerkin3d_per_f64(const double *so, int ns, const double *ci, double *soc,
                                     int iif, int jjf, int kkf, int iic, int jjc, int kkc, int jpn,
                                     void *st) {
    if (!so || !ci || !soc || !sten3_ok(ns) || !coarse3_ok(iif, jjf, kkf, iic, jjc, kkc) ||
        !jpn3_ok(jpn) || !even_per3(jpn, iif, jjf, kkf))
        return CDR_ERR_ARG;
    k3::setup_galerkin_per(so, ns, ci, soc, iif, jjf, kkf, iic, jjc, kkc, jpn, S(st));
    return launched();
}

int cdr_hip_setup_cg3d_per_f64(const double *so, int ns, double *abd, int ii, int jj, int kk,
                               int jpn, int *info_dev, void *st) {
    if (!so || !abd || !info_dev || !ext3_ok(ii, jj, kk) || !sten3_ok(ns) || !jpn3_ok(jpn) ||
        (long long)(ii - 2) * (jj - 2) * (kk - 2) > kDenseMax3)
        return CDR_ERR_ARG;
    k2::fill<int>(info_dev, 0, 1, S(st));
    k3::setup_cg_per(so, ns, abd, ii, jj, kk, jpn, info_dev, S(st));
    if (hipGetLastError() != hipSuccess) return CDR_ERR_HIP;
    int info = 0;
    if (hipMemcpyAsync(&info, info_dev, sizeof(int), hipMemcpyDeviceToHost, S(st)) != hipSuccess)
        return CDR_ERR_HIP;
    if (hipStreamSynchronize(S(st)) != hipSuccess) return CDR_ERR_HIP;
    return info != 0 ? CDR_ERR_NOT_SPD : CDR_OK;
}

int cdr_hip_solve_cg3d_per_f64(double *q, const double *qf, const double *abd, double *bbd, int ii,
                               int jj, int kk, int jpn, void *st) {
    return solve_cg3_per_t<double>(q, qf, abd, bbd, ii, jj, kk, jpn, st);
}
int cdr_hip_solve_cg3d_per_f32(float *q, const float *qf, const double *abd, double *bbd, int ii,
                               int jj, int kk, int jpn, void *st) {
    return solve_cg3_per_t<float>(q, qf, abd, bbd, ii, jj, kk, jpn, st);
}

}  // extern "C"
