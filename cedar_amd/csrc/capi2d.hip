// C ABI of the 2D kernels: argument checks, launch, status.  See include/cedar_hip.h.
#include "../../include/cedar_hip.h"

#include <hip/hip_runtime.h>

#include "bmg2_launch.hpp"

namespace {

using namespace cdr;

inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline int launched() { return hipGetLastError() == hipSuccess ? CDR_OK : CDR_ERR_HIP; }

inline bool ext_ok(int ii, int jj) { return ii >= 3 && jj >= 3 && (long long)ii * jj < (1LL << 31); }
inline bool sten_ok(int ns) { return ns == 3 || ns == 5; }
inline bool dir_ok(int ud) { return ud == CDR_DOWN || ud == CDR_UP; }
inline bool coarse_ok(int iif, int jjf, int iic, int jjc) {
    // coarse interior = (fine interior - 1)/2 + 1, Cedar's coarsening rule
    return ext_ok(iif, jjf) && ext_ok(iic, jjc) && iic - 2 == (iif - 3) / 2 + 1 &&
           jjc - 2 == (jjf - 3) / 2 + 1;
}

template <typename T>
int residual_t(const T *so, int ns, const T *qf, const T *q, T *res, int ii, int jj, void *st) {
    if (!so || !qf || !q || !res || !ext_ok(ii, jj) || !sten_ok(ns)) return CDR_ERR_ARG;
    k2::residual<T>(so, ns, qf, q, res, ii, jj, S(st));
    return launched();
}

template <typename T>
int gs_t(const T *so, int ns, const T *qf, T *q, const T *sor, int ii, int jj, int ud, void *st) {
    if (!so || !qf || !q || !sor || !ext_ok(ii, jj) || !sten_ok(ns) || !dir_ok(ud))
        return CDR_ERR_ARG;
    k2::relax_gs<T>(so, ns, qf, q, sor, ii, jj, ud, S(st));
    return launched();
}

template <typename T, int DIR>
int lines_t(const T *so, int ns, const T *qf, T *q, const T *sor, int ii, int jj, int ud, void *st) {
    if (!so || !qf || !q || !sor || !ext_ok(ii, jj) || !sten_ok(ns) || !dir_ok(ud))
        return CDR_ERR_ARG;
    if (DIR == 0)
        k2::relax_lines_x<T>(so, ns, qf, q, sor, ii, jj, ud, S(st));
    else
        k2::relax_lines_y<T>(so, ns, qf, q, sor, ii, jj, ud, S(st));
    return launched();
}

template <typename T>
int restrict_t(const T *ci, const T *q, T *qc, int iif, int jjf, int iic, int jjc, void *st) {
    if (!ci || !q || !qc || !coarse_ok(iif, jjf, iic, jjc)) return CDR_ERR_ARG;
    k2::restrict_<T>(ci, q, qc, iif, jjf, iic, jjc, S(st));
    return launched();
}

template <typename T>
int interp_t(T *q, const T *qc, const T *res, const T *so, int ns, const T *ci, int iic, int jjc,
             int iif, int jjf, void *st) {
    if (!q || !qc || !res || !so || !ci || !sten_ok(ns) || !coarse_ok(iif, jjf, iic, jjc))
        return CDR_ERR_ARG;
    k2::interp_add<T>(q, qc, res, so, ns, ci, iic, jjc, iif, jjf, S(st));
    return launched();
}

template <typename T>
int solve_cg_t(T *q, const T *qf, const double *abd, double *bbd, int ii, int jj, void *st) {
    if (!q || !qf || !abd || !bbd || !ext_ok(ii, jj)) return CDR_ERR_ARG;
    k2::solve_cg<T>(q, qf, abd, bbd, ii, jj, S(st));
    return launched();
}

// runs a factorisation kernel, then reads its device info word (setup-time synchronisation)
template <typename F>
int with_info(int *info_dev, void *st, F launch) {
    if (!info_dev) return CDR_ERR_ARG;
    const int big = 0x7fffffff;
    if (hipMemcpyAsync(info_dev, &big, sizeof(int), hipMemcpyHostToDevice, S(st)) != hipSuccess)
        return CDR_ERR_HIP;
    launch();
    if (hipGetLastError() != hipSuccess) return CDR_ERR_HIP;
    int info = 0;
    if (hipMemcpyAsync(&info, info_dev, sizeof(int), hipMemcpyDeviceToHost, S(st)) != hipSuccess)
        return CDR_ERR_HIP;
    if (hipStreamSynchronize(S(st)) != hipSuccess) return CDR_ERR_HIP;
    return (info != big && info != 0) ? CDR_ERR_NOT_SPD : CDR_OK;
}

}  // namespace

extern "C" {

const char *cdr_hip_status_string(int s) {
    switch (s) {
    case CDR_OK: return "ok";
    case CDR_ERR_ARG: return "invalid argument";
    case CDR_ERR_HIP: return "HIP runtime error";
    case CDR_ERR_NOT_SPD: return "factorisation failed: operator not positive definite";
    case CDR_ERR_NONFINITE: return "non-finite residual norm";
    case CDR_ERR_NO_DEVICE: return "no GPU available";
    case CDR_ERR_LEVELS: return "too many levels specified";
    case CDR_ERR_STALLED: return "residual stopped decreasing";
    default: return "unknown status";
    }
}

const char *cdr_hip_version(void) { return "cedar_amd 0.1 gfx950"; }

size_t cdr_hip_norm_work2d(int ii, int jj) {
    return ext_ok(ii, jj) ? k2::norm_partials(ii, jj) : 0;
}

int cdr_hip_residual2d_f64(const double *so, int ns, const double *qf, const double *q,
                           double *res, int ii, int jj, void *st) {
    return residual_t<double>(so, ns, qf, q, res, ii, jj, st);
}
int cdr_hip_residual2d_f32(const float *so, int ns, const float *qf, const float *q, float *res,
                           int ii, int jj, void *st) {
    return residual_t<float>(so, ns, qf, q, res, ii, jj, st);
}

int cdr_hip_residual_norm2d_f64(const double *so, int ns, const double *qf, const double *q,
                                double *res, int ii, int jj, double *work, double *sumsq_dev,
                                void *st) {
    if (!so || !qf || !q || !res || !work || !sumsq_dev || !ext_ok(ii, jj) || !sten_ok(ns))
        return CDR_ERR_ARG;
    k2::residual_norm<double>(so, ns, qf, q, res, ii, jj, work, sumsq_dev, S(st));
    return launched();
}

int cdr_hip_sumsq2d_f64(const double *x, int ii, int jj, double *work, double *sumsq_dev,
                        void *st) {
    if (!x || !work || !sumsq_dev || !ext_ok(ii, jj)) return CDR_ERR_ARG;
    k2::sumsq<double>(x, ii, jj, work, sumsq_dev, S(st));
    return launched();
}

int cdr_hip_setup_recip2d_f64(const double *so, double *sor, int ii, int jj, int ns, void *st) {
    if (!so || !sor || !ext_ok(ii, jj) || !sten_ok(ns)) return CDR_ERR_ARG;
    k2::setup_recip(so, sor, ii, jj, S(st));
    return launched();
}

int cdr_hip_relax_gs2d_f64(const double *so, int ns, const double *qf, double *q,
                           const double *sor, int ii, int jj, int ud, void *st) {
    return gs_t<double>(so, ns, qf, q, sor, ii, jj, ud, st);
}
int cdr_hip_relax_gs2d_f32(const float *so, int ns, const float *qf, float *q, const float *sor,
                           int ii, int jj, int ud, void *st) {
    return gs_t<float>(so, ns, qf, q, sor, ii, jj, ud, st);
}

int cdr_hip_setup_lines_x2d_f64(const double *so, double *sor, int ii, int jj, int ns,
                                int *info_dev, void *st) {
    if (!so || !sor || !ext_ok(ii, jj) || !sten_ok(ns)) return CDR_ERR_ARG;
    return with_info(info_dev, st, [&] { k2::setup_lines_x(so, sor, ii, jj, info_dev, S(st)); });
}
int cdr_hip_setup_lines_y2d_f64(const double *so, double *sor, int ii, int jj, int ns,
                                int *info_dev, void *st) {
    if (!so || !sor || !ext_ok(ii, jj) || !sten_ok(ns)) return CDR_ERR_ARG;
    return with_info(info_dev, st, [&] { k2::setup_lines_y(so, sor, ii, jj, info_dev, S(st)); });
}

int cdr_hip_relax_lines_x2d_f64(const double *so, int ns, const double *qf, double *q,
                                const double *sor, double *b, int ii, int jj, int ud, void *st) {
    (void)b;
    return lines_t<double, 0>(so, ns, qf, q, sor, ii, jj, ud, st);
}
int cdr_hip_relax_lines_y2d_f64(const double *so, int ns, const double *qf, double *q,
                                const double *sor, double *b, int ii, int jj, int ud, void *st) {
    (void)b;
    return lines_t<double, 1>(so, ns, qf, q, sor, ii, jj, ud, st);
}
int cdr_hip_relax_lines_x2d_f32(const float *so, int ns, const float *qf, float *q,
                                const float *sor, float *b, int ii, int jj, int ud, void *st) {
    (void)b;
    return lines_t<float, 0>(so, ns, qf, q, sor, ii, jj, ud, st);
}
int cdr_hip_relax_lines_y2d_f32(const float *so, int ns, const float *qf, float *q,
                                const float *sor, float *b, int ii, int jj, int ud, void *st) {
    (void)b;
    return lines_t<float, 1>(so, ns, qf, q, sor, ii, jj, ud, st);
}

int cdr_hip_restrict2d_f64(const double *ci, const double *q, double *qc, int iif, int jjf,
                           int iic, int jjc, void *st) {
    return restrict_t<double>(ci, q, qc, iif, jjf, iic, jjc, st);
}
int cdr_hip_restrict2d_f32(const float *ci, const float *q, float *qc, int iif, int jjf, int iic,
                           int jjc, void *st) {
    return restrict_t<float>(ci, q, qc, iif, jjf, iic, jjc, st);
}

int cdr_hip_interp_add2d_f64(double *q, const double *qc, const double *res, const double *so,
                             int ns, const double *ci, int iic, int jjc, int iif, int jjf,
                             void *st) {
    return interp_t<double>(q, qc, res, so, ns, ci, iic, jjc, iif, jjf, st);
}
int cdr_hip_interp_add2d_f32(float *q, const float *qc, const float *res, const float *so, int ns,
                             const float *ci, int iic, int jjc, int iif, int jjf, void *st) {
    return interp_t<float>(q, qc, res, so, ns, ci, iic, jjc, iif, jjf, st);
}

int cdr_hip_setup_interp2d_f64(const double *so, int ns, double *ci, int iif, int jjf, int iic,
                               int jjc, void *st) {
    if (!so || !ci || !sten_ok(ns) || !coarse_ok(iif, jjf, iic, jjc)) return CDR_ERR_ARG;
    k2::setup_interp(so, ns, ci, iif, jjf, iic, jjc, S(st));
    return launched();
}

int cdr_hip_setup_galerkin2d_f64(const double *so, int ns, const double *ci, double *soc, int iif,
                                 int jjf, int iic, int jjc, void *st) {
    if (!so || !ci || !soc || !sten_ok(ns) || !coarse_ok(iif, jjf, iic, jjc)) return CDR_ERR_ARG;
    k2::setup_galerkin(so, ns, ci, soc, iif, jjf, iic, jjc, S(st));
    return launched();
}

int cdr_hip_setup_cg2d_f64(const double *so, int ns, double *abd, int ii, int jj, int *info_dev,
                           void *st) {
    if (!so || !abd || !ext_ok(ii, jj) || !sten_ok(ns)) return CDR_ERR_ARG;
    return with_info(info_dev, st, [&] { k2::setup_cg(so, ns, abd, ii, jj, info_dev, S(st)); });
}

int cdr_hip_solve_cg2d_f64(double *q, const double *qf, const double *abd, double *bbd, int ii,
                           int jj, void *st) {
    return solve_cg_t<double>(q, qf, abd, bbd, ii, jj, st);
}
int cdr_hip_solve_cg2d_f32(float *q, const float *qf, const double *abd, double *bbd, int ii,
                           int jj, void *st) {
    return solve_cg_t<float>(q, qf, abd, bbd, ii, jj, st);
}

}  // extern "C"
