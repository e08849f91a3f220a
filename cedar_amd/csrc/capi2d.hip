// C ABI of the 2D kernels: argument checks, launch, status.  See include/cedar_hip.h.
#include "../../include/cedar_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "bmg2_launch.hpp"
#include "capi2d_impl.hpp"

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

inline bool nb_ok(int nb) { return nb >= 1 && nb <= 65535; }
// rep problems per operator: a positive divisor of the batch
inline bool rep_ok(int nb, int rep) { return rep >= 1 && nb % rep == 0; }

// runs a factorisation kernel, then reads its device info word (setup-time synchronisation)
template <typename F>
int with_info(int *info_dev, hipStream_t st, F launch) {
    if (!info_dev) return CDR_ERR_ARG;
    k2::fill<int>(info_dev, 0, 1, st);  // kernels raise it with atomicMax
    launch();
    if (hipGetLastError() != hipSuccess) return CDR_ERR_HIP;
    int info = 0;
    if (hipMemcpyAsync(&info, info_dev, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess)
        return CDR_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return CDR_ERR_HIP;
    return info != 0 ? CDR_ERR_NOT_SPD : CDR_OK;
}

}  // namespace

namespace cdr {
namespace api2 {

template <typename T>
int residual(const T *so, int ns, const T *qf, const T *q, T *res, int ii, int jj, hipStream_t s,
             int nb, int rep) {
    if (!so || !qf || !q || !res || !ext_ok(ii, jj) || !sten_ok(ns) || !nb_ok(nb) ||
        !rep_ok(nb, rep))
        return CDR_ERR_ARG;
    k2::residual<T>(so, ns, qf, q, res, ii, jj, s, nb, rep);
    return launched();
}

int residual_norm(const double *so, int ns, const double *qf, const double *q, double *res, int ii,
                  int jj, double *work, double *sumsq_dev, hipStream_t s, int nb, int rep) {
    if (!so || !qf || !q || !res || !work || !sumsq_dev || !ext_ok(ii, jj) || !sten_ok(ns) ||
        !nb_ok(nb) || !rep_ok(nb, rep))
        return CDR_ERR_ARG;
    k2::residual_norm<double>(so, ns, qf, q, res, ii, jj, work, sumsq_dev, s, nb, rep);
    return launched();
}

int sumsq(const double *x, int ii, int jj, double *work, double *sumsq_dev, hipStream_t s,
          int nb) {
    if (!x || !work || !sumsq_dev || !ext_ok(ii, jj) || !nb_ok(nb)) return CDR_ERR_ARG;
    k2::sumsq<double>(x, ii, jj, work, sumsq_dev, s, nb);
    return launched();
}

int setup_recip(const double *so, double *sor, int ii, int jj, int ns, hipStream_t s, int nb) {
    if (!so || !sor || !ext_ok(ii, jj) || !sten_ok(ns) || !nb_ok(nb)) return CDR_ERR_ARG;
    k2::setup_recip(so, ns, sor, ii, jj, s, nb);
    return launched();
}

template <typename T>
int relax_gs(const T *so, int ns, const T *qf, T *q, const T *sor, int ii, int jj, int ud,
             hipStream_t s, int nb, int rep) {
    if (!so || !qf || !q || !sor || !ext_ok(ii, jj) || !sten_ok(ns) || !dir_ok(ud) || !nb_ok(nb) ||
        !rep_ok(nb, rep))
        return CDR_ERR_ARG;
    k2::relax_gs<T>(so, ns, qf, q, sor, ii, jj, ud, s, nb, rep);
    return launched();
}

int setup_lines(int dir, const double *so, double *sor, int ii, int jj, int ns, int *info_dev,
                hipStream_t s, int nb) {
    if (!so || !sor || !ext_ok(ii, jj) || !sten_ok(ns) || !nb_ok(nb) || (dir != 0 && dir != 1))
        return CDR_ERR_ARG;
    return with_info(info_dev, s, [&] {
        if (dir == 0)
            k2::setup_lines_x(so, ns, sor, ii, jj, info_dev, s, nb);
        else
            k2::setup_lines_y(so, ns, sor, ii, jj, info_dev, s, nb);
    });
}

template <typename T>
int relax_lines(int dir, const T *so, int ns, const T *qf, T *q, const T *sor, int ii, int jj,
                int ud, hipStream_t s, int nb, int rep) {
    if (!so || !qf || !q || !sor || !ext_ok(ii, jj) || !sten_ok(ns) || !dir_ok(ud) ||
        !nb_ok(nb) || !rep_ok(nb, rep) || (dir != 0 && dir != 1))
        return CDR_ERR_ARG;
    if (dir == 0)
        k2::relax_lines_x<T>(so, ns, qf, q, sor, ii, jj, ud, s, nb, rep);
    else
        k2::relax_lines_y<T>(so, ns, qf, q, sor, ii, jj, ud, s, nb, rep);
    return launched();
}

template <typename T>
int relax_lines_xy(const T *so, int ns, const T *qf, T *q, const T *sorx, const T *sory, T *work,
                   int ii, int jj, int ud, hipStream_t s, int nb, int rep) {
    if (!so || !qf || !q || !sorx || !sory || !ext_ok(ii, jj) || !sten_ok(ns) || !dir_ok(ud) ||
        !nb_ok(nb) || !rep_ok(nb, rep))
        return CDR_ERR_ARG;
    k2::relax_lines_xy<T>(so, ns, qf, q, sorx, sory, work, ii, jj, ud, s, nb, rep);
    return launched();
}

template <typename T>
int restrict_(const T *ci, const T *q, T *qc, int iif, int jjf, int iic, int jjc, hipStream_t s,
              int nb, int rep) {
    if (!ci || !q || !qc || !coarse_ok(iif, jjf, iic, jjc) || !nb_ok(nb) || !rep_ok(nb, rep))
        return CDR_ERR_ARG;
    k2::restrict_<T>(ci, q, qc, iif, jjf, iic, jjc, s, nb, rep);
    return launched();
}

template <typename T>
int interp_add(T *q, const T *qc, const T *res, const T *so, int ns, const T *ci, int iic, int jjc,
               int iif, int jjf, hipStream_t s, int nb, int rep) {
    if (!q || !qc || !res || !so || !ci || !sten_ok(ns) || !coarse_ok(iif, jjf, iic, jjc) ||
        !nb_ok(nb) || !rep_ok(nb, rep))
        return CDR_ERR_ARG;
    k2::interp_add<T>(q, qc, res, so, ns, ci, iic, jjc, iif, jjf, s, nb, rep);
    return launched();
}

int setup_interp(const double *so, int ns, double *ci, int iif, int jjf, int iic, int jjc,
                 hipStream_t s, int nb) {
    if (!so || !ci || !sten_ok(ns) || !coarse_ok(iif, jjf, iic, jjc) || !nb_ok(nb))
        return CDR_ERR_ARG;
    k2::setup_interp(so, ns, ci, iif, jjf, iic, jjc, s, nb);
    return launched();
}

int setup_galerkin(const double *so, int ns, const double *ci, double *soc, int iif, int jjf,
                   int iic, int jjc, hipStream_t s, int nb) {
    if (!so || !ci || !soc || !sten_ok(ns) || !coarse_ok(iif, jjf, iic, jjc) || !nb_ok(nb))
        return CDR_ERR_ARG;
    k2::setup_galerkin(so, ns, ci, soc, iif, jjf, iic, jjc, s, nb);
    return launched();
}

int setup_cg(const double *so, int ns, double *abd, int ii, int jj, int *info_dev, hipStream_t s,
             int nb) {
    if (!so || !abd || !ext_ok(ii, jj) || !sten_ok(ns) || !nb_ok(nb)) return CDR_ERR_ARG;
    return with_info(info_dev, s, [&] { k2::setup_cg(so, ns, abd, ii, jj, info_dev, s, nb); });
}

template <typename T>
int solve_cg(T *q, const T *qf, const double *abd, double *bbd, int ii, int jj, hipStream_t s,
             int nb, int rep) {
    if (!q || !qf || !abd || !bbd || !ext_ok(ii, jj) || !nb_ok(nb) || !rep_ok(nb, rep))
        return CDR_ERR_ARG;
    k2::solve_cg<T>(q, qf, abd, bbd, ii, jj, s, nb, rep);
    return launched();
}

// ---- periodic boundaries (jpn = BMG_get_bc(per_mask): 0, 1, 2 or 3) -------------------------
inline bool jpn_ok(int jpn) { return jpn >= 0 && jpn <= 3; }

template <typename T>
int relax_gs_per(const T *so, int ns, const T *qf, T *q, const T *sor, int ii, int jj, int ud,
                 int jpn, hipStream_t s, int nb) {
    if (!so || !qf || !q || !sor || !ext_ok(ii, jj) || !sten_ok(ns) || !dir_ok(ud) ||
        !nb_ok(nb) || !jpn_ok(jpn))
        return CDR_ERR_ARG;
    if (jpn == 0)
        k2::relax_gs<T>(so, ns, qf, q, sor, ii, jj, ud, s, nb);
    else
        k2::relax_gs_per<T>(so, ns, qf, q, sor, ii, jj, ud, jpn, s, nb);
    return launched();
}

template <typename T>
int restrict_per(const T *ci, T *q, T *qc, int iif, int jjf, int iic, int jjc, int jpn,
                 hipStream_t s, int nb) {
    if (!ci || !q || !qc || !coarse_ok(iif, jjf, iic, jjc) || !nb_ok(nb) || !jpn_ok(jpn))
        return CDR_ERR_ARG;
    k2::restrict_per<T>(ci, q, qc, iif, jjf, iic, jjc, jpn, s, nb);
    return launched();
}

template <typename T>
int interp_add_per(T *q, const T *qc, const T *res, const T *so, int ns, const T *ci, int iic,
                   int jjc, int iif, int jjf, int jpn, hipStream_t s, int nb) {
    if (!q || !qc || !res || !so || !ci || !sten_ok(ns) || !coarse_ok(iif, jjf, iic, jjc) ||
        !nb_ok(nb) || !jpn_ok(jpn))
        return CDR_ERR_ARG;
    k2::interp_add_per<T>(q, qc, res, so, ns, ci, iic, jjc, iif, jjf, jpn, s, nb);
    return launched();
}

int setup_interp_per(const double *so, int ns, double *ci, int iif, int jjf, int iic, int jjc,
                     int jpn, hipStream_t s, int nb) {
    if (!so || !ci || !sten_ok(ns) || !coarse_ok(iif, jjf, iic, jjc) || !nb_ok(nb) ||
        !jpn_ok(jpn))
        return CDR_ERR_ARG;
    if (jpn == 0)
        k2::setup_interp(so, ns, ci, iif, jjf, iic, jjc, s, nb);
    else
        k2::setup_interp_per(so, ns, ci, iif, jjf, iic, jjc, jpn, s, nb);
    return launched();
}

int setup_galerkin_per(const double *so, int ns, const double *ci, double *soc, int iif, int jjf,
                       int iic, int jjc, int jpn, hipStream_t s, int nb) {
    if (!so || !ci || !soc || !sten_ok(ns) || !coarse_ok(iif, jjf, iic, jjc) || !nb_ok(nb) ||
        !jpn_ok(jpn))
        return CDR_ERR_ARG;
    if (jpn == 0)
        k2::setup_galerkin(so, ns, ci, soc, iif, jjf, iic, jjc, s, nb);
    else
        k2::setup_galerkin_per(so, ns, ci, soc, iif, jjf, iic, jjc, jpn, s, nb);
    return launched();
}

int setup_lines_per(int dir, const double *so, double *sor, int ii, int jj, int ns, int jpn,
                    int *info_dev, hipStream_t s, int nb) {
    if (!so || !sor || !ext_ok(ii, jj) || !sten_ok(ns) || !nb_ok(nb) || (dir != 0 && dir != 1) ||
        !jpn_ok(jpn))
        return CDR_ERR_ARG;
    return with_info(info_dev, s, [&] {
        k2::setup_lines_per(dir, so, ns, sor, ii, jj, jpn, info_dev, s, nb);
    });
}

template <typename T>
int relax_lines_per(int dir, const T *so, int ns, const T *qf, T *q, const T *sor, T *work, int ii,
                    int jj, int ud, int jpn, hipStream_t s, int nb) {
    if (!so || !qf || !q || !sor || !work || !ext_ok(ii, jj) || !sten_ok(ns) || !dir_ok(ud) ||
        !nb_ok(nb) || (dir != 0 && dir != 1) || !jpn_ok(jpn))
        return CDR_ERR_ARG;
    k2::relax_lines_per<T>(dir, so, ns, qf, q, sor, work, ii, jj, ud, jpn, s, nb);
    return launched();
}

// the dense periodic matrix needs at least 2 x 2 points (all of its writes then stay inside it)
inline bool per_cg_ok(int ii, int jj) {
    return ii >= 4 && jj >= 4 && (long long)(ii - 2) * (jj - 2) <= 2048;
}

int setup_cg_per(const double *so, int ns, double *abd, int ii, int jj, int jpn, int *info_dev,
                 hipStream_t s, int nb) {
    if (!so || !abd || !ext_ok(ii, jj) || !sten_ok(ns) || !nb_ok(nb) || !jpn_ok(jpn))
        return CDR_ERR_ARG;
    if (jpn == 0) return setup_cg(so, ns, abd, ii, jj, info_dev, s, nb);
    if (!per_cg_ok(ii, jj)) return CDR_ERR_ARG;
    return with_info(info_dev, s, [&] { k2::setup_cg_per(so, ns, abd, ii, jj, jpn, info_dev, s, nb); });
}

template <typename T>
int solve_cg_per(T *q, const T *qf, const double *abd, double *bbd, int ii, int jj, int jpn,
                 hipStream_t s, int nb) {
    if (!q || !qf || !abd || !bbd || !ext_ok(ii, jj) || !nb_ok(nb) || !jpn_ok(jpn))
        return CDR_ERR_ARG;
    if (jpn == 0) return solve_cg<T>(q, qf, abd, bbd, ii, jj, s, nb);
    if (!per_cg_ok(ii, jj)) return CDR_ERR_ARG;
    k2::solve_cg_per<T>(q, qf, abd, bbd, ii, jj, jpn, s, nb);
    return launched();
}

#define CDR_API_INST(T)                                                                            \
    template int relax_gs_per<T>(const T *, int, const T *, T *, const T *, int, int, int, int,    \
                                 hipStream_t, int);                                                \
    template int restrict_per<T>(const T *, T *, T *, int, int, int, int, int, hipStream_t, int);  \
    template int interp_add_per<T>(T *, const T *, const T *, const T *, int, const T *, int, int,  \
                                   int, int, int, hipStream_t, int);                               \
    template int solve_cg_per<T>(T *, const T *, const double *, double *, int, int, int,          \
                                 hipStream_t, int);                                                \
    template int relax_lines_per<T>(int, const T *, int, const T *, T *, const T *, T *, int, int, \
                                    int, int, hipStream_t, int);                                   \
    template int residual<T>(const T *, int, const T *, const T *, T *, int, int, hipStream_t,     \
                             int, int);                                                            \
    template int relax_gs<T>(const T *, int, const T *, T *, const T *, int, int, int, hipStream_t, \
                             int, int);                                                            \
    template int relax_lines<T>(int, const T *, int, const T *, T *, const T *, int, int, int,      \
                                hipStream_t, int, int);                                            \
    template int relax_lines_xy<T>(const T *, int, const T *, T *, const T *, const T *, T *, int,  \
                                   int, int, hipStream_t, int, int);                               \
    template int restrict_<T>(const T *, const T *, T *, int, int, int, int, hipStream_t, int,     \
                              int);                                                                \
    template int interp_add<T>(T *, const T *, const T *, const T *, int, const T *, int, int, int, \
                               int, hipStream_t, int, int);                                        \
    template int solve_cg<T>(T *, const T *, const double *, double *, int, int, hipStream_t, int,  \
                             int);
CDR_API_INST(double)
CDR_API_INST(float)
#undef CDR_API_INST

}  // namespace api2
}  // namespace cdr

namespace api2 = cdr::api2;

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
    case CDR_ERR_TOO_LARGE: return "input larger than the library's fixed-size device buffer";
    default: return "unknown status";
    }
}

const char *cdr_hip_version(void) { return "cedar_amd 0.1 gfx950"; }

size_t cdr_hip_norm_work2d(int ii, int jj) {
    return ext_ok(ii, jj) ? k2::norm_partials(ii, jj) : 0;
}

int cdr_hip_residual2d_f64(const double *so, int ns, const double *qf, const double *q,
                           double *res, int ii, int jj, void *st) {
    return api2::residual<double>(so, ns, qf, q, res, ii, jj, S(st));
}
int cdr_hip_residual2d_f32(const float *so, int ns, const float *qf, const float *q, float *res,
                           int ii, int jj, void *st) {
    return api2::residual<float>(so, ns, qf, q, res, ii, jj, S(st));
}

int cdr_hip_residual_norm2d_f64(const double *so, int ns, const double *qf, const double *q,
                                double *res, int ii, int jj, double *work, double *sumsq_dev,
                                void *st) {
    return api2::residual_norm(so, ns, qf, q, res, ii, jj, work, sumsq_dev, S(st));
}

int cdr_hip_sumsq2d_f64(const double *x, int ii, int jj, double *work, double *sumsq_dev,
                        void *st) {
    return api2::sumsq(x, ii, jj, work, sumsq_dev, S(st));
}

int cdr_hip_setup_recip2d_f64(const double *so, double *sor, int ii, int jj, int ns, void *st) {
    return api2::setup_recip(so, sor, ii, jj, ns, S(st));
}

int cdr_hip_relax_gs2d_f64(const double *so, int ns, const double *qf, double *q,
                           const double *sor, int ii, int jj, int ud, void *st) {
    return api2::relax_gs<double>(so, ns, qf, q, sor, ii, jj, ud, S(st));
}
int cdr_hip_relax_gs2d_f32(const float *so, int ns, const float *qf, float *q, const float *sor,
                           int ii, int jj, int ud, void *st) {
    return api2::relax_gs<float>(so, ns, qf, q, sor, ii, jj, ud, S(st));
}

int cdr_hip_setup_lines_x2d_f64(const double *so, double *sor, int ii, int jj, int ns,
                                int *info_dev, void *st) {
    return api2::setup_lines(0, so, sor, ii, jj, ns, info_dev, S(st));
}
int cdr_hip_setup_lines_y2d_f64(const double *so, double *sor, int ii, int jj, int ns,
                                int *info_dev, void *st) {
    return api2::setup_lines(1, so, sor, ii, jj, ns, info_dev, S(st));
}

int cdr_hip_relax_lines_x2d_f64(const double *so, int ns, const double *qf, double *q,
                                const double *sor, double *b, int ii, int jj, int ud, void *st) {
    (void)b;
    return api2::relax_lines<double>(0, so, ns, qf, q, sor, ii, jj, ud, S(st));
}
int cdr_hip_relax_lines_y2d_f64(const double *so, int ns, const double *qf, double *q,
                                const double *sor, double *b, int ii, int jj, int ud, void *st) {
    (void)b;
    return api2::relax_lines<double>(1, so, ns, qf, q, sor, ii, jj, ud, S(st));
}
int cdr_hip_relax_lines_x2d_f32(const float *so, int ns, const float *qf, float *q,
                                const float *sor, float *b, int ii, int jj, int ud, void *st) {
    (void)b;
    return api2::relax_lines<float>(0, so, ns, qf, q, sor, ii, jj, ud, S(st));
}
int cdr_hip_relax_lines_y2d_f32(const float *so, int ns, const float *qf, float *q,
                                const float *sor, float *b, int ii, int jj, int ud, void *st) {
    (void)b;
    return api2::relax_lines<float>(1, so, ns, qf, q, sor, ii, jj, ud, S(st));
}

int cdr_hip_relax_lines2d_batched_f64(int dir, const double *so, int ns, const double *qf,
                                      double *q, const double *sor, int ii, int jj, int nb,
                                      int ud, void *st) {
    return api2::relax_lines<double>(dir, so, ns, qf, q, sor, ii, jj, ud, S(st), nb);
}
int cdr_hip_relax_lines2d_batched_f32(int dir, const float *so, int ns, const float *qf,
                                      float *q, const float *sor, int ii, int jj, int nb, int ud,
                                      void *st) {
    return api2::relax_lines<float>(dir, so, ns, qf, q, sor, ii, jj, ud, S(st), nb);
}

int cdr_hip_relax_lines_xy2d_f64(const double *so, int ns, const double *qf, double *q,
                                 const double *sorx, const double *sory, double *work, int ii,
                                 int jj, int ud, void *st) {
    return api2::relax_lines_xy<double>(so, ns, qf, q, sorx, sory, work, ii, jj, ud, S(st));
}
int cdr_hip_relax_lines_xy2d_f32(const float *so, int ns, const float *qf, float *q,
                                 const float *sorx, const float *sory, float *work, int ii, int jj,
                                 int ud, void *st) {
    return api2::relax_lines_xy<float>(so, ns, qf, q, sorx, sory, work, ii, jj, ud, S(st));
}

int cdr_hip_autotune(int on) { return cdr::k2::set_autotune(on); }

int cdr_hip_line_shapes2d(int dir, int fp_bits, int nstncl, int ii, int jj) {
    if (dir < 0 || dir > 2 || (fp_bits != 32 && fp_bits != 64) || !sten_ok(nstncl) ||
        !ext_ok(ii, jj))
        return -CDR_ERR_ARG;
    return cdr::k2::line_shape_count(dir, fp_bits / 8, nstncl, ii, jj);
}

int cdr_hip_set_line_shape2d(int dir, int fp_bits, int nstncl, int ii, int jj, int nb,
                             int shape) {
    if (dir < 0 || dir > 2 || (fp_bits != 32 && fp_bits != 64) || !sten_ok(nstncl) ||
        !ext_ok(ii, jj) || !nb_ok(nb))
        return CDR_ERR_ARG;
    if (shape >= cdr::k2::line_shape_count(dir, fp_bits / 8, nstncl, ii, jj)) return CDR_ERR_ARG;
    cdr::k2::set_line_shape(dir, fp_bits / 8, nstncl, ii, jj, nb, shape);
    return CDR_OK;
}

int cdr_hip_autotune_report(char *buf, int len) {
    const std::string r = cdr::k2::tune_report();
    if (buf && len > 0) {
        const std::size_t n = std::min(r.size(), (std::size_t)len - 1);
        std::memcpy(buf, r.data(), n);
        buf[n] = '\0';
    }
    return (int)r.size();
}

int cdr_hip_restrict2d_f64(const double *ci, const double *q, double *qc, int iif, int jjf,
                           int iic, int jjc, void *st) {
    return api2::restrict_<double>(ci, q, qc, iif, jjf, iic, jjc, S(st));
}
int cdr_hip_restrict2d_f32(const float *ci, const float *q, float *qc, int iif, int jjf, int iic,
                           int jjc, void *st) {
    return api2::restrict_<float>(ci, q, qc, iif, jjf, iic, jjc, S(st));
}

int cdr_hip_interp_add2d_f64(double *q, const double *qc, const double *res, const double *so,
                             int ns, const double *ci, int iic, int jjc, int iif, int jjf,
                             void *st) {
    return api2::interp_add<double>(q, qc, res, so, ns, ci, iic, jjc, iif, jjf, S(st));
}
int cdr_hip_interp_add2d_f32(float *q, const float *qc, const float *res, const float *so, int ns,
                             const float *ci, int iic, int jjc, int iif, int jjf, void *st) {
    return api2::interp_add<float>(q, qc, res, so, ns, ci, iic, jjc, iif, jjf, S(st));
}

int cdr_hip_setup_interp2d_f64(const double *so, int ns, double *ci, int iif, int jjf, int iic,
                               int jjc, void *st) {
    return api2::setup_interp(so, ns, ci, iif, jjf, iic, jjc, S(st));
}

int cdr_hip_setup_galerkin2d_f64(const double *so, int ns, const double *ci, double *soc, int iif,
                                 int jjf, int iic, int jjc, void *st) {
    return api2::setup_galerkin(so, ns, ci, soc, iif, jjf, iic, jjc, S(st));
}

int cdr_hip_setup_cg2d_f64(const double *so, int ns, double *abd, int ii, int jj, int *info_dev,
                           void *st) {
    return api2::setup_cg(so, ns, abd, ii, jj, info_dev, S(st));
}

int cdr_hip_solve_cg2d_f64(double *q, const double *qf, const double *abd, double *bbd, int ii,
                           int jj, void *st) {
    return api2::solve_cg<double>(q, qf, abd, bbd, ii, jj, S(st));
}
int cdr_hip_solve_cg2d_f32(float *q, const float *qf, const double *abd, double *bbd, int ii,
                           int jj, void *st) {
    return api2::solve_cg<float>(q, qf, abd, bbd, ii, jj, S(st));
}

// ---- periodic boundaries ------------------------------------------------------------------
int cdr_hip_relax_gs2d_per_f64(const double *so, int ns, const double *qf, double *q,
                               const double *sor, int ii, int jj, int ud, int jpn, void *st) {
    return api2::relax_gs_per<double>(so, ns, qf, q, sor, ii, jj, ud, jpn, S(st));
}
int cdr_hip_relax_gs2d_per_f32(const float *so, int ns, const float *qf, float *q,
                               const float *sor, int ii, int jj, int ud, int jpn, void *st) {
    return api2::relax_gs_per<float>(so, ns, qf, q, sor, ii, jj, ud, jpn, S(st));
}
int cdr_hip_setup_lines_x2d_per_f64(const double *so, double *sor, int ii, int jj, int ns, int jpn,
                                    int *info_dev, void *st) {
    return api2::setup_lines_per(0, so, sor, ii, jj, ns, jpn, info_dev, S(st));
}
int cdr_hip_setup_lines_y2d_per_f64(const double *so, double *sor, int ii, int jj, int ns, int jpn,
                                    int *info_dev, void *st) {
    return api2::setup_lines_per(1, so, sor, ii, jj, ns, jpn, info_dev, S(st));
}
int cdr_hip_relax_lines_x2d_per_f64(const double *so, int ns, const double *qf, double *q,
                                    const double *sor, double *work, int ii, int jj, int ud,
                                    int jpn, void *st) {
    return api2::relax_lines_per<double>(0, so, ns, qf, q, sor, work, ii, jj, ud, jpn, S(st));
}
int cdr_hip_relax_lines_y2d_per_f64(const double *so, int ns, const double *qf, double *q,
                                    const double *sor, double *work, int ii, int jj, int ud,
                                    int jpn, void *st) {
    return api2::relax_lines_per<double>(1, so, ns, qf, q, sor, work, ii, jj, ud, jpn, S(st));
}
int cdr_hip_relax_lines_x2d_per_f32(const float *so, int ns, const float *qf, float *q,
                                    const float *sor, float *work, int ii, int jj, int ud, int jpn,
                                    void *st) {
    return api2::relax_lines_per<float>(0, so, ns, qf, q, sor, work, ii, jj, ud, jpn, S(st));
}
int cdr_hip_relax_lines_y2d_per_f32(const float *so, int ns, const float *qf, float *q,
                                    const float *sor, float *work, int ii, int jj, int ud, int jpn,
                                    void *st) {
    return api2::relax_lines_per<float>(1, so, ns, qf, q, sor, work, ii, jj, ud, jpn, S(st));
}
int cdr_hip_restrict2d_per_f64(const double *ci, double *q, double *qc, int iif, int jjf, int iic,
                               int jjc, int jpn, void *st) {
    return api2::restrict_per<double>(ci, q, qc, iif, jjf, iic, jjc, jpn, S(st));
}
int cdr_hip_restrict2d_per_f32(const float *ci, float *q, float *qc, int iif, int jjf, int iic,
                               int jjc, int jpn, void *st) {
    return api2::restrict_per<float>(ci, q, qc, iif, jjf, iic, jjc, jpn, S(st));
}
int cdr_hip_interp_add2d_per_f64(double *q, const double *qc, const double *res, const double *so,
                                 int ns, const double *ci, int iic, int jjc, int iif, int jjf,
                                 int jpn, void *st) {
    return api2::interp_add_per<double>(q, qc, res, so, ns, ci, iic, jjc, iif, jjf, jpn, S(st));
}
int cdr_hip_interp_add2d_per_f32(float *q, const float *qc, const float *res, const float *so,
                                 int ns, const float *ci, int iic, int jjc, int iif, int jjf,
                                 int jpn, void *st) {
    return api2::interp_add_per<float>(q, qc, res, so, ns, ci, iic, jjc, iif, jjf, jpn, S(st));
}
int cdr_hip_setup_interp2d_per_f64(const double *so, int ns, double *ci, int iif, int jjf, int iic,
                                   int jjc, int jpn, void *st) {
    return api2::setup_interp_per(so, ns, ci, iif, jjf, iic, jjc, jpn, S(st));
}
int cdr_hip_setup_galerkin2d_per_f64(const double *so, int ns, const double *ci, double *soc,
                                     int iif, int jjf, int iic, int jjc, int jpn, void *st) {
    return api2::setup_galerkin_per(so, ns, ci, soc, iif, jjf, iic, jjc, jpn, S(st));
}
int cdr_hip_setup_cg2d_per_f64(const double *so, int ns, double *abd, int ii, int jj, int jpn,
                               int *info_dev, void *st) {
    return api2::setup_cg_per(so, ns, abd, ii, jj, jpn, info_dev, S(st));
}
int cdr_hip_solve_cg2d_per_f64(double *q, const double *qf, const double *abd, double *bbd, int ii,
                               int jj, int jpn, void *st) {
    return api2::solve_cg_per<double>(q, qf, abd, bbd, ii, jj, jpn, S(st));
}
int cdr_hip_solve_cg2d_per_f32(float *q, const float *qf, const double *abd, double *bbd, int ii,
                               int jj, int jpn, void *st) {
    return api2::solve_cg_per<float>(q, qf, abd, bbd, ii, jj, jpn, S(st));
}

}  // extern "C"
