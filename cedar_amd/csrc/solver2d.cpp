// "hip" implementations of the 2D kernel interfaces (registered in the kernel manager) and the
// solver half of the 2D C ABI; the solver itself is in solver2d.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "bmg2_launch.hpp"
#include "capi2d_impl.hpp"
#include "cedar_hip.hpp"
#include "solver2d.hpp"

namespace cedar {
namespace hip {

namespace {

template <typename T>
constexpr bool is_f64 = std::is_same<T, double>::value;

// ------------------------------------------------------------------------------------------
// "hip" implementations: forwards to the checked entry points behind the C ABI, with the batch
// count of the operands' extents
// ------------------------------------------------------------------------------------------
template <typename T>
struct hip_residual : kernels::residual<T> {
    void run(const stencil_op<T> &so, const grid_func<T> &x, const grid_func<T> &b,
             grid_func<T> &r) override {
        check(api2::residual<T>(so.p, so.nstncl, b.p, x.p, r.p, so.ext.ii, so.ext.jj, this->stream, x.ext.nb, so.ext.rep), "residual");
    }
};

template <typename T>
struct hip_rbgs : kernels::point_relax<T> {
    void setup(const stencil_op<double> &so, relax_stencil<double> &sor) override {
        check(api2::setup_recip(so.p, sor.p, so.ext.ii, so.ext.jj, so.nstncl, this->stream, so.ext.nb), "setup_recip");
    }
    void run(const stencil_op<T> &so, grid_func<T> &x, const grid_func<T> &b,
             const relax_stencil<T> &sor, Dir d) override {
        if (this->jpn)
            check(api2::relax_gs_per<T>(so.p, so.nstncl, b.p, x.p, sor.p, so.ext.ii, so.ext.jj, (int)d, this->jpn, this->stream, so.ext.nb), "relax_gs");
        else
            check(api2::relax_gs<T>(so.p, so.nstncl, b.p, x.p, sor.p, so.ext.ii, so.ext.jj, (int)d, this->stream, x.ext.nb, so.ext.rep), "relax_gs");
    }
};

template <typename T, relax_dir RD>
struct hip_lines : kernels::line_relax<T, RD> {
    static constexpr int dir = RD == relax_dir::x ? 0 : 1;
    explicit hip_lines(int *info) : info_(info) {}
    void setup(const stencil_op<double> &so, relax_stencil<double> &sor) override {
        if (this->jpn)
            check(api2::setup_lines_per(dir, so.p, sor.p, so.ext.ii, so.ext.jj, so.nstncl, this->jpn, info_, this->stream, so.ext.nb), "setup_lines");
        else
            check(api2::setup_lines(dir, so.p, sor.p, so.ext.ii, so.ext.jj, so.nstncl, info_, this->stream, so.ext.nb), "setup_lines");
    }
    void run(const stencil_op<T> &so, grid_func<T> &x, const grid_func<T> &b,
             const relax_stencil<T> &sor, grid_func<T> &res, Dir d) override {
        // res is the reference's line scratch; the scan kernels keep their intermediates on chip,
        // the periodic kernel keeps the Sherman-Morrison vector there
        if (this->jpn)
            check(api2::relax_lines_per<T>(dir, so.p, so.nstncl, b.p, x.p, sor.p, res.p, so.ext.ii, so.ext.jj, (int)d, this->jpn, this->stream, so.ext.nb), "relax_lines");
        else
            check(api2::relax_lines<T>(dir, so.p, so.nstncl, b.p, x.p, sor.p, so.ext.ii, so.ext.jj, (int)d, this->stream, x.ext.nb, so.ext.rep), "relax_lines");
    }

private:
    int *info_;
};

struct hip_setup_interp : kernels::setup_interp {
    void run(const stencil_op<double> &fop, const stencil_op<double> &cop,
             prolong_op<double> &P) override {
        P.fine_op = &fop;
        if (this->jpn)
            check(api2::setup_interp_per(fop.p, fop.nstncl, P.p, fop.ext.ii, fop.ext.jj, cop.ext.ii, cop.ext.jj, this->jpn, this->stream, fop.ext.nb), "setup_interp");
        else
            check(api2::setup_interp(fop.p, fop.nstncl, P.p, fop.ext.ii, fop.ext.jj, cop.ext.ii, cop.ext.jj, this->stream, fop.ext.nb), "setup_interp");
    }
};

struct hip_galerkin : kernels::coarsen_op {
    void run(const prolong_op<double> &P, const stencil_op<double> &fop,
             stencil_op<double> &cop) override {
        if (this->jpn)
            check(api2::setup_galerkin_per(fop.p, fop.nstncl, P.p, cop.p, fop.ext.ii, fop.ext.jj, cop.ext.ii, cop.ext.jj, this->jpn, this->stream, fop.ext.nb), "galerkin");
        else
            check(api2::setup_galerkin(fop.p, fop.nstncl, P.p, cop.p, fop.ext.ii, fop.ext.jj, cop.ext.ii, cop.ext.jj, this->stream, fop.ext.nb), "galerkin");
    }
};

template <typename T>
struct hip_restrict : kernels::restriction<T> {
    void run(const prolong_op<T> &R, const grid_func<T> &fine, grid_func<T> &coarse) override {
        // periodic: the images of the fine grid are refreshed in place, as the reference does
        // (src/2d/restrict.cc:23-27 passes the fine grid function as mutable)
        if (this->jpn)
            check(api2::restrict_per<T>(R.p, fine.p, coarse.p, fine.ext.ii, fine.ext.jj, coarse.ext.ii, coarse.ext.jj, this->jpn, this->stream, fine.ext.nb), "restrict");
        else
            check(api2::restrict_<T>(R.p, fine.p, coarse.p, fine.ext.ii, fine.ext.jj, coarse.ext.ii, coarse.ext.jj, this->stream, fine.ext.nb, R.ext.rep), "restrict");
    }
};

template <typename T>
struct hip_interp : kernels::interp_add<T> {
    void run(const prolong_op<T> &P, const grid_func<T> &coarse, const grid_func<T> &residual,
             grid_func<T> &fine) override {
        const stencil_op<T> &fop = *P.fine_op;
        if (this->jpn)
            check(api2::interp_add_per<T>(fine.p, coarse.p, residual.p, fop.p, fop.nstncl, P.p, coarse.ext.ii, coarse.ext.jj, fine.ext.ii, fine.ext.jj, this->jpn, this->stream, fine.ext.nb), "interp_add");
        else
            check(api2::interp_add<T>(fine.p, coarse.p, residual.p, fop.p, fop.nstncl, P.p, coarse.ext.ii, coarse.ext.jj, fine.ext.ii, fine.ext.jj, this->stream, fine.ext.nb, P.ext.rep), "interp_add");
    }
};

template <typename T>
struct hip_solve_cg : kernels::solve_cg<T> {
    explicit hip_solve_cg(int *info) : info_(info) {}
    void setup(const stencil_op<double> &so, grid_func<double> &abd) override {
        (void)abd.ext;
        if (this->jpn)
            check(api2::setup_cg_per(so.p, so.nstncl, abd.p, so.ext.ii, so.ext.jj, this->jpn, info_, this->stream, so.ext.nb), "setup_cg");
        else
            check(api2::setup_cg(so.p, so.nstncl, abd.p, so.ext.ii, so.ext.jj, info_, this->stream, so.ext.nb), "setup_cg");
    }
    void run(grid_func<T> &x, const grid_func<T> &b, const grid_func<double> &abd,
             double *bbd) override {
        if (this->jpn)
            check(api2::solve_cg_per<T>(x.p, b.p, abd.p, bbd, x.ext.ii, x.ext.jj, this->jpn, this->stream, x.ext.nb), "solve_cg");
        else
            check(api2::solve_cg<T>(x.p, b.p, abd.p, bbd, x.ext.ii, x.ext.jj, this->stream, x.ext.nb, abd.ext.rep), "solve_cg");
    }

private:
    int *info_;
};

}  // namespace

void register_hip_kernels(kernel_manager &km, int *info) {
    km.add<kernels::residual<double>, hip_residual<double>>("hip");
    km.add<kernels::residual<float>, hip_residual<float>>("hip");
    km.add<kernels::point_relax<double>, hip_rbgs<double>>("hip");
    km.add<kernels::point_relax<float>, hip_rbgs<float>>("hip");
    km.add<kernels::line_relax<double, relax_dir::x>, hip_lines<double, relax_dir::x>>("hip", info);
    km.add<kernels::line_relax<double, relax_dir::y>, hip_lines<double, relax_dir::y>>("hip", info);
    km.add<kernels::line_relax<float, relax_dir::x>, hip_lines<float, relax_dir::x>>("hip", info);
    km.add<kernels::line_relax<float, relax_dir::y>, hip_lines<float, relax_dir::y>>("hip", info);
    km.add<kernels::setup_interp, hip_setup_interp>("hip");
    km.add<kernels::coarsen_op, hip_galerkin>("hip");
    km.add<kernels::restriction<double>, hip_restrict<double>>("hip");
    km.add<kernels::restriction<float>, hip_restrict<float>>("hip");
    km.add<kernels::interp_add<double>, hip_interp<double>>("hip");
    km.add<kernels::interp_add<float>, hip_interp<float>>("hip");
    km.add<kernels::solve_cg<double>, hip_solve_cg<double>>("hip", info);
    km.add<kernels::solve_cg<float>, hip_solve_cg<float>>("hip", info);
}

}  // namespace hip
}  // namespace cedar

// ==========================================================================================
// C ABI: solver handle
// ==========================================================================================
struct cdr_hip_solver2d {
    std::unique_ptr<cedar::hip::solver2d> impl;
};

namespace {
template <typename F>
int guarded(F f) {
    try {
        return f();
    } catch (const cedar::hip::error &e) {
        std::fprintf(stderr, "cedar_hip: %s\n", e.what());
        return e.status;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "cedar_hip: %s\n", e.what());
        return CDR_ERR_HIP;
    }
}
}  // namespace

extern "C" {

void cdr_hip_default_settings(cdr_solver_settings *s) {
    if (!s) return;
    s->relaxation = CDR_RELAX_POINT;
    s->nrelax_pre = 2;
    s->nrelax_post = 1;
    s->num_levels = -1;
    s->maxiter = 10;
    s->tol = 1e-8;
    s->min_coarse = 3;
    s->precision = CDR_PREC_F64;
    s->graph = 1;
    s->cycle = CDR_CYCLE_V;
    s->fused_coarse = 1;
    s->periodic = 0;
    s->stall_cycles = 0;
}

int cdr_hip_stop_rule(const double *rho, int n, double tol, int stall_cycles, int *ncycles) {
    if ((!rho && n > 0) || n < 0 || !ncycles || !(tol >= 0.0) || stall_cycles < 0)
        return CDR_ERR_ARG;
    cedar::hip::stop_rule rule(tol, stall_cycles);
    for (int i = 0; i < n; ++i) {
        const cedar::hip::stop_rule::verdict v = rule.step(rho[i]);
        if (v != cedar::hip::stop_rule::GO) {
            *ncycles = i + 1;
            return cedar::hip::stop_rule::status(v);
        }
    }
    *ncycles = n;
    return CDR_OK;
}

int cdr_hip_solver2d_create(const double *so_dev, int nstncl, int nx, int ny,
                            const cdr_solver_settings *s, void *stream, cdr_hip_solver2d **out) {
    return cdr_hip_solver2d_create_batched(so_dev, nstncl, nx, ny, 1, s, stream, out);
}

int cdr_hip_solver2d_create_batched(const double *so_dev, int nstncl, int nx, int ny, int nb,
                                    const cdr_solver_settings *s, void *stream,
                                    cdr_hip_solver2d **out) {
    return cdr_hip_solver2d_create_shared(so_dev, nstncl, nx, ny, nb, 1, s, stream, out);
}

int cdr_hip_solver2d_create_shared(const double *so_dev, int nstncl, int nx, int ny, int nops,
                                   int nrhs, const cdr_solver_settings *s, void *stream,
                                   cdr_hip_solver2d **out) {
    const long long nb = (long long)nops * nrhs;
    if (!so_dev || !s || !out || nx < 1 || ny < 1 || nops < 1 || nrhs < 1 || nb > 65535 ||
        (nstncl != 3 && nstncl != 5) ||
        s->relaxation < CDR_RELAX_POINT || s->relaxation > CDR_RELAX_LINE_XY ||
        s->nrelax_pre < 0 || s->nrelax_post < 0 || s->maxiter < 0 || s->min_coarse < 2 ||
        (s->precision != CDR_PREC_F64 && s->precision != CDR_PREC_MIXED) ||
        (s->graph != 0 && s->graph != 1) || (s->cycle != CDR_CYCLE_V && s->cycle != CDR_CYCLE_F) ||
        (s->fused_coarse != 0 && s->fused_coarse != 1) || s->periodic < 0 || s->periodic > 3 ||
        s->stall_cycles < 0)
        return CDR_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CDR_ERR_NO_DEVICE;
    return guarded([&] {
        auto h = std::make_unique<cdr_hip_solver2d>();
        h->impl = std::make_unique<cedar::hip::solver2d>(so_dev, nstncl, nx, ny, nops, *s,
                                                         reinterpret_cast<hipStream_t>(stream),
                                                         nrhs);
        *out = h.release();
        return (int)CDR_OK;
    });
}

int cdr_hip_solver2d_solve(cdr_hip_solver2d *slv, const double *b_dev, double *x_dev,
                           double *hist, int maxhist, int *niter) {
    if (!slv || !b_dev || !x_dev || maxhist < 0) return CDR_ERR_ARG;
    return guarded([&] { return slv->impl->solve(b_dev, x_dev, hist, maxhist, niter); });
}

int cdr_hip_solver2d_vcycle(cdr_hip_solver2d *slv, const double *b_dev, double *x_dev) {
    if (!slv || !b_dev || !x_dev) return CDR_ERR_ARG;
    return guarded([&] {
        slv->impl->vcycle<double>(b_dev, x_dev);
        return (int)CDR_OK;
    });
}

int cdr_hip_solver2d_nlevels(const cdr_hip_solver2d *slv) {
    return slv ? slv->impl->nlevels() : 0;
}

int cdr_hip_solver2d_nbatch(const cdr_hip_solver2d *slv) {
    return slv ? slv->impl->nbatch() : 0;
}

int cdr_hip_solver2d_level(const cdr_hip_solver2d *slv, int lvl, int *nx, int *ny,
                           const double **so, const double **ci, const double **sor0,
                           const double **sor1) {
    if (!slv || lvl < 0 || lvl >= slv->impl->nlevels()) return CDR_ERR_ARG;
    const auto &L = slv->impl->level(lvl);
    if (nx) *nx = L.ext.nx();
    if (ny) *ny = L.ext.ny();
    if (so) *so = L.A.p;
    if (ci) *ci = lvl > 0 ? L.P.p : nullptr;
    if (sor0) *sor0 = L.SOR[0].p;
    if (sor1) *sor1 = L.SOR[1].p;
    return CDR_OK;
}

int cdr_hip_solver2d_times(const cdr_hip_solver2d *slv, double *setup_s, double *solve_s) {
    if (!slv) return CDR_ERR_ARG;
    if (setup_s) *setup_s = slv->impl->setup_seconds();
    if (solve_s) *solve_s = slv->impl->solve_seconds();
    return CDR_OK;
}

int cdr_hip_solver2d_timers(cdr_hip_solver2d *slv, int on) {
    if (!slv || (on != 0 && on != 1)) return CDR_ERR_ARG;
    slv->impl->set_timers(on != 0);
    return CDR_OK;
}

int cdr_hip_solver2d_save_timers(cdr_hip_solver2d *slv, const char *path) {
    if (!slv || !path) return CDR_ERR_ARG;
    return guarded([&] {
        slv->impl->save_timers(path);
        return CDR_OK;
    });
}

int cdr_hip_stencil2d_write(const double *so_dev, int nstncl, int ii, int jj, const char *path) {
    if (!so_dev || !path || (nstncl != 3 && nstncl != 5) || ii < 3 || jj < 3) return CDR_ERR_ARG;
    return guarded([&] {
        std::vector<double> so((std::size_t)nstncl * ii * jj);
        cedar::hip::check_hip(hipMemcpy(so.data(), so_dev, so.size() * sizeof(double),
                                        hipMemcpyDeviceToHost),
                              "hipMemcpy");
        const std::string text = cedar::hip::stencil_text(so.data(), nstncl, ii, jj);
        std::FILE *f = std::fopen(path, "w");
        if (!f) throw cedar::hip::error(CDR_ERR_ARG, std::string("cannot open ") + path);
        std::fwrite(text.data(), 1, text.size(), f);
        std::fclose(f);
        return CDR_OK;
    });
}

int cdr_hip_solver2d_kernels(cdr_hip_solver2d *slv, char *buf, int len) {
    if (!slv || len < 0 || (len > 0 && !buf)) return -CDR_ERR_ARG;
    const std::string d = slv->impl->kernels().describe();
    if (len > 0) {
        const std::size_t n = std::min<std::size_t>(d.size(), (std::size_t)len - 1);
        std::memcpy(buf, d.data(), n);
        buf[n] = '\0';
    }
    return (int)d.size();
}

int cdr_hip_solver2d_set_kernel(cdr_hip_solver2d *slv, const char *iface, const char *impl) {
    if (!slv || !iface || !impl) return CDR_ERR_ARG;
    return guarded([&] {
        slv->impl->kernels().set_by_name(iface, impl);
        return (int)CDR_OK;
    });
}

void cdr_hip_solver2d_destroy(cdr_hip_solver2d *slv) { delete slv; }

}  // extern "C"
