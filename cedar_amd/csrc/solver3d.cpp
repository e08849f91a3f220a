// The cdr3 solver on the GPU (FP64): level hierarchy, setup, V-cycle with point or plane
// relaxation, the outer solve loop, and the 3D solver half of the C ABI.
//
// Plane relaxation (include/cedar/3d/relax_planes.h:35-234, src/3d/relax_planes.cc) relaxes the
// planes of one family in two colours; within a colour the planes are independent, so each
// colour is one batched 2D V-cycle over all its planes (solver2d with nb = number of planes):
// gather x and the plane right-hand sides (copy32, copy_rhs), one cycle, scatter (copy23).
// Each plane keeps its own 2D hierarchy, built at setup from its own operator (copy_coeff), as
// the reference's per-plane cdr2 solvers do.
//
// Control flow: include/cedar/3d/solver.h:17-128 (levels), include/cedar/multilevel.h:95-298
// (setup order, smoother sequencing, solve loop), include/cedar/cycle/vcycle.h:57-115.  The
// solver calls every kernel by interface through its kernel_manager (the "hip" implementations
// below, registered by register_hip_kernels3), as multilevel.h does through kman.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "bmg2_launch.hpp"
#include "bmg3_launch.hpp"
#include "cedar_hip.hpp"
#include "cedar_hip3.hpp"
#include "solver2d.hpp"
#include "stop_rule.hpp"

namespace cedar {
namespace hip {

namespace k3 = cdr::k3;

struct ext3 {
    int ii = 0, jj = 0, kk = 0;
    std::size_t points() const { return (std::size_t)ii * jj * kk; }
};

// The planes of one family (xy, xz or yz) of one level: a batched 2D solver per colour.  In
// mixed precision the plane solvers also hold FP32 hierarchies and the sweep runs in FP32.
class plane_family {
public:
    plane_family(const double *so3, int ns3, ext3 e, int dir, const cdr_solver_settings &s2,
                 hipStream_t st)
        : dir_(dir), ns3_(ns3), e_(e), stream_(st) {
        n1_ = dir == k3::YZ ? e.jj : e.ii;
        n2_ = dir == k3::XY ? e.jj : e.kk;
        const int nplanes = dir == k3::XY ? e.kk - 2 : dir == k3::XZ ? e.jj - 2 : e.ii - 2;
        const int ns2 = ns3 == 14 ? 5 : 3;
        const std::size_t np2 = (std::size_t)n1_ * n2_;
        maxiter_ = s2.maxiter;
        for (int c = 0; c < 2; ++c) {
            const int first = 1 + c;
            const int npl = first > nplanes ? 0 : (nplanes - first) / 2 + 1;
            col_[c].npl = npl;
            if (npl == 0) continue;
            col_[c].so2 = dbuf<double>(np2 * ns2 * npl, st);
            col_[c].x2 = dbuf<double>(np2 * npl, st);
            col_[c].b2 = dbuf<double>(np2 * npl, st);
            if (s2.precision == CDR_PREC_MIXED) {
                col_[c].x2f = dbuf<float>(np2 * npl, st);
                col_[c].b2f = dbuf<float>(np2 * npl, st);
            }
            k3::copy_coeff(dir, so3, ns3, col_[c].so2.data(), first, npl, e.ii, e.jj, e.kk, st);
            check_hip(hipGetLastError(), "copy_coeff");
            col_[c].slv = std::make_unique<solver2d>(col_[c].so2.data(), ns2, n1_ - 2, n2_ - 2, npl,
                                                     s2, st);
        }
    }

    // relax_planes: DOWN relaxes the odd planes (1, 3, ...) first, UP the even ones first
    template <typename T>
    void run(const T *so3, T *x, const T *b, Dir d) {
        for (int t = 0; t < 2; ++t) {
            const int c = d == Dir::DOWN ? t : 1 - t;
            auto &C = col_[c];
            if (C.npl == 0) continue;
            const int first = 1 + c;
            T *x2 = C.template xbuf<T>(), *b2 = C.template bbuf<T>();
            k3::copy32<T>(dir_, x, x2, first, C.npl, e_.ii, e_.jj, e_.kk, stream_);
            k3::copy_rhs<T>(dir_, so3, ns3_, x, b, b2, first, C.npl, e_.ii, e_.jj, e_.kk, stream_);
            check_hip(hipGetLastError(), "plane gather");
            // the nested solve: its residual norms only feed its log and stop test, which with
            // max-iter cycles and no early stop leave x unchanged
            for (int it = 0; it < maxiter_; ++it) C.slv->template vcycle<T>(b2, x2);
            k3::copy23<T>(dir_, x2, x, first, C.npl, e_.ii, e_.jj, e_.kk, stream_);
            check_hip(hipGetLastError(), "plane scatter");
        }
    }

private:
    struct colour {
        int npl = 0;
        dbuf<double> so2, x2, b2;
        dbuf<float> x2f, b2f;
        std::unique_ptr<solver2d> slv;
        template <typename T>
        T *xbuf() {
            if constexpr (std::is_same<T, double>::value)
                return x2.data();
            else
                return x2f.data();
        }
        template <typename T>
        T *bbuf() {
            if constexpr (std::is_same<T, double>::value)
                return b2.data();
            else
                return b2f.data();
        }
    };
    int dir_, ns3_;
    ext3 e_;
    hipStream_t stream_;
    int n1_ = 0, n2_ = 0, maxiter_ = 1;
    colour col_[2];
};

// ------------------------------------------------------------------------------------------
// "hip" implementations of the 3D interfaces
// ------------------------------------------------------------------------------------------
namespace {

inline void launched3(const char *what) { check_hip(hipGetLastError(), what); }

template <typename T>
struct hip_residual3 : kernels3::residual<T> {
    void run(const stencil_op3<T> &so, const grid_func3<T> &x, const grid_func3<T> &b,
             grid_func3<T> &r) override {
        const extents3 e = so.ext;
        k3::residual<T>(so.p, so.nstncl, b.p, x.p, r.p, e.ii, e.jj, e.kk, this->stream);
        launched3("residual3");
    }
};

// jpn: the BoxMG boundary id of the periodic mask (0: Dirichlet); periodic problems run the
// reference's JPN != 0 branches
template <typename T>
struct hip_rbgs3 : kernels3::point_relax<T> {
    explicit hip_rbgs3(int j = 0) : jpn(j) {}
    int jpn;
    void setup(const stencil_op3<double> &so, relax_stencil3<double> &sor) override {
        const extents3 e = so.ext;
        k3::setup_recip(so.p, so.nstncl, sor.p, e.ii, e.jj, e.kk, this->stream);
        launched3("setup_recip3");
    }
    void run(const stencil_op3<T> &so, grid_func3<T> &x, const grid_func3<T> &b,
             const relax_stencil3<T> &sor, Dir d) override {
        const extents3 e = so.ext;
        if (jpn)
            k3::relax_gs_per<T>(so.p, so.nstncl, b.p, x.p, sor.p, e.ii, e.jj, e.kk, (int)d, jpn,
                                this->stream);
        else
            k3::relax_gs<T>(so.p, so.nstncl, b.p, x.p, sor.p, e.ii, e.jj, e.kk, (int)d, this->stream);
        launched3("relax_gs3");
    }
};

// The plane solvers of one family, per level (keyed by the level's extents, which differ from
// level to level; the reference keys them by the first extent, include/cedar/3d/relax_planes.h
// :157-161).  Shared by the FP64 and FP32 registrations of the family: a mixed-precision
// family holds both hierarchies.
struct plane_store {
    cdr_solver_settings s2;
    std::map<std::tuple<int, int, int>, std::unique_ptr<plane_family>> fam;
    plane_family *find(const extents3 &e) {
        const auto it = fam.find(std::make_tuple(e.ii, e.jj, e.kk));
        if (it == fam.end()) throw error(CDR_ERR_ARG, "plane relaxation: level not set up");
        return it->second.get();
    }
};

template <typename T, plane_dir PD>
struct hip_planes : kernels3::plane_relax<T, PD> {
    explicit hip_planes(std::shared_ptr<plane_store> st) : store(std::move(st)) {}
    std::shared_ptr<plane_store> store;
    void setup(const stencil_op3<double> &so) override {
        const extents3 e = so.ext;
        store->fam[std::make_tuple(e.ii, e.jj, e.kk)] = std::make_unique<plane_family>(
            so.p, so.nstncl, ext3{e.ii, e.jj, e.kk}, (int)PD, store->s2, this->stream);
    }
    void run(const stencil_op3<T> &so, grid_func3<T> &x, const grid_func3<T> &b,
             Dir d) override {
        store->find(so.ext)->template run<T>(so.p, x.p, b.p, d);
    }
};

struct hip_setup_interp3 : kernels3::setup_interp {
    explicit hip_setup_interp3(int j = 0) : jpn(j) {}
    int jpn;
    void run(const stencil_op3<double> &fop, const stencil_op3<double> &cop,
             prolong_op3<double> &P) override {
        const extents3 f = fop.ext, c = cop.ext;
        if (jpn)  // the periodic loops write only what the reference writes: start from zero
            k2::fill<double>(P.p, 0.0, (std::size_t)26 * c.ii * c.jj * c.kk, this->stream);
        if (jpn)
            k3::setup_interp_per(fop.p, fop.nstncl, P.p, f.ii, f.jj, f.kk, c.ii, c.jj, c.kk, jpn,
                                 this->stream);
        else
            k3::setup_interp(fop.p, fop.nstncl, P.p, f.ii, f.jj, f.kk, c.ii, c.jj, c.kk,
                             this->stream);
        P.fine_op = &fop;
        launched3("setup_interp3");
    }
};

struct hip_galerkin3 : kernels3::coarsen_op {
    explicit hip_galerkin3(int j = 0) : jpn(j) {}
    int jpn;
    void run(const prolong_op3<double> &P, const stencil_op3<double> &fop,
             stencil_op3<double> &cop) override {
        const extents3 f = fop.ext, c = cop.ext;
        if (jpn)
            k3::setup_galerkin_per(fop.p, fop.nstncl, P.p, cop.p, f.ii, f.jj, f.kk, c.ii, c.jj,
                                   c.kk, jpn, this->stream);
        else
            k3::setup_galerkin(fop.p, fop.nstncl, P.p, cop.p, f.ii, f.jj, f.kk, c.ii, c.jj, c.kk,
                               this->stream);
        launched3("setup_galerkin3");
    }
};

template <typename T>
struct hip_restrict3 : kernels3::restriction<T> {
    explicit hip_restrict3(int j = 0) : jpn(j) {}
    int jpn;
    void run(const prolong_op3<T> &R, const grid_func3<T> &fine, grid_func3<T> &coarse) override {
        const extents3 f = fine.ext, c = coarse.ext;
        if (jpn)  // refreshes the images of the fine input in place, as the reference does
            k3::restrict_per<T>(R.p, fine.p, coarse.p, f.ii, f.jj, f.kk, c.ii, c.jj, c.kk, jpn,
                                this->stream);
        else
            k3::restrict_<T>(R.p, fine.p, coarse.p, f.ii, f.jj, f.kk, c.ii, c.jj, c.kk, this->stream);
        launched3("restrict3");
    }
};

template <typename T>
struct hip_interp3 : kernels3::interp_add<T> {
    explicit hip_interp3(int j = 0) : jpn(j) {}
    int jpn;
    void run(const prolong_op3<T> &P, const grid_func3<T> &coarse, const grid_func3<T> &residual,
             grid_func3<T> &fine) override {
        if (!P.fine_op) throw error(CDR_ERR_ARG, "interp_add: prolongation without fine operator");
        const extents3 f = fine.ext, c = coarse.ext;
        if (jpn)
            k3::interp_add_per<T>(fine.p, coarse.p, residual.p, P.fine_op->p, P.fine_op->nstncl,
                                  P.p, c.ii, c.jj, c.kk, f.ii, f.jj, f.kk, jpn, this->stream);
        else
            k3::interp_add<T>(fine.p, coarse.p, residual.p, P.fine_op->p, P.fine_op->nstncl, P.p,
                              c.ii, c.jj, c.kk, f.ii, f.jj, f.kk, this->stream);
        launched3("interp_add3");
    }
};

template <typename T>
struct hip_solve_cg3 : kernels3::solve_cg<T> {
    explicit hip_solve_cg3(int *info, int j = 0) : info_dev(info), jpn(j) {}
    int *info_dev;
    int jpn;  // periodic: a dense factor (ldab = n), include/cedar/3d/solver.h:123-126
    void setup(const stencil_op3<double> &so, double *abd, int ldab) override {
        const extents3 e = so.ext;
        k2::fill<int>(info_dev, 0, 1, this->stream);
        if (jpn)
            k3::setup_cg_per(so.p, so.nstncl, abd, e.ii, e.jj, e.kk, jpn, info_dev, this->stream);
        else
            k3::setup_cg(so.p, so.nstncl, abd, ldab, e.ii, e.jj, e.kk, info_dev, this->stream);
        launched3("setup_cg3");
        int info = 0;
        check_hip(hipMemcpyAsync(&info, info_dev, sizeof(int), hipMemcpyDeviceToHost,
                                 this->stream), "read info");
        check_hip(hipStreamSynchronize(this->stream), "hipStreamSynchronize");
        if (info != 0) throw error(CDR_ERR_NOT_SPD, "coarse-grid Cholesky failed");
    }
    void run(grid_func3<T> &x, const grid_func3<T> &b, const double *abd, int ldab,
             double *bbd) override {
        const extents3 e = x.ext;
        if (jpn)
            k3::solve_cg_per<T>(x.p, b.p, abd, bbd, e.ii, e.jj, e.kk, jpn, this->stream);
        else
            k3::solve_cg<T>(x.p, b.p, abd, ldab, bbd, e.ii, e.jj, e.kk, this->stream);
        launched3("solve_cg3");
    }
};

template <plane_dir PD>
void add_planes(kernel_manager &km, const cdr_solver_settings &s2) {
    auto st = std::make_shared<plane_store>();
    st->s2 = s2;
    km.add<kernels3::plane_relax<double, PD>, hip_planes<double, PD>>("hip", st);
    km.add<kernels3::plane_relax<float, PD>, hip_planes<float, PD>>("hip", st);
}

}  // namespace

void register_hip_kernels3(kernel_manager &km, int *info_dev, const cdr_solver_settings &plane,
                           int precision, int jpn) {
    km.add<kernels3::residual<double>, hip_residual3<double>>("hip");
    km.add<kernels3::residual<float>, hip_residual3<float>>("hip");
    km.add<kernels3::point_relax<double>, hip_rbgs3<double>>("hip", jpn);
    km.add<kernels3::point_relax<float>, hip_rbgs3<float>>("hip", jpn);
    cdr_solver_settings s2 = plane;
    s2.precision = precision;
    s2.graph = 0;
    add_planes<plane_dir::xy>(km, s2);
    add_planes<plane_dir::xz>(km, s2);
    add_planes<plane_dir::yz>(km, s2);
    km.add<kernels3::setup_interp, hip_setup_interp3>("hip", jpn);
    km.add<kernels3::coarsen_op, hip_galerkin3>("hip", jpn);
    km.add<kernels3::restriction<double>, hip_restrict3<double>>("hip", jpn);
    km.add<kernels3::restriction<float>, hip_restrict3<float>>("hip", jpn);
    km.add<kernels3::interp_add<double>, hip_interp3<double>>("hip", jpn);
    km.add<kernels3::interp_add<float>, hip_interp3<float>>("hip", jpn);
    km.add<kernels3::solve_cg<double>, hip_solve_cg3<double>>("hip", info_dev, jpn);
    km.add<kernels3::solve_cg<float>, hip_solve_cg3<float>>("hip", info_dev, jpn);
}

// one precision's views of a level (device arrays owned by level3)
template <typename T>
struct lvl3v {
    stencil_op3<T> A;
    prolong_op3<T> P;  // on this (coarse) level; fine_op points at the finer level's A
    grid_func3<T> x, b, res;
    relax_stencil3<T> sor;
};

struct level3 {
    ext3 e;
    int ns = 14;
    dbuf<double> so_buf, ci_buf, x_buf, b_buf, res_buf, sor_buf;
    dbuf<float> so_f, ci_f, x_f, b_f, res_f, sor_f;
    lvl3v<double> d;
    lvl3v<float> f;
    template <typename T>
    lvl3v<T> &v() {
        if constexpr (std::is_same<T, double>::value)
            return d;
        else
            return f;
    }
};

static bool is_plane(int r) { return r >= CDR_RELAX_PLANE_XY && r <= CDR_RELAX_PLANE_XYZ; }

class solver3d {
public:
    solver3d(const double *so_dev, int ns, int nx, int ny, int nz, const cdr_solver3_settings &s,
             hipStream_t stream)
        : set_(s), stream_(stream), km_(stream) {
        check_hip(hipEventCreate(&ev0_), "hipEventCreate");
        check_hip(hipEventCreate(&ev1_), "hipEventCreate");
        info_ = dbuf<int>(1, stream_);
        jpn_ = k3::bc_from_mask3(set_.periodic);
        register_hip_kernels3(km_, info_.data(), set_.plane, set_.precision, jpn_);
        build(so_dev, ns, nx, ny, nz);
    }
    ~solver3d() {
        (void)hipEventDestroy(ev0_);
        (void)hipEventDestroy(ev1_);
        if (host_norm_) (void)hipHostFree(host_norm_);
    }

    int nlevels() const { return (int)L_.size(); }
    const level3 &level(int l) const { return L_[l]; }
    double setup_seconds() const { return setup_s_; }
    double solve_seconds() const { return solve_s_; }
    bool mixed() const { return set_.precision == CDR_PREC_MIXED; }
    kernel_manager &kernels() { return km_; }

    // include/cedar/3d/solver.h:59-77
    static int compute_num_levels(int nx, int ny, int nz, int min_coarse) {
        float nxc, nyc, nzc;
        int ng = 0;
        do {
            ng++;
            nxc = (float)((unsigned)(nx - 1) / (unsigned)(1 << ng) + 1);
            nyc = (float)((unsigned)(ny - 1) / (unsigned)(1 << ng) + 1);
            nzc = (float)((unsigned)(nz - 1) / (unsigned)(1 << ng) + 1);
            float m = nxc < nyc ? nxc : nyc;
            if (nzc < m) m = nzc;
            if (!(m >= (float)min_coarse)) break;
        } while (true);
        return ng;
    }

    // one FP64 cycle of the configured type (V, or F: include/cedar/cycle/fcycle.h:38-83)
    void vcycle(const double *b_dev, double *x_dev) { cycle<double>(x_dev, b_dev); }

    // one FP64 sweep of the planes of family dir on level lvl, through the registry
    void relax_planes(int lvl, int dir, double *x, const double *b, Dir d) {
        const int r = set_.relaxation;
        const bool has = r == CDR_RELAX_PLANE_XYZ || (dir == 0 && r == CDR_RELAX_PLANE_XY) ||
                         (dir == 1 && r == CDR_RELAX_PLANE_XZ) ||
                         (dir == 2 && r == CDR_RELAX_PLANE_YZ);
        if (lvl < 0 || lvl >= nlevels() - 1 || dir < 0 || dir > 2 || !has)
            throw error(CDR_ERR_ARG, "no planes of this family on this level");
        level3 &L = L_[lvl];
        grid_func3<double> xg{x, L.d.A.ext}, bg{const_cast<double *>(b), L.d.A.ext};
        switch (dir) {
        case 0: km_.run<kernels3::plane_relax<double, plane_dir::xy>>(L.d.A, xg, bg, d); break;
        case 1: km_.run<kernels3::plane_relax<double, plane_dir::xz>>(L.d.A, xg, bg, d); break;
        default: km_.run<kernels3::plane_relax<double, plane_dir::yz>>(L.d.A, xg, bg, d);
        }
    }

    // multilevel::solve (include/cedar/multilevel.h:278-298); mixed precision: FP64 iterative
    // refinement around the FP32 cycle, x <- x + e and the next FP64 residual in one kernel
    int solve(const double *b_dev, double *x_dev, double *hist, int maxhist, int *niter) {
        level3 &F = L_[0];
        const ext3 e = F.e;
        check_hip(hipEventRecord(ev0_, stream_), "hipEventRecord");
        double *xa = nullptr, *xb = nullptr;
        if (mixed()) {
            xa = xa_.data();
            xb = xb_.data();
            k3::residual_ir(F.d.A.p, F.ns, b_dev, x_dev, nullptr, xa, F.f.b.p, e.ii, e.jj, e.kk,
                            work_.data(), sumsq_.data(), stream_);
            // the outer iterate keeps its periodic images (residual_ir writes the interior)
            if (jpn_) k3::images<double>(xa, e.ii, e.jj, e.kk, jpn_, 0, stream_);
        } else {
            k3::residual_norm(F.d.A.p, F.ns, b_dev, x_dev, F.d.res.p, e.ii, e.jj, e.kk,
                              work_.data(), sumsq_.data(), stream_);
        }
        check_hip(hipGetLastError(), "residual");
        const double res0 = std::sqrt(read_norm2());
        int it = 0, status = CDR_OK;
        stop_rule rule(set_.tol, set_.stall_cycles);
        for (int i = 0; i < set_.maxiter; ++i) {
            if (mixed()) {
                k2::fill<float>(F.f.x.p, 0.0f, e.points(), stream_);
                cycle<float>(F.f.x.p, F.f.b.p);
                k3::residual_ir(F.d.A.p, F.ns, b_dev, xa, F.f.x.p, xb, F.f.b.p, e.ii, e.jj, e.kk,
                                work_.data(), sumsq_.data(), stream_);
                if (jpn_) k3::images<double>(xb, e.ii, e.jj, e.kk, jpn_, 0, stream_);
                std::swap(xa, xb);
            } else {
                cycle<double>(x_dev, b_dev);
                k3::residual_norm(F.d.A.p, F.ns, b_dev, x_dev, F.d.res.p, e.ii, e.jj, e.kk,
                                  work_.data(), sumsq_.data(), stream_);
            }
            check_hip(hipGetLastError(), "residual");
            const double rel = std::sqrt(read_norm2()) / res0;
            if (hist && i < maxhist) hist[i] = rel;
            it = i + 1;
            const stop_rule::verdict v = rule.step(rel);
            if (v != stop_rule::GO) {
                status = stop_rule::status(v);
                break;
            }
        }
        if (mixed())
            check_hip(hipMemcpyAsync(x_dev, xa, e.points() * sizeof(double),
                                     hipMemcpyDeviceToDevice, stream_), "copy x");
        check_hip(hipEventRecord(ev1_, stream_), "hipEventRecord");
        check_hip(hipEventSynchronize(ev1_), "hipEventSynchronize");
        float ms = 0.f;
        check_hip(hipEventElapsedTime(&ms, ev0_, ev1_), "hipEventElapsedTime");
        solve_s_ = ms * 1e-3;
        if (niter) *niter = it;
        return status;
    }

private:
    cdr_solver3_settings set_;
    hipStream_t stream_;
    kernel_manager km_;
    int jpn_ = 0;  // BoxMG boundary id of set_.periodic (BMG_get_bc)
    std::vector<level3> L_;
    dbuf<double> abd_, bbd_, work_, sumsq_, xa_, xb_;
    dbuf<int> info_;
    int ldab_ = 0;
    double *host_norm_ = nullptr;
    hipEvent_t ev0_{}, ev1_{};
    double setup_s_ = 0, solve_s_ = 0;

    double read_norm2() {
        check_hip(hipMemcpyAsync(host_norm_, sumsq_.data(), sizeof(double), hipMemcpyDeviceToHost,
                                 stream_), "read norm");
        check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        return *host_norm_;
    }

    template <typename T>
    static void set_views(lvl3v<T> &v, const ext3 &e, int ns, T *so, T *ci, T *x, T *b, T *res,
                          T *sor) {
        const extents3 x3{e.ii, e.jj, e.kk};
        v.A = stencil_op3<T>{so, x3, ns};
        v.P = prolong_op3<T>{ci, x3, nullptr};
        v.x = grid_func3<T>{x, x3};
        v.b = grid_func3<T>{b, x3};
        v.res = grid_func3<T>{res, x3};
        v.sor = relax_stencil3<T>{sor, x3};
    }

    void build(const double *so_dev, int ns, int nx, int ny, int nz) {
        if (!so_dev || (ns != 4 && ns != 14) || nx < 1 || ny < 1 || nz < 1)
            throw error(CDR_ERR_ARG, "invalid fine operator");
        int nl = compute_num_levels(nx, ny, nz, set_.min_coarse);
        if (set_.num_levels > 0) {
            if (set_.num_levels > nl) throw error(CDR_ERR_LEVELS, "too many levels specified");
            nl = set_.num_levels;
        }
        if (jpn_) {
            // a periodic direction of n points coarsens to n/2: every coarsened level needs even
            // periodic extents (odd ones put two C-points side by side across the seam)
            const int m = k3::per_mask3(jpn_);
            int cx = nx, cy = ny, cz = nz;
            for (int l = 0; l + 1 < nl; ++l) {
                if (((m & 1) && cx % 2) || ((m & 2) && cy % 2) || ((m & 4) && cz % 2))
                    throw error(CDR_ERR_ARG, "periodic extents must be even on every coarsened level");
                cx = (int)((cx - 1) / 2.0 + 1);
                cy = (int)((cy - 1) / 2.0 + 1);
                cz = (int)((cz - 1) / 2.0 + 1);
            }
            if ((long long)cx * cy * cz > 2048)
                throw error(CDR_ERR_ARG, "periodic coarsest grid above 2048 points (dense factor)");
        }
        // setup_space (include/cedar/3d/solver.h:80-128)
        L_.resize(nl);
        int cx = nx, cy = ny, cz = nz;
        for (int l = 0; l < nl; ++l) {
            level3 &L = L_[l];
            L.e = ext3{cx + 2, cy + 2, cz + 2};
            const std::size_t np = L.e.points();
            double *so = nullptr;
            if (l == 0) {
                L.ns = ns;
                so = const_cast<double *>(so_dev);  // held by reference
            } else {
                L.ns = 14;
                L.so_buf = dbuf<double>(np * 14, stream_);
                L.ci_buf = dbuf<double>(np * 26, stream_);
                L.x_buf = dbuf<double>(np, stream_);
                L.b_buf = dbuf<double>(np, stream_);
                so = L.so_buf.data();
            }
            L.res_buf = dbuf<double>(np, stream_);
            if (set_.relaxation == CDR_RELAX_POINT) L.sor_buf = dbuf<double>(np * 2, stream_);
            set_views<double>(L.d, L.e, L.ns, so, L.ci_buf.data(), L.x_buf.data(), L.b_buf.data(),
                              L.res_buf.data(), L.sor_buf.data());
            cx = (int)((cx - 1) / 2.0 + 1);
            cy = (int)((cy - 1) / 2.0 + 1);
            cz = (int)((cz - 1) / 2.0 + 1);
        }
        for (int l = 1; l < nl; ++l) L_[l].d.P.fine_op = &L_[l - 1].d.A;
        const ext3 ec = L_[nl - 1].e;
        const int kd = k3::cg_bandwidth(ec.ii, ec.jj);
        const std::size_t n = (std::size_t)(ec.ii - 2) * (ec.jj - 2) * (ec.kk - 2);
        ldab_ = jpn_ ? (int)n : kd + 1;  // periodic: dense n x n (include/cedar/3d/solver.h:124)
        abd_ = dbuf<double>((std::size_t)ldab_ * n, stream_);
        bbd_ = dbuf<double>(n, stream_);
        work_ = dbuf<double>(k3::norm_partials(L_[0].e.ii, L_[0].e.jj, L_[0].e.kk), stream_);
        sumsq_ = dbuf<double>(1, stream_);
        check_hip(hipHostMalloc(&host_norm_, sizeof(double)), "hipHostMalloc");

        check_hip(hipEventRecord(ev0_, stream_), "hipEventRecord");
        // multilevel::setup (include/cedar/multilevel.h:243-265): interp, RAP, relaxation per
        // level, then the coarse factorisation
        for (int l = 0; l < nl - 1; ++l) {
            level3 &F = L_[l];
            level3 &C = L_[l + 1];
            km_.run<kernels3::setup_interp>(F.d.A, C.d.A, C.d.P);
            km_.run<kernels3::coarsen_op>(C.d.P, F.d.A, C.d.A);
            setup_relax(F);
        }
        km_.setup<kernels3::solve_cg<double>>(L_[nl - 1].d.A, abd_.data(), ldab_);
        if (mixed()) build_f32();
        check_hip(hipEventRecord(ev1_, stream_), "hipEventRecord");
        check_hip(hipEventSynchronize(ev1_), "hipEventSynchronize");
        float ms = 0.f;
        check_hip(hipEventElapsedTime(&ms, ev0_, ev1_), "hipEventElapsedTime");
        setup_s_ = ms * 1e-3;
    }

    // FP32 copy of the hierarchy for the mixed-precision cycle (operators, CI, 1/SO_p); the
    // coarsest band stays FP64
    void build_f32() {
        const std::size_t npf = L_[0].e.points();
        xa_ = dbuf<double>(npf, stream_);
        xb_ = dbuf<double>(npf, stream_);
        for (int l = 0; l < nlevels(); ++l) {
            level3 &L = L_[l];
            const std::size_t np = L.e.points();
            L.so_f = dbuf<float>(np * L.ns, stream_);
            k2::convert<double, float>(L.d.A.p, L.so_f.data(), np * L.ns, stream_);
            if (l > 0) {
                L.ci_f = dbuf<float>(np * 26, stream_);
                k2::convert<double, float>(L.d.P.p, L.ci_f.data(), np * 26, stream_);
            }
            L.x_f = dbuf<float>(np, stream_);
            L.b_f = dbuf<float>(np, stream_);
            L.res_f = dbuf<float>(np, stream_);
            if (L.d.sor.p) {
                L.sor_f = dbuf<float>(np * 2, stream_);
                k2::convert<double, float>(L.d.sor.p, L.sor_f.data(), np * 2, stream_);
            }
            set_views<float>(L.f, L.e, L.ns, L.so_f.data(), L.ci_f.data(), L.x_f.data(),
                             L.b_f.data(), L.res_f.data(), L.sor_f.data());
        }
        for (int l = 1; l < nlevels(); ++l) L_[l].f.P.fine_op = &L_[l - 1].f.A;
        check_hip(hipGetLastError(), "convert");
    }

    // multilevel::setup_relax_helper (include/cedar/multilevel.h:135-157): plane-xyz sets up
    // xy, xz, yz
    void setup_relax(level3 &F) {
        const int r = set_.relaxation;
        if (r == CDR_RELAX_POINT) {
            km_.setup<kernels3::point_relax<double>>(F.d.A, F.d.sor);
            return;
        }
        if (r == CDR_RELAX_PLANE_XY || r == CDR_RELAX_PLANE_XYZ)
            km_.setup<kernels3::plane_relax<double, plane_dir::xy>>(F.d.A);
        if (r == CDR_RELAX_PLANE_XZ || r == CDR_RELAX_PLANE_XYZ)
            km_.setup<kernels3::plane_relax<double, plane_dir::xz>>(F.d.A);
        if (r == CDR_RELAX_PLANE_YZ || r == CDR_RELAX_PLANE_XYZ)
            km_.setup<kernels3::plane_relax<double, plane_dir::yz>>(F.d.A);
    }

    template <typename T, plane_dir PD>
    void planes(level3 &L, grid_func3<T> &x, const grid_func3<T> &b, Dir d) {
        km_.run<kernels3::plane_relax<T, PD>>(L.v<T>().A, x, b, d);
    }

    // smoother sequencing (include/cedar/multilevel.h:165-222)
    template <typename T>
    void smooth(level3 &L, grid_func3<T> &x, const grid_func3<T> &b, Dir d) {
        const int nrep = d == Dir::DOWN ? set_.nrelax_pre : set_.nrelax_post;
        lvl3v<T> &V = L.v<T>();
        for (int rep = 0; rep < nrep; ++rep) {
            switch (set_.relaxation) {
            case CDR_RELAX_POINT: km_.run<kernels3::point_relax<T>>(V.A, x, b, V.sor, d); break;
            case CDR_RELAX_PLANE_XY: planes<T, plane_dir::xy>(L, x, b, d); break;
            case CDR_RELAX_PLANE_XZ: planes<T, plane_dir::xz>(L, x, b, d); break;
            case CDR_RELAX_PLANE_YZ: planes<T, plane_dir::yz>(L, x, b, d); break;
            default:
                if (d == Dir::DOWN) {
                    planes<T, plane_dir::xy>(L, x, b, d);
                    planes<T, plane_dir::yz>(L, x, b, d);
                    planes<T, plane_dir::xz>(L, x, b, d);
                } else {
                    planes<T, plane_dir::xz>(L, x, b, d);
                    planes<T, plane_dir::yz>(L, x, b, d);
                    planes<T, plane_dir::xy>(L, x, b, d);
                }
            }
        }
    }

    template <typename T>
    void coarse_solve(grid_func3<T> &x, const grid_func3<T> &b) {
        km_.run<kernels3::solve_cg<T>>(x, b, abd_.data(), ldab_, bbd_.data());
    }

    template <typename T>
    void cycle(T *x_dev, const T *b_dev) {
        const extents3 e = L_[0].v<T>().A.ext;
        grid_func3<T> x{x_dev, e}, b{const_cast<T *>(b_dev), e};
        if (L_.size() == 1)
            coarse_solve<T>(x, b);
        else if (set_.cycle == CDR_CYCLE_F)
            fmg<T>(0, x, b);
        else
            ncycle<T>(0, x, b);
    }

    // fcycle::fmg_cycle (include/cedar/cycle/fcycle.h:49-83)
    template <typename T>
    void fmg(int l, grid_func3<T> &x, const grid_func3<T> &b) {
        if (l == nlevels() - 1) {
            coarse_solve<T>(x, b);
            return;
        }
        level3 &F = L_[l];
        level3 &C = L_[l + 1];
        lvl3v<T> &fv = F.v<T>();
        lvl3v<T> &cv = C.v<T>();
        if (jpn_) {  // the periodic restriction refreshes its input's images: restrict a copy
            check_hip(hipMemcpyAsync(fv.res.p, b.p, F.e.points() * sizeof(T),
                                     hipMemcpyDeviceToDevice, stream_), "copy b");
            km_.run<kernels3::restriction<T>>(cv.P, fv.res, cv.b);
        } else {
            km_.run<kernels3::restriction<T>>(cv.P, b, cv.b);
        }
        fmg<T>(l + 1, cv.x, cv.b);
        k2::fill<T>(x.p, T(0), F.e.points(), stream_);
        k2::fill<T>(fv.res.p, T(0), F.e.points(), stream_);
        km_.run<kernels3::interp_add<T>>(cv.P, cv.x, fv.res, x);
        ncycle<T>(l, x, b);
    }

    // vcycle::ncycle_helper (include/cedar/cycle/vcycle.h:57-115)
    template <typename T>
    void ncycle(int l, grid_func3<T> &x, const grid_func3<T> &b) {
        level3 &F = L_[l];
        level3 &C = L_[l + 1];
        lvl3v<T> &fv = F.v<T>();
        lvl3v<T> &cv = C.v<T>();
        smooth<T>(F, x, b, Dir::DOWN);
        km_.run<kernels3::residual<T>>(fv.A, x, b, fv.res);
        km_.run<kernels3::restriction<T>>(cv.P, fv.res, cv.b);
        k2::fill<T>(cv.x.p, T(0), C.e.points(), stream_);
        if (l + 1 == nlevels() - 1)
            coarse_solve<T>(cv.x, cv.b);
        else
            ncycle<T>(l + 1, cv.x, cv.b);
        km_.run<kernels3::interp_add<T>>(cv.P, cv.x, fv.res, x);
        smooth<T>(F, x, b, Dir::UP);
    }
};

}  // namespace hip
}  // namespace cedar

// ==========================================================================================
// C ABI: 3D solver handle
// ==========================================================================================
struct cdr_hip_solver3d {
    std::unique_ptr<cedar::hip::solver3d> impl;
};

namespace {
template <typename F>
int guarded3(F f) {
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

void cdr_hip_default_settings3(cdr_solver3_settings *s) {
    if (!s) return;
    s->relaxation = CDR_RELAX_POINT;
    s->nrelax_pre = 2;
    s->nrelax_post = 1;
    s->num_levels = -1;
    s->maxiter = 10;
    s->tol = 1e-8;
    s->min_coarse = 3;
    s->cycle = CDR_CYCLE_V;
    s->precision = CDR_PREC_F64;
    s->stall_cycles = 0;
    s->periodic = 0;
    // plane-config defaults (src/kernel_params.cc:83-89): line-xy, max-iter 1
    cdr_hip_default_settings(&s->plane);
    s->plane.relaxation = CDR_RELAX_LINE_XY;
    s->plane.maxiter = 1;
    s->plane.graph = 0;
}

int cdr_hip_solver3d_create(const double *so_dev, int nstncl, int nx, int ny, int nz,
                            const cdr_solver3_settings *s, void *stream, cdr_hip_solver3d **out) {
    if (!so_dev || !s || !out || nx < 1 || ny < 1 || nz < 1 || (nstncl != 4 && nstncl != 14) ||
        !(s->relaxation == CDR_RELAX_POINT || cedar::hip::is_plane(s->relaxation)) ||
        s->nrelax_pre < 0 || s->nrelax_post < 0 || s->maxiter < 0 || s->min_coarse < 2 ||
        s->plane.relaxation < CDR_RELAX_POINT || s->plane.relaxation > CDR_RELAX_LINE_XY ||
        s->plane.maxiter < 0 || s->plane.nrelax_pre < 0 || s->plane.nrelax_post < 0 ||
        s->plane.min_coarse < 2 || (s->cycle != CDR_CYCLE_V && s->cycle != CDR_CYCLE_F) ||
        (s->plane.fused_coarse != 0 && s->plane.fused_coarse != 1) ||
        (s->precision != CDR_PREC_F64 && s->precision != CDR_PREC_MIXED) || s->stall_cycles < 0 ||
        s->periodic < 0 || s->periodic > 7 || (s->periodic != 0 && s->relaxation != CDR_RELAX_POINT))
        return CDR_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CDR_ERR_NO_DEVICE;
    return guarded3([&] {
        auto h = std::make_unique<cdr_hip_solver3d>();
        h->impl = std::make_unique<cedar::hip::solver3d>(so_dev, nstncl, nx, ny, nz, *s,
                                                         reinterpret_cast<hipStream_t>(stream));
        *out = h.release();
        return (int)CDR_OK;
    });
}

int cdr_hip_solver3d_solve(cdr_hip_solver3d *slv, const double *b_dev, double *x_dev,
                           double *hist, int maxhist, int *niter) {
    if (!slv || !b_dev || !x_dev || maxhist < 0) return CDR_ERR_ARG;
    return guarded3([&] { return slv->impl->solve(b_dev, x_dev, hist, maxhist, niter); });
}

int cdr_hip_solver3d_vcycle(cdr_hip_solver3d *slv, const double *b_dev, double *x_dev) {
    if (!slv || !b_dev || !x_dev) return CDR_ERR_ARG;
    return guarded3([&] {
        slv->impl->vcycle(b_dev, x_dev);
        return (int)CDR_OK;
    });
}

int cdr_hip_solver3d_relax_planes(cdr_hip_solver3d *slv, int lvl, int family, double *x_dev,
                                  const double *b_dev, int updown) {
    if (!slv || !b_dev || !x_dev || (updown != CDR_DOWN && updown != CDR_UP)) return CDR_ERR_ARG;
    return guarded3([&] {
        slv->impl->relax_planes(lvl, family, x_dev, b_dev,
                                updown == CDR_UP ? cedar::hip::Dir::UP : cedar::hip::Dir::DOWN);
        return (int)CDR_OK;
    });
}

int cdr_hip_solver3d_nlevels(const cdr_hip_solver3d *slv) {
    return slv ? slv->impl->nlevels() : 0;
}

int cdr_hip_solver3d_level(const cdr_hip_solver3d *slv, int lvl, int *nx, int *ny, int *nz,
                           const double **so, const double **ci) {
    if (!slv || lvl < 0 || lvl >= slv->impl->nlevels()) return CDR_ERR_ARG;
    const auto &L = slv->impl->level(lvl);
    if (nx) *nx = L.e.ii - 2;
    if (ny) *ny = L.e.jj - 2;
    if (nz) *nz = L.e.kk - 2;
    if (so) *so = L.d.A.p;
    if (ci) *ci = L.d.P.p;
    return CDR_OK;
}

int cdr_hip_solver3d_times(const cdr_hip_solver3d *slv, double *setup_s, double *solve_s) {
    if (!slv) return CDR_ERR_ARG;
    if (setup_s) *setup_s = slv->impl->setup_seconds();
    if (solve_s) *solve_s = slv->impl->solve_seconds();
    return CDR_OK;
}

int cdr_hip_solver3d_kernels(cdr_hip_solver3d *slv, char *buf, int len) {
    if (!slv || len < 0 || (len > 0 && !buf)) return -CDR_ERR_ARG;
    const std::string d = slv->impl->kernels().describe();
    if (len > 0) {
        const std::size_t n = std::min<std::size_t>(d.size(), (std::size_t)len - 1);
        std::memcpy(buf, d.data(), n);
        buf[n] = '\0';
    }
    return (int)d.size();
}

int cdr_hip_solver3d_set_kernel(cdr_hip_solver3d *slv, const char *iface, const char *impl) {
    if (!slv || !iface || !impl) return CDR_ERR_ARG;
    return guarded3([&] {
        slv->impl->kernels().set_by_name(iface, impl);
        return (int)CDR_OK;
    });
}

void cdr_hip_solver3d_destroy(cdr_hip_solver3d *slv) { delete slv; }

}  // extern "C"
