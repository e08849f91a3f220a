// The cdr2 solver on the GPU (class solver2d): level hierarchy, setup, V-cycle and the outer
// solve loop (FP64, or an FP32 V-cycle preconditioning FP64 iterative refinement), for one
// problem or a batch of problems of the same extents.  Used by the 2D C ABI (solver2d.cpp) and
// by the 3D plane relaxation, whose plane solves are batches of 2D problems (solver3d.cpp).
//
// Control flow follows Cedar: multilevel::setup (include/cedar/multilevel.h:243-265),
// multilevel::solve (include/cedar/multilevel.h:278-298), vcycle::ncycle_helper
// (include/cedar/cycle/vcycle.h:57-115), smoother sequencing (include/cedar/multilevel.h:165-222),
// level sizes (include/cedar/2d/solver.h:63-123).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

#include "bmg2_launch.hpp"
#include "capi2d_impl.hpp"
#include "cedar_hip.hpp"
#include "stop_rule.hpp"

namespace cedar {
namespace hip {

namespace k2 = cdr::k2;
namespace api2 = cdr::api2;

// ------------------------------------------------------------------------------------------
// level storage (one precision)
// ------------------------------------------------------------------------------------------
template <typename T>
struct level_data {
    extents ext;
    dbuf<T> so_buf, ci_buf, x_buf, b_buf, res_buf, sor_buf[2];
    // line-xy smoothing: the operator of the transposed grid and transposed copies of x and b,
    // for y-line sweeps run as x-line sweeps of the transposed grid (empty unless line-xy)
    dbuf<T> soT_buf, xT_buf, bT_buf;
    stencil_op<T> A;
    prolong_op<T> P;  // stored on this (coarse) level; unused on level 0
    grid_func<T> x, b, res;
    relax_stencil<T> SOR[2];
};

class solver2d {
public:
    // nb operators of the same extents (so_dev holds them back to back), each with rep
    // right-hand sides: nb * rep problems, problem b using operator b / rep
    solver2d(const double *so_dev, int nstncl, int nx, int ny, int nb,
             const cdr_solver_settings &s, hipStream_t stream, int rep = 1)
        : set_(s), stream_(stream), cur_(stream), km_(stream) {
        check_hip(hipEventCreate(&ev0_), "hipEventCreate");
        check_hip(hipEventCreate(&ev1_), "hipEventCreate");
        info_ = dbuf<int>(1, stream_);
        register_hip_kernels(km_, info_.data());
        // BMG_get_bc (src/2d/ftn/BMG_get_bc.f90): per_mask 1 (x) -> 2, 2 (y) -> 1, 3 -> 3
        static const int bc_of_mask[4] = {0, 2, 1, 3};
        if (s.periodic < 0 || s.periodic > 3) throw error(CDR_ERR_ARG, "invalid periodic mask");
        jpn_ = bc_of_mask[s.periodic];
        km_.set_bc(jpn_);
        build(so_dev, nstncl, nx, ny, nb, rep);
    }

    ~solver2d() {
        for (auto *g : {&g64_, &g32_})
            if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (cap_stream_) (void)hipStreamDestroy(cap_stream_);
        (void)hipEventDestroy(ev0_);
        (void)hipEventDestroy(ev1_);
        if (host_norm_) (void)hipHostFree(host_norm_);
        if (host_active_) (void)hipHostFree(host_active_);
    }

    int nlevels() const { return (int)L64_.size(); }
    int nbatch() const { return L64_[0].ext.nb; }  // problems (operators x right-hand sides)
    int nrep() const { return rep_; }
    const level_data<double> &level(int l) const { return L64_[l]; }
    double setup_seconds() const { return setup_s_; }
    double solve_seconds() const { return solve_s_; }
    kernel_manager &kernels() { return km_; }

    // include/cedar/2d/solver.h:63-79: float coarse sizes, unsigned division
    static int compute_num_levels(int nx, int ny, int min_coarse) {
        float nxc, nyc;
        int ng = 0;
        do {
            ng++;
            nxc = (float)((unsigned)(nx - 1) / (unsigned)(1 << ng) + 1);
            nyc = (float)((unsigned)(ny - 1) / (unsigned)(1 << ng) + 1);
        } while ((nxc < nyc ? nxc : nyc) >= (float)min_coarse);
        return ng;
    }

    // one V-cycle on x (all problems of the batch), in the precision of T; the FP32 hierarchy
    // exists only for mixed-precision solvers
    template <typename T>
    void vcycle(const T *b_dev, T *x_dev) {
        auto &L = levels<T>();
        if (L.empty()) throw error(CDR_ERR_ARG, "no hierarchy in this precision");
        grid_func<T> x{x_dev, L[0].ext}, b{const_cast<T *>(b_dev), L[0].ext};
        run_cycle<T>(L, x, b);
    }

    template <typename T>
    std::vector<level_data<T>> &levels() {
        if constexpr (std::is_same<T, double>::value)
            return L64_;
        else
            return L32_;
    }

    // kernels go to `s` from now on (the caller's stream, or a capture stream)
    void set_stream(hipStream_t s) {
        km_.set_stream(s);
        cur_ = s;
    }

    // Cedar's solve loop (include/cedar/multilevel.h:278-298) for every problem of the batch:
    // problem b records rho_i in hist[b * maxhist + i] and its cycle count in niter[b], and stops
    // (keeps its iterate) once rho_i < tol, exactly as if it were solved alone.
    int solve(const double *b_dev, double *x_dev, double *hist, int maxhist, int *niter) {
        const extents e = L64_[0].ext;
        const int nb = e.nb;
        const stencil_op<double> &A = L64_[0].A;
        check_hip(hipEventRecord(ev0_, stream_), "hipEventRecord");
        std::vector<double> res0(nb);
        std::vector<int> it(nb, 0);
        int status = CDR_OK, nactive = nb;
        for (int p = 0; p < nb; ++p) host_active_[p] = 1;
        bool mask_dirty = false;
        // per-problem bookkeeping after a norm; returns false when every problem has stopped
        std::vector<stop_rule> rule(nb, stop_rule(set_.tol, set_.stall_cycles));
        auto record = [&](int i) {
            for (int p = 0; p < nb; ++p) {
                if (!host_active_[p]) continue;
                const double rel = std::sqrt(host_norm_[p]) / res0[p];
                if (i < maxhist && hist) hist[(std::size_t)p * maxhist + i] = rel;
                it[p] = i + 1;
                const stop_rule::verdict v = rule[p].step(rel);
                if (v == stop_rule::GO) continue;
                // a non-finite norm outranks a stall in the returned status
                const int st = stop_rule::status(v);
                if (st == CDR_ERR_NONFINITE || (st != CDR_OK && status == CDR_OK)) status = st;
                host_active_[p] = 0;
                --nactive;
                mask_dirty = true;
            }
            return nactive > 0;
        };
        auto push_mask = [&] {
            if (!mask_dirty) return;
            check_hip(hipMemcpyAsync(active_.data(), host_active_, nb * sizeof(int),
                                     hipMemcpyHostToDevice, stream_), "copy mask");
            mask_dirty = false;
        };
        if (set_.precision == CDR_PREC_F64) {
            grid_func<double> x{x_dev, e}, b{const_cast<double *>(b_dev), e};
            // b is fixed through the solve: its transpose for the finest level's y sweeps is made
            // once here, and every cycle starts with it in place
            double *bT0 = bT_target<double>(L64_[0]);
            if (bT0) k2::transpose<double>(b.p, bT0, e.jj, e.ii, stream_, nb);
            check(api2::residual_norm(A.p, A.nstncl, b.p, x.p, L64_[0].res.p, e.ii, e.jj,
                                      work_.data(), sumsq_.data(), stream_, nb, rep_),
                  "residual_norm");
            read_norms();
            for (int p = 0; p < nb; ++p) res0[p] = std::sqrt(host_norm_[p]);
            for (int i = 0; i < set_.maxiter; ++i) {
                const bool masked = nactive < nb;
                if (masked) {
                    // problems that have stopped keep their iterate through this cycle (xprev_
                    // is allocated with the hierarchy whenever nb > 1)
                    check_hip(hipMemcpyAsync(xprev_.data(), x.p, e.total() * sizeof(double),
                                             hipMemcpyDeviceToDevice, stream_), "save x");
                    push_mask();
                }
                run_maybe_graph(g64_, x.p, b.p,
                                [&] { run_cycle<double>(L64_, x, b, bT0 != nullptr); });
                if (masked) k2::keep_inactive(x.p, xprev_.data(), active_.data(), e.points(), nb, stream_);
                check(api2::residual_norm(A.p, A.nstncl, b.p, x.p, L64_[0].res.p, e.ii, e.jj,
                                          work_.data(), sumsq_.data(), stream_, nb, rep_),
                      "residual_norm");
                read_norms();
                if (!record(i)) break;
            }
        } else {
            // FP64 iterative refinement around an FP32 V-cycle:
            //   r = b - A x (FP64), e = Vcycle32(r), x <- x + e, fused into one residual pass.
            double *xa = xa_.data(), *xb = xb_.data();
            level_data<float> &F = L32_[0];
            // the outer step also writes the FP32 right-hand side transposed where the finest
            // level's y sweeps run through the transposed grid
            float *r32T = bT_target<float>(F);
            k2::residual_ir(A.p, A.nstncl, b_dev, x_dev, nullptr, xa, F.b.p, e.ii, e.jj,
                            work_.data(), sumsq_.data(), stream_, nb, nullptr, rep_, r32T);
            check_hip(hipGetLastError(), "residual_ir");
            periodic_images(xa, e, nb);
            read_norms();
            for (int p = 0; p < nb; ++p) res0[p] = std::sqrt(host_norm_[p]);
            const bool zs = zero_start_ok();
            for (int i = 0; i < set_.maxiter; ++i) {
                run_maybe_graph(g32_, F.x.p, F.b.p, [&] {
                    if (!zs) k2::fill<float>(F.x.p, 0.0f, e.total(), cur_);
                    run_cycle<float>(L32_, F.x, F.b, r32T != nullptr, zs);
                });
                push_mask();
                k2::residual_ir(A.p, A.nstncl, b_dev, xa, F.x.p, xb, F.b.p, e.ii, e.jj,
                                work_.data(), sumsq_.data(), stream_, nb,
                                nactive < nb ? active_.data() : nullptr, rep_, r32T);
                check_hip(hipGetLastError(), "residual_ir");
                std::swap(xa, xb);
                periodic_images(xa, e, nb);
                read_norms();
                if (!record(i)) break;
            }
            check_hip(hipMemcpyAsync(x_dev, xa, e.total() * sizeof(double),
                                     hipMemcpyDeviceToDevice, stream_), "copy x");
        }
        check_hip(hipEventRecord(ev1_, stream_), "hipEventRecord");
        check_hip(hipEventSynchronize(ev1_), "hipEventSynchronize");
        float ms = 0.f;
        check_hip(hipEventElapsedTime(&ms, ev0_, ev1_), "hipEventElapsedTime");
        solve_s_ = ms * 1e-3;
        log_time(0, "solve", solve_s_);
        if (niter)
            for (int p = 0; p < nb; ++p) niter[p] = it[p];
        return status;
    }

private:
    // Periodic boundaries: the FP64 iterate of the mixed-precision loop is updated in the interior
    // (x += e), so its periodic images are refreshed after each update from the interior values;
    // the residual of the next step then reads x + e at the images, the image of the updated
    // iterate.  (The FP64 cycle keeps its images through the periodic kernels themselves.)
    void periodic_images(double *x, const extents &e, int nb) {
        if (jpn_ == 0) return;
        k2::periodic_images<double>(x, e.ii, e.jj, jpn_, stream_, nb);
        check_hip(hipGetLastError(), "periodic_images");
    }

    int jpn_ = 0;
    int rep_ = 1;  // right-hand sides per operator
    cdr_solver_settings set_;
    hipStream_t stream_;
    hipStream_t cur_;                  // stream the cycle kernels go to (capture stream while capturing)
    hipStream_t cap_stream_ = nullptr;
    kernel_manager km_;
    struct graph_t {
        hipGraphExec_t exec = nullptr;
        const void *k1 = nullptr, *k2 = nullptr;  // the x / b buffers the graph was captured on
    } g64_, g32_;
    std::vector<level_data<double>> L64_;
    std::vector<level_data<float>> L32_;
    dbuf<double> abd_, bbd_, work_, sumsq_, xa_, xb_, xprev_;
    dbuf<int> info_, active_;
    grid_func<double> abd_g_;
    double *host_norm_ = nullptr;  // nb squared norms (pinned)
    // fused coarse-level cycle: first level (-1: none), largest side, device level descriptors
    int fuse_lvl_ = -1, fuse_side_ = 0;
    // CEDAR_AMD_ZERO_START=0 at construction turns the zero start of the FP32 cycles off (the
    // correction is then zero-filled before each cycle; the results are the same bit for bit)
    bool zs_env_ = [] {
        const char *v = std::getenv("CEDAR_AMD_ZERO_START");
        return !(v && v[0] == '0');
    }();
    std::size_t fuse_lds_ = 0;  // LDS bytes of the on-chip fused cycle (0: arrays in memory)
    int fuse_lds_mode_ = 0;     // its lds_mode (k2::coarse_cycle)
    dbuf<unsigned char> fuse64_, fuse32_;
    int *host_active_ = nullptr;   // nb flags (pinned)
    hipEvent_t ev0_{}, ev1_{};
    double setup_s_ = 0, solve_s_ = 0;

    // Enqueues body() on stream_, either directly or by replaying a graph captured from it.  The
    // graph is captured on a private stream (capture is not allowed on the null stream) and
    // re-captured when the buffers it was captured on change.  Captured cycles contain kernel
    // nodes only (fills are kernels, no memset nodes).
    template <typename F>
    void run_maybe_graph(graph_t &g, const void *k1, const void *k2, F body) {
        if (!set_.graph || timing_) {
            body();
            return;
        }
        if (!g.exec || g.k1 != k1 || g.k2 != k2) {
            if (g.exec) {
                check_hip(hipGraphExecDestroy(g.exec), "hipGraphExecDestroy");
                g.exec = nullptr;
            }
            if (!cap_stream_)
                check_hip(hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking),
                          "hipStreamCreateWithFlags");
            check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
            set_stream(cap_stream_);
            check_hip(hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeThreadLocal),
                      "hipStreamBeginCapture");
            hipGraph_t graph = nullptr;
            try {
                body();
            } catch (...) {
                (void)hipStreamEndCapture(cap_stream_, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                set_stream(stream_);
                throw;
            }
            check_hip(hipStreamEndCapture(cap_stream_, &graph), "hipStreamEndCapture");
            set_stream(stream_);
            const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            check_hip(e, "hipGraphInstantiate");
            g.k1 = k1;
            g.k2 = k2;
        }
        check_hip(hipGraphLaunch(g.exec, stream_), "hipGraphLaunch");
    }

    void read_norms() {
        check_hip(hipMemcpyAsync(host_norm_, sumsq_.data(), sumsq_.size() * sizeof(double),
                                 hipMemcpyDeviceToHost, stream_), "read norm");
        check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    }

    template <typename T>
    // e: extents of the level's grid functions (e.nb problems); the operator-side arrays hold
    // e.nb / rep entries, one per operator
    static void alloc_level(level_data<T> &L, extents e, int nstncl, bool own_op, bool need_ci,
                            bool need_xb, hipStream_t s, int rep) {
        L.ext = e;
        const std::size_t np = e.total();
        extents eo = e;
        eo.nb = e.nb / rep;
        eo.rep = rep;
        const std::size_t npo = eo.total();
        if (own_op) {
            L.so_buf = dbuf<T>(npo * (std::size_t)nstncl, s);
            L.A = stencil_op<T>{L.so_buf.data(), eo, nstncl};
        }
        if (need_ci) {
            L.ci_buf = dbuf<T>(npo * 8, s);
            L.P = prolong_op<T>{L.ci_buf.data(), eo, nullptr};
        }
        if (need_xb) {
            L.x_buf = dbuf<T>(np, s);
            L.b_buf = dbuf<T>(np, s);
            L.x = grid_func<T>{L.x_buf.data(), e};
            L.b = grid_func<T>{L.b_buf.data(), e};
        }
        L.res_buf = dbuf<T>(np, s);
        L.res = grid_func<T>{L.res_buf.data(), e};
        for (int k = 0; k < 2; ++k) {
            L.sor_buf[k] = dbuf<T>(npo * 2, s);
            L.SOR[k] = relax_stencil<T>{L.sor_buf[k].data(), eo};
        }
    }

    void build(const double *so_dev, int nstncl, int nx, int ny, int nops, int rep) {
        if (!so_dev || (nstncl != 3 && nstncl != 5) || nx < 1 || ny < 1 || nops < 1 || rep < 1 ||
            (long long)nops * rep > 65535)
            throw error(CDR_ERR_ARG, "invalid fine operator");
        if (rep > 1 && jpn_ != 0)
            throw error(CDR_ERR_ARG, "shared operators are not supported with periodic boundaries");
        rep_ = rep;
        const int nb = nops * rep;  // problems
        int nl = compute_num_levels(nx, ny, set_.min_coarse);
        if (set_.num_levels > 0) {
            if (set_.num_levels > nl) throw error(CDR_ERR_LEVELS, "too many levels specified");
            nl = set_.num_levels;
        }
        // setup_space (include/cedar/2d/solver.h:81-123)
        L64_.resize(nl);
        int cx = nx, cy = ny;
        for (int l = 0; l < nl; ++l) {
            extents e{cx + 2, cy + 2, nb};
            if (l == 0) {
                alloc_level(L64_[0], e, nstncl, false, false, false, stream_, rep);
                extents eo{cx + 2, cy + 2, nops, rep};
                L64_[0].A = stencil_op<double>{const_cast<double *>(so_dev), eo, nstncl};
            } else {
                alloc_level(L64_[l], e, 5, true, true, true, stream_, rep);
            }
            cx = (int)((cx - 1) / 2.0 + 1);
            cy = (int)((cy - 1) / 2.0 + 1);
        }
        const extents ec = L64_[nl - 1].ext;
        // the coarse matrix: band ABD(ii, n), or dense ABD(n, n) when periodic
        // (include/cedar/2d/solver.h:116-121)
        const std::size_t ncg = (std::size_t)ec.nx() * ec.ny();
        abd_ = dbuf<double>((jpn_ ? ncg : (std::size_t)ec.ii) * ncg * nops, stream_);
        bbd_ = dbuf<double>((std::size_t)ec.nx() * ec.ny() * nb, stream_);
        work_ = dbuf<double>(cdr_hip_norm_work2d(L64_[0].ext.ii, L64_[0].ext.jj) * nb, stream_);
        sumsq_ = dbuf<double>(nb, stream_);
        active_ = dbuf<int>(nb, stream_);
        // the FP64 loop of a batch saves the iterate of a cycle in which some problems have
        // stopped, so that they keep theirs; allocated here, solve() allocates nothing
        if (nb > 1 && set_.precision == CDR_PREC_F64)
            xprev_ = dbuf<double>(L64_[0].ext.total(), stream_);
        check_hip(hipHostMalloc(&host_norm_, nb * sizeof(double)), "hipHostMalloc");
        check_hip(hipHostMalloc(&host_active_, nb * sizeof(int)), "hipHostMalloc");

        check_hip(hipEventRecord(ev0_, stream_), "hipEventRecord");
        // setup: interp -> Galerkin -> relaxation per level, then the coarse factorisation
        for (int l = 0; l < nl - 1; ++l) {
            auto &F = L64_[l];
            auto &C = L64_[l + 1];
            km_.run<kernels::setup_interp>(F.A, C.A, C.P);
            km_.run<kernels::coarsen_op>(C.P, F.A, C.A);
            setup_relax(F);
        }
        km_.setup<kernels::solve_cg<double>>(L64_[nl - 1].A, abd_grid());

        if (set_.precision == CDR_PREC_MIXED) {
            build_f32();
            tune_lines(L32_);
            setup_fused(L32_, fuse32_);
        } else {
            tune_lines(L64_);
            setup_fused(L64_, fuse64_);
        }
        check_hip(hipEventRecord(ev1_, stream_), "hipEventRecord");
        check_hip(hipEventSynchronize(ev1_), "hipEventSynchronize");
        float ms = 0.f;
        check_hip(hipEventElapsedTime(&ms, ev0_, ev1_), "hipEventElapsedTime");
        setup_s_ = ms * 1e-3;
        log_time(0, "setup", setup_s_);
    }

    grid_func<double> &abd_grid() {
        const extents ec = L64_.back().ext;
        const int ncg = ec.nx() * ec.ny();
        abd_g_ = grid_func<double>{abd_.data(), extents{jpn_ ? ncg : ec.ii, ncg, ec.nb / rep_, rep_}};
        return abd_g_;
    }

    // The fused coarse-level cycle starts at the first level below the finest whose sides are at
    // most 8, 16, 32, 64 or 128 points and covers it and every coarser level (it needs at least
    // two of them).  One workgroup per problem is latency-bound on larger levels, where one launch
    // per step over the whole GPU is faster (and a batch of fewer problems than the GPU has
    // compute units leaves most of it idle); where the candidate starting levels differ, the
    // cycle from the highest of them down is timed with each (captured graphs, replayed) and the
    // fastest is kept.  The fused cycle is bitwise the unfused one, so the choice never changes a
    // result.
    template <typename T>
    void setup_fused(std::vector<level_data<T>> &L, dbuf<unsigned char> &dst) {
        fuse_lvl_ = -1;
        if (!set_.fused_coarse || jpn_ != 0) return;  // the fused cycle has no periodic images
        const int nl = (int)L.size();
        std::vector<int> cand;
        for (int side : {8, 16, 32, 64, 128}) {
            for (int l = 1; l + 1 < nl; ++l)
                if (std::max(L[l].ext.nx(), L[l].ext.ny()) <= side) {
                    if (std::find(cand.begin(), cand.end(), l) == cand.end()) cand.push_back(l);
                    break;
                }
        }
        if (cand.empty()) return;
        int lt = cand[0];
        if (cand.size() > 1 && k2::autotune_enabled()) {
            const int top = *std::min_element(cand.begin(), cand.end());
            float tbest = -1.f;
            for (int c : cand) {
                build_fused<T>(L, c, dst);
                k2::fill<T>(L[top].b.p, T(0), L[top].ext.total(), stream_);
                // the faster of two timings (one timing alone let the noise pick among
                // near-equal candidates)
                float t = -1.f;
                for (int round = 0; round < 2; ++round) {
                    const float tr = time_graph([&] { ncycle<T>(L, top, L[top].x, L[top].b); });
                    if (tr >= 0.f && (t < 0.f || tr < t)) t = tr;
                }
                if (t >= 0.f && (tbest < 0.f || t < tbest)) {
                    tbest = t;
                    lt = c;
                }
            }
        }
        build_fused<T>(L, lt, dst);
    }

    // level descriptors of the fused cycle from level lt down (fuse_lvl_ = lt)
    template <typename T>
    void build_fused(std::vector<level_data<T>> &L, int lt, dbuf<unsigned char> &dst) {
        const int nl = (int)L.size();
        std::vector<k2::coarse_level<T>> v;
        for (int l = lt; l < nl; ++l) {
            const level_data<T> &D = L[l];
            v.push_back(k2::coarse_level<T>{D.A.p, D.x.p, D.b.p, D.res.p, D.SOR[0].p, D.SOR[1].p,
                                            l == lt ? nullptr : D.P.p, D.ext.ii, D.ext.jj});
        }
        const std::size_t bytes = v.size() * sizeof(k2::coarse_level<T>);
        dst = dbuf<unsigned char>(bytes, stream_);
        check_hip(hipMemcpyAsync(dst.data(), v.data(), bytes, hipMemcpyHostToDevice, stream_),
                  "copy level descriptors");
        check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        fuse_lvl_ = lt;
        fuse_side_ = std::max(L[lt].ext.nx(), L[lt].ext.ny());
        // the fused levels run out of LDS where all their arrays fit in one workgroup's share
        std::vector<int> ii, jj;
        for (int l = lt; l < nl; ++l) {
            ii.push_back(L[l].ext.ii);
            jj.push_back(L[l].ext.jj);
        }
        // every array on chip where they fit, else all but the first level's read-only arrays,
        // else all but everything of the first level except its x, else none
        fuse_lds_mode_ = 0;
        fuse_lds_ = 0;
        for (int mode = 1; mode <= 3 && !fuse_lds_mode_; ++mode) {
            const std::size_t need = k2::fused_cycle_lds_bytes((int)sizeof(T), ii.data(),
                                                               jj.data(), nl - lt, mode);
            if (need <= kFusedLdsMax) {
                fuse_lds_mode_ = mode;
                fuse_lds_ = need;
            }
        }
    }
    // of the 160 KiB per workgroup; the kernel's static 8 KiB staging area of the coarsest
    // solve comes on top
    static constexpr std::size_t kFusedLdsMax = 150 * 1024;

    // Milliseconds per replay of what body() launches, captured as a graph on the capture stream
    // and replayed 3 times on the solver's stream after a warm-up; -1 if capture fails.
    template <typename F>
    float time_graph(F body) {
        if (!cap_stream_)
            check_hip(hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking),
                      "hipStreamCreateWithFlags");
        check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        set_stream(cap_stream_);
        if (hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            set_stream(stream_);
            (void)hipGetLastError();
            return -1.f;
        }
        hipGraph_t graph = nullptr;
        try {
            body();
        } catch (...) {
            (void)hipStreamEndCapture(cap_stream_, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            set_stream(stream_);
            throw;
        }
        const hipError_t ec = hipStreamEndCapture(cap_stream_, &graph);
        set_stream(stream_);
        if (ec != hipSuccess || !graph) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();
            return -1.f;
        }
        hipGraphExec_t exec = nullptr;
        float ms = -1.f;
        if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
            hipEvent_t a, b;
            check_hip(hipEventCreate(&a), "hipEventCreate");
            check_hip(hipEventCreate(&b), "hipEventCreate");
            check_hip(hipGraphLaunch(exec, stream_), "hipGraphLaunch");
            check_hip(hipEventRecord(a, stream_), "hipEventRecord");
            for (int r = 0; r < 3; ++r) check_hip(hipGraphLaunch(exec, stream_), "hipGraphLaunch");
            check_hip(hipEventRecord(b, stream_), "hipEventRecord");
            check_hip(hipEventSynchronize(b), "hipEventSynchronize");
            check_hip(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
            ms /= 3.f;
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            (void)hipGraphExecDestroy(exec);
        }
        (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return ms;
    }

    // Picks the line-kernel launch shapes of every smoothing level by timing (the shapes give
    // identical results); a no-op for point relaxation and for shapes already recorded.
    template <typename T>
    void tune_lines(std::vector<level_data<T>> &L) {
        const int r = set_.relaxation;
        if (r != CDR_RELAX_LINE_X && r != CDR_RELAX_LINE_Y && r != CDR_RELAX_LINE_XY) return;
        if (jpn_ != 0) return;  // periodic lines run one lane per line, nothing to tune
        const std::size_t np = L[0].ext.total();
        dbuf<T> q(np, stream_), f(np, stream_);
        k2::fill<T>(q.data(), T(0), np, stream_);
        k2::fill<T>(f.data(), T(0), np, stream_);
        for (std::size_t l = 0; l + 1 < L.size(); ++l) {
            level_data<T> &D = L[l];
            for (int dir = 0; dir < 2; ++dir) {
                if ((dir == 0 && r == CDR_RELAX_LINE_Y) || (dir == 1 && r == CDR_RELAX_LINE_X)) continue;
                const relax_stencil<T> &S = D.SOR[r == CDR_RELAX_LINE_XY ? dir : 0];
                if (k2::tune_lines<T>(dir, D.A.p, D.A.nstncl, f.data(), q.data(), S.p, D.ext.ii,
                                      D.ext.jj, stream_, D.ext.nb, rep_) < 0)
                    throw error(CDR_ERR_HIP, "line-kernel tuning failed");
            }
            if (r != CDR_RELAX_LINE_XY) continue;
            // y-lines through the transposed grid: kept only where it is faster
            const std::size_t nl = D.ext.total();
            D.soT_buf = dbuf<T>(D.A.ext.total() * (std::size_t)D.A.nstncl, stream_);
            D.xT_buf = dbuf<T>(nl, stream_);
            D.bT_buf = dbuf<T>(nl, stream_);
            k2::transpose_op<T>(D.A.p, D.soT_buf.data(), D.A.nstncl, D.ext.ii, D.ext.jj, stream_,
                                D.A.ext.nb);
            const int c = k2::tune_lines_ytr<T>(D.A.p, D.soT_buf.data(), D.A.nstncl, f.data(),
                                                q.data(), D.SOR[1].p, D.xT_buf.data(),
                                                D.bT_buf.data(), D.ext.ii, D.ext.jj, stream_,
                                                D.ext.nb, set_.nrelax_pre + set_.nrelax_post,
                                                rep_);
            if (c == -2) throw error(CDR_ERR_HIP, "line-kernel tuning failed");
            if (c < 0) {  // direct y-lines: the transposed workspace is not needed
                D.soT_buf = dbuf<T>();
                D.xT_buf = dbuf<T>();
                D.bT_buf = dbuf<T>();
            }
        }
    }

    void setup_relax(level_data<double> &F) {
        switch (set_.relaxation) {
        case CDR_RELAX_POINT: km_.setup<kernels::point_relax<double>>(F.A, F.SOR[0]); break;
        case CDR_RELAX_LINE_X: km_.setup<kernels::line_relax<double, relax_dir::x>>(F.A, F.SOR[0]); break;
        case CDR_RELAX_LINE_Y: km_.setup<kernels::line_relax<double, relax_dir::y>>(F.A, F.SOR[0]); break;
        case CDR_RELAX_LINE_XY:
            km_.setup<kernels::line_relax<double, relax_dir::x>>(F.A, F.SOR[0]);
            km_.setup<kernels::line_relax<double, relax_dir::y>>(F.A, F.SOR[1]);
            break;
        default: throw error(CDR_ERR_ARG, "invalid relaxation");
        }
    }

    // FP32 copy of the hierarchy (operators, interpolation, smoother factors)
    void build_f32() {
        const int nl = nlevels();
        L32_.resize(nl);
        xa_ = dbuf<double>(L64_[0].ext.total(), stream_);
        xb_ = dbuf<double>(L64_[0].ext.total(), stream_);
        for (int l = 0; l < nl; ++l) {
            auto &D = L64_[l];
            auto &F = L32_[l];
            alloc_level(F, D.ext, D.A.nstncl, true, l > 0, true, stream_, rep_);
            const std::size_t npo = D.A.ext.total();  // operator-side entries
            k2::convert<double, float>(D.A.p, F.A.p, npo * D.A.nstncl, stream_);
            if (l > 0) k2::convert<double, float>(D.P.p, F.P.p, npo * 8, stream_);
            for (int k = 0; k < 2; ++k) k2::convert<double, float>(D.SOR[k].p, F.SOR[k].p, npo * 2, stream_);
        }
        for (int l = 1; l < nl; ++l) L32_[l].P.fine_op = &L32_[l - 1].A;
        check_hip(hipGetLastError(), "convert");
    }

    // bT_ok: the level's transposed right-hand side (L.bT_buf) already holds b; set when a y sweep
    // through the transposed grid has transposed it (b does not change within one visit of a
    // level, so the DOWN and UP smoothing of a visit share it).
    // xT (optional): the iterate lives in L.xT_buf, transposed, and x itself is stale.  Going in,
    // a y sweep through the transposed grid then starts from L.xT_buf; going out, a DOWN
    // smoothing whose last sweep is such a y sweep leaves the iterate there (the residual and
    // interp_add of the visit read it in that layout) and an UP smoothing always ends with x
    // current.  Without xT every sweep leaves x current.
    // zero_start: x is zero on entry to this (DOWN) smoothing, and its first sweep is an x sweep
    // that reads x only there (see zero_start_ok)
    template <typename T>
    void smooth(level_data<T> &L, grid_func<T> &x, const grid_func<T> &b, Dir d, bool &bT_ok,
                bool *xT = nullptr, bool zero_start = false) {
        const int nrep = d == Dir::DOWN ? set_.nrelax_pre : set_.nrelax_post;
        for (int r = 0; r < nrep; ++r) {
            switch (set_.relaxation) {
            case CDR_RELAX_POINT: km_.run<kernels::point_relax<T>>(L.A, x, b, L.SOR[0], d); break;
            case CDR_RELAX_LINE_X: km_.run<kernels::line_relax<T, relax_dir::x>>(L.A, x, b, L.SOR[0], L.res, d); break;
            case CDR_RELAX_LINE_Y: km_.run<kernels::line_relax<T, relax_dir::y>>(L.A, x, b, L.SOR[0], L.res, d); break;
            default:
                // x then y going down, y then x going up (include/cedar/multilevel.h:175-178,
                // 204-207); the y sweep runs through the transposed grid where that was faster
                for (int half = 0; half < 2; ++half) {
                    const bool xs = (half == 0) == (d == Dir::DOWN);
                    if (jpn_ != 0) {  // periodic: one lane per line, L.res as line scratch
                        check(api2::relax_lines_per<T>(xs ? 0 : 1, L.A.p, L.A.nstncl, b.p, x.p,
                                                       L.SOR[xs ? 0 : 1].p, L.res.p, L.ext.ii,
                                                       L.ext.jj, (int)d, jpn_, cur_, L.ext.nb),
                              "relax_lines_per");
                    } else if (xs) {
                        if (xT && *xT) {  // the iterate back from the transposed layout
                            k2::transpose<T>(L.xT_buf.data(), x.p, L.ext.ii, L.ext.jj, cur_,
                                             L.ext.nb);
                            *xT = false;
                        }
                        if (zero_start && r == 0 && half == 0) {
                            k2::relax_lines_x<T>(L.A.p, L.A.nstncl, b.p, x.p, L.SOR[0].p,
                                                 L.ext.ii, L.ext.jj, (int)d, cur_, L.ext.nb, rep_,
                                                 1);
                            check_hip(hipGetLastError(), "relax_lines_x");
                        } else {
                            check(api2::relax_lines<T>(0, L.A.p, L.A.nstncl, b.p, x.p,
                                                       L.SOR[0].p, L.ext.ii, L.ext.jj, (int)d,
                                                       cur_, L.ext.nb, rep_),
                                  "relax_lines_x");
                        }
                    } else {
                        // the last sweep of a DOWN smoothing may leave the iterate transposed
                        const bool last = r == nrep - 1 && half == 1;
                        const bool keep = xT && d == Dir::DOWN && last;
                        const bool in_T = xT && *xT;
                        bT_ok = k2::relax_lines_y_auto<T>(
                            L.A.p, L.soT_buf.data(), L.A.nstncl, b.p, x.p, L.SOR[1].p,
                            L.xT_buf.data(), L.bT_buf.data(), L.ext.ii, L.ext.jj, (int)d, cur_,
                            L.ext.nb, bT_ok, in_T, keep, rep_);
                        if (xT) *xT = bT_ok && keep;
                        check_hip(hipGetLastError(), "relax_lines_y");
                    }
                }
            }
        }
    }

    // The FP32 V-cycle of the mixed solve starts from a zero correction.  Instead of a fill of
    // the finest level's iterate before each cycle, the first colour of its first x sweep takes
    // the zeros as literal operands (bitwise the same sweep).  That covers every read of the old
    // iterate when: line-xy relaxation (the first sweep is an x sweep, red-black by lines, and
    // the first colour reads only lines of the other colour and the boundary rows), at least one
    // pre-smoothing sweep, a V-cycle, non-periodic, the finest level not inside the fused coarse
    // cycle, and odd extents (interp_add then writes no boundary point, so the boundary of the
    // iterate keeps the zeros of its allocation; every sweep writes interior points only).
    bool zero_start_ok() const {
        if (!zs_env_ || set_.relaxation != CDR_RELAX_LINE_XY || jpn_ != 0 ||
            set_.cycle == CDR_CYCLE_F || set_.nrelax_pre < 1 || L32_.size() < 2)
            return false;
        if (fuse_lvl_ == 0 && !timing_ && fuse32_.data()) return false;
        const extents &e = L32_[0].ext;
        return (e.ii - 2) % 2 == 1 && (e.jj - 2) % 2 == 1 && k2::lines_zero_start(e.ii);
    }

    // bT_ready: the finest level's transposed right-hand side (L[0].bT_buf) already holds b
    // zero_start: x is zero on entry (zero_start_ok() holds; the caller skips the fill)
    template <typename T>
    void run_cycle(std::vector<level_data<T>> &L, grid_func<T> &x, const grid_func<T> &b,
                   bool bT_ready = false, bool zero_start = false) {
        if (L.size() == 1) {
            km_.run<kernels::solve_cg<T>>(x, b, abd_grid(), bbd_.data());
            return;
        }
        if (set_.cycle == CDR_CYCLE_F)
            fmg<T>(L, 0, x, b);
        else
            ncycle<T>(L, 0, x, b, bT_ready, zero_start);
    }

    // the level's transposed right-hand side buffer when its y sweeps run through the transposed
    // grid (so that a producer of b can write it there as well), else null
    template <typename T>
    T *bT_target(level_data<T> &D) {
        if (jpn_ != 0 || set_.relaxation != CDR_RELAX_LINE_XY || !D.soT_buf.data()) return nullptr;
        return k2::uses_transposed_y<T>(D.soT_buf.data(), D.A.nstncl, D.ext.ii, D.ext.jj,
                                        D.ext.nb) >= 0
                   ? D.bT_buf.data()
                   : nullptr;
    }

    // fcycle::fmg_cycle (include/cedar/cycle/fcycle.h:49-83): restrict b itself, recurse, then
    // x = P x_coarse (x and res zeroed first) and one V-cycle from this level.  x is overwritten,
    // so every F-cycle of a solve starts from zero, as in the reference.
    template <typename T>
    void fmg(std::vector<level_data<T>> &L, int lvl, grid_func<T> &x, const grid_func<T> &b) {
        if (lvl == (int)L.size() - 1) {
            km_.run<kernels::solve_cg<T>>(x, b, abd_grid(), bbd_.data());
            return;
        }
        auto &F = L[lvl];
        auto &C = L[lvl + 1];
        if (jpn_ != 0 && lvl == 0) {
            // the periodic restriction refreshes the images of its input in place; the caller's
            // b is restricted through a copy (same values, b left untouched)
            k2::copy<T>(b.p, F.res.p, F.ext.total(), cur_);
            km_.run<kernels::restriction<T>>(C.P, F.res, C.b);
        } else {
            km_.run<kernels::restriction<T>>(C.P, b, C.b);
        }
        fmg<T>(L, lvl + 1, C.x, C.b);
        k2::fill<T>(x.p, T(0), F.ext.total(), cur_);
        k2::fill<T>(F.res.p, T(0), F.ext.total(), cur_);
        km_.run<kernels::interp_add<T>>(C.P, C.x, F.res, x);
        ncycle<T>(L, lvl, x, b);
    }

    // bT_ready: L[lvl].bT_buf already holds the transpose of b (its producer wrote it)
    template <typename T>
    void ncycle(std::vector<level_data<T>> &L, int lvl, grid_func<T> &x, const grid_func<T> &b,
                bool bT_ready = false, bool zero_start = false) {
        if (lvl == fuse_lvl_ && !timing_ && x.p == L[lvl].x.p && b.p == L[lvl].b.p) {
            // this level and all coarser ones in one launch (bitwise the same V-cycle)
            const dbuf<unsigned char> &d = std::is_same<T, double>::value ? fuse64_ : fuse32_;
            if (d.data()) {
                k2::coarse_cycle<T>(reinterpret_cast<const k2::coarse_level<T> *>(d.data()),
                                    (int)L.size() - lvl, fuse_side_, set_.relaxation,
                                    set_.nrelax_pre, set_.nrelax_post, abd_.data(), bbd_.data(),
                                    cur_, L[lvl].ext.nb, rep_, fuse_lds_, fuse_lds_mode_);
                check_hip(hipGetLastError(), "coarse_cycle");
                return;
            }
        }
        auto &F = L[lvl];
        auto &C = L[lvl + 1];
        bool bT_ok = bT_ready;  // b is fixed for this visit of the level
        // xT: between the smoothings the iterate may stay in the transposed layout of the y
        // sweeps (F.xT_buf), read there by the residual and updated there by interp_add
        bool xT = false;
        // timer labels and levels as in include/cedar/cycle/vcycle.h:61-109
        timed(lvl, "relaxation", [&] { smooth<T>(F, x, b, Dir::DOWN, bT_ok, &xT, zero_start); });
        timed(lvl, "residual", [&] {
            if (xT)
                k2::residual_qT<T>(F.A.p, F.A.nstncl, b.p, F.xT_buf.data(), F.res.p, F.ext.ii,
                                   F.ext.jj, cur_, F.ext.nb, rep_);
            else
                km_.run<kernels::residual<T>>(F.A, x, b, F.res);
        });
        // restriction, the coarse iterate's zero start and (where the coarse level's y sweeps run
        // through the transposed grid) the transposed coarse right-hand side in one launch
        bool cbT = false;
        timed(lvl, "restrict", [&] {
            if (jpn_ == 0) {
                T *qcT = (lvl + 1 < (int)L.size() - 1 && lvl + 1 != fuse_lvl_) ? bT_target<T>(C)
                                                                                : nullptr;
                k2::restrict_cycle<T>(C.P.p, F.res.p, C.b.p, C.x.p, qcT, F.ext.ii, F.ext.jj,
                                      C.ext.ii, C.ext.jj, cur_, F.ext.nb, rep_);
                check_hip(hipGetLastError(), "restrict");
                cbT = qcT != nullptr;
            } else {
                km_.run<kernels::restriction<T>>(C.P, F.res, C.b);
                k2::fill<T>(C.x.p, T(0), C.ext.total(), cur_);
            }
        });
        if (lvl + 1 == (int)L.size() - 1)
            timed(lvl + 1, "coarse-solve", [&] {
                km_.run<kernels::solve_cg<T>>(C.x, C.b, abd_grid(), bbd_.data());
            });
        else
            ncycle<T>(L, lvl + 1, C.x, C.b, cbT);
        timed(lvl, "interp-add", [&] {
            if (xT)
                k2::interp_add_qT<T>(F.xT_buf.data(), C.x.p, F.res.p, F.A.p, F.A.nstncl, C.P.p,
                                     C.ext.ii, C.ext.jj, F.ext.ii, F.ext.jj, cur_, F.ext.nb,
                                     rep_);
            else
                km_.run<kernels::interp_add<T>>(C.P, C.x, F.res, x);
        });
        timed(lvl, "relaxation", [&] { smooth<T>(F, x, b, Dir::UP, bT_ok, &xT); });
        if (xT) {  // an UP smoothing without x sweeps: the iterate back to its own layout
            k2::transpose<T>(F.xT_buf.data(), x.p, F.ext.ii, F.ext.jj, cur_, F.ext.nb);
            xT = false;
        }
    }

public:
    // ---- timers (include/cedar/util/time_log.h) -------------------------------------------
    // With timers on, cycles run eagerly (no graph replay, no fused coarse levels) and every
    // phase of the V-cycle is bracketed by events on the solver's stream; save_timers writes the
    // accumulated device times per level and label in the reference's serial JSON layout
    // (src/util/time_log.cc:103-129) and resets them.
    void set_timers(bool on) { timing_ = on; }

    void save_timers(const std::string &path) {
        collect_timers();
        std::ostringstream os;
        os << "{\n    \"levels\": [\n";
        for (int i = (int)tlog_.size() - 1; i >= 0; --i) {
            os << "        {";
            bool first = true;
            for (const auto &kv : tlog_[i]) {
                os << (first ? "\n" : ",\n") << "            \"" << kv.first << "\": {\n"
                   << "                \"time\": \"" << fmt(kv.second.first) << "\",\n"
                   << "                \"count\": \"" << kv.second.second << "\"\n"
                   << "            }";
                first = false;
            }
            os << (first ? "}" : "\n        }") << (i > 0 ? ",\n" : "\n");
        }
        os << "    ]\n}\n";
        std::ofstream f(path);
        if (!f) throw error(CDR_ERR_ARG, "cannot open " + path);
        f << os.str();
        tlog_.clear();
    }

private:
    struct timed_phase {
        int level;
        const char *label;
        hipEvent_t a, b;
    };
    bool timing_ = false;
    std::vector<timed_phase> phases_;
    std::vector<std::map<std::string, std::pair<double, int>>> tlog_;  // level -> label -> (s, n)

    template <typename F>
    void timed(int level, const char *label, F &&f) {
        if (!timing_) {
            f();
            return;
        }
        timed_phase t{level, label, nullptr, nullptr};
        check_hip(hipEventCreate(&t.a), "hipEventCreate");
        check_hip(hipEventCreate(&t.b), "hipEventCreate");
        check_hip(hipEventRecord(t.a, cur_), "hipEventRecord");
        f();
        check_hip(hipEventRecord(t.b, cur_), "hipEventRecord");
        phases_.push_back(t);
    }

    void log_time(int level, const char *label, double seconds) {
        if ((int)tlog_.size() <= level) tlog_.resize(level + 1);
        auto &e = tlog_[level][label];
        e.first += seconds;
        e.second += 1;
    }

    void collect_timers() {
        for (auto &t : phases_) {
            float ms = 0.f;
            if (hipEventSynchronize(t.b) == hipSuccess &&
                hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess)
                log_time(t.level, t.label, ms * 1e-3);
            (void)hipEventDestroy(t.a);
            (void)hipEventDestroy(t.b);
        }
        phases_.clear();
    }

    static std::string fmt(double v) {
        char buf[40];
        std::snprintf(buf, sizeof(buf), "%.17g", v);
        return buf;
    }
};

}  // namespace hip
}  // namespace cedar
