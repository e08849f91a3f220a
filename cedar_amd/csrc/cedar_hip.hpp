// C++ host layer of the MI355X BoxMG path: Cedar's cdr2 programming model on device memory.
//
// Mirrors the reference's types and plugin registry:
//   grid_func / stencil_op / prolong_op / relax_stencil   (include/cedar/2d/*.h)
//   kernels::{residual, point_relax, line_relax, setup_interp, coarsen_op, restriction,
//             interp_add, solve_cg}                        (include/cedar/kernels/*.h)
//   kernel_manager: add<I>(name) / set<I>(name) / setup<I>(...) / run<I>(...)
//                                                           (include/cedar/kernel_manager.h:38-72)
// The "hip" implementations forward to the C ABI in include/cedar_hip.h; the solver in
// solver2d.cpp drives them in Cedar's order (include/cedar/multilevel.h, cycle/vcycle.h).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <typeindex>
#include <utility>
#include <vector>

#include "../../include/cedar_hip.h"

namespace cedar {
namespace hip {

// status -> exception inside the C++ layer; the C ABI converts back to a status code
struct error : std::runtime_error {
    int status;
    explicit error(int s, const std::string &what) : std::runtime_error(what), status(s) {}
};

inline void check(int status, const char *what) {
    if (status != CDR_OK) throw error(status, std::string(what) + ": " + cdr_hip_status_string(status));
}

inline void check_hip(hipError_t e, const char *what) {
    if (e != hipSuccess) throw error(CDR_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// owning device buffer (zero-initialised, like the reference's std::vector storage)
template <typename T>
class dbuf {
public:
    dbuf() = default;
    explicit dbuf(std::size_t n, hipStream_t s = nullptr) : n_(n) {
        if (n_ == 0) return;
        check_hip(hipMalloc(&p_, n_ * sizeof(T)), "hipMalloc");
        check_hip(hipMemsetAsync(p_, 0, n_ * sizeof(T), s), "hipMemsetAsync");
    }
    dbuf(const dbuf &) = delete;
    dbuf &operator=(const dbuf &) = delete;
    dbuf(dbuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    dbuf &operator=(dbuf &&o) noexcept {
        if (this != &o) {
            reset();
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    ~dbuf() { reset(); }
    void reset() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T *data() const { return p_; }
    std::size_t size() const { return n_; }

private:
    T *p_ = nullptr;
    std::size_t n_ = 0;
};

// Text dump of a 2D stencil (host copy, ghost-inclusive ii x jj per plane) in the reference's
// layout (src/2d/stencil_op.cc:5-51): per interior point the running index, i, j (1-based with
// the ghost layer) and the neighbour coefficients as a 3x3 (9-point) or cross (5-point) row,
// off-diagonals with their sign restored, 7 significant digits.
inline std::string stencil_text(const double *so, int nstncl, int ii, int jj) {
    const std::size_t P = (std::size_t)ii * jj;
    auto S = [&](int i, int j, int k) { return so[P * k + (std::size_t)j * ii + i]; };
    std::string out;
    char buf[64];
    auto num = [&](double v) {
        std::snprintf(buf, sizeof(buf), "%.7e", v);
        out += buf;
    };
    for (int j = 1; j <= jj - 2; ++j)
        for (int i = 1; i <= ii - 2; ++i) {
            std::snprintf(buf, sizeof(buf), "%4d %4d, %4d, ", (j - 1) * ii + i - 1, i + 1, j + 1);
            out += buf;
            if (nstncl == 3) {  // c, w, s
                num(-S(i, j + 1, 2));
                num(-S(i, j, 1));
                out += ' ';
                num(S(i, j, 0));
                num(-S(i + 1, j, 1));
                num(-S(i, j, 2));
            } else {  // c, w, s, sw, nw
                num(-S(i, j + 1, 4));
                num(-S(i, j + 1, 2));
                num(-S(i + 1, j + 1, 3));
                num(-S(i, j, 1));
                out += ' ';
                num(S(i, j, 0));
                num(-S(i + 1, j, 1));
                num(-S(i, j, 3));
                num(-S(i, j, 2));
                num(-S(i + 1, j, 4));
            }
            out += '\n';
        }
    return out;
}

enum class Dir { DOWN = CDR_DOWN, UP = CDR_UP };  // cycle::Dir
enum class relax_dir { x, y };

struct five_pt { static constexpr int ndirs = 3; };
struct nine_pt { static constexpr int ndirs = 5; };

// ghost-inclusive extents of a 2D grid quantity; nb problems of these extents are stored back
// to back (a batch), each kernel runs on all of them.  Operator-side quantities (stencils,
// interpolation, relaxation factors, coarse factor) may be shared by groups of rep consecutive
// problems: they then store nb entries for nb * rep problems (rep = 1 for grid functions).
struct extents {
    int ii = 0, jj = 0;
    int nb = 1;
    int rep = 1;
    std::size_t points() const { return (std::size_t)ii * jj; }  // one problem
    std::size_t total() const { return points() * (std::size_t)nb; }
    int nx() const { return ii - 2; }
    int ny() const { return jj - 2; }
};

// Non-owning device views: a level container owns the storage (include/cedar/2d/level_container.h).
template <typename T>
struct grid_func {
    T *p = nullptr;
    extents ext;
    T *data() const { return p; }
    int len(int d) const { return d == 0 ? ext.ii : ext.jj; }
};

template <typename T>
struct stencil_op {
    T *p = nullptr;
    extents ext;
    int nstncl = 5;  // 3: five_pt {c,w,s}, 5: nine_pt {c,w,s,sw,nw}
    T *data() const { return p; }
    int len(int d) const { return d == 0 ? ext.ii : ext.jj; }
};

template <typename T>
struct prolong_op {      // CI on the coarse grid, 8 planes; R = P^T uses the same data
    T *p = nullptr;
    extents ext;         // coarse extents
    const stencil_op<T> *fine_op = nullptr;  // P.fine_op_* (src/2d/interp.cc:61-62,91-92)
    T *data() const { return p; }
};

template <typename T>
struct relax_stencil {   // 2 planes
    T *p = nullptr;
    extents ext;
    T *data() const { return p; }
};

// ------------------------------------------------------------------------------------------
// kernel interfaces (include/cedar/kernels/*.h), templated on the arithmetic type
// ------------------------------------------------------------------------------------------
struct kernel_base {
    virtual ~kernel_base() = default;
    hipStream_t stream = nullptr;
    // boundary id BMG_get_bc(per_mask) of the kernel parameters (include/cedar/kernel_params.h
    // :35-42): 0 Dirichlet, 2 periodic in x, 1 in y, 3 in both
    int jpn = 0;
};

namespace kernels {

template <typename T>
struct residual : kernel_base {
    static const char *name() { return "residual"; }
    virtual void run(const stencil_op<T> &so, const grid_func<T> &x, const grid_func<T> &b,
                     grid_func<T> &r) = 0;
};

template <typename T>
struct point_relax : kernel_base {
    static const char *name() { return "point-relax"; }
    virtual void setup(const stencil_op<double> &so, relax_stencil<double> &sor) = 0;
    virtual void run(const stencil_op<T> &so, grid_func<T> &x, const grid_func<T> &b,
                     const relax_stencil<T> &sor, Dir d) = 0;
};

template <typename T, relax_dir RD>
struct line_relax : kernel_base {
    static const char *name() { return RD == relax_dir::x ? "line-relax-x" : "line-relax-y"; }
    virtual void setup(const stencil_op<double> &so, relax_stencil<double> &sor) = 0;
    virtual void run(const stencil_op<T> &so, grid_func<T> &x, const grid_func<T> &b,
                     const relax_stencil<T> &sor, grid_func<T> &res, Dir d) = 0;
};

struct setup_interp : kernel_base {
    static const char *name() { return "setup-interp"; }
    virtual void run(const stencil_op<double> &fop, const stencil_op<double> &cop,
                     prolong_op<double> &P) = 0;
};

struct coarsen_op : kernel_base {
    static const char *name() { return "coarsen-op"; }
    virtual void run(const prolong_op<double> &P, const stencil_op<double> &fop,
                     stencil_op<double> &cop) = 0;
};

template <typename T>
struct restriction : kernel_base {
    static const char *name() { return "restriction"; }
    virtual void run(const prolong_op<T> &R, const grid_func<T> &fine, grid_func<T> &coarse) = 0;
};

template <typename T>
struct interp_add : kernel_base {
    static const char *name() { return "interp-add"; }
    virtual void run(const prolong_op<T> &P, const grid_func<T> &coarse,
                     const grid_func<T> &residual, grid_func<T> &fine) = 0;
};

template <typename T>
struct solve_cg : kernel_base {
    static const char *name() { return "solve-cg"; }
    // ABD band is FP64 in both precisions (the coarsest solve stays FP64)
    virtual void setup(const stencil_op<double> &so, grid_func<double> &abd) = 0;
    virtual void run(grid_func<T> &x, const grid_func<T> &b, const grid_func<double> &abd,
                     double *bbd) = 0;
};

}  // namespace kernels

// ------------------------------------------------------------------------------------------
// name-keyed registry with one default implementation per interface
// (include/cedar/type_list.h:38-136, include/cedar/kernel_manager.h:11-112)
// ------------------------------------------------------------------------------------------
class kernel_manager {
public:
    explicit kernel_manager(hipStream_t s) : stream_(s) {}

    template <class I, class Impl, class... Args>
    void add(const std::string &name, Args &&...args) {
        auto k = std::make_shared<Impl>(std::forward<Args>(args)...);
        k->stream = stream_;
        auto &slot = table_[std::type_index(typeid(I))];
        slot.iface = I::name();
        slot.impls[name] = k;
        if (slot.def.empty()) slot.def = name;
    }

    // set<I>(name) by the interface's name (I::name()); applies to every registered interface
    // of that name (the FP64 and FP32 forms share it).  Unknown interface or implementation:
    // CDR_ERR_ARG, nothing changed.
    void set_by_name(const std::string &iface, const std::string &name) {
        bool found = false;
        for (auto &t : table_)
            if (t.second.iface == iface) {
                if (!t.second.impls.count(name))
                    throw error(CDR_ERR_ARG, "kernel not found: " + iface + "/" + name);
                found = true;
            }
        if (!found) throw error(CDR_ERR_ARG, "no such kernel interface: " + iface);
        for (auto &t : table_)
            if (t.second.iface == iface) t.second.def = name;
    }

    // {"interface": {"impls": [...], "default": "..."}, ...} over the registered interfaces
    std::string describe() const {
        std::map<std::string, std::pair<std::vector<std::string>, std::string>> by;
        for (const auto &t : table_) {
            auto &e = by[t.second.iface];
            for (const auto &kv : t.second.impls)
                if (std::find(e.first.begin(), e.first.end(), kv.first) == e.first.end())
                    e.first.push_back(kv.first);
            e.second = t.second.def;
        }
        std::string out = "{";
        for (const auto &kv : by) {
            out += (out.size() > 1 ? "," : "") + std::string("\"") + kv.first + "\": {\"impls\": [";
            for (std::size_t i = 0; i < kv.second.first.size(); ++i)
                out += (i ? ",\"" : "\"") + kv.second.first[i] + "\"";
            out += "], \"default\": \"" + kv.second.second + "\"}";
        }
        return out + "}";
    }

    template <class I>
    void set(const std::string &name) {
        auto &slot = table_.at(std::type_index(typeid(I)));
        if (!slot.impls.count(name))
            throw error(CDR_ERR_ARG, std::string("kernel not found: ") + I::name() + "/" + name);
        slot.def = name;
    }

    template <class I>
    I &get() {
        auto it = table_.find(std::type_index(typeid(I)));
        if (it == table_.end() || it->second.def.empty())
            throw error(CDR_ERR_ARG, std::string("kernel not found: ") + I::name());
        return *static_cast<I *>(it->second.impls.at(it->second.def).get());
    }

    template <class I, class... Args>
    void setup(Args &&...args) { get<I>().setup(std::forward<Args>(args)...); }

    template <class I, class... Args>
    void run(Args &&...args) { get<I>().run(std::forward<Args>(args)...); }

    hipStream_t stream() const { return stream_; }

    // sets the boundary id of every registered implementation
    void set_bc(int jpn) {
        for (auto &t : table_)
            for (auto &kv : t.second.impls) kv.second->jpn = jpn;
    }

    // points every registered implementation at stream s (used while capturing a graph)
    void set_stream(hipStream_t s) {
        stream_ = s;
        for (auto &t : table_)
            for (auto &kv : t.second.impls) kv.second->stream = s;
    }

private:
    struct slot_t {
        std::string iface;  // I::name()
        std::map<std::string, std::shared_ptr<kernel_base>> impls;
        std::string def;
    };
    std::map<std::type_index, slot_t> table_;
    hipStream_t stream_;
};

// registers the "hip" implementations of every interface (both precisions)
void register_hip_kernels(kernel_manager &km, int *info_dev);

}  // namespace hip
}  // namespace cedar
