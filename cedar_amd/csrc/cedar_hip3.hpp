// 3D half of the C++ host layer: Cedar's cdr3 objects on device memory and the 3D kernel
// interfaces of the plugin registry.
//
// Mirrors include/cedar/3d/{grid_func,stencil_op,prolong_op,relax_stencil}.h and the serial
// kernel list ser_kernels (include/cedar/kernel_registry.h:23-35) for 3D: residual, point_relax,
// plane_relax<xy|xz|yz> (include/cedar/kernels/plane_relax.h:11-31), setup_interp, coarsen_op,
// restriction, interp_add and solve_cg, each templated on the arithmetic type of the cycle.
// They are registered in the same kernel_manager as the 2D interfaces (cedar_hip.hpp), whose
// table is keyed by interface type; the "hip" implementations (solver3d.cpp) forward to the 3D
// launchers behind include/cedar_hip.h.
#pragma once

#include "cedar_hip.hpp"

namespace cedar {
namespace hip {

enum class plane_dir { xy = 0, xz = 1, yz = 2 };  // relax_dir::{xy, xz, yz}

// ghost-inclusive extents of a 3D grid quantity (ii = nx+2, jj = ny+2, kk = nz+2)
struct extents3 {
    int ii = 0, jj = 0, kk = 0;
    std::size_t points() const { return (std::size_t)ii * jj * kk; }
    int nx() const { return ii - 2; }
    int ny() const { return jj - 2; }
    int nz() const { return kk - 2; }
};

template <typename T>
struct grid_func3 {
    T *p = nullptr;
    extents3 ext;
};

template <typename T>
struct stencil_op3 {     // 4 planes (seven_pt {p, pw, ps, b}) or 14 (xxvii_pt)
    T *p = nullptr;
    extents3 ext;
    int nstncl = 14;
};

template <typename T>
struct prolong_op3 {     // CI on the coarse grid, 26 planes; R = P^T uses the same data
    T *p = nullptr;
    extents3 ext;        // coarse extents
    const stencil_op3<T> *fine_op = nullptr;
};

template <typename T>
struct relax_stencil3 {  // 2 planes, 1/SO_p in the second
    T *p = nullptr;
    extents3 ext;
};

namespace kernels3 {

template <typename T>
struct residual : kernel_base {
    static const char *name() { return "residual (3D)"; }
    virtual void run(const stencil_op3<T> &so, const grid_func3<T> &x, const grid_func3<T> &b,
                     grid_func3<T> &r) = 0;
};

template <typename T>
struct point_relax : kernel_base {
    static const char *name() { return "point-relax (3D)"; }
    virtual void setup(const stencil_op3<double> &so, relax_stencil3<double> &sor) = 0;
    virtual void run(const stencil_op3<T> &so, grid_func3<T> &x, const grid_func3<T> &b,
                     const relax_stencil3<T> &sor, Dir d) = 0;
};

// include/cedar/kernels/plane_relax.h:11-31: setup(so) builds the plane solvers of the
// operator's level; run(so, x, b, dir) relaxes that level's planes of this family
template <typename T, plane_dir PD>
struct plane_relax : kernel_base {
    static const char *name() {
        return PD == plane_dir::xy ? "plane-relax-xy" : PD == plane_dir::xz ? "plane-relax-xz"
                                                                             : "plane-relax-yz";
    }
    virtual void setup(const stencil_op3<double> &so) = 0;
    virtual void run(const stencil_op3<T> &so, grid_func3<T> &x, const grid_func3<T> &b,
                     Dir d) = 0;
};

struct setup_interp : kernel_base {
    static const char *name() { return "setup-interp (3D)"; }
    virtual void run(const stencil_op3<double> &fop, const stencil_op3<double> &cop,
                     prolong_op3<double> &P) = 0;
};

struct coarsen_op : kernel_base {
    static const char *name() { return "coarsen-op (3D)"; }
    virtual void run(const prolong_op3<double> &P, const stencil_op3<double> &fop,
                     stencil_op3<double> &cop) = 0;
};

template <typename T>
struct restriction : kernel_base {
    static const char *name() { return "restriction (3D)"; }
    virtual void run(const prolong_op3<T> &R, const grid_func3<T> &fine,
                     grid_func3<T> &coarse) = 0;
};

template <typename T>
struct interp_add : kernel_base {
    static const char *name() { return "interp-add (3D)"; }
    virtual void run(const prolong_op3<T> &P, const grid_func3<T> &coarse,
                     const grid_func3<T> &residual, grid_func3<T> &fine) = 0;
};

// the band ABD(ldab, nx*ny*nz) is FP64 in both precisions
template <typename T>
struct solve_cg : kernel_base {
    static const char *name() { return "solve-cg (3D)"; }
    virtual void setup(const stencil_op3<double> &so, double *abd, int ldab) = 0;
    virtual void run(grid_func3<T> &x, const grid_func3<T> &b, const double *abd, int ldab,
                     double *bbd) = 0;
};

}  // namespace kernels3

// registers the "hip" implementations of the 3D interfaces (both precisions); plane solves use
// `plane` (the plane-config) with the given precision; jpn is the BoxMG boundary id of the
// periodic mask (0: Dirichlet)
void register_hip_kernels3(kernel_manager &km, int *info_dev, const cdr_solver_settings &plane,
                           int precision, int jpn = 0);

}  // namespace hip
}  // namespace cedar
