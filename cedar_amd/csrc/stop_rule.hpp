// Stopping rule of the solve loops, one instance per problem.
//
// Cedar's rule (include/cedar/multilevel.h:283-297): after cycle i the relative residual
// rho_i = ||r_i|| / ||r_0|| is computed and the loop stops once rho_i < tol, or after max-iter
// cycles.  Two guards are added, both of which end a loop the reference would keep running:
//   * a rho that is NaN or Inf stops the problem with CDR_ERR_NONFINITE;
//   * with stall_cycles = k > 0, a problem whose rho has not gone below its smallest earlier
//     value (rho_0 = 1 before the first cycle) for k consecutive cycles stops with
//     CDR_ERR_STALLED.  k = 0 (the default) is the reference's behaviour: no stall test.
// Host code only; the solvers apply it to the norms they read back after each cycle.
#pragma once

#include <cmath>

#include "../../include/cedar_hip.h"

namespace cedar {
namespace hip {

struct stop_rule {
    enum verdict { GO = 0, CONVERGED = 1, NONFINITE = 2, STALLED = 3 };

    double tol = 0.0;
    int stall_cycles = 0;
    double best = 1.0;  // smallest rho so far (rho_0 = 1)
    int since = 0;      // cycles since best last decreased

    stop_rule() = default;
    stop_rule(double t, int k) : tol(t), stall_cycles(k) {}

    // rho of the cycle just run -> what the loop does with this problem
    verdict step(double rho) {
        if (!std::isfinite(rho)) return NONFINITE;
        if (rho < tol) return CONVERGED;
        if (rho < best) {
            best = rho;
            since = 0;
        } else if (stall_cycles > 0 && ++since >= stall_cycles) {
            return STALLED;
        }
        return GO;
    }

    // status code of a verdict that stops the problem (CDR_OK for convergence)
    static int status(verdict v) {
        return v == NONFINITE ? CDR_ERR_NONFINITE : v == STALLED ? CDR_ERR_STALLED : CDR_OK;
    }
};

}  // namespace hip
}  // namespace cedar
