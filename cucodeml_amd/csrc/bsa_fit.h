// Internal header: the bounded quasi-Newton minimiser behind bsa_fit.
#pragma once
#include <functional>

#include "bsa.h"

namespace bsa {

// objective: f and its gradient g at x; returns a bsa status code
using ObjFn = std::function<int(const double *x, double *f, double *g)>;

struct FitResult {
    double f = 0;
    int iterations = 0, evaluations = 0, stop = BSA_STOP_LIMIT;
};

// minimise fg over the box [lo, hi] from x (moved into the box); x holds
// the minimiser on return
// verify: where the quasi-Newton phase ends by the stall rule or without a
// descent step, and opts asks for no Newton iterations, up to four Newton
// iterations check whether x is the maximum (BSA_STOP_NEWTON if so); for
// exact FP64 gradients (the adjoint on FP64 partials), whose differences give
// the Hessian
int minimize_box(const ObjFn &fg, int n, double *x, const double *lo, const double *hi,
                 const bsa_fit_opts *opts, FitResult *res, bool verify = false);

}  // namespace bsa
