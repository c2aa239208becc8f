// Internal header: the bounded quasi-Newton minimiser behind bsa_fit.
#pragma once
#include <algorithm>
#include <functional>
#include <memory>
#include <vector>

#include "bsa.h"

namespace bsa {

// objective: f and its gradient g at x; returns a bsa status code
using ObjFn = std::function<int(const double *x, double *f, double *g)>;

// objective in two calls, for an optimiser that needs the gradient only at
// the points it accepts: value(x, &f) makes x the current point, grad(g)
// writes the gradient at the current point; both return bsa status codes.
// both (optional): value and gradient at x in one call, where the caller
// knows it will want both (the start, a Hessian column); it too makes x
// the current point.
struct Objective {
    std::function<int(const double *x, double *f)> value;
    std::function<int(double *g)> grad;
    ObjFn both;
};

// the two calls from one combined function: value computes both and keeps
// the gradient for grad
inline Objective combined(const ObjFn &fg, int n)
{
    auto buf = std::make_shared<std::vector<double>>(n);
    return {[fg, buf](const double *x, double *f) { return fg(x, f, buf->data()); },
            [buf](double *g) {
                std::copy(buf->begin(), buf->end(), g);
                return BSA_OK;
            }};
}

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
// nblock: the last nblock variables (in codeml's x, the model parameters
// after the branch lengths), whose Hessian block Newton steps on more than
// 160 free variables form exactly (their conjugate-gradient preconditioner)
// The gradient is asked for at the start point, at every accepted point
// and for the Newton finish's Hessian products, never at a rejected trial
// point; res->evaluations counts the value calls.
int minimize_box(const Objective &obj, int n, double *x, const double *lo, const double *hi,
                 const bsa_fit_opts *opts, FitResult *res, bool verify = false,
                 int nblock = 0);
inline int minimize_box(const ObjFn &fg, int n, double *x, const double *lo, const double *hi,
                        const bsa_fit_opts *opts, FitResult *res, bool verify = false,
                        int nblock = 0)
{
    return minimize_box(combined(fg, n), n, x, lo, hi, opts, res, verify, nblock);
}

}  // namespace bsa
