// Derivative-free, bound-constrained minimisers with a batch objective, standing in for
// the NLopt calls of the reference's Optimizers* loop (optimizers.cu:298, :385, :854, :937,
// :1473, :1548). NLopt itself is not available to this build, so these are written from
// the published methods:
//   direct_l  -- DIRECT with Gablonsky & Kelley's locally-biased selection (DIRECT-L,
//                J. Global Optim. 21, 2001): one potentially optimal box per size class,
//                size = longest side, trisection along the longest sides in order of the
//                sampled values. Every iteration's new centres form ONE batch.
//   qmtr      -- a bound-constrained quadratic-model trust-region method in the BOBYQA
//                family: per iteration a 2n-point stencil (one batch; one-sided at a bound)
//                fits gradient and diagonal curvature; the box-constrained model step at
//                three scalings plus two pattern moves along the last accepted step form a
//                second batch. The radius shrinks 4x when nothing improves and doubles after
//                a step that gains 3/4 of the model's prediction. Stops at radius
//                < max(10 xtol_rel, 1e-9) in variables scaled by 0.1 (ub - lb), or maxeval.
// Both call the objective with many points at once so the caller can evaluate them in
// one fused GPU launch.
#pragma once

#include <functional>
#include <string>
#include <vector>

namespace lc2opt {

// Evaluate f at n points of dimension d (row-major X[n*d]) into F[n].
using BatchObjective = std::function<void(int n, int d, const double* X, double* F)>;

struct Result {
    std::vector<double> x;
    double f = 0.0;
    int nevals = 0;
    int code = 0;  // 1 success (tolerance), 5 maxeval reached, <0 failure
    std::string message;
};

Result direct_l(const BatchObjective& f, const std::vector<double>& lb,
                const std::vector<double>& ub, int maxeval, double eps = 1e-4);

Result qmtr(const BatchObjective& f, const std::vector<double>& x0, const std::vector<double>& lb,
            const std::vector<double>& ub, double xtol_rel, int maxeval);

}  // namespace lc2opt
