// Derivative-free, bound-constrained minimisers with a batch objective, standing in for
// the NLopt calls of the reference's Optimizers* loop (optimizers.cu:298, :385, :854, :937,
// :1473, :1548). NLopt itself is not available to this build, so these are written from
// the published methods:
//   direct_l  -- DIRECT with Gablonsky & Kelley's locally-biased selection (DIRECT-L,
//                J. Global Optim. 21, 2001): one potentially optimal box per size class,
//                size = longest side, trisection along the longest sides in order of the
//                sampled values. Every iteration's new centres form ONE batch.
//   bobyqa    -- Powell's BOBYQA (bobyqa.cpp), the LN_BOBYQA rounds;
//   qmtr      -- a bound-constrained trust-region method with diagonal quadratic models,
//                kept as an option: per iteration a 2n-point stencil (one batch;
//                one-sided at a bound) fits gradient and diagonal curvature; the
//                box-constrained model step at
//                three scalings plus two pattern moves along the last accepted step form a
//                second batch. The radius shrinks 4x when nothing improves and doubles after
//                a step that gains 3/4 of the model's prediction. Stops at radius
//                < max(10 xtol_rel, 1e-9) in variables scaled by 0.1 (ub - lb), or maxeval.
// All call the objective with many points at once where the method allows, so the caller
// can evaluate them in one fused GPU launch.
#pragma once

#include <functional>
#include <string>
#include <utility>
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

// DIRECT-L. original = true follows Gablonsky's own code (NLopt's GN_ORIG_DIRECT_L,
// optimizers.cu:854): the budget is checked between iterations, so the last iteration is
// completed and may pass maxeval. original = false follows NLopt's re-implementation
// (GN_DIRECT_L, optimizers.cu:298): evaluation stops at maxeval exactly, the boxes of the
// last iteration taken largest first.
Result direct_l(const BatchObjective& f, const std::vector<double>& lb,
                const std::vector<double>& ub, int maxeval, double eps = 1e-4,
                bool original = false);

Result qmtr(const BatchObjective& f, const std::vector<double>& x0, const std::vector<double>& lb,
            const std::vector<double>& ub, double xtol_rel, int maxeval);

// Powell's BOBYQA (bobyqa.cpp): x0 inside [lb, ub], step = initial step per variable (the
// trust region starts at one step along every axis), stop when the trust-region radius
// falls below xtol_rel * step (NLopt's rho_end) or after maxeval evaluations (<= 0: none).
// The 2n+1 initial points are evaluated in one batch.
Result bobyqa(const BatchObjective& f, const std::vector<double>& x0, const std::vector<double>& lb,
              const std::vector<double>& ub, const std::vector<double>& step, double xtol_rel,
              int maxeval);
// NLopt's default initial step: a quarter of the bound range, at most 3/4 of the distance
// to the nearer bound.
std::vector<double> default_initial_step(const std::vector<double>& x, const std::vector<double>& lb,
                                         const std::vector<double>& ub);

// The nlopt::opt interface the reference's Optimizers* use (optimizers.cu:298-313,
// :385-407, :854-882, :937-965): algorithm and dimension, a min or max objective with
// NLopt's vfunc signature (or a batch objective), bounds, maxeval, xtol_rel, optimize().
enum algorithm { GN_DIRECT_L = 0, GN_ORIG_DIRECT_L = 1, LN_BOBYQA = 2 };
enum result {
    FAILURE = -1,
    INVALID_ARGS = -2,
    SUCCESS = 1,
    XTOL_REACHED = 4,
    MAXEVAL_REACHED = 5
};
using vfunc = double (*)(const std::vector<double>& x, std::vector<double>& grad, void* data);

class opt {
public:
    opt(algorithm a, unsigned n) : alg_(a), n_(n) {}
    void set_min_objective(vfunc f, void* data) { set_objective(f, data, false); }
    void set_max_objective(vfunc f, void* data) { set_objective(f, data, true); }
    // batch form: all points of a DIRECT iteration or of BOBYQA's initial set in one call
    void set_batch_objective(BatchObjective f, bool maximize = false) {
        batch_ = std::move(f);
        maximize_ = maximize;
    }
    void set_lower_bounds(const std::vector<double>& lb) { lb_ = lb; }
    void set_upper_bounds(const std::vector<double>& ub) { ub_ = ub; }
    void set_maxeval(int n) { maxeval_ = n; }
    void set_xtol_rel(double t) { xtol_rel_ = t; }
    void set_initial_step(const std::vector<double>& dx) { step_ = dx; }
    // x: start on entry (BOBYQA), optimum on return; opt_f: the objective there
    result optimize(std::vector<double>& x, double& opt_f);
    int get_numevals() const { return nevals_; }
    algorithm get_algorithm() const { return alg_; }
    unsigned get_dimension() const { return n_; }

private:
    void set_objective(vfunc f, void* data, bool maximize);
    algorithm alg_;
    unsigned n_;
    BatchObjective batch_;
    bool maximize_ = false;
    std::vector<double> lb_, ub_, step_;
    int maxeval_ = 0;
    double xtol_rel_ = 0.0;
    int nevals_ = 0;
};

}  // namespace lc2opt
