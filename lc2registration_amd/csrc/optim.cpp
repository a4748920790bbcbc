// Batched derivative-free minimisers (see optim.hpp).
#include "optim.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

namespace lc2opt {

namespace {

struct Box {
    std::vector<double> c;  // centre in the unit cube
    std::vector<int> lev;   // side along i = 3^-lev[i]
    double f;
};

int min_level(const Box& b) { return *std::min_element(b.lev.begin(), b.lev.end()); }

}  // namespace

Result direct_l(const BatchObjective& f, const std::vector<double>& lb,
                const std::vector<double>& ub, int maxeval, double eps, bool original) {
    Result res;
    const int n = (int)lb.size();
    if (n == 0 || ub.size() != lb.size()) {
        res.code = -2;
        res.message = "invalid bounds";
        return res;
    }
    for (int i = 0; i < n; ++i)
        if (!(ub[i] > lb[i])) {
            res.code = -2;
            res.message = "upper bound must exceed lower bound";
            return res;
        }
    auto denorm = [&](const std::vector<double>& c, double* x) {
        for (int i = 0; i < n; ++i) x[i] = lb[i] + c[i] * (ub[i] - lb[i]);
    };
    std::vector<Box> boxes;
    Box b0{std::vector<double>(n, 0.5), std::vector<int>(n, 0), 0.0};
    std::vector<double> X(n);
    denorm(b0.c, X.data());
    f(1, n, X.data(), &b0.f);
    res.nevals = 1;
    boxes.push_back(b0);
    size_t best = 0;

    while (res.nevals < maxeval) {
        // one representative (lowest f) per size class; size = longest side 3^-minlevel
        std::vector<int> rep_of_level;  // index by level
        std::vector<size_t> rep;
        for (size_t k = 0; k < boxes.size(); ++k) {
            const int L = min_level(boxes[k]);
            if ((int)rep_of_level.size() <= L) rep_of_level.resize(L + 1, -1);
            const int cur = rep_of_level[L];
            if (cur < 0 || boxes[k].f < boxes[cur].f) rep_of_level[L] = (int)k;
        }
        for (int L = 0; L < (int)rep_of_level.size(); ++L)
            if (rep_of_level[L] >= 0) rep.push_back((size_t)rep_of_level[L]);
        const double fmin = boxes[best].f;
        std::vector<size_t> sel;
        for (size_t a = 0; a < rep.size(); ++a) {
            const Box& bj = boxes[rep[a]];
            const double sj = std::pow(3.0, -min_level(bj));
            double klo = 0.0, khi = std::numeric_limits<double>::infinity();
            bool ok = true;
            for (size_t q = 0; q < rep.size() && ok; ++q) {
                if (q == a) continue;
                const Box& bi = boxes[rep[q]];
                const double si = std::pow(3.0, -min_level(bi));
                if (si < sj) klo = std::max(klo, (bj.f - bi.f) / (sj - si));
                else if (si > sj) khi = std::min(khi, (bi.f - bj.f) / (si - sj));
            }
            if (klo > khi) ok = false;
            if (ok && std::isfinite(khi) && !(bj.f - khi * sj <= fmin - eps * std::fabs(fmin)))
                ok = false;
            if (ok) sel.push_back(rep[a]);
        }
        if (sel.empty()) sel.push_back(best);
        // larger boxes first when the evaluation budget truncates the iteration
        std::sort(sel.begin(), sel.end(), [&](size_t p, size_t q) {
            return min_level(boxes[p]) < min_level(boxes[q]);
        });
        // sample c +- delta e_i along every longest side of every selected box: one batch
        struct Split {
            size_t box;
            std::vector<int> dims;
        };
        std::vector<Split> splits;
        std::vector<double> pts;
        int budget = maxeval - res.nevals;
        for (size_t s : sel) {
            const int L = min_level(boxes[s]);
            Split sp{s, {}};
            for (int i = 0; i < n; ++i)
                if (boxes[s].lev[i] == L) sp.dims.push_back(i);
            if (!original && (int)sp.dims.size() * 2 > budget) break;
            budget -= (int)sp.dims.size() * 2;
            const double d = std::pow(3.0, -(L + 1));
            for (int i : sp.dims)
                for (int sgn = -1; sgn <= 1; sgn += 2) {
                    std::vector<double> c = boxes[s].c;
                    c[i] += sgn * d;
                    denorm(c, X.data());
                    pts.insert(pts.end(), X.begin(), X.end());
                }
            splits.push_back(sp);
        }
        if (splits.empty()) break;
        const int np = (int)(pts.size() / n);
        std::vector<double> F(np);
        f(np, n, pts.data(), F.data());
        res.nevals += np;
        int k = 0;
        for (auto& sp : splits) {
            const size_t s = sp.box;
            const int L = min_level(boxes[s]);
            const double d = std::pow(3.0, -(L + 1));
            std::vector<std::pair<double, int>> order;
            std::vector<double> fm(n), fp(n);
            for (int i : sp.dims) {
                fm[i] = F[k++];
                fp[i] = F[k++];
                order.push_back({std::min(fm[i], fp[i]), i});
            }
            std::sort(order.begin(), order.end());
            for (auto& o : order) {
                const int i = o.second;
                boxes[s].lev[i] += 1;
                for (int sgn = -1; sgn <= 1; sgn += 2) {
                    Box nb{boxes[s].c, boxes[s].lev, sgn < 0 ? fm[i] : fp[i]};
                    nb.c[i] += sgn * d;
                    boxes.push_back(nb);
                    if (nb.f < boxes[best].f) best = boxes.size() - 1;
                }
            }
        }
    }
    res.x.resize(n);
    denorm(boxes[best].c, res.x.data());
    res.f = boxes[best].f;
    res.code = 5;
    res.message = "maxeval reached";
    return res;
}

Result qmtr(const BatchObjective& f, const std::vector<double>& x0, const std::vector<double>& lb,
            const std::vector<double>& ub, double xtol_rel, int maxeval) {
    Result res;
    const int n = (int)x0.size();
    if (n == 0 || lb.size() != x0.size() || ub.size() != x0.size()) {
        res.code = -2;
        res.message = "invalid dimensions";
        return res;
    }
    std::vector<double> s(n), y(n), ylo(n), yhi(n);
    for (int i = 0; i < n; ++i) {
        const double range = ub[i] - lb[i];
        s[i] = range > 0 ? 0.1 * range : 1.0;
        ylo[i] = lb[i] / s[i];
        yhi[i] = ub[i] / s[i];
        y[i] = std::min(std::max(x0[i], lb[i]), ub[i]) / s[i];
    }
    auto eval = [&](const std::vector<std::vector<double>>& Ys, std::vector<double>& F) {
        int m = (int)Ys.size();
        if (m > maxeval - res.nevals) m = std::max(0, maxeval - res.nevals);
        F.assign(Ys.size(), std::numeric_limits<double>::infinity());
        if (m == 0) return;
        std::vector<double> X((size_t)m * n);
        for (int k = 0; k < m; ++k)
            for (int i = 0; i < n; ++i) X[(size_t)k * n + i] = Ys[k][i] * s[i];
        f(m, n, X.data(), F.data());
        for (int k = 0; k < m; ++k)
            if (std::isnan(F[k])) F[k] = std::numeric_limits<double>::infinity();
        res.nevals += m;
    };
    std::vector<double> F0;
    eval({y}, F0);
    double f0 = F0[0];
    const double rho_end = std::max(10.0 * xtol_rel, 1e-9);
    const double rho_max = 2.0;
    double rho = 1.0;
    std::vector<double> last_move(n, 0.0);  // previous accepted step (pattern direction)
    res.code = 1;
    res.message = "radius below tolerance";
    while (rho >= rho_end) {
        if (res.nevals >= maxeval) {
            res.code = 5;
            res.message = "maxeval reached";
            break;
        }
        // 1. stencil: two points per coordinate, one batch
        std::vector<std::vector<double>> Ys;
        std::vector<double> t1(n, 0.0), t2(n, 0.0);
        std::vector<char> active(n, 0);
        for (int i = 0; i < n; ++i) {
            if (!(yhi[i] > ylo[i])) continue;
            const double up = std::min(rho, yhi[i] - y[i]), dn = std::min(rho, y[i] - ylo[i]);
            if (up >= 0.25 * rho && dn >= 0.25 * rho) {
                t1[i] = up;
                t2[i] = -dn;
            } else if (up < 0.25 * rho) {
                t1[i] = -dn;
                t2[i] = -std::min(2.0 * rho, y[i] - ylo[i]);
            } else {
                t1[i] = up;
                t2[i] = std::min(2.0 * rho, yhi[i] - y[i]);
            }
            if (std::fabs(t1[i]) < 1e-15 || std::fabs(t2[i]) < 1e-15 || std::fabs(t1[i] - t2[i]) < 1e-15)
                continue;
            active[i] = 1;
            std::vector<double> a = y, b = y;
            a[i] += t1[i];
            b[i] += t2[i];
            Ys.push_back(a);
            Ys.push_back(b);
        }
        if (Ys.empty()) break;
        std::vector<double> Fs;
        eval(Ys, Fs);
        // 2. diagonal quadratic model through (0, f0), (t1, f1), (t2, f2)
        std::vector<double> g(n, 0.0), h(n, 0.0);
        int k = 0;
        double fbest = f0;
        std::vector<double> ybest = y;
        for (int i = 0; i < n; ++i) {
            if (!active[i]) continue;
            const double f1 = Fs[k] - f0, f2 = Fs[k + 1] - f0;
            if (Fs[k] < fbest) { fbest = Fs[k]; ybest = Ys[k]; }
            if (Fs[k + 1] < fbest) { fbest = Fs[k + 1]; ybest = Ys[k + 1]; }
            k += 2;
            if (!std::isfinite(f1) || !std::isfinite(f2)) continue;
            const double a = t1[i], b = t2[i];
            g[i] = (f1 * b * b - f2 * a * a) / (a * b * (b - a));
            h[i] = 2.0 * (f1 * b - f2 * a) / (a * b * (a - b));
        }
        // 3. box-constrained model step within radius 2*rho, and two scalings: one batch
        std::vector<double> st(n, 0.0);
        double pred = 0.0;
        for (int i = 0; i < n; ++i) {
            if (!active[i]) continue;
            double si;
            if (h[i] > 0.0) si = -g[i] / h[i];
            else si = g[i] > 0 ? -2.0 * rho : (g[i] < 0 ? 2.0 * rho : 0.0);
            si = std::max(-2.0 * rho, std::min(2.0 * rho, si));
            si = std::max(ylo[i] - y[i], std::min(yhi[i] - y[i], si));
            st[i] = si;
            pred -= g[i] * si + 0.5 * h[i] * si * si;
        }
        double snorm = 0.0;
        for (double v : st) snorm = std::max(snorm, std::fabs(v));
        int best_cand = -1;
        double mnorm = 0.0;
        for (double v : last_move) mnorm = std::max(mnorm, std::fabs(v));
        if (snorm > 0.0 || mnorm > 0.0) {
            // model step at three scalings, and the pattern move (the previous accepted step
            // repeated, from here and after the model step) for curved valleys
            std::vector<std::vector<double>> C;
            auto clip = [&](std::vector<double> c) {
                for (int i = 0; i < n; ++i) c[i] = std::max(ylo[i], std::min(yhi[i], c[i]));
                return c;
            };
            for (double sc : {1.0, 0.5, 2.0}) {
                std::vector<double> c = y;
                for (int i = 0; i < n; ++i) c[i] += sc * st[i];
                C.push_back(clip(c));
            }
            if (mnorm > 0.0) {
                std::vector<double> c = y, c2 = y;
                for (int i = 0; i < n; ++i) {
                    c[i] += last_move[i];
                    c2[i] += last_move[i] + st[i];
                }
                C.push_back(clip(c));
                C.push_back(clip(c2));
            }
            std::vector<double> Fc;
            eval(C, Fc);
            for (int q = 0; q < (int)C.size(); ++q)
                if (Fc[q] < fbest) {
                    fbest = Fc[q];
                    ybest = C[q];
                    best_cand = q;
                }
        }
        if (fbest < f0) {
            const double gain = f0 - fbest;
            for (int i = 0; i < n; ++i) last_move[i] = ybest[i] - y[i];
            y = ybest;
            f0 = fbest;
            if (best_cand >= 0 && pred > 0 && gain > 0.75 * pred && snorm >= rho)
                rho = std::min(2.0 * rho, rho_max);
        } else {
            std::fill(last_move.begin(), last_move.end(), 0.0);
            rho *= 0.25;
        }
    }
    res.x.resize(n);
    for (int i = 0; i < n; ++i) res.x[i] = y[i] * s[i];
    res.f = f0;
    return res;
}

void opt::set_objective(vfunc f, void* data, bool maximize) {
    maximize_ = maximize;
    batch_ = [f, data](int k, int d, const double* X, double* F) {
        std::vector<double> grad;  // derivative-free: NLopt passes an empty gradient
        for (int j = 0; j < k; ++j) {
            std::vector<double> x(X + (size_t)j * d, X + (size_t)j * d + d);
            F[j] = f(x, grad, data);
        }
    };
}

result opt::optimize(std::vector<double>& x, double& opt_f) {
    nevals_ = 0;
    if (!batch_ || lb_.size() != n_ || ub_.size() != n_ || (alg_ == LN_BOBYQA && x.size() != n_))
        return INVALID_ARGS;
    const bool mx = maximize_;
    const BatchObjective inner = batch_;
    // minimise -f for a max objective, as NLopt does
    BatchObjective obj = [inner, mx](int k, int d, const double* X, double* F) {
        inner(k, d, X, F);
        if (mx)
            for (int j = 0; j < k; ++j) F[j] = -F[j];
    };
    Result r;
    if (alg_ == LN_BOBYQA) {
        const std::vector<double> step =
            step_.size() == n_ ? step_ : default_initial_step(x, lb_, ub_);
        r = bobyqa(obj, x, lb_, ub_, step, xtol_rel_, maxeval_);
    } else {
        if (maxeval_ <= 0) return INVALID_ARGS;  // DIRECT needs a budget here
        r = direct_l(obj, lb_, ub_, maxeval_, 1e-4, alg_ == GN_ORIG_DIRECT_L);
    }
    nevals_ = r.nevals;
    if (r.code < 0) return r.code == -2 ? INVALID_ARGS : FAILURE;
    x = r.x;
    opt_f = mx ? -r.f : r.f;
    return r.code == 5 ? MAXEVAL_REACHED : r.code == 4 ? XTOL_REACHED : SUCCESS;
}

}  // namespace lc2opt
