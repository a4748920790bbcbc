// BOBYQA: bound-constrained optimisation by quadratic approximation (M. J. D. Powell, "The
// BOBYQA algorithm for bound constrained optimization without derivatives", Cambridge
// NA Report NA2009/06), the method behind the reference's nlopt::LN_BOBYQA rounds
// (optimizers.cu:385-407, :937-965, :1548-1569). Written from the published method:
//
//  * m = 2n+1 interpolation points: x0 and x0 +- rho_beg e_i (one-sided next to a bound),
//    all evaluated in ONE batch; the initial model takes its gradient and diagonal
//    curvature from them.
//  * Each later model is the quadratic that interpolates all m points and whose Hessian
//    differs least, in the Frobenius norm, from the previous one. Powell updates the
//    inverse of that KKT system in O(m^2) per step; here the (m+n+1)^2 system is solved
//    afresh (LU with partial pivoting, points scaled by rho) -- for n <= 12 that is a few
//    tens of microseconds, far below one LC2 evaluation, and it keeps no drifting state.
//  * Trust-region step: truncated conjugate gradients on the model inside the ball of
//    radius delta and the bounds, with variables that reach a bound fixed (TRSBOX's
//    active-set scheme, without its final boundary search).
//  * The point replaced by a new one maximises |Lagrange function at the new point|
//    weighted by max(1, (distance/delta)^4), Powell's choice of KNEW.
//  * Geometry steps (ALTMOV's line searches): when a step fails and some point lies further
//    than 2 delta from the best point, that point moves to where its Lagrange function is
//    largest along the lines through the best point and the other points, or along the
//    projected gradient of that Lagrange function.
//  * delta: halved, kept or doubled by the ratio of actual to predicted reduction (0.1 and
//    0.7 thresholds); rho falls by 10, to sqrt(rho rho_end), or to rho_end (Powell's
//    schedule) once steps stop succeeding at delta = rho with the points close.
// Variables are scaled so that the initial step is 1 along every axis (as NLopt rescales by
// its initial step); rho_end = xtol_rel * rho_beg, as NLopt derives it from set_xtol_rel.
// Objective values that are not finite are replaced by the largest finite value seen plus
// a margin, so the models stay finite (the LC2 objective -log(LC2) is +inf where LC2 = 0).
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "optim.hpp"

namespace lc2opt {

namespace {

using Vec = std::vector<double>;

double dot(const Vec& a, const Vec& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

// Solve A x = b (A: k x k row-major, overwritten) with partial pivoting; false if singular.
bool lu_solve(std::vector<double> A, int k, std::vector<double>& b) {
    std::vector<int> piv(k);
    for (int c = 0; c < k; ++c) {
        int p = c;
        double best = std::fabs(A[(size_t)c * k + c]);
        for (int r = c + 1; r < k; ++r) {
            const double v = std::fabs(A[(size_t)r * k + c]);
            if (v > best) {
                best = v;
                p = r;
            }
        }
        if (!(best > 0.0) || !std::isfinite(best)) return false;
        if (p != c) {
            for (int j = 0; j < k; ++j) std::swap(A[(size_t)c * k + j], A[(size_t)p * k + j]);
            std::swap(b[c], b[p]);
        }
        const double d = A[(size_t)c * k + c];
        for (int r = c + 1; r < k; ++r) {
            const double f = A[(size_t)r * k + c] / d;
            if (f == 0.0) continue;
            for (int j = c; j < k; ++j) A[(size_t)r * k + j] -= f * A[(size_t)c * k + j];
            b[r] -= f * b[c];
        }
    }
    for (int r = k - 1; r >= 0; --r) {
        double s = b[r];
        for (int j = r + 1; j < k; ++j) s -= A[(size_t)r * k + j] * b[j];
        b[r] = s / A[(size_t)r * k + r];
    }
    for (double v : b)
        if (!std::isfinite(v)) return false;
    return true;
}

// Quadratic model Q(x) = c + g.(x - xm) + 1/2 (x - xm)' H (x - xm).
struct Model {
    Vec xm, g, H;  // H: n x n row-major
    double c = 0.0;
    int n = 0;
    double value(const Vec& x) const {
        double v = c;
        Vec d(n);
        for (int i = 0; i < n; ++i) d[i] = x[i] - xm[i];
        for (int i = 0; i < n; ++i) {
            v += g[i] * d[i];
            double hd = 0.0;
            for (int j = 0; j < n; ++j) hd += H[(size_t)i * n + j] * d[j];
            v += 0.5 * d[i] * hd;
        }
        return v;
    }
    Vec grad(const Vec& x) const {
        Vec out = g;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) out[i] += H[(size_t)i * n + j] * (x[j] - xm[j]);
        return out;
    }
};

class Bobyqa {
public:
    Bobyqa(const BatchObjective& f, const Vec& lo, const Vec& hi, int maxeval)
        : f_(f), lo_(lo), hi_(hi), n_((int)lo.size()), m_(2 * (int)lo.size() + 1),
          maxeval_(maxeval) {}

    Result run(const Vec& x0, double rhobeg, double rhoend);

private:
    // evaluations (all in the scaled space; the caller's f sees the same coordinates)
    void eval(const std::vector<Vec>& X, Vec& F);
    // the KKT system of the interpolation points about base b, scaled by s
    bool build_kkt(const Vec& b, double s, std::vector<double>& W) const;
    bool update_model(const Model& old);
    bool lagrange_values(const Vec& x, Vec& ell) const;
    bool lagrange_function(int t, Model& L) const;
    Vec trust_step(const Vec& xopt, double delta, double* pred) const;
    Vec geometry_step(int t, double dgeo) const;
    int best_index() const;

    const BatchObjective& f_;
    Vec lo_, hi_;
    int n_, m_;
    int maxeval_;
    int nevals_ = 0;
    double fmax_finite_ = -std::numeric_limits<double>::infinity();
    std::vector<Vec> Y_;  // interpolation points
    Vec F_;
    Model Q_;
    // the KKT matrix of the current points, factorised on demand (about xopt, scale rho)
    Vec kkt_base_;
    double kkt_scale_ = 1.0;
    std::vector<double> kkt_;
    bool kkt_ok_ = false;
    void refresh_kkt(const Vec& base, double s) {
        kkt_base_ = base;
        kkt_scale_ = s;
        kkt_ok_ = build_kkt(base, s, kkt_);
    }
};

void Bobyqa::eval(const std::vector<Vec>& X, Vec& F) {
    const int k = (int)X.size();
    std::vector<double> flat((size_t)k * n_);
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < n_; ++j) flat[(size_t)i * n_ + j] = X[i][j];
    F.assign(k, 0.0);
    f_(k, n_, flat.data(), F.data());
    nevals_ += k;
    for (double v : F)
        if (std::isfinite(v)) fmax_finite_ = std::max(fmax_finite_, v);
    for (double& v : F)
        if (!std::isfinite(v))
            v = std::isfinite(fmax_finite_) ? fmax_finite_ + 10.0 * (1.0 + std::fabs(fmax_finite_))
                                            : 1e30;
}

int Bobyqa::best_index() const {
    int k = 0;
    for (int i = 1; i < m_; ++i)
        if (F_[i] < F_[k]) k = i;
    return k;
}

// W = [[A, E], [E', 0]], A_jk = 1/2 (z_j . z_k)^2, E_j = [1, z_j], z = (y - b)/s.
bool Bobyqa::build_kkt(const Vec& b, double s, std::vector<double>& W) const {
    const int k = m_ + n_ + 1;
    W.assign((size_t)k * k, 0.0);
    std::vector<Vec> Z(m_, Vec(n_));
    for (int j = 0; j < m_; ++j)
        for (int i = 0; i < n_; ++i) Z[j][i] = (Y_[j][i] - b[i]) / s;
    for (int j = 0; j < m_; ++j) {
        for (int l = 0; l < m_; ++l) {
            const double d = dot(Z[j], Z[l]);
            W[(size_t)j * k + l] = 0.5 * d * d;
        }
        W[(size_t)j * k + m_] = 1.0;
        W[(size_t)m_ * k + j] = 1.0;
        for (int i = 0; i < n_; ++i) {
            W[(size_t)j * k + m_ + 1 + i] = Z[j][i];
            W[(size_t)(m_ + 1 + i) * k + j] = Z[j][i];
        }
    }
    return true;
}

// Minimum-Frobenius-norm change of the Hessian that interpolates all points.
bool Bobyqa::update_model(const Model& old) {
    const int k = m_ + n_ + 1;
    const Vec& b = kkt_base_;
    const double s = kkt_scale_;
    Vec rhs(k, 0.0);
    for (int j = 0; j < m_; ++j) rhs[j] = F_[j] - old.value(Y_[j]);
    if (!kkt_ok_ || !lu_solve(kkt_, k, rhs)) return false;
    Model q;
    q.n = n_;
    q.xm = b;
    q.c = old.value(b) + rhs[m_];
    q.g = old.grad(b);
    for (int i = 0; i < n_; ++i) q.g[i] += rhs[m_ + 1 + i] / s;
    q.H = old.H;
    const double s4 = s * s * s * s;
    for (int j = 0; j < m_; ++j) {
        const double lam = rhs[j] / s4;
        if (lam == 0.0) continue;
        for (int a = 0; a < n_; ++a) {
            const double za = Y_[j][a] - b[a];
            for (int c = 0; c < n_; ++c) q.H[(size_t)a * n_ + c] += lam * za * (Y_[j][c] - b[c]);
        }
    }
    Q_ = q;
    return true;
}

// Values of all m Lagrange functions at x: W^-1 [w(x); 1; z] (W is symmetric).
bool Bobyqa::lagrange_values(const Vec& x, Vec& ell) const {
    const int k = m_ + n_ + 1;
    const double s = kkt_scale_;
    Vec z(n_);
    for (int i = 0; i < n_; ++i) z[i] = (x[i] - kkt_base_[i]) / s;
    Vec v(k, 0.0);
    for (int j = 0; j < m_; ++j) {
        double d = 0.0;
        for (int i = 0; i < n_; ++i) d += (Y_[j][i] - kkt_base_[i]) / s * z[i];
        v[j] = 0.5 * d * d;
    }
    v[m_] = 1.0;
    for (int i = 0; i < n_; ++i) v[m_ + 1 + i] = z[i];
    if (!kkt_ok_ || !lu_solve(kkt_, k, v)) return false;
    ell.assign(v.begin(), v.begin() + m_);
    return true;
}

// The t-th Lagrange function as a quadratic model.
bool Bobyqa::lagrange_function(int t, Model& L) const {
    const int k = m_ + n_ + 1;
    Vec rhs(k, 0.0);
    rhs[t] = 1.0;
    if (!kkt_ok_ || !lu_solve(kkt_, k, rhs)) return false;
    const double s = kkt_scale_, s4 = s * s * s * s;
    L.n = n_;
    L.xm = kkt_base_;
    L.c = rhs[m_];
    L.g.assign(n_, 0.0);
    for (int i = 0; i < n_; ++i) L.g[i] = rhs[m_ + 1 + i] / s;
    L.H.assign((size_t)n_ * n_, 0.0);
    for (int j = 0; j < m_; ++j) {
        const double lam = rhs[j] / s4;
        for (int a = 0; a < n_; ++a)
            for (int c = 0; c < n_; ++c)
                L.H[(size_t)a * n_ + c] += lam * (Y_[j][a] - kkt_base_[a]) * (Y_[j][c] - kkt_base_[c]);
    }
    return true;
}

// Truncated CG on Q(xopt + d), |d| <= delta, lo <= xopt + d <= hi; variables that reach a
// bound are fixed and CG restarts on the others. *pred = Q(xopt) - Q(xopt + d).
Vec Bobyqa::trust_step(const Vec& xopt, double delta, double* pred) const {
    Vec d(n_, 0.0), g = Q_.grad(xopt);
    std::vector<char> fixed(n_, 0);
    // variables on a bound whose gradient pushes outwards start fixed
    for (int i = 0; i < n_; ++i)
        if ((xopt[i] <= lo_[i] && g[i] > 0.0) || (xopt[i] >= hi_[i] && g[i] < 0.0)) fixed[i] = 1;
    const double gnorm0 = std::sqrt(dot(g, g));
    Vec p(n_, 0.0);
    bool restart = true;
    double rr_old = 0.0;
    for (int it = 0; it < 4 * n_ + 4; ++it) {
        Vec r(n_, 0.0);
        for (int i = 0; i < n_; ++i) r[i] = fixed[i] ? 0.0 : -g[i];
        const double rr = dot(r, r);
        if (rr <= 1e-24 * std::max(1.0, gnorm0 * gnorm0)) break;
        if (restart) {
            p = r;
            restart = false;
        } else {
            const double beta = rr / rr_old;
            for (int i = 0; i < n_; ++i) p[i] = fixed[i] ? 0.0 : r[i] + beta * p[i];
        }
        rr_old = rr;
        // Hp
        Vec hp(n_, 0.0);
        for (int a = 0; a < n_; ++a)
            for (int c = 0; c < n_; ++c) hp[a] += Q_.H[(size_t)a * n_ + c] * p[c];
        const double php = dot(p, hp);
        // step to the trust-region boundary
        const double dd = dot(d, d), dp = dot(d, p), pp = dot(p, p);
        if (pp <= 0.0) break;
        const double disc = dp * dp + pp * (delta * delta - dd);
        const double a_tr = disc > 0.0 ? (-dp + std::sqrt(disc)) / pp : 0.0;
        // step to the first bound
        double a_b = std::numeric_limits<double>::infinity();
        int ib = -1;
        for (int i = 0; i < n_; ++i) {
            if (fixed[i] || p[i] == 0.0) continue;
            const double lim = p[i] > 0.0 ? (hi_[i] - xopt[i] - d[i]) / p[i]
                                          : (lo_[i] - xopt[i] - d[i]) / p[i];
            if (lim < a_b) {
                a_b = std::max(lim, 0.0);
                ib = i;
            }
        }
        const double a_cg = php > 0.0 ? rr / php : std::numeric_limits<double>::infinity();
        const double a = std::min(a_cg, std::min(a_tr, a_b));
        for (int i = 0; i < n_; ++i) {
            d[i] += a * p[i];
            g[i] += a * hp[i];
        }
        if (a == a_b && ib >= 0 && a_b <= a_tr) {
            d[ib] = (p[ib] > 0.0 ? hi_[ib] : lo_[ib]) - xopt[ib];
            fixed[ib] = 1;
            restart = true;
            continue;
        }
        if (a >= a_tr) break;  // on the trust-region boundary
    }
    for (int i = 0; i < n_; ++i) d[i] = std::min(std::max(d[i], lo_[i] - xopt[i]), hi_[i] - xopt[i]);
    Vec xn(n_);
    for (int i = 0; i < n_; ++i) xn[i] = xopt[i] + d[i];
    *pred = Q_.value(xopt) - Q_.value(xn);
    return d;
}

// Move point t to where its Lagrange function is largest in modulus: along the lines
// through xopt and each other point, and along the projected gradient of the function,
// at most dgeo from xopt and inside the bounds (ALTMOV's candidates).
Vec Bobyqa::geometry_step(int t, double dgeo) const {
    const int kopt = best_index();
    const Vec& xo = Y_[kopt];
    Model L;
    Vec best = xo;
    if (!lagrange_function(t, L)) return best;
    double best_val = -1.0;
    auto consider = [&](const Vec& dir) {
        const double nn = std::sqrt(dot(dir, dir));
        if (!(nn > 0.0)) return;
        // largest step in [-amax, amax] along dir inside the bounds and the ball
        double amax = dgeo / nn, amin = -dgeo / nn;
        for (int i = 0; i < n_; ++i) {
            if (dir[i] > 0.0) {
                amax = std::min(amax, (hi_[i] - xo[i]) / dir[i]);
                amin = std::max(amin, (lo_[i] - xo[i]) / dir[i]);
            } else if (dir[i] < 0.0) {
                amax = std::min(amax, (lo_[i] - xo[i]) / dir[i]);
                amin = std::max(amin, (hi_[i] - xo[i]) / dir[i]);
            }
        }
        if (!(amax >= amin)) return;
        // L is quadratic along the line: check both ends and the stationary point
        const Vec g0 = L.grad(xo);
        Vec hd(n_, 0.0);
        for (int a = 0; a < n_; ++a)
            for (int c = 0; c < n_; ++c) hd[a] += L.H[(size_t)a * n_ + c] * dir[c];
        const double l0 = L.value(xo), l1 = dot(g0, dir), l2 = dot(dir, hd);
        std::vector<double> cand = {amin, amax};
        if (l2 != 0.0) {
            const double as = -l1 / l2;
            if (as > amin && as < amax) cand.push_back(as);
        }
        for (double a : cand) {
            if (a == 0.0) continue;
            const double v = std::fabs(l0 + a * l1 + 0.5 * a * a * l2);
            if (v > best_val) {
                best_val = v;
                for (int i = 0; i < n_; ++i)
                    best[i] = std::min(std::max(xo[i] + a * dir[i], lo_[i]), hi_[i]);
            }
        }
    };
    for (int k = 0; k < m_; ++k) {
        if (k == kopt) continue;
        Vec dir(n_);
        for (int i = 0; i < n_; ++i) dir[i] = Y_[k][i] - xo[i];
        consider(dir);
    }
    Vec gl = L.grad(xo);
    for (int i = 0; i < n_; ++i)
        if ((xo[i] <= lo_[i] && gl[i] < 0.0) || (xo[i] >= hi_[i] && gl[i] > 0.0)) gl[i] = 0.0;
    consider(gl);
    return best;
}

Result Bobyqa::run(const Vec& x0in, double rhobeg, double rhoend) {
    Result res;
    const int n = n_;
    // starting point: inside the bounds, and either on a bound or rhobeg away from it
    Vec x0 = x0in;
    std::vector<int> side(n, 0);  // 0 both ways, +1 only upwards, -1 only downwards
    for (int i = 0; i < n; ++i) {
        x0[i] = std::min(std::max(x0[i], lo_[i]), hi_[i]);
        if (x0[i] - lo_[i] < rhobeg) {
            if (x0[i] - lo_[i] <= 0.5 * rhobeg) {
                x0[i] = lo_[i];
                side[i] = +1;
            } else {
                x0[i] = lo_[i] + rhobeg;
            }
        } else if (hi_[i] - x0[i] < rhobeg) {
            if (hi_[i] - x0[i] <= 0.5 * rhobeg) {
                x0[i] = hi_[i];
                side[i] = -1;
            } else {
                x0[i] = hi_[i] - rhobeg;
            }
        }
    }
    // 2n+1 initial points in one batch
    Y_.assign(m_, x0);
    for (int i = 0; i < n; ++i) {
        if (side[i] == 0) {
            Y_[1 + 2 * i][i] += rhobeg;
            Y_[2 + 2 * i][i] -= rhobeg;
        } else {
            Y_[1 + 2 * i][i] += side[i] * rhobeg;
            Y_[2 + 2 * i][i] += side[i] * 2.0 * rhobeg;
        }
    }
    eval(Y_, F_);
    // initial model: gradient and diagonal curvature through the three points per axis
    Q_.n = n;
    Q_.xm = x0;
    Q_.c = F_[0];
    Q_.g.assign(n, 0.0);
    Q_.H.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        const double a = Y_[1 + 2 * i][i] - x0[i], b = Y_[2 + 2 * i][i] - x0[i];
        const double fa = F_[1 + 2 * i] - F_[0], fb = F_[2 + 2 * i] - F_[0];
        // through (0, 0), (a, fa), (b, fb)
        Q_.g[i] = (fa * b * b - fb * a * a) / (a * b * (b - a));
        Q_.H[(size_t)i * n + i] = 2.0 * (fa * b - fb * a) / (a * b * (a - b));
    }
    double rho = rhobeg, delta = rhobeg;
    res.code = 4;  // XTOL_REACHED
    res.message = "trust region radius below tolerance";
    // farthest interpolation point from the best one
    auto farthest = [&](int kopt, double* dist) {
        int kf = -1;
        double df = 0.0;
        for (int k = 0; k < m_; ++k) {
            double s2 = 0.0;
            for (int i = 0; i < n; ++i) s2 += (Y_[k][i] - Y_[kopt][i]) * (Y_[k][i] - Y_[kopt][i]);
            if (s2 > df) {
                df = s2;
                kf = k;
            }
        }
        *dist = std::sqrt(df);
        return kf;
    };
    // Powell's rho schedule; false when rho_end is already reached (stop)
    auto reduce_rho = [&]() {
        if (rho <= rhoend) return false;
        const double rnew = rho > 250.0 * rhoend ? 0.1 * rho
                            : rho > 16.0 * rhoend ? std::sqrt(rho * rhoend)
                                                  : rhoend;
        delta = std::max(0.5 * rho, rnew);
        rho = rnew;
        return true;
    };
    bool geometry_next = false;
    while (true) {
        if (nevals_ >= maxeval_) {
            res.code = 5;
            res.message = "maxeval reached";
            break;
        }
        const int kopt = best_index();
        const Vec xopt = Y_[kopt];
        refresh_kkt(xopt, rho);
        double dfar = 0.0;
        const int kfar = farthest(kopt, &dfar);
        if (geometry_next) {
            // a poorly placed point moves to where its Lagrange function is largest
            geometry_next = false;
            const double dgeo = std::max(std::min(0.1 * dfar, delta), rho);
            const Vec xg = geometry_step(kfar, dgeo);
            Vec Fg;
            eval({xg}, Fg);
            const Model old = Q_;
            Y_[kfar] = xg;
            F_[kfar] = Fg[0];
            refresh_kkt(Y_[best_index()], rho);
            if (!update_model(old)) {
                res.code = -1;
                res.message = "interpolation system singular";
                break;
            }
            continue;
        }
        double pred = 0.0;
        const Vec d = trust_step(xopt, delta, &pred);
        const double dnorm = std::sqrt(dot(d, d));
        if (dnorm < 0.5 * rho || !(pred > 0.0)) {
            // no useful step at this scale: improve the geometry if a point is far,
            // otherwise refine rho
            delta = std::max(0.5 * delta, rho);
            if (kfar >= 0 && dfar > 10.0 * rho) {
                geometry_next = true;
                continue;
            }
            if (!reduce_rho()) break;
            continue;
        }
        Vec xn(n);
        for (int i = 0; i < n; ++i) xn[i] = xopt[i] + d[i];
        Vec Fn;
        eval({xn}, Fn);
        const double fnew = Fn[0], fopt = F_[kopt];
        const double ratio = (fopt - fnew) / pred;
        if (ratio <= 0.1) delta = std::min(0.5 * delta, dnorm);
        else if (ratio <= 0.7) delta = std::max(0.5 * delta, dnorm);
        else delta = std::max(0.5 * delta, 2.0 * dnorm);
        if (delta <= 1.5 * rho) delta = rho;
        // the point to replace: largest weighted |Lagrange value| at the new point
        Vec ell;
        if (!lagrange_values(xn, ell)) {
            res.code = -1;
            res.message = "interpolation system singular";
            break;
        }
        const bool better = fnew < fopt;
        int knew = -1;
        double score = -1.0;
        for (int k = 0; k < m_; ++k) {
            if (k == kopt && !better) continue;  // a failed step keeps the best point
            double q = 0.0;
            for (int i = 0; i < n; ++i) {
                const double dv = Y_[k][i] - (better ? xn[i] : xopt[i]);
                q += dv * dv;
            }
            q /= delta * delta;
            const double sc = std::max(1.0, q * q) * std::fabs(ell[k]);
            if (sc > score) {
                score = sc;
                knew = k;
            }
        }
        if (knew < 0) break;
        const Model old = Q_;
        Y_[knew] = xn;
        F_[knew] = fnew;
        refresh_kkt(Y_[best_index()], rho);
        if (!update_model(old)) {
            res.code = -1;
            res.message = "interpolation system singular";
            break;
        }
        if (ratio > 0.1) continue;
        // a poor step: move a far point first, else keep trying at this rho while the
        // step could still succeed, else refine rho
        double df2 = 0.0;
        farthest(best_index(), &df2);
        if (df2 > std::max(2.0 * delta, 10.0 * rho)) {
            geometry_next = true;
            continue;
        }
        if (ratio > 0.0 || std::max(delta, dnorm) > rho) continue;
        if (!reduce_rho()) break;
    }
    const int kb = best_index();
    res.x = Y_[kb];
    res.f = F_[kb];
    res.nevals = nevals_;
    if (res.code < 0) res.code = -1;
    return res;
}

}  // namespace

Result bobyqa(const BatchObjective& f, const std::vector<double>& x0, const std::vector<double>& lb,
              const std::vector<double>& ub, const std::vector<double>& step, double xtol_rel,
              int maxeval) {
    Result res;
    const int n = (int)x0.size();
    if (n == 0 || lb.size() != x0.size() || ub.size() != x0.size() || step.size() != x0.size()) {
        res.code = -2;
        res.message = "invalid dimensions";
        return res;
    }
    // scaled variables u = x / step: the initial step is 1 along every axis
    std::vector<double> s(n), lo(n), hi(n), u0(n);
    double range_min = std::numeric_limits<double>::infinity();
    for (int i = 0; i < n; ++i) {
        if (!(ub[i] >= lb[i]) || !(step[i] > 0.0)) {
            res.code = -2;
            res.message = "invalid bounds or step";
            return res;
        }
        s[i] = step[i];
        lo[i] = lb[i] / s[i];
        hi[i] = ub[i] / s[i];
        u0[i] = x0[i] / s[i];
        range_min = std::min(range_min, hi[i] - lo[i]);
    }
    // BOBYQA needs every range to hold two initial steps
    const double rhobeg = std::min(1.0, 0.5 * range_min);
    if (!(rhobeg > 0.0)) {
        res.code = -2;
        res.message = "a lower bound equals its upper bound";
        return res;
    }
    const double rhoend = std::max(xtol_rel * rhobeg, 1e-12 * rhobeg);
    auto fs = [&](int k, int d, const double* U, double* F) {
        std::vector<double> X((size_t)k * d);
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < d; ++i) X[(size_t)j * d + i] = U[(size_t)j * d + i] * s[i];
        f(k, d, X.data(), F);
    };
    BatchObjective bf = fs;
    Bobyqa b(bf, lo, hi, maxeval > 0 ? maxeval : std::numeric_limits<int>::max());
    res = b.run(u0, rhobeg, rhoend);
    for (int i = 0; i < n; ++i) res.x[i] *= s[i];
    return res;
}

std::vector<double> default_initial_step(const std::vector<double>& x, const std::vector<double>& lb,
                                         const std::vector<double>& ub) {
    std::vector<double> dx(x.size());
    for (size_t i = 0; i < x.size(); ++i) {
        double step = std::numeric_limits<double>::infinity();
        const bool flo = std::isfinite(lb[i]), fhi = std::isfinite(ub[i]);
        if (flo && fhi && ub[i] > lb[i]) step = 0.25 * (ub[i] - lb[i]);
        if (fhi && ub[i] - x[i] < step && ub[i] > x[i]) step = 0.75 * (ub[i] - x[i]);
        if (flo && x[i] - lb[i] < step && x[i] > lb[i]) step = 0.75 * (x[i] - lb[i]);
        if (!std::isfinite(step)) step = x[i] != 0.0 ? std::fabs(x[i]) : 1.0;
        if (!(step > 0.0)) step = 1.0;
        dx[i] = step;
    }
    return dx;
}

}  // namespace lc2opt
