// C ABI of the registration layer (include/lc2reg.h).
#include "lc2reg.h"

#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "lc2_registration.hpp"
#include "lc2e_geometry.hpp"
#include "optim.hpp"
#include "yaml_lite.hpp"

namespace {

thread_local std::string g_err;

int fail(const std::exception& e) {
    g_err = e.what();
    return -1;
}

template <class Interp, class Opt>
int run(const char* yaml, int write_images, int nparams, int mdim, double* out_params,
        double* out_lc2, int* out_nevals, float* out_matrix, bool igtl = false) {
    Interp interp = [&]() -> Interp {
        if constexpr (std::is_same<Interp, Interpolator3D>::value) return Interp{std::string(yaml), igtl};
        else return Interp{std::string(yaml)};
    }();
    interp.InitiateMappers();
    interp.InitiateRegistrationHandle();
    Opt opt{std::string(yaml)};
    opt.LinkToInterpolator(&interp);
    opt.Optimize();
    opt.GetOptimalTransform();
    opt.ExportResults();
    std::vector<float> m(mdim * mdim);
    opt.GetInternalMatrix(m.data());
    if (write_images) interp.WriteOut(m.data());
    if (out_params)
        for (int i = 0; i < nparams; ++i) out_params[i] = opt.GetOptimalSolution()[i];
    if (out_lc2) *out_lc2 = opt.GetLastSimilarity();
    if (out_nevals) *out_nevals = opt.GetEvaluationCount();
    if (out_matrix) std::memcpy(out_matrix, m.data(), m.size() * sizeof(float));
    return 0;
}

lc2opt::BatchObjective wrap(lc2reg_batch_fn f, void* user) {
    return [f, user](int n, int d, const double* X, double* F) { f(n, d, X, F, user); };
}

}  // namespace

extern "C" {

const char* lc2reg_last_error(void) { return g_err.c_str(); }

int lc2reg_pair_geometry(const lc2e_geom* fixed, const lc2e_geom* moving,
                         const float resampled_spacing[3], int dim, int center_overlaid,
                         lc2e_pair_info* out) {
    if (!fixed || !moving || !resampled_spacing || !out || (dim != 2 && dim != 3)) {
        g_err = "lc2reg_pair_geometry: invalid argument";
        return -1;
    }
    const int src_ndim = moving->dims[2] > 1 ? 3 : 2;
    const lc2e::PairGeometry pg = lc2e::make_pair_geometry(*fixed, *moving, resampled_spacing, dim,
                                                           src_ndim, center_overlaid != 0);
    *out = lc2e_pair_info{};
    out->virt = pg.virt;
    for (int a = 0; a < 3; ++a) {
        out->fixed_center[a] = pg.fixed_center[a];
        out->shift[a] = pg.shift[a];
    }
    out->grid_ndim = dim;
    out->src_ndim = src_ndim;
    out->n_frames = dim == 2 ? pg.virt.dims[2] : 1;
    return 0;
}

int lc2reg_pose_to_matrix(const lc2e_pair_info* info, const double* params, int n_params,
                          float out12[12]) {
    if (!info || !params || !out12 || lc2e::pose_to_matrix(*info, params, n_params, out12) != 0) {
        g_err = "lc2reg_pose_to_matrix: invalid argument";
        return -1;
    }
    return 0;
}

int lc2reg_moving_to_fixed(const lc2e_pair_info* info, const double* params, int n_params,
                           float* out) {
    float m[12];
    if (lc2reg_pose_to_matrix(info, params, n_params, m) != 0) return -1;
    if (!lc2e::moving_to_fixed(m, info->src_ndim == 3 ? 4 : 3, out)) {
        g_err = "lc2reg_moving_to_fixed: singular transform";
        return -1;
    }
    return 0;
}

int lc2reg_lc2_class_metric(int x, int y, int z, const float* us, const float* hr, int patch_size,
                            double* out_lc2, float* out_gradient, float* out_weighted_map) {
    if (x <= 0 || y <= 0 || z <= 0 || !us || !hr || patch_size < 1) {
        g_err = "lc2reg_lc2_class_metric: invalid argument";
        return -1;
    }
    try {
        const size_t n = (size_t)x * y * z;
        std::vector<float> U(us, us + n), H(hr, hr + n);
        std::unique_ptr<LC2_3D_class> c;
        if (z == 1) c = std::make_unique<LC2_2D_class>(x, y, U, H);
        else c = std::make_unique<LC2_3D_class>(x, y, z, U, H);
        c->PrepareGradientFilter();
        c->CalculateGradient();
        const double l = z == 1 ? static_cast<LC2_2D_class*>(c.get())->GetSimilarityMetric(16, 16, patch_size)
                                : c->GetSimilarityMetric(8, 8, 8, patch_size);
        if (out_lc2) *out_lc2 = l;
        if (out_gradient) {
            const std::vector<float> g = c->GetHighResolutionGradientBuffer();
            std::memcpy(out_gradient, g.data(), n * sizeof(float));
        }
        if (out_weighted_map) {
            const std::vector<float> w = c->GetWeightedMapBuffer();
            std::memcpy(out_weighted_map, w.data(), n * sizeof(float));
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int lc2reg_register_yaml(const char* mode, const char* yaml_path, int write_images,
                         double* out_params, double* out_lc2, int* out_nevals, float* out_matrix) {
    return lc2reg_register_yaml_igtl(mode, yaml_path, write_images, 0, out_params, out_lc2,
                                     out_nevals, out_matrix);
}

int lc2reg_register_yaml_igtl(const char* mode, const char* yaml_path, int write_images,
                              int igtl_port, double* out_params, double* out_lc2, int* out_nevals,
                              float* out_matrix) {
    try {
        const std::string m = mode ? mode : "";
        if (igtl_port > 0) lc2reg::set_igtl_port(igtl_port);
        if (m == "3D")
            return run<Interpolator3D, Optimizers3D>(yaml_path, write_images, 12, 4, out_params,
                                                     out_lc2, out_nevals, out_matrix, igtl_port > 0);
        if (m == "2D")
            return run<Interpolator2D, Optimizers2D>(yaml_path, write_images, 6, 3, out_params,
                                                     out_lc2, out_nevals, out_matrix);
        if (m == "2DTo3D")
            return run<Interpolator2DTo3D, Optimizers2DTo3D>(yaml_path, write_images, 12, 4,
                                                             out_params, out_lc2, out_nevals,
                                                             out_matrix);
        g_err = "mode must be 2D, 3D or 2DTo3D";
        return -1;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int lc2reg_evaluate_yaml(const char* mode, const char* yaml_path, const double* params, int n,
                         double* out_lc2) {
    try {
        const std::string m = mode ? mode : "";
        std::unique_ptr<lc2reg::InterpolatorBase> ip;
        // one evaluation streams nothing: no OpenIGTLink server (the constructor's default
        // opens one and waits for its client)
        if (m == "3D") ip.reset(new Interpolator3D(std::string(yaml_path), false));
        else if (m == "2D") ip.reset(new Interpolator2D(std::string(yaml_path)));
        else if (m == "2DTo3D") ip.reset(new Interpolator2DTo3D(std::string(yaml_path)));
        else {
            g_err = "mode must be 2D, 3D or 2DTo3D";
            return -1;
        }
        ip->InitiateMappers();
        ip->InitiateRegistrationHandle();
        std::vector<double> l;
        ip->EvaluateBatch({std::vector<double>(params, params + n)}, l);
        if (out_lc2) *out_lc2 = l[0];
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int lc2reg_minimize_nlopt(int algorithm, int d, const double* x0, const double* lb,
                          const double* ub, const double* step, double xtol_rel, int maxeval,
                          lc2reg_batch_fn f, void* user, double* x_out, double* f_out,
                          int* nevals) {
    if (d <= 0 || !lb || !ub || !f || algorithm < 0 || algorithm > 2 || (algorithm == 2 && !x0)) {
        g_err = "lc2reg_minimize_nlopt: invalid argument";
        return -2;
    }
    try {
        lc2opt::opt o((lc2opt::algorithm)algorithm, (unsigned)d);
        o.set_batch_objective(wrap(f, user));
        o.set_lower_bounds(std::vector<double>(lb, lb + d));
        o.set_upper_bounds(std::vector<double>(ub, ub + d));
        o.set_xtol_rel(xtol_rel);
        if (maxeval > 0) o.set_maxeval(maxeval);
        if (step) o.set_initial_step(std::vector<double>(step, step + d));
        std::vector<double> x = x0 ? std::vector<double>(x0, x0 + d) : std::vector<double>(d, 0.0);
        double fx = 0.0;
        const lc2opt::result rc = o.optimize(x, fx);
        if (nevals) *nevals = o.get_numevals();
        if (rc < 0) {
            g_err = "lc2reg_minimize_nlopt: optimizer failed";
            return (int)rc;
        }
        if (x_out) std::memcpy(x_out, x.data(), d * sizeof(double));
        if (f_out) *f_out = fx;
        return (int)rc;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int lc2reg_minimize_direct(int d, const double* lb, const double* ub, int maxeval,
                           lc2reg_batch_fn f, void* user, double* x_out, double* f_out,
                           int* nevals) {
    try {
        lc2opt::Result r = lc2opt::direct_l(wrap(f, user), std::vector<double>(lb, lb + d),
                                            std::vector<double>(ub, ub + d), maxeval);
        if (r.code < 0) {
            g_err = r.message;
            return -1;
        }
        for (int i = 0; i < d; ++i) x_out[i] = r.x[i];
        if (f_out) *f_out = r.f;
        if (nevals) *nevals = r.nevals;
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int lc2reg_minimize_qmtr(int d, const double* x0, const double* lb, const double* ub,
                         double xtol_rel, int maxeval, lc2reg_batch_fn f, void* user,
                         double* x_out, double* f_out, int* nevals) {
    try {
        lc2opt::Result r = lc2opt::qmtr(wrap(f, user), std::vector<double>(x0, x0 + d),
                                        std::vector<double>(lb, lb + d),
                                        std::vector<double>(ub, ub + d), xtol_rel, maxeval);
        if (r.code < 0) {
            g_err = r.message;
            return -1;
        }
        for (int i = 0; i < d; ++i) x_out[i] = r.x[i];
        if (f_out) *f_out = r.f;
        if (nevals) *nevals = r.nevals;
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int lc2reg_yaml_roundtrip(const char* in_path, const char* out_path) {
    try {
        lc2yaml::save_file(lc2yaml::load_file(in_path), out_path);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

}  // extern "C"
