// Registration facade (include/lc2_registration.hpp) over the engine's C ABI.
#include "lc2_registration.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <iostream>
#include <limits>
#include <stdexcept>

#include "image_io.hpp"
#include "lc2e_geometry.hpp"
#include "optim.hpp"
#include "yaml_lite.hpp"

namespace lc2reg {

namespace {

void check(int rc, const char* what) {
    if (rc < 0) throw std::runtime_error(std::string(what) + ": " + lc2e_last_error());
}

const lc2yaml::Node& req(const lc2yaml::Node& root, const std::string& path) {
    const lc2yaml::Node* n = lc2yaml::find(root, path);
    if (!n || n->is_null()) throw std::runtime_error("config: missing key " + path);
    return *n;
}

const lc2yaml::Node* opt(const lc2yaml::Node& root, const std::string& path) {
    const lc2yaml::Node* n = lc2yaml::find(root, path);
    return (n && !n->is_null()) ? n : nullptr;
}

void read_xyz(const lc2yaml::Node& root, const std::string& path, float* out, int n) {
    const char* k[3] = {"x", "y", "z"};
    for (int a = 0; a < n; ++a) {
        const lc2yaml::Node* v = opt(root, path + "/" + k[a]);
        if (v) out[a] = (float)v->as_double();
    }
}

void read_xyz_int(const lc2yaml::Node& root, const std::string& path, int* out, int n) {
    const char* k[3] = {"x", "y", "z"};
    for (int a = 0; a < n; ++a) {
        const lc2yaml::Node* v = opt(root, path + "/" + k[a]);
        if (v) out[a] = v->as_int();
    }
}

std::string base_name(const std::string& path) {
    const size_t s = path.find_last_of("/\\");
    std::string b = s == std::string::npos ? path : path.substr(s + 1);
    const size_t d = b.find_last_of('.');
    return d == std::string::npos ? b : b.substr(0, d);
}

}  // namespace

lc2e_engine* shared_engine() {
    static lc2e_engine* e = [] {
        lc2e_engine* h = nullptr;
        size_t budget = 0;
        if (const char* b = std::getenv("LC2E_HBM_BUDGET")) budget = (size_t)std::strtoull(b, nullptr, 10);
        check(lc2e_create(0, budget, &h), "lc2e_create");
        return h;
    }();
    return e;
}

// ---------------------------------------------------------------------------
// Interpolators
// ---------------------------------------------------------------------------

namespace {
int g_igtl_port = 23333;
}  // namespace

void set_igtl_port(int port) { g_igtl_port = port; }
int igtl_port() { return g_igtl_port; }

// interpolator.cu:359-388: server on the port, then wait (polling every 200 ms) for the
// client before anything else happens.
void InterpolatorBase::start_igtl() {
    if (lc2igtl_server_create(g_igtl_port, 0, &igtl_) < 0)
        throw std::runtime_error(std::string("Cannot create a server socket: ") + lc2igtl_last_error());
    std::cout << "Registration Server: server created, localhost:" << lc2igtl_port(igtl_) << std::endl;
    if (lc2igtl_wait_client(igtl_, -1.0) != 1)
        throw std::runtime_error(std::string("OpenIGTLink: ") + lc2igtl_last_error());
    std::cout << "Registration Server: client connected" << std::endl;
}

// interpolator.cu:545-573: the current transform (GetMyselfInversed of VirtualToMoving) as
// TRANSFORM, device "RegistrationServer", header version 2.
void InterpolatorBase::stream_pose(const float m12[12]) {
    if (!igtl_) return;
    float inv[16];
    if (!lc2e::invert_affine(m12, inv)) return;
    if (lc2igtl_send_transform(igtl_, "RegistrationServer", inv) < 0)
        std::cerr << "OpenIGTLink: " << lc2igtl_last_error() << std::endl;
}

InterpolatorBase::~InterpolatorBase() {
    if (igtl_) lc2igtl_server_destroy(igtl_);
    if (pair_) lc2e_pair_destroy(pair_);
    if (fixed_.vol) lc2e_volume_destroy(fixed_.vol);
    if (moving_.vol) lc2e_volume_destroy(moving_.vol);
}

// Interpolator3D/2D/2DTo3D YAML constructors (interpolator.cu:234-391, :923-1103,
// :1538-1633): image paths and geometry, HighResolutionModality, ResampledSpacing,
// FinalRoundExportRootPath, CenterOverlaid, DisplayPattern.
void InterpolatorBase::load_from_yaml(const std::string& path, int grid_ndim, bool raw_only) {
    InputParameterPath = path;
    grid_ndim_ = grid_ndim;
    const lc2yaml::Node y = lc2yaml::load_file(path);
    const bool is2d = (grid_ndim == 2 && !raw_only);
    const std::string fk = is2d ? "InputParameters/FixedImage" : "InputParameters/FixedVolume";
    const std::string mk = is2d ? "InputParameters/MovingImage" : "InputParameters/MovingVolume";
    const int nd = is2d ? 2 : 3;
    auto load = [&](const std::string& key, LoadedImage& out) {
        int dims[3] = {1, 1, 1};
        float sp[3] = {1, 1, 1}, org[3] = {0, 0, 0};
        read_xyz_int(y, key + "/Dimension", dims, nd);
        read_xyz(y, key + "/Spacing", sp, nd);
        read_xyz(y, key + "/Origin", org, nd);
        lc2e_dtype dt = LC2E_F32;
        if (const lc2yaml::Node* t = opt(y, key + "/DataType")) dt = lc2io::parse_dtype(t->str());
        out.path = req(y, key + "/FilePath").str();
        lc2io::Image im = lc2io::read_image(out.path, dims, sp, org, dt);
        // YAML geometry, when given, overrides what a self-describing file says
        // (interpolator.cu:1013-1060)
        for (int a = 0; a < nd; ++a) {
            const char* k[3] = {"x", "y", "z"};
            if (opt(y, key + "/Spacing/" + k[a])) im.geom.spacing[a] = sp[a];
            if (opt(y, key + "/Origin/" + k[a])) im.geom.origin[a] = org[a];
        }
        out.geom = im.geom;
        check(lc2e_volume_upload(shared_engine(), im.bytes.data(), im.dtype, &im.geom,
                                 LC2E_STORE_AUTO, &out.vol),
              "volume upload");
    };
    load(fk, fixed_);
    load(mk, moving_);
    FixedFilePath = fixed_.path;
    MovingFilePath = moving_.path;
    HighResolutionModality = req(y, "InputParameters/RegistrationParameters/HighResolutionModality").str();
    read_xyz(y, "InputParameters/RegistrationParameters/ResampledSpacing", resampled_, nd);
    if (nd == 2) resampled_[2] = fixed_.geom.spacing[2];
    FinalExportPath = req(y, "InputParameters/RegistrationParameters/FinalRoundExportRootPath").str();
    Centered = req(y, "InputParameters/RegistrationParameters/CenterOverlaid").as_bool();
    if (const lc2yaml::Node* d = opt(y, "InputParameters/RegistrationParameters/DisplayPattern"))
        DisplayPattern = d->as_int();
    if (const lc2yaml::Node* p = opt(y, "InputParameters/RegistrationParameters/PatchSize"))
        patch_size_ = p->as_int();
    if (HighResolutionModality != "Fixed" && HighResolutionModality != "Moving")
        throw std::runtime_error("config: HighResolutionModality must be Fixed or Moving");
}

void InterpolatorBase::load_from_buffers(const float* fixed, const lc2e_geom& fg,
                                         const float* moving, const lc2e_geom& mg,
                                         const float resampled[3], bool high_res_moving,
                                         bool centered, int patch, int grid_ndim,
                                         const std::string& export_path) {
    if (!fixed || !moving) throw std::runtime_error("configuration has no image buffers");
    fixed_.geom = fg;
    moving_.geom = mg;
    check(lc2e_volume_upload(shared_engine(), fixed, LC2E_F32, &fg, LC2E_STORE_AUTO, &fixed_.vol),
          "volume upload");
    check(lc2e_volume_upload(shared_engine(), moving, LC2E_F32, &mg, LC2E_STORE_AUTO, &moving_.vol),
          "volume upload");
    for (int a = 0; a < 3; ++a) resampled_[a] = resampled[a];
    HighResolutionModality = high_res_moving ? "Moving" : "Fixed";
    Centered = centered;
    patch_size_ = patch;
    grid_ndim_ = grid_ndim;
    FinalExportPath = export_path;
}

// InitiateMappers (interpolator.cu:495-529): centres and shift are fixed at pair creation.
void InterpolatorBase::InitiateMappers() {}

// InitiateRegistrationHandle (interpolator.cu:531-533): the engine pair replaces the
// LC2 class buffers.
void InterpolatorBase::InitiateRegistrationHandle() {
    if (pair_) lc2e_pair_destroy(pair_);
    pair_ = nullptr;
    if (patch_size_ == 0 && grid_ndim_ != 3)
        throw std::runtime_error("patchSize 0 (mutual information) needs a 3D grid");
    // a mutual-information pair still carries LC2 statistics; radius 1 keeps them cheap
    check(lc2e_pair_create(shared_engine(), fixed_.vol, moving_.vol, resampled_,
                           HighResolutionModality == "Moving", Centered ? 1 : 0,
                           patch_size_ > 0 ? patch_size_ : 1, grid_ndim_, &pair_),
          "pair create");
    check(lc2e_pair_get_info(pair_, &info_), "pair info");
    // Graph replay only on request (LC2E_GRAPHS=1): on MI355X with ROCm 7.2 a replayed
    // batch took 5-13 us longer than the same batch launched eagerly, at every batch size.
    const char* g = std::getenv("LC2E_GRAPHS");
    if (g && std::atoi(g) > 0) check(lc2e_pair_use_graph(pair_, 1), "graph mode");
}

void InterpolatorBase::set_patch_size(int patchSize) {
    if (patchSize < 0) throw std::runtime_error("negative patch size");
    if (pair_ && patchSize == patch_size_) return;
    patch_size_ = patchSize;
    InitiateRegistrationHandle();
}

void InterpolatorBase::set_pose(const double* params, int n) {
    if (!pair_) InitiateRegistrationHandle();
    check(lc2e_pose_to_matrix(pair_, params, n, current_), "pose");
    stream_pose(current_);
    ++counter;
}

// GetSimilarityMeasure (interpolator.cu:617-631): LC2 of the current pose, or the
// normalized mutual information for patchSize 0 (interpolator.cu:620-622).
double InterpolatorBase::similarity(int patchSize) {
    set_patch_size(patchSize);
    double l = 0.0;
    if (patchSize == 0) check(lc2e_eval_mi_batch(pair_, 1, current_, 256, 256, &l), "mutual information");
    else check(lc2e_eval(pair_, current_, &l), "eval");
    return l;
}

int InterpolatorBase::EvaluateBatch(const std::vector<std::vector<double>>& params,
                                    std::vector<double>& lc2) {
    if (!pair_) InitiateRegistrationHandle();
    std::vector<float> mats(params.size() * 12);
    for (size_t i = 0; i < params.size(); ++i) {
        check(lc2e_pose_to_matrix(pair_, params[i].data(), (int)params[i].size(), &mats[i * 12]),
              "pose");
        stream_pose(&mats[i * 12]);  // the reference sends at every GenerateMappers
    }
    lc2.assign(params.size(), 0.0);
    const int rc = patch_size_ == 0
                       ? lc2e_eval_mi_batch(pair_, (int)params.size(), mats.data(), 256, 256, lc2.data())
                       : lc2e_eval_batch(pair_, (int)params.size(), mats.data(), lc2.data(), nullptr, nullptr);
    check(rc, "eval batch");
    return rc;
}

std::vector<float> InterpolatorBase::FetchStage(lc2e_stage stage) {
    if (!pair_) InitiateRegistrationHandle();
    const size_t V = (size_t)info_.virt.dims[0] * info_.virt.dims[1] * info_.virt.dims[2];
    if (stage == LC2E_STAGE_W || stage == LC2E_STAGE_MEASURE) {
        std::vector<double> d(V);
        check(lc2e_debug_fetch(pair_, stage, current_, d.data()), "fetch");
        return std::vector<float>(d.begin(), d.end());
    }
    std::vector<float> out(V);
    check(lc2e_debug_fetch(pair_, stage, current_, out.data()), "fetch");
    return out;
}

// GetMovingToFixedTransform (interpolator.cu:633-679, :1382-1425): VirtualToFixed (identity)
// times inverse(VirtualToMoving), bottom row forced.
void InterpolatorBase::moving_to_fixed(const double* params, int n, float* out, int out_dim) {
    if (!pair_) InitiateRegistrationHandle();
    float m[12];
    check(lc2e_pose_to_matrix(pair_, params, n, m), "pose");
    if (!lc2e::moving_to_fixed(m, out_dim, out)) throw std::runtime_error("singular transform");
}

// WriteOut (interpolator.cu:681-920): both images on the virtual grid as NRRD, and the
// moving-to-fixed matrix as text.
void InterpolatorBase::write_out(const float* matrix, int mdim) {
    const std::vector<float> hr = FetchStage(LC2E_STAGE_WARP_I);
    const std::vector<float> us = FetchStage(LC2E_STAGE_WARP_US);
    const bool hr_is_fixed = HighResolutionModality == "Fixed";
    const std::string dir = FinalExportPath.empty() ? "." : FinalExportPath;
    lc2io::write_nrrd(dir + "/" + base_name(FixedFilePath.empty() ? "fixed" : FixedFilePath) + ".nrrd",
                      (hr_is_fixed ? hr : us).data(), info_.virt);
    lc2io::write_nrrd(dir + "/" + base_name(MovingFilePath.empty() ? "moving" : MovingFilePath) +
                          "_Registered.nrrd",
                      (hr_is_fixed ? us : hr).data(), info_.virt);
    if (matrix) {
        FILE* f = std::fopen((dir + "/MovingToFixed.mat").c_str(), "w");
        if (!f) throw std::runtime_error("cannot write MovingToFixed.mat");
        for (int r = 0; r < mdim; ++r) {
            for (int c = 0; c < mdim; ++c) std::fprintf(f, "%.9g%s", matrix[r * mdim + c], c + 1 < mdim ? " " : "\n");
        }
        std::fclose(f);
    }
}

}  // namespace lc2reg

// ---- Interpolator3D ----
Interpolator3D::Interpolator3D(const std::string yaml_filePath, bool useIGTL) {
    load_from_yaml(yaml_filePath, 3, false);
    if (useIGTL) start_igtl();
}

Interpolator3D::Interpolator3D(const LC2Configuration3D* c) {
    if (!c || !c->IsConfigured) throw std::runtime_error("LC2Configuration3D is not configured");
    lc2e_geom fg{{c->FixedDimension.x, c->FixedDimension.y, c->FixedDimension.z},
                 {c->FixedSpacing.x, c->FixedSpacing.y, c->FixedSpacing.z},
                 {c->FixedOrigin.x, c->FixedOrigin.y, c->FixedOrigin.z}};
    lc2e_geom mg{{c->MovingDimension.x, c->MovingDimension.y, c->MovingDimension.z},
                 {c->MovingSpacing.x, c->MovingSpacing.y, c->MovingSpacing.z},
                 {c->MovingOrigin.x, c->MovingOrigin.y, c->MovingOrigin.z}};
    const float rs[3] = {c->SamplerSpacing.x, c->SamplerSpacing.y, c->SamplerSpacing.z};
    // interpolator.cu:430-436: IsFixedHighResolution takes precedence
    load_from_buffers(c->FixedBuffer, fg, c->MovingBuffer, mg, rs,
                      !c->IsFixedHighResolution && c->IsMovingHighResolution, c->IsCenterOverlaid,
                      c->PatchSize, 3, c->ExportPath);
    FixedFilePath = c->FixedImageFilePath;
    MovingFilePath = c->MovingImageFilePath;
}

void Interpolator3D::GenerateMappers(float3 R, float3 T) {
    const double p[6] = {T.x, T.y, T.z, R.x, R.y, R.z};
    set_pose(p, 6);
}

void Interpolator3D::GenerateMappers(float3 R, float3 T, float3 S, float3 H) {
    const double p[12] = {T.x, T.y, T.z, R.x, R.y, R.z, S.x, S.y, S.z, H.x, H.y, H.z};
    set_pose(p, 12);
}

void Interpolator3D::GetCurrentTransform(float t[16]) {
    if (!lc2e::invert_affine(current_, t)) throw std::runtime_error("singular transform");
}

void Interpolator3D::Interpolate(const int, const int, const int) {}

double Interpolator3D::GetSimilarityMeasure(const int patchSize) { return similarity(patchSize); }

void Interpolator3D::GetMovingToFixedTransform(float3 R, float3 T, float* out) {
    const double p[6] = {T.x, T.y, T.z, R.x, R.y, R.z};
    moving_to_fixed(p, 6, out, 4);
}

void Interpolator3D::GetMovingToFixedTransform(float3 R, float3 T, float3 S, float3 H, float* out) {
    const double p[12] = {T.x, T.y, T.z, R.x, R.y, R.z, S.x, S.y, S.z, H.x, H.y, H.z};
    moving_to_fixed(p, 12, out, 4);
}

void Interpolator3D::WriteOut() { write_out(nullptr, 4); }
void Interpolator3D::WriteOut(float* m) { write_out(m, 4); }

// ---- Interpolator2D ----
Interpolator2D::Interpolator2D(const std::string yaml_filePath) { load_from_yaml(yaml_filePath, 2, false); }

Interpolator2D::Interpolator2D(const LC2Configuration2D* c) {
    if (!c || !c->IsConfigured) throw std::runtime_error("LC2Configuration2D is not configured");
    lc2e_geom fg{{c->FixedDimension.x, c->FixedDimension.y, 1},
                 {c->FixedSpacing.x, c->FixedSpacing.y, 1.0f},
                 {c->FixedOrigin.x, c->FixedOrigin.y, 0.0f}};
    lc2e_geom mg{{c->MovingDimension.x, c->MovingDimension.y, 1},
                 {c->MovingSpacing.x, c->MovingSpacing.y, 1.0f},
                 {c->MovingOrigin.x, c->MovingOrigin.y, 0.0f}};
    const float rs[3] = {c->SamplerSpacing.x, c->SamplerSpacing.y, 1.0f};
    load_from_buffers(c->FixedBuffer, fg, c->MovingBuffer, mg, rs, !c->IsFixedHighResolution && c->IsMovingHighResolution,
                      c->IsCenterOverlaid, c->PatchSize, 2, c->ExportPath);
    FixedFilePath = c->FixedImageFilePath;
    MovingFilePath = c->MovingImageFilePath;
}

void Interpolator2D::GenerateMappers(float R, float2 T) {
    const double p[3] = {T.x, T.y, R};
    set_pose(p, 3);
}

void Interpolator2D::GenerateMappers(float R, float2 T, float2 S, float H) {
    const double p[6] = {T.x, T.y, R, S.x, S.y, H};
    set_pose(p, 6);
}

void Interpolator2D::GetCurrentTransform(float t[9]) {
    float inv[16];
    if (!lc2e::invert_affine(current_, inv)) throw std::runtime_error("singular transform");
    const float o[9] = {inv[0], inv[1], inv[3], inv[4], inv[5], inv[7], 0, 0, 1};
    for (int i = 0; i < 9; ++i) t[i] = o[i];
}

void Interpolator2D::Interpolate(const int, const int) {}

double Interpolator2D::GetSimilarityMeasure(const int patchSize) { return similarity(patchSize); }

void Interpolator2D::GetMovingToFixedTransform(float R, float2 T, float* out) {
    const double p[3] = {T.x, T.y, R};
    moving_to_fixed(p, 3, out, 3);
}

void Interpolator2D::GetMovingToFixedTransform(float R, float2 T, float2 S, float H, float* out) {
    const double p[6] = {T.x, T.y, R, S.x, S.y, H};
    moving_to_fixed(p, 6, out, 3);
}

void Interpolator2D::WriteOut() { write_out(nullptr, 3); }
void Interpolator2D::WriteOut(float* m) { write_out(m, 3); }

// ---- Interpolator2DTo3D ----
Interpolator2DTo3D::Interpolator2DTo3D(const std::string yaml_filePath) { load_from_yaml(yaml_filePath, 2, true); }

// Declared by the reference (interpolator.cuh:219) but never defined there: the fixed image
// is the 2D slice (FixedDimension.z = 1), the moving one the 3D volume; 2D LC2 on the slice.
Interpolator2DTo3D::Interpolator2DTo3D(const LC2Configuration2DTo3D* c) {
    if (!c || !c->IsConfigured) throw std::runtime_error("LC2Configuration2DTo3D is not configured");
    lc2e_geom fg{{c->FixedDimension.x, c->FixedDimension.y, std::max(1, c->FixedDimension.z)},
                 {c->FixedSpacing.x, c->FixedSpacing.y, c->FixedSpacing.z},
                 {c->FixedOrigin.x, c->FixedOrigin.y, c->FixedOrigin.z}};
    lc2e_geom mg{{c->MovingDimension.x, c->MovingDimension.y, c->MovingDimension.z},
                 {c->MovingSpacing.x, c->MovingSpacing.y, c->MovingSpacing.z},
                 {c->MovingOrigin.x, c->MovingOrigin.y, c->MovingOrigin.z}};
    const float rs[3] = {c->SamplerSpacing.x, c->SamplerSpacing.y, c->SamplerSpacing.z};
    load_from_buffers(c->FixedBuffer, fg, c->MovingBuffer, mg, rs,
                      !c->IsFixedHighResolution && c->IsMovingHighResolution, c->IsCenterOverlaid,
                      c->PatchSize, 2, c->ExportPath);
    FixedFilePath = c->FixedImageFilePath;
    MovingFilePath = c->MovingImageFilePath;
}

void Interpolator2DTo3D::GenerateMappers(float3 R, float3 T) {
    const double p[6] = {T.x, T.y, T.z, R.x, R.y, R.z};
    set_pose(p, 6);
}

void Interpolator2DTo3D::GenerateMappers(float3 R, float3 T, float3 S, float3 H) {
    const double p[12] = {T.x, T.y, T.z, R.x, R.y, R.z, S.x, S.y, S.z, H.x, H.y, H.z};
    set_pose(p, 12);
}

void Interpolator2DTo3D::GetCurrentTransform(float t[16]) {
    if (!lc2e::invert_affine(current_, t)) throw std::runtime_error("singular transform");
}

void Interpolator2DTo3D::Interpolate(const int, const int, const int) {}

double Interpolator2DTo3D::GetSimilarityMeasure(const int patchSize) { return similarity(patchSize); }

void Interpolator2DTo3D::GetMovingToFixedTransform(float3 R, float3 T, float* out) {
    const double p[6] = {T.x, T.y, T.z, R.x, R.y, R.z};
    moving_to_fixed(p, 6, out, 4);
}

void Interpolator2DTo3D::GetMovingToFixedTransform(float3 R, float3 T, float3 S, float3 H, float* out) {
    const double p[12] = {T.x, T.y, T.z, R.x, R.y, R.z, S.x, S.y, S.z, H.x, H.y, H.z};
    moving_to_fixed(p, 12, out, 4);
}

void Interpolator2DTo3D::WriteOut() { write_out(nullptr, 4); }
void Interpolator2DTo3D::WriteOut(float* m) { write_out(m, 4); }

// ---------------------------------------------------------------------------
// LC2 classes on images that already share a grid
// ---------------------------------------------------------------------------
LC2_3D_class::LC2_3D_class(int x, int y, int z, std::vector<float>& US, std::vector<float>& HR) {
    dims_[0] = x;
    dims_[1] = y;
    dims_[2] = z;
    us_ = US;
    hr_ = HR;
}

LC2_3D_class::~LC2_3D_class() {
    if (pair_) lc2e_pair_destroy(pair_);
    if (vus_) lc2e_volume_destroy(vus_);
    if (vhr_) lc2e_volume_destroy(vhr_);
}

void LC2_3D_class::PortedFromInterpolator(int x, int y, int z) {
    dims_[0] = x;
    dims_[1] = y;
    dims_[2] = z;
    us_.assign((size_t)x * y * z, 0.0f);
    hr_.assign((size_t)x * y * z, 0.0f);
}

// Both images on a unit grid: the identity pose reproduces them exactly, so the engine's
// LC2 at identity is the LC2 of the given grids.
void LC2_3D_class::ensure(int patchSize) {
    if (pair_ && patch_ == patchSize) return;
    if (pair_) lc2e_pair_destroy(pair_);
    pair_ = nullptr;
    const size_t n = (size_t)dims_[0] * dims_[1] * dims_[2];
    if (us_.size() != n || hr_.size() != n) throw std::runtime_error("LC2 class: buffer sizes differ from dims");
    if (!vus_) {
        lc2e_geom g{{dims_[0], dims_[1], dims_[2]}, {1, 1, 1}, {0, 0, 0}};
        lc2reg::shared_engine();
        if (lc2e_volume_upload(lc2reg::shared_engine(), us_.data(), LC2E_F32, &g, LC2E_F32, &vus_) < 0 ||
            lc2e_volume_upload(lc2reg::shared_engine(), hr_.data(), LC2E_F32, &g, LC2E_F32, &vhr_) < 0)
            throw std::runtime_error(lc2e_last_error());
    }
    const float rs[3] = {1, 1, 1};
    if (lc2e_pair_create(lc2reg::shared_engine(), vus_, vhr_, rs, 1, 0, patchSize, ndim_, &pair_) < 0)
        throw std::runtime_error(lc2e_last_error());
    patch_ = patchSize;
}

double LC2_3D_class::GetSimilarityMetric(const int, const int, const int, const int patchSize) {
    ensure(patchSize);
    const float ident[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    double l = 0.0;
    if (lc2e_eval(pair_, ident, &l) < 0) throw std::runtime_error(lc2e_last_error());
    return l;
}

double LC2_3D_class::GetMutualInformationSimilarityMetric(const int, const int, const int,
                                                         const int bins_usVolume,
                                                         const int bins_highResoVolume) {
    ensure(patch_ > 0 ? patch_ : 3);
    const float ident[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    double v = 0.0;
    if (lc2e_eval_mi_batch(pair_, 1, ident, bins_usVolume, bins_highResoVolume, &v) < 0)
        throw std::runtime_error(lc2e_last_error());
    return v;
}

std::vector<float> LC2_3D_class::GetUltrasoundBuffer() { return us_; }
std::vector<float> LC2_3D_class::GetHighResolutionBuffer() { return hr_; }

std::vector<float> LC2_3D_class::GetHighResolutionGradientBuffer() {
    ensure(patch_ > 0 ? patch_ : 3);
    const float ident[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    std::vector<float> g(hr_.size());
    if (lc2e_debug_fetch(pair_, LC2E_STAGE_GRAD, ident, g.data()) < 0) throw std::runtime_error(lc2e_last_error());
    return g;
}

std::vector<float> LC2_3D_class::GetWeightedMapBuffer() {
    ensure(patch_ > 0 ? patch_ : 3);
    const float ident[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    std::vector<double> w(hr_.size()), m(hr_.size());
    if (lc2e_debug_fetch(pair_, LC2E_STAGE_W, ident, w.data()) < 0 ||
        lc2e_debug_fetch(pair_, LC2E_STAGE_MEASURE, ident, m.data()) < 0)
        throw std::runtime_error(lc2e_last_error());
    std::vector<float> out(hr_.size(), 0.0f);
    for (size_t i = 0; i < out.size(); ++i)
        if (w[i] > 0.0) out[i] = (float)(m[i] / w[i]);
    return out;
}

LC2_2D_class::LC2_2D_class(int x, int y, std::vector<float>& US, std::vector<float>& HR)
    : LC2_3D_class(x, y, 1, US, HR) {
    set_ndim(2);
}

double LC2_2D_class::GetSimilarityMetric(const int, const int, const int patchSize) {
    set_ndim(2);
    return LC2_3D_class::GetSimilarityMetric(8, 8, 8, patchSize);
}

// ---------------------------------------------------------------------------
// Optimizers
// ---------------------------------------------------------------------------
namespace lc2reg {

namespace {
const char* kKeys3D[12] = {"TX", "TY", "TZ", "RX", "RY", "RZ", "ScaleX", "ScaleY", "ScaleZ",
                           "ShearX", "ShearY", "ShearZ"};
const char* kKeys2D[6] = {"TX", "TY", "R", "ScaleX", "ScaleY", "Shear"};
const char* kEntryNames[4] = {"Translation", "Rotation", "Scale", "Shear"};
}  // namespace

OptimizerBase::OptimizerBase(int nparams, int mdim) : nparams_(nparams), mdim_(mdim) {
    MovingToFixed_Optimal.assign(mdim * mdim, 0.0f);
    for (int i = 0; i < mdim; ++i) MovingToFixed_Optimal[i * mdim + i] = 1.0f;
}

OptimizerBase::~OptimizerBase() = default;

// Optimizers*::Optimizers* (optimizers.cu:23-106, :484-582, :1109-1199): initial values,
// bounds (translation and rotation bounds relative to the initial values, :49-57,
// :516-524), DIRECT settings, BOBYQA rounds "0".."k-1", PatchSize.
void OptimizerBase::load_yaml(const std::string& path) {
    YAML_ParameterPath = path;
    yaml_ = std::make_shared<lc2yaml::Node>(lc2yaml::load_file(path));
    const lc2yaml::Node& y = *yaml_;
    const char** keys = nparams_ == 12 ? kKeys3D : kKeys2D;
    auto vec = [&](const std::string& tag) {
        std::vector<double> v(nparams_);
        for (int i = 0; i < nparams_; ++i)
            v[i] = req(y, "InputParameters/OptimizerSettings/" + tag + "/" + keys[i]).as_double();
        return v;
    };
    InitialValues = vec("InitialValues");
    OptimalSolution = InitialValues;
    LowerBounds = vec("LowerBounds");
    UpperBounds = vec("UpperBounds");
    const int nrel = nparams_ == 12 ? 6 : 3;
    for (int i = 0; i < nrel; ++i) {
        LowerBounds[i] += InitialValues[i];
        UpperBounds[i] += InitialValues[i];
    }
    cfg_.direct_enabled = req(y, "InputParameters/OptimizerSettings/DIRECT/Enabled").as_bool();
    for (int e = 0; e < 4; ++e)
        cfg_.direct_entries[e] = req(y, std::string("InputParameters/OptimizerSettings/DIRECT/OptimizedEntries/") + kEntryNames[e]).as_bool();
    cfg_.direct_maxeval = req(y, "InputParameters/OptimizerSettings/DIRECT/MaxEvaluations").as_int();
    const lc2yaml::Node& b = req(y, "InputParameters/OptimizerSettings/BOBYQA");
    cfg_.rounds.clear();
    for (size_t i = 0; i < b.size(); ++i) {
        const std::string k = std::to_string(i);
        if (!b.has(k)) throw std::runtime_error("config: BOBYQA rounds are not numbered 0..k-1");
        OptimizerSettings::Round r;
        r.enabled = b.at(k).at("Enabled").as_bool();
        for (int e = 0; e < 4; ++e) r.entries[e] = b.at(k).at("OptimizedEntries").at(kEntryNames[e]).as_bool();
        r.abs_tol = b.at(k).at("AbsTol").as_double();
        cfg_.rounds.push_back(r);
    }
    cfg_.patch_size = req(y, "InputParameters/RegistrationParameters/PatchSize").as_int();
}

int OptimizerBase::entry_width(int e) const { return nparams_ == 12 ? 3 : (e == 0 || e == 2 ? 2 : 1); }
int OptimizerBase::entry_offset(int e) const {
    if (nparams_ == 12) return 3 * e;
    static const int off[4] = {0, 2, 3, 5};
    return off[e];
}

// operator() (optimizers.cu:130-177, :650-728): x overlays the enabled entries onto the
// current optimum.
std::vector<double> OptimizerBase::full_params(const double* x) const {
    std::vector<double> p = OptimalSolution;
    int k = 0;
    for (int e : CurrentHolding_OptEntries)
        for (int j = 0; j < entry_width(e); ++j) p[entry_offset(e) + j] = (float)x[k++];
    for (auto& v : p) v = (float)v;  // the reference hands float parameters to the mappers
    return p;
}

void OptimizerBase::bounds_for(std::vector<double>& l, std::vector<double>& u, std::vector<double>& x) const {
    l.clear();
    u.clear();
    x.clear();
    for (int e : CurrentHolding_OptEntries)
        for (int j = 0; j < entry_width(e); ++j) {
            l.push_back(LowerBounds[entry_offset(e) + j]);
            u.push_back(UpperBounds[entry_offset(e) + j]);
            x.push_back(OptimalSolution[entry_offset(e) + j]);
        }
}

void OptimizerBase::assign_back(const std::vector<double>& x) {
    int k = 0;
    for (int e : CurrentHolding_OptEntries)
        for (int j = 0; j < entry_width(e); ++j) OptimalSolution[entry_offset(e) + j] = x[k++];
}

// RegistrationInfo.log header: the 3D layout (optimizers.cu:610) or the 2D->3D layout with
// its "Initial step" line (:1263-1264).
void OptimizerBase::log_header() {
    if (!m_logFileId.is_open()) return;
    static const char* kCols =
        "Translation_x,Translation_y,Translation_z,Rotation_x,Rotation_y,Rotation_z,Scale_x,"
        "Scale_y,Scale_z,Shear_x,Shear_y,Shear_z,Similarity";
    if (log_layout_ == LogLayout::Slice)
        m_logFileId << kCols << ",IteratorCounter\nInitial step\n";
    else
        m_logFileId << "Step,Transform[16]," << kCols << ",-log(sim),IteratorCounter\n";
}

// One CSV line per evaluation. 3D (optimizers.cu:709-720): step marker, the 16 entries of
// GetCurrentTransform (the moving-to-fixed matrix), the 12 parameters as floats, LC2,
// -log(LC2), evaluation count; the starting-pose line of ShowCurrentResults (:613-623) has
// no -log column. 2D->3D (:1335-1340, :1266-1272): parameters, LC2, evaluation count.
void OptimizerBase::log_eval(const std::vector<double>& p, double sim, bool initial) {
    if (!m_logFileId.is_open()) return;
    if (log_layout_ == LogLayout::Volume) {
        float m12[12], inv[16];
        m_logFileId << m_stepMarker << ",";
        if (lc2e_pose_to_matrix(interp_->pair(), p.data(), (int)p.size(), m12) == 0 &&
            lc2e::invert_affine(m12, inv))
            for (int i = 0; i < 16; ++i) m_logFileId << inv[i] << ",";
    }
    for (double v : p) m_logFileId << (float)v << ",";
    m_logFileId << sim << ",";
    if (log_layout_ == LogLayout::Volume && !initial) m_logFileId << -std::log(sim) << ",";
    m_logFileId << EvalCounter << "\n";
}

void OptimizerBase::evaluate_batch(int n, int d, const double* X, double* F) {
    if (!interp_) throw std::runtime_error("optimizer is not linked to an interpolator");
    for (int base = 0; base < n; base += max_batch_) {
        const int m = std::min(max_batch_, n - base);
        std::vector<std::vector<double>> params(m);
        for (int i = 0; i < m; ++i) params[i] = full_params(X + (size_t)(base + i) * d);
        std::vector<double> lc2;
        interp_->EvaluateBatch(params, lc2);
        for (int i = 0; i < m; ++i) {
            ++EvalCounter;
            SimilarityMeasure = lc2[i];
            // optimizers.cu:722-727: mutual information is maximised as is, LC2 as -log
            if (interp_->patch_size() == 0) F[base + i] = -lc2[i];
            else F[base + i] = lc2[i] > 0.0 ? -std::log(lc2[i]) : std::numeric_limits<double>::infinity();
            log_eval(params[i], lc2[i]);
        }
    }
}

double OptimizerBase::operator()(const std::vector<double>& x) {
    double f = 0.0;
    evaluate_batch(1, (int)x.size(), x.data(), &f);
    return f;
}

double OptimizerBase::wrapper(const std::vector<double>& x, std::vector<double>&, void* data) {
    return reinterpret_cast<OptimizerBase*>(data)->operator()(x);
}

// ShowCurrentResults (optimizers.cu:588-648): one evaluation at the starting pose, logged.
void OptimizerBase::ShowCurrentResults() {
    std::vector<std::vector<double>> p = {full_params(nullptr)};
    CurrentHolding_OptEntries.clear();
    p[0] = full_params(nullptr);
    std::vector<double> l;
    interp_->EvaluateBatch(p, l);
    SimilarityMeasure = l[0];
    log_header();
    log_eval(p[0], l[0], true);
}

// Optimize (optimizers.cu:183-410, :734-970): DIRECT stage over its entries, then each
// enabled BOBYQA round over its entries, minimising -log(LC2) (or -NMI for patchSize 0);
// each stage starts from and writes back to the running optimum.
void OptimizerBase::Optimize() {
    if (!interp_) throw std::runtime_error("optimizer is not linked to an interpolator");
    // the optimizer's PatchSize (read at optimizers.cu:577) is what it evaluates with (:705)
    interp_->set_patch_size(cfg_.patch_size);
    auto sim_of = [this](double f) { return interp_->patch_size() == 0 ? -f : std::exp(-f); };
    // optimizers.cu:736 (3D); the 2D optimizer writes no log (the reference's has none), the
    // 2D->3D log goes next to the other outputs as well (the reference: the working
    // directory, :1351)
    const std::string logp = interp_->GetOutputPath().empty() ? std::string("./RegistrationInfo.log")
                                                              : interp_->GetOutputPath() + "/RegistrationInfo.log";
    if (log_layout_ != LogLayout::None) m_logFileId.open(logp, std::ofstream::out);
    m_stepMarker = -1;
    ShowCurrentResults();
    auto obj = [this](int n, int d, const double* X, double* F) { evaluate_batch(n, d, X, F); };
    if (cfg_.direct_enabled) {
        m_stepMarker = 0;
        CurrentHolding_OptEntries.clear();
        for (int e = 0; e < 4; ++e)
            if (cfg_.direct_entries[e]) CurrentHolding_OptEntries.push_back(e);
        if (!CurrentHolding_OptEntries.empty()) {
            std::vector<double> l, u, x;
            bounds_for(l, u, x);
            if (!DirectLower.empty()) {  // the programmatic configuration's DIRECT bounds
                int k = 0;
                for (int e : CurrentHolding_OptEntries)
                    for (int j = 0; j < entry_width(e); ++j, ++k) {
                        l[k] = DirectLower[entry_offset(e) + j];
                        u[k] = DirectUpper[entry_offset(e) + j];
                    }
            }
            // GN_ORIG_DIRECT_L in 3D and 2D->3D (optimizers.cu:854, :1473), GN_DIRECT_L in 2D (:298)
            lc2opt::opt o(direct_original_ ? lc2opt::GN_ORIG_DIRECT_L : lc2opt::GN_DIRECT_L,
                          (unsigned)x.size());
            o.set_batch_objective(obj);
            o.set_lower_bounds(l);
            o.set_upper_bounds(u);
            o.set_maxeval(cfg_.direct_maxeval);
            double minf = 0.0;
            const lc2opt::result rc = o.optimize(x, minf);
            if (rc < 0) std::clog << "optimizer failed: code " << (int)rc << std::endl;
            else {
                assign_back(x);
                SimilarityMeasure = sim_of(minf);
                std::clog << "Optimization stage-DIRECT finished: similarity -> " << SimilarityMeasure
                          << ", accu evaluations -> " << EvalCounter << std::endl;
            }
        }
    }
    for (size_t k = 0; k < cfg_.rounds.size(); ++k) {
        const auto& rd = cfg_.rounds[k];
        if (!rd.enabled) continue;
        m_stepMarker = (int)k + 1;
        CurrentHolding_OptEntries.clear();
        for (int e = 0; e < 4; ++e)
            if (rd.entries[e]) CurrentHolding_OptEntries.push_back(e);
        if (CurrentHolding_OptEntries.empty()) continue;
        std::vector<double> l, u, x;
        bounds_for(l, u, x);
        double minf = 0.0;
        if (local_method_ == LocalMethod::Qmtr) {
            lc2opt::Result r = lc2opt::qmtr(obj, x, l, u, rd.abs_tol, 100000);
            if (r.code < 0) {
                std::clog << "optimizer failed: " << r.message << std::endl;
                continue;
            }
            x = r.x;
            minf = r.f;
        } else {
            // nlopt::LN_BOBYQA with set_xtol_rel(AbsTol) and NLopt's default initial step
            // (optimizers.cu:937-946)
            lc2opt::opt o(lc2opt::LN_BOBYQA, (unsigned)x.size());
            o.set_batch_objective(obj);
            o.set_lower_bounds(l);
            o.set_upper_bounds(u);
            o.set_xtol_rel(rd.abs_tol);
            if (rd.max_eval > 0) o.set_maxeval(rd.max_eval);
            if (rd.rho_begin > 0.0) {
                // the programmatic configuration's RhoBegin, as a fraction of each range
                std::vector<double> dx(x.size());
                for (size_t i = 0; i < x.size(); ++i) dx[i] = rd.rho_begin * (u[i] - l[i]);
                o.set_initial_step(dx);
            }
            const lc2opt::result rc = o.optimize(x, minf);
            if (rc < 0) {
                std::clog << "optimizer failed: code " << (int)rc << std::endl;
                continue;
            }
        }
        assign_back(x);
        SimilarityMeasure = sim_of(minf);
        std::clog << "Optimization stage-BOBYQA finished: similarity -> " << SimilarityMeasure
                  << ", accu evaluations -> " << EvalCounter << std::endl;
    }
    m_logFileId.close();
}

void OptimizerBase::GetInternalMatrix(float* out) {
    for (int i = 0; i < mdim_ * mdim_; ++i) out[i] = MovingToFixed_Optimal[i];
}

// "<config>_result.<ext>" next to the configuration. The reference inserts at the FIRST
// '.' of the whole path (optimizers.cu:461), which breaks "./rel/cfg.yaml"; this uses
// the extension's dot.
std::string OptimizerBase::result_path() const {
    const size_t slash = YAML_ParameterPath.find_last_of("/\\");
    const size_t dot = YAML_ParameterPath.find_last_of('.');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash))
        return YAML_ParameterPath + "_result";
    std::string p = YAML_ParameterPath;
    p.insert(dot, "_result");
    return p;
}

// ExportResults (optimizers.cu:1040-1106, :435-480): the configuration with the optimum as
// new InitialValues, DIRECT and BOBYQA round 0 disabled (so the file resumes the run),
// plus Outputs.
void OptimizerBase::ExportResults() {
    if (!yaml_) throw std::runtime_error("ExportResults needs a YAML-configured optimizer");
    lc2yaml::Node& y = *yaml_;
    if (update_initial_on_export_) {
        const char** keys = nparams_ == 12 ? kKeys3D : kKeys2D;
        lc2yaml::Node& iv = y["InputParameters"]["OptimizerSettings"]["InitialValues"];
        for (int i = 0; i < nparams_; ++i) iv[keys[i]] = lc2yaml::Node::number((float)OptimalSolution[i], 9);
        y["InputParameters"]["OptimizerSettings"]["DIRECT"]["Enabled"] = lc2yaml::Node::boolean(false);
        y["InputParameters"]["OptimizerSettings"]["BOBYQA"]["0"]["Enabled"] = lc2yaml::Node::boolean(false);
    }
    lc2yaml::Node& out = y["Outputs"];
    out["IsRegistered"] = lc2yaml::Node::boolean(true);
    lc2yaml::Node tm = lc2yaml::Node::seq(false);
    for (int r = 0; r < mdim_; ++r) {
        lc2yaml::Node row = lc2yaml::Node::seq(false);
        for (int c = 0; c < mdim_; ++c) row.push_back(lc2yaml::Node::number(MovingToFixed_Optimal[r * mdim_ + c], 9));
        tm.push_back(row);
    }
    out["TransformMatrix"] = tm;
    out["LastSimilarity"] = lc2yaml::Node::number(SimilarityMeasure, 17);
    out["NumEvaluations"] = lc2yaml::Node::number(EvalCounter, 10);
    const std::time_t t = std::chrono::system_clock::to_time_t(std::chrono::system_clock::now());
    out["RegistrationDateTime"] = lc2yaml::Node::scalar(std::string(std::ctime(&t)), true);
    lc2yaml::save_file(y, result_path());
    std::clog << "Last Similarity Estimate: " << SimilarityMeasure << ", with " << EvalCounter
              << " evaluations. " << std::endl;
}

}  // namespace lc2reg

Optimizers3D::Optimizers3D(std::string FilePath) : OptimizerBase(12, 4) { load_yaml(FilePath); }

// Optimizers*(const LC2Configuration*): declared by the reference (optimizers.cuh:102, :178)
// but never defined there. The configuration's optimizer members map onto the YAML's
// OptimizerSettings as follows (this build's reading of them):
//   InitialValues = init{Translation, Rotation, Scale, Shear};
//   bounds (BOBYQA rounds): translation +- initFRE * boundTranslation_BOBYQA (mm), rotation
//     +- initFRE * boundRotation_BOBYQA (degrees), scale +- 0.2 * boundScale_BOBYQA, shear
//     +- 0.2 * boundShear_BOBYQA, about the initial values;
//   DIRECT stage (enabled when MaxEval > 0): translation only, +- initFRE *
//     boundTranslation_DIRECT, MaxEval evaluations;
//   BOBYQA round 0: translation and rotation; round 1 (IsAffine only): all parameters; each
//     with xtol_rel = Tol, at most MaxIterations evaluations and an initial step of RhoBegin
//     times each range.
void lc2reg::OptimizerBase::apply_config(const std::vector<double>& initial,
                                 const std::vector<double>& halfwidth,
                                 const std::vector<double>& direct_halfwidth, int direct_maxeval,
                                 bool affine, double rho_begin, double tol, int max_iterations,
                                 int patch) {
    InitialValues = initial;
    OptimalSolution = initial;
    LowerBounds.resize(nparams_);
    UpperBounds.resize(nparams_);
    DirectLower.assign(nparams_, 0.0);
    DirectUpper.assign(nparams_, 0.0);
    for (int i = 0; i < nparams_; ++i) {
        LowerBounds[i] = initial[i] - halfwidth[i];
        UpperBounds[i] = initial[i] + halfwidth[i];
        DirectLower[i] = initial[i] - direct_halfwidth[i];
        DirectUpper[i] = initial[i] + direct_halfwidth[i];
    }
    cfg_.direct_enabled = direct_maxeval > 0;
    cfg_.direct_entries = {true, false, false, false};
    cfg_.direct_maxeval = direct_maxeval;
    cfg_.rounds.clear();
    for (int k = 0; k < (affine ? 2 : 1); ++k) {
        lc2reg::OptimizerSettings::Round rd;
        rd.enabled = true;
        rd.entries = k == 0 ? std::vector<bool>{true, true, false, false}
                            : std::vector<bool>{true, true, true, true};
        rd.abs_tol = tol;
        rd.max_eval = max_iterations;
        rd.rho_begin = rho_begin;
        cfg_.rounds.push_back(rd);
    }
    cfg_.patch_size = patch;
}

namespace {
void config_vectors_3d(const LC2Configuration3D* c, std::vector<double>& x, std::vector<double>& h,
                       std::vector<double>& hd) {
    const double F = c->initFRE;
    x = {c->initTranslation.x, c->initTranslation.y, c->initTranslation.z, c->initRotation.x,
         c->initRotation.y, c->initRotation.z, c->initScale.x, c->initScale.y, c->initScale.z,
         c->initShear.x, c->initShear.y, c->initShear.z};
    h = {F * c->boundTranslation_BOBYQA.x, F * c->boundTranslation_BOBYQA.y,
         F * c->boundTranslation_BOBYQA.z, F * c->boundRotation_BOBYQA.x,
         F * c->boundRotation_BOBYQA.y, F * c->boundRotation_BOBYQA.z,
         0.2 * c->boundScale_BOBYQA.x, 0.2 * c->boundScale_BOBYQA.y, 0.2 * c->boundScale_BOBYQA.z,
         0.2 * c->boundShear_BOBYQA.x, 0.2 * c->boundShear_BOBYQA.y, 0.2 * c->boundShear_BOBYQA.z};
    hd = h;
    hd[0] = F * c->boundTranslation_DIRECT.x;
    hd[1] = F * c->boundTranslation_DIRECT.y;
    hd[2] = F * c->boundTranslation_DIRECT.z;
}
}  // namespace

Optimizers3D::Optimizers3D(const LC2Configuration3D* c) : OptimizerBase(12, 4) {
    if (!c) throw std::runtime_error("null configuration");
    std::vector<double> x, h, hd;
    config_vectors_3d(c, x, h, hd);
    apply_config(x, h, hd, c->MaxEval, c->IsAffine, c->RhoBegin, c->Tol, (int)c->MaxIterations,
                 c->PatchSize);
}

void Optimizers3D::GetOptimalTransform(float* out) {
    static_cast<Interpolator3D*>(interp_)->GetMovingToFixedTransform(
        make_float3((float)OptimalSolution[3], (float)OptimalSolution[4], (float)OptimalSolution[5]),
        make_float3((float)OptimalSolution[0], (float)OptimalSolution[1], (float)OptimalSolution[2]),
        make_float3((float)OptimalSolution[6], (float)OptimalSolution[7], (float)OptimalSolution[8]),
        make_float3((float)OptimalSolution[9], (float)OptimalSolution[10], (float)OptimalSolution[11]),
        out);
}
void Optimizers3D::GetOptimalTransform() { GetOptimalTransform(MovingToFixed_Optimal.data()); }

Optimizers2D::Optimizers2D(std::string FilePath) : OptimizerBase(6, 3) {
    load_yaml(FilePath);
    direct_original_ = false;        // GN_DIRECT_L (optimizers.cu:298)
    update_initial_on_export_ = false;  // the 2D export keeps the inputs (optimizers.cu:451-480)
    log_layout_ = LogLayout::None;
}

// Not declared by the reference; the 2D counterpart of Optimizers3D(const
// LC2Configuration3D*), with the same mapping of the configuration.
Optimizers2D::Optimizers2D(const LC2Configuration2D* c) : OptimizerBase(6, 3) {
    if (!c) throw std::runtime_error("null configuration");
    const double F = c->initFRE;
    const std::vector<double> x = {c->initTranslation.x, c->initTranslation.y, c->initRotation,
                                   c->initScale.x, c->initScale.y, c->initShear};
    const std::vector<double> h = {F * c->boundTranslation_BOBYQA.x, F * c->boundTranslation_BOBYQA.y,
                                   F * c->boundRotation_BOBYQA, 0.2 * c->boundScale_BOBYQA.x,
                                   0.2 * c->boundScale_BOBYQA.y, 0.2 * c->boundShear_BOBYQA};
    std::vector<double> hd = h;
    hd[0] = F * c->boundTranslation_DIRECT.x;
    hd[1] = F * c->boundTranslation_DIRECT.y;
    apply_config(x, h, hd, c->MaxEval, c->IsAffine, c->RhoBegin, c->Tol, (int)c->MaxIterations,
                 c->PatchSize);
    direct_original_ = false;
    update_initial_on_export_ = false;
    log_layout_ = LogLayout::None;
}

void Optimizers2D::GetOptimalTransform(float* out) {
    static_cast<Interpolator2D*>(interp_)->GetMovingToFixedTransform(
        (float)OptimalSolution[2], make_float2((float)OptimalSolution[0], (float)OptimalSolution[1]),
        make_float2((float)OptimalSolution[3], (float)OptimalSolution[4]), (float)OptimalSolution[5],
        out);
}
void Optimizers2D::GetOptimalTransform() { GetOptimalTransform(MovingToFixed_Optimal.data()); }

Optimizers2DTo3D::Optimizers2DTo3D(std::string FilePath) : OptimizerBase(12, 4) {
    load_yaml(FilePath);
    log_layout_ = LogLayout::Slice;
}

Optimizers2DTo3D::Optimizers2DTo3D(const LC2Configuration3D* c) : OptimizerBase(12, 4) {
    if (!c) throw std::runtime_error("null configuration");
    std::vector<double> x, h, hd;
    config_vectors_3d(c, x, h, hd);
    apply_config(x, h, hd, c->MaxEval, c->IsAffine, c->RhoBegin, c->Tol, (int)c->MaxIterations,
                 c->PatchSize);
    log_layout_ = LogLayout::Slice;
}

void Optimizers2DTo3D::GetOptimalTransform(float* out) {
    static_cast<Interpolator2DTo3D*>(interp_)->GetMovingToFixedTransform(
        make_float3((float)OptimalSolution[3], (float)OptimalSolution[4], (float)OptimalSolution[5]),
        make_float3((float)OptimalSolution[0], (float)OptimalSolution[1], (float)OptimalSolution[2]),
        make_float3((float)OptimalSolution[6], (float)OptimalSolution[7], (float)OptimalSolution[8]),
        make_float3((float)OptimalSolution[9], (float)OptimalSolution[10], (float)OptimalSolution[11]),
        out);
}
void Optimizers2DTo3D::GetOptimalTransform() { GetOptimalTransform(MovingToFixed_Optimal.data()); }
