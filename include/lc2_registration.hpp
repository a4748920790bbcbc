// C++ registration facade over the LC2 engine (lc2e.h): the reference's public classes
// with their names and call sequence, so a program written against the reference's
// interpolator.cuh / cu_lc2_*.cuh / optimizers.cuh / RegistrationParametersConfig.h
// links against this library instead:
//
//   Interpolator3D interp(cfg_yaml);          // interpolator.cuh:43-130
//   interp.InitiateMappers();                 // centres, shift, virtual grid
//   interp.InitiateRegistrationHandle();      // engine pair (resident volumes, LC2 buffers)
//   Optimizers3D opt(cfg_yaml);               // optimizers.cuh:100-177
//   opt.LinkToInterpolator(&interp);
//   opt.Optimize();                           // DIRECT stage + BOBYQA rounds, -log(LC2)
//   opt.GetOptimalTransform(); opt.ExportResults();  interp.WriteOut(matrix);
//
// Differences from the reference, by design:
//  * Interpolate() only records the pose: the warp is fused into the similarity kernel;
//    GetSimilarityMeasure() runs warp + Sobel + LC2 in one launch.
//  * Buffer getters return host copies (the reference's 3D getters return device
//    vectors by value, cu_lc2_3d.cu:1335-1341, which makes its 3D LC2 read zeros).
//  * patchSize 0 selects the normalized mutual information (cu_lc2_3d.cu:1196-1301,
//    256 x 256 bins) for 3D grids, as the reference does; the optimizers then maximise it.
//    2D grids reject patchSize 0 (the reference runs LC2 with an empty patch there).
//  * The optimizers' NLopt calls are served by lc2opt::opt (optim.hpp): DIRECT-L (both
//    NLopt variants) and Powell's BOBYQA; DIRECT iterations and BOBYQA's initial
//    interpolation set are evaluated in batches (one fused launch each).
//  * Images come from raw files or NRRD (the reference reads through ITK).
//  * Interpolator3D(yaml, useIGTL = true) opens the OpenIGTLink server and waits for its
//    client as the reference does; the pose of every evaluation (the optimizers' batches
//    included) is streamed, in evaluation order (lc2igtl.h).
#pragma once

#include <hip/hip_vector_types.h>

#include <fstream>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "lc2e.h"
#include "lc2igtl.h"

// ---- RegistrationParametersConfig.h (programmatic configuration) ----
struct regIntPoints3D { int x, y, z; };
struct regFloatingPoints3D { float x, y, z; };
struct regIntPoints2D { int x, y; };
struct regFloatingPoints2D { float x, y; };

// The reference's programmatic configuration (RegistrationParametersConfig.h:30-431): the
// same members with the same defaults. The image fields feed the Interpolator* constructors
// (interpolator.cu:393-492, :1105-1196); the optimizer fields feed Optimizers*(const
// LC2Configuration*) -- declared by the reference (optimizers.cuh:102, :178) but never
// defined there; here they map onto the YAML's OptimizerSettings as lc2reg::settings_from
// documents.
struct LC2Configuration2D {
    regIntPoints2D FixedDimension{0, 0};
    regFloatingPoints2D FixedSpacing{1.0f, 1.0f};
    regFloatingPoints2D FixedOrigin{0.0f, 0.0f};
    std::string FixedImageFilePath;
    regIntPoints2D MovingDimension{0, 0};
    regFloatingPoints2D MovingSpacing{1.0f, 1.0f};
    regFloatingPoints2D MovingOrigin{0.0f, 0.0f};
    std::string MovingImageFilePath;
    bool IsFixedHighResolution = true;
    bool IsMovingHighResolution = false;
    regFloatingPoints2D SamplerSpacing{1.0f, 1.0f};
    int PatchSize = 9;
    std::string ExportPath;
    bool IsCenterOverlaid = true;
    bool IsAffine = false;
    // optimizers
    float initFRE = 50.0f;
    regFloatingPoints2D initTranslation{0.0f, 0.0f};
    float initRotation = 0.0f;
    regFloatingPoints2D initScale{1.0f, 1.0f};
    float initShear = 0.0f;
    regFloatingPoints2D boundTranslation_DIRECT{0.8f, 0.5f};
    int MaxEval = 25;
    regFloatingPoints2D boundTranslation_BOBYQA{1.0f, 1.0f};
    float boundRotation_BOBYQA = 1.0f;
    regFloatingPoints2D boundScale_BOBYQA{1.0f, 1.0f};
    float boundShear_BOBYQA = 1.0f;
    float RhoBegin = 0.1f;
    float Tol = 1e-4f;
    float MaxIterations = 1000;
    bool IsConfigured = false;
    float* MovingBuffer = nullptr;  // caller-owned float32 images (x fastest)
    float* FixedBuffer = nullptr;
    // results
    bool IsConverged = false;
    float LastSimilarity = 0.0f;
    int NumberOfEvaluations = 0;
    std::vector<float> MovingToFixedMatrix = std::vector<float>{1, 0, 0, 0, 1, 0, 0, 0, 1};
    void DumpRegistrtationResults(const bool conv, const float sim, const int nev, const float* m) {
        IsConverged = conv;
        LastSimilarity = sim;
        NumberOfEvaluations = nev;
        MovingToFixedMatrix.assign(m, m + 9);
    }
};

// The 3D and 2D->3D configurations have the same members and defaults.
struct LC2Configuration3D {
    regIntPoints3D FixedDimension{0, 0, 0};
    regFloatingPoints3D FixedSpacing{1.0f, 1.0f, 1.0f};
    regFloatingPoints3D FixedOrigin{0.0f, 0.0f, 0.0f};
    std::string FixedImageFilePath;
    regIntPoints3D MovingDimension{0, 0, 0};
    regFloatingPoints3D MovingSpacing{1.0f, 1.0f, 1.0f};
    regFloatingPoints3D MovingOrigin{0.0f, 0.0f, 0.0f};
    std::string MovingImageFilePath;
    bool IsFixedHighResolution = false;
    bool IsMovingHighResolution = false;
    regFloatingPoints3D SamplerSpacing{1.0f, 1.0f, 1.0f};
    int PatchSize = 3;
    std::string ExportPath;
    bool IsCenterOverlaid = true;
    bool IsAffine = false;
    // optimizers
    float initFRE = 30.0f;
    regFloatingPoints3D initTranslation{0.0f, 0.0f, 0.0f};
    regFloatingPoints3D initRotation{0.0f, 0.0f, 0.0f};
    regFloatingPoints3D initScale{1.0f, 1.0f, 1.0f};
    regFloatingPoints3D initShear{0.0f, 0.0f, 0.0f};
    regFloatingPoints3D boundTranslation_DIRECT{0.5f, 0.5f, 0.5f};
    int MaxEval = 100;
    regFloatingPoints3D boundTranslation_BOBYQA{1.0f, 1.0f, 1.0f};
    regFloatingPoints3D boundRotation_BOBYQA{1.0f, 1.0f, 1.0f};
    regFloatingPoints3D boundScale_BOBYQA{1.0f, 1.0f, 1.0f};
    regFloatingPoints3D boundShear_BOBYQA{1.0f, 1.0f, 1.0f};
    float RhoBegin = 0.01f;
    float Tol = 1e-6f;
    float MaxIterations = 1000;
    bool IsConfigured = false;
    float* MovingBuffer = nullptr;
    float* FixedBuffer = nullptr;
    // results
    bool IsConverged = false;
    float LastSimilarity = 0.0f;
    int NumberOfEvaluations = 0;
    std::vector<float> MovingToFixedMatrix = std::vector<float>{1, 0, 0, 0, 0, 1, 0, 0,
                                                                0, 0, 1, 0, 0, 0, 0, 1};
    void DumpRegistrtationResults(const bool conv, const float sim, const int nev, const float* m) {
        IsConverged = conv;
        LastSimilarity = sim;
        NumberOfEvaluations = nev;
        MovingToFixedMatrix.assign(m, m + 16);
    }
};

struct LC2Configuration2DTo3D : LC2Configuration3D {};

namespace lc2yaml {
class Node;
}

namespace lc2reg {

// Shared engine state of one process (one device); created on first use.
lc2e_engine* shared_engine();
// TCP port of the OpenIGTLink server of Interpolator3D(yaml, useIGTL = true); the
// reference's is 23333 (interpolator.cu:366).
void set_igtl_port(int port);
int igtl_port();

// Geometry + data of one image as resident engine volume.
struct LoadedImage {
    lc2e_geom geom{};
    lc2e_volume* vol = nullptr;
    std::string path;
};

class InterpolatorBase {
public:
    virtual ~InterpolatorBase();
    void InitiateMappers();
    void InitiateRegistrationHandle();
    std::string GetOutputPath() const { return FinalExportPath; }
    // The engine pair behind the facade (valid after InitiateRegistrationHandle).
    lc2e_pair* pair() const { return pair_; }
    int patch_size() const { return patch_size_; }
    // Patch radius of later evaluations (0: mutual information, 3D grids only).
    void set_patch_size(int patchSize);
    // Current VirtualToMoving as 3x4 row-major.
    const float* current_matrix() const { return current_; }
    // One batch of full parameter vectors (12 for 3D sources, 6 for 2D) -> LC2 values (NMI
    // when the patch radius is 0).
    int EvaluateBatch(const std::vector<std::vector<double>>& params, std::vector<double>& lc2);
    // Warped images on the virtual grid at the current pose (host copies).
    std::vector<float> FetchStage(lc2e_stage stage);
    const lc2e_pair_info& info() const { return info_; }

protected:
    InterpolatorBase() = default;
    void load_from_yaml(const std::string& path, int grid_ndim, bool raw_only);
    void load_from_buffers(const float* fixed, const lc2e_geom& fg, const float* moving,
                           const lc2e_geom& mg, const float resampled[3], bool high_res_moving,
                           bool centered, int patch, int grid_ndim, const std::string& export_path);
    double similarity(int patchSize);
    void set_pose(const double* params, int n);
    void moving_to_fixed(const double* params, int n, float* out, int out_dim);
    void write_out(const float* matrix, int mdim);

    std::string InputParameterPath, FixedFilePath, MovingFilePath, FinalExportPath;
    std::string HighResolutionModality = "Fixed";
    bool Centered = false;
    int DisplayPattern = 0;
    int grid_ndim_ = 3;
    int patch_size_ = 3;
    float resampled_[3] = {1, 1, 1};
    LoadedImage fixed_, moving_;
    lc2e_pair* pair_ = nullptr;
    lc2e_pair_info info_{};
    float current_[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    int counter = 0;
    // OpenIGTLink pose stream (3D with useIGTL): every evaluated pose is sent as TRANSFORM
    lc2igtl_server* igtl_ = nullptr;
    void start_igtl();
    void stream_pose(const float m12[12]);
};

}  // namespace lc2reg

// ---- Interpolator (interpolator.cuh:43-292) ----
class Interpolator3D : public lc2reg::InterpolatorBase {
public:
    Interpolator3D(const std::string yaml_filePath, bool useIGTL = true);
    Interpolator3D(const LC2Configuration3D* ExternalConfig);
    void GenerateMappers(float3 RotationAngle, float3 Translation);
    void GenerateMappers(float3 RotationAngle, float3 Translation, float3 scale, float3 shear);
    void GetCurrentTransform(float _t[16]);
    void Interpolate(const int Blk_Dim_x = 8, const int Blk_Dim_y = 8, const int Blk_Dim_z = 8);
    double GetSimilarityMeasure(const int patchSize = 3);
    void GetMovingToFixedTransform(float3 RotationAngle, float3 Translation, float* outMovingToFixed);
    void GetMovingToFixedTransform(float3 RotationAngle, float3 Translation, float3 scale,
                                   float3 shear, float* outMovingToFixed);
    void WriteOut();
    void WriteOut(float* OptimizedMatrix);
};

class Interpolator2D : public lc2reg::InterpolatorBase {
public:
    Interpolator2D(const std::string yaml_filePath);
    Interpolator2D(const LC2Configuration2D* ExternalConfig);
    void GenerateMappers(float RotationAngle, float2 Translation);
    void GenerateMappers(float RotationAngle, float2 Translation, float2 scale, float shear);
    void GetCurrentTransform(float _t[9]);
    void Interpolate(const int Blk_Dim_x = 16, const int Blk_Dim_y = 16);
    double GetSimilarityMeasure(const int patchSize = 9);
    void GetMovingToFixedTransform(float RotationAngle, float2 Translation, float* outMovingToFixed);
    void GetMovingToFixedTransform(float RotationAngle, float2 Translation, float2 scale, float shear,
                                   float* outMovingToFixed);
    void WriteOut();
    void WriteOut(float* OptimizedMatrix);
};

class Interpolator2DTo3D : public lc2reg::InterpolatorBase {
public:
    Interpolator2DTo3D(const std::string yaml_filePath);
    Interpolator2DTo3D(const LC2Configuration2DTo3D* ExternalConfig);
    void GenerateMappers(float3 RotationAngle, float3 Translation);
    void GenerateMappers(float3 RotationAngle, float3 Translation, float3 scale, float3 shear);
    void GetCurrentTransform(float _t[16]);
    void Interpolate(const int Blk_Dim_x = 8, const int Blk_Dim_y = 8, const int Blk_Dim_z = 8);
    double GetSimilarityMeasure(const int patchSize = 9);
    void GetMovingToFixedTransform(float3 RotationAngle, float3 Translation, float* outMovingToFixed);
    void GetMovingToFixedTransform(float3 RotationAngle, float3 Translation, float3 scale,
                                   float3 shear, float* outMovingToFixed);
    void WriteOut();
    void WriteOut(float* OptimizedMatrix);
};

// ---- LC2 classes (cu_lc2_3d.cuh:10-93, cu_lc2_2d.cuh:10-81) over images already on a grid ----
class LC2_3D_class {
public:
    LC2_3D_class(int x, int y, int z, std::vector<float>& UltrasoundVolume_Extern,
                 std::vector<float>& HighResoVolume_Extern);
    LC2_3D_class() = default;
    ~LC2_3D_class();
    void PortedFromInterpolator(int x, int y, int z);
    void PrepareGradientFilter() {}
    void CalculateGradient(const int = 8, const int = 8, const int = 8) {}
    double GetSimilarityMetric(const int = 8, const int = 8, const int = 8, const int patchSize = 3);
    double GetMutualInformationSimilarityMetric(const int = 8, const int = 8, const int = 8,
                                                const int = 256, const int = 256);
    std::vector<float> GetUltrasoundBuffer();
    std::vector<float> GetHighResolutionBuffer();
    std::vector<float> GetHighResolutionGradientBuffer();
    std::vector<float> GetWeightedMapBuffer();  // measure / w per voxel (0 where w = 0)
    void ShowImages(int = 0) {}

private:
    void ensure(int patchSize);
    int dims_[3] = {0, 0, 0};
    std::vector<float> us_, hr_;
    lc2e_volume* vus_ = nullptr;
    lc2e_volume* vhr_ = nullptr;
    lc2e_pair* pair_ = nullptr;
    int patch_ = -1;
    int ndim_ = 3;

protected:
    void set_ndim(int n) { ndim_ = n; }
};

class LC2_2D_class : public LC2_3D_class {
public:
    LC2_2D_class(int x, int y, std::vector<float>& Ultrasound_Extern,
                 std::vector<float>& HighReso_Extern);
    LC2_2D_class() = default;
    void PortedFromInterpolator(int x, int y) { LC2_3D_class::PortedFromInterpolator(x, y, 1); }
    double GetSimilarityMetric(const int = 16, const int = 16, const int patchSize = 9);
};

// ---- Optimizers (optimizers.cuh:28-247) ----
enum OptimizationEntry { Translation, Rotation, Scale, Shear };

namespace lc2reg {

// Settings of one registration run (the YAML's OptimizerSettings).
struct OptimizerSettings {
    std::vector<double> initial, lower, upper;  // lower/upper of T and R are relative to initial
    bool direct_enabled = false;
    std::vector<bool> direct_entries = {true, false, false, false};
    int direct_maxeval = 50;
    struct Round {
        bool enabled = false;
        std::vector<bool> entries = {true, true, false, false};
        double abs_tol = 1e-4;
        int max_eval = 0;         // 0: no limit (the YAML rounds set none)
        double rho_begin = 0.0;   // initial step as a fraction of each range (0: NLopt's default)
    };
    std::vector<Round> rounds;
    int patch_size = 3;
};

class OptimizerBase {
public:
    virtual ~OptimizerBase();
    void ShowCurrentResults();
    double operator()(const std::vector<double>& x);
    static double wrapper(const std::vector<double>& x, std::vector<double>& grad, void* data);
    void Optimize();
    void GetInternalMatrix(float* OutMatrix);
    void ExportResults();
    const std::vector<double>& GetOptimalSolution() const { return OptimalSolution; }
    int GetEvaluationCount() const { return EvalCounter; }
    double GetLastSimilarity() const { return SimilarityMeasure; }
    // largest number of candidate poses per fused launch (default 64)
    void SetMaxBatch(int n) { max_batch_ = n > 0 ? n : 1; }
    // method of the BOBYQA rounds: Powell's BOBYQA (default, the reference's LN_BOBYQA) or
    // the batched diagonal-model trust region (more evaluations, fewer launches)
    enum class LocalMethod { Bobyqa, Qmtr };
    void SetLocalMethod(LocalMethod m) { local_method_ = m; }

protected:
    OptimizerBase(int nparams, int mdim);
    void load_yaml(const std::string& path);
    void evaluate_batch(int n, int d, const double* X, double* F);
    std::vector<double> full_params(const double* x) const;
    int entry_width(int entry) const;
    int entry_offset(int entry) const;
    void bounds_for(std::vector<double>& l, std::vector<double>& u, std::vector<double>& x) const;
    void assign_back(const std::vector<double>& x);
    std::string result_path() const;
    void log_header();
    void log_eval(const std::vector<double>& p, double sim, bool initial = false);
    void apply_config(const std::vector<double>& initial, const std::vector<double>& halfwidth,
                      const std::vector<double>& direct_halfwidth, int direct_maxeval,
                      bool affine, double rho_begin, double tol, int max_iterations, int patch);

    InterpolatorBase* interp_ = nullptr;
    OptimizerSettings cfg_;
    std::vector<double> InitialValues, LowerBounds, UpperBounds, OptimalSolution;
    std::vector<double> DirectLower, DirectUpper;  // DIRECT-stage bounds (configuration structs)
    std::vector<int> CurrentHolding_OptEntries;
    int nparams_, mdim_;
    int EvalCounter = 0;
    double SimilarityMeasure = 0.0;
    std::vector<float> MovingToFixed_Optimal;
    std::string YAML_ParameterPath;
    std::shared_ptr<lc2yaml::Node> yaml_;
    std::ofstream m_logFileId;
    int m_stepMarker = -1;
    bool direct_original_ = true;  // GN_ORIG_DIRECT_L (3D) vs GN_DIRECT_L (2D)
    // RegistrationInfo.log: none (2D, as the reference), the 3D layout (optimizers.cu:610,
    // :709-720) or the 2D->3D layout (:1263-1272, :1335-1340)
    enum class LogLayout { None, Volume, Slice };
    LogLayout log_layout_ = LogLayout::Volume;
    bool update_initial_on_export_ = true;
    int max_batch_ = 64;
    LocalMethod local_method_ = LocalMethod::Bobyqa;
};

}  // namespace lc2reg

class Optimizers3D : public lc2reg::OptimizerBase {
public:
    Optimizers3D(std::string FilePath);
    Optimizers3D(const LC2Configuration3D* ExternalConfig);
    void LinkToInterpolator(Interpolator3D* external_Interpolator) { interp_ = external_Interpolator; }
    void GetOptimalTransform(float* outMatrix);
    void GetOptimalTransform();
};

class Optimizers2D : public lc2reg::OptimizerBase {
public:
    Optimizers2D(std::string FilePath);
    Optimizers2D(const LC2Configuration2D* ExternalConfig);
    void LinkToInterpolator(Interpolator2D* external_Interpolator) { interp_ = external_Interpolator; }
    void GetOptimalTransform(float* outMatrix);
    void GetOptimalTransform();
};

class Optimizers2DTo3D : public lc2reg::OptimizerBase {
public:
    Optimizers2DTo3D(std::string FilePath);
    Optimizers2DTo3D(const LC2Configuration3D* ExternalConfig);
    void LinkToInterpolator(Interpolator2DTo3D* external_Interpolator) { interp_ = external_Interpolator; }
    void GetOptimalTransform(float* outMatrix);
    void GetOptimalTransform();
};
