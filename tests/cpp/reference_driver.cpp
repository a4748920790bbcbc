// A registration driver written against the reference's public interface only
// (RegistrationParametersConfig.h, interpolator.cuh, optimizers.cuh): the programmatic
// configuration structs, the Interpolator*/Optimizers* constructors that take them, and the
// call sequence of main.cu. It compiles against include/lc2_registration.hpp unchanged,
// which is what a caller switching from the reference needs.
//
//   reference_driver <dir>
// <dir>/case.txt holds, one per line: fixed dims (3 ints), spacing, origin (3 floats each),
// moving dims, spacing, origin, sampler spacing, patch radius, initial pose (12 floats);
// <dir>/fixed.raw and <dir>/moving.raw hold the float32 images. Results are printed as
// "key values..." lines; the optimizers write RegistrationInfo.log under <dir>/3d and
// <dir>/slice.
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "lc2_registration.hpp"

static std::vector<float> read_raw(const std::string& path, size_t n) {
    std::vector<float> v(n);
    std::ifstream f(path, std::ios::binary);
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(float)));
    if (!f) throw std::runtime_error("cannot read " + path);
    return v;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: reference_driver <dir>\n";
        return 2;
    }
    const std::string dir = argv[1];
    std::ifstream in(dir + "/case.txt");
    int fd[3], md[3], patch;
    float fs[3], fo[3], ms[3], mo[3], rs[3], pose[12];
    for (int& v : fd) in >> v;
    for (float& v : fs) in >> v;
    for (float& v : fo) in >> v;
    for (int& v : md) in >> v;
    for (float& v : ms) in >> v;
    for (float& v : mo) in >> v;
    for (float& v : rs) in >> v;
    in >> patch;
    for (float& v : pose) in >> v;
    if (!in) {
        std::cerr << "bad case.txt\n";
        return 2;
    }
    std::vector<float> fixed = read_raw(dir + "/fixed.raw", (size_t)fd[0] * fd[1] * fd[2]);
    std::vector<float> moving = read_raw(dir + "/moving.raw", (size_t)md[0] * md[1] * md[2]);
    try {
        // ---- 3D: RegistrationParametersConfig.h:164-297, interpolator.cuh:47, optimizers.cuh:102
        const std::string out3 = dir + "/3d";
        mkdir(out3.c_str(), 0755);
        LC2Configuration3D cfg;
        cfg.FixedDimension = {fd[0], fd[1], fd[2]};
        cfg.FixedSpacing = {fs[0], fs[1], fs[2]};
        cfg.FixedOrigin = {fo[0], fo[1], fo[2]};
        cfg.MovingDimension = {md[0], md[1], md[2]};
        cfg.MovingSpacing = {ms[0], ms[1], ms[2]};
        cfg.MovingOrigin = {mo[0], mo[1], mo[2]};
        cfg.IsFixedHighResolution = false;
        cfg.IsMovingHighResolution = true;
        cfg.SamplerSpacing = {rs[0], rs[1], rs[2]};
        cfg.PatchSize = patch;
        cfg.ExportPath = out3;
        cfg.IsCenterOverlaid = false;
        cfg.IsAffine = true;
        cfg.initFRE = 4.0f;
        cfg.initTranslation = {pose[0], pose[1], pose[2]};
        cfg.initRotation = {pose[3], pose[4], pose[5]};
        cfg.initScale = {pose[6], pose[7], pose[8]};
        cfg.initShear = {pose[9], pose[10], pose[11]};
        cfg.MaxEval = 0;  // no DIRECT stage
        cfg.RhoBegin = 0.1f;
        cfg.Tol = 1e-4f;
        cfg.MaxIterations = 2000;
        cfg.FixedBuffer = fixed.data();
        cfg.MovingBuffer = moving.data();
        cfg.IsConfigured = true;

        Interpolator3D interp(&cfg);
        interp.InitiateMappers();
        interp.InitiateRegistrationHandle();
        const float3 T = make_float3(pose[0], pose[1], pose[2]), R = make_float3(pose[3], pose[4], pose[5]);
        const float3 S = make_float3(pose[6], pose[7], pose[8]), H = make_float3(pose[9], pose[10], pose[11]);
        interp.GenerateMappers(R, T, S, H);
        interp.Interpolate();
        const double l_default = interp.GetSimilarityMeasure();  // interpolator.cuh:55: 3
        const double l_patch = interp.GetSimilarityMeasure(patch);
        float cur[16], m2f[16];
        interp.GetCurrentTransform(cur);
        interp.GetMovingToFixedTransform(R, T, S, H, m2f);
        std::printf("lc2_3d_default %.17g\nlc2_3d %.17g\n", l_default, l_patch);
        std::printf("current_equals_m2f %d\n", (int)std::equal(cur, cur + 16, m2f));

        Optimizers3D opt(&cfg);
        opt.LinkToInterpolator(&interp);
        opt.Optimize();
        opt.GetOptimalTransform();
        float M[16];
        opt.GetInternalMatrix(M);
        cfg.DumpRegistrtationResults(true, (float)opt.GetLastSimilarity(), opt.GetEvaluationCount(), M);
        std::printf("opt_3d");
        for (double v : opt.GetOptimalSolution()) std::printf(" %.9g", v);
        std::printf("\nnev_3d %d\nlc2_3d_final %.17g\n", cfg.NumberOfEvaluations, (double)cfg.LastSimilarity);

        // ---- 2D->3D: the middle plane of the fixed grid as a slice (interpolator.cuh:218-219,
        // optimizers.cuh:178)
        const std::string outs = dir + "/slice";
        mkdir(outs.c_str(), 0755);
        const int z = fd[2] / 2;
        std::vector<float> slice(fixed.begin() + (size_t)z * fd[0] * fd[1],
                                 fixed.begin() + (size_t)(z + 1) * fd[0] * fd[1]);
        LC2Configuration2DTo3D c2;
        c2.FixedDimension = {fd[0], fd[1], 1};
        c2.FixedSpacing = {fs[0], fs[1], fs[2]};
        c2.FixedOrigin = {fo[0], fo[1], fo[2] + z * fs[2]};
        c2.MovingDimension = cfg.MovingDimension;
        c2.MovingSpacing = cfg.MovingSpacing;
        c2.MovingOrigin = cfg.MovingOrigin;
        c2.IsMovingHighResolution = true;
        c2.SamplerSpacing = {rs[0], rs[1], rs[2]};
        c2.PatchSize = 3;
        c2.ExportPath = outs;
        c2.IsCenterOverlaid = false;
        c2.FixedBuffer = slice.data();
        c2.MovingBuffer = moving.data();
        c2.IsConfigured = true;
        Interpolator2DTo3D islice(&c2);
        islice.InitiateMappers();
        islice.InitiateRegistrationHandle();
        islice.GenerateMappers(R, T, S, H);
        islice.Interpolate();
        const double s_default = islice.GetSimilarityMeasure();  // interpolator.cuh:225: 9
        const double s9 = islice.GetSimilarityMeasure(9);
        std::printf("lc2_slice_default %.17g\nlc2_slice_9 %.17g\n", s_default, s9);
        LC2Configuration3D co = cfg;
        co.ExportPath = outs;
        co.PatchSize = 3;
        co.IsAffine = false;
        co.MaxIterations = 300;
        Optimizers2DTo3D oslice(&co);
        oslice.LinkToInterpolator(&islice);
        oslice.Optimize();
        std::printf("nev_slice %d\n", oslice.GetEvaluationCount());
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
