// Command-line registration tool with the reference's contract (main.cu:20-43, :728-827):
//   lc2_registration -m 2D|3D|2DTo3D -c parameters.yaml [--avoid-igtl] [--igtl-port N]
//                    [--no-images]
// Runs Interpolator* + Optimizers* from the YAML, prints the optimum and timing, writes
// <yaml stem>_result.<ext> and (unless --no-images) the registered images.
// In 3D mode, as the reference's (main.cu:757-770), an OpenIGTLink server (port 23333, or
// --igtl-port) waits for a client first and then streams every evaluated pose;
// --avoid-igtl skips it.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "lc2reg.h"

static void usage() {
    std::printf(
        "usage: lc2_registration -m 2D|3D|2DTo3D -c parameters.yaml [--avoid-igtl] "
        "[--igtl-port N] [--no-images]\n");
}

int main(int argc, char** argv) {
    std::string mode, cfg;
    bool images = true, igtl = true;
    int igtl_port = 23333;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if ((a == "-m" || a == "--mode") && i + 1 < argc) mode = argv[++i];
        else if ((a == "-c" || a == "--config") && i + 1 < argc) cfg = argv[++i];
        else if (a == "--avoid-igtl") igtl = false;
        else if (a == "--igtl-port" && i + 1 < argc) igtl_port = std::atoi(argv[++i]);
        else if (a == "--no-images") images = false;
        else if (a == "-h" || a == "--help") {
            usage();
            return 0;
        } else {
            std::fprintf(stderr, "unknown argument: %s\n", a.c_str());
            usage();
            return 2;
        }
    }
    if (mode.empty() || cfg.empty()) {
        usage();
        return 2;
    }
    const bool is2d = mode == "2D";
    double params[12] = {0};
    double lc2 = 0.0;
    int nev = 0;
    float m[16] = {0};
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = lc2reg_register_yaml_igtl(mode.c_str(), cfg.c_str(), images ? 1 : 0,
                                             mode == "3D" && igtl ? igtl_port : 0, params, &lc2,
                                             &nev, m);
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != 0) {
        std::fprintf(stderr, "registration failed: %s\n", lc2reg_last_error());
        return 1;
    }
    const int np = is2d ? 6 : 12, md = is2d ? 3 : 4;
    std::printf("Optimized parameters:");
    for (int i = 0; i < np; ++i) std::printf(" %.6g", params[i]);
    std::printf("\nMoving-to-fixed matrix:\n");
    for (int r = 0; r < md; ++r) {
        for (int c = 0; c < md; ++c) std::printf(" %12.7g", m[r * md + c]);
        std::printf("\n");
    }
    std::printf("Last similarity: %.9f after %d evaluations, %.3f s\n", lc2, nev, secs);
    return 0;
}
