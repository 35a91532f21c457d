// bsar_proc -- drop-in replacement for the reference's proc_test_script
// (proc_test_script.cpp:184-904): reads the stage-boundary file set
// build/output/proc/ (rootData_complex.bin, grid_{x,y,z}Axis.csv, DEM_X_vec.csv,
// DEM_Y_vec.csv, DEM_Z_matrix.csv, Tx_positions.csv, Rx_positions.csv,
// environment.txt), back-projects every voxel on the GPU and writes
// output_3D_matrix.bin plus output_3D_matrix_grid_{x,y,z}.csv in the same
// formats.  Radar constants default to the reference's compiled-in values
// (radar_init.cpp:9,17,44,74): fc = 10 GHz, fs = 60 MHz, c = 3e8, factor 4
// (proc_test_script.cpp:220).  Exit status 1 on any load error
// (proc_test_script.cpp:898-901), 2 on a malformed command line.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bsar.h"
#include "bsar_facade.hpp"
#include "proc_inputs.hpp"

namespace fa = bsar::facade;
using namespace bsar_tools;

static void usage()
{
    std::printf(
        "usage: bsar_proc [--dir DIR] [--fc HZ] [--fs HZ] [--c M_S] [--factor K] [--precision fp64|mixed]\n"
        "                 [--device N] [--check]\n"
        "  DIR defaults to build/output/proc (relative to the working directory, as the reference).\n"
        "  --check parses the inputs and prints their shapes without using the GPU.\n");
}

int main(int argc, char** argv)
{
    std::string dir = "build/output/proc";
    double fc = 10e9, fs = 60e6, c = 3e8;
    int factor = 4, precision = BSAR_FP64;
    bool check_only = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&](const char* what) -> std::string {
            if (i + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", what); std::exit(2); }
            return argv[++i];
        };
        if (a == "--dir") dir = next("--dir");
        else if (a == "--fc") fc = option_number(next("--fc"), "--fc");
        else if (a == "--fs") fs = option_number(next("--fs"), "--fs");
        else if (a == "--c") c = option_number(next("--c"), "--c");
        else if (a == "--factor") factor = option_integer(next("--factor"), "--factor");
        else if (a == "--precision") {
            const std::string p = next("--precision");
            if (p != "fp64" && p != "mixed") { std::fprintf(stderr, "--precision must be fp64 or mixed\n"); return 2; }
            precision = p == "mixed" ? BSAR_MIXED : BSAR_FP64;
        }
        else if (a == "--device") setenv("BSAR_DEVICE", next("--device").c_str(), 1);
        else if (a == "--check") check_only = true;
        else if (a == "--help" || a == "-h") { usage(); return 0; }
        else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); usage(); return 2; }
    }
    try {
        using clk = std::chrono::steady_clock;
        const auto since = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };
        const auto t_read = clk::now();
        ProcInputs in;
        const std::string bad = read_proc_inputs(dir, in);
        if (!bad.empty()) {
            std::fprintf(stderr, "[ERROR] %s\n", bad.c_str());
            return 1;
        }
        print_proc_inputs(in);
        if (check_only) return 0;
        if (in.S < 2 || in.Nd < 1) { std::fprintf(stderr, "[ERROR] rootData needs S >= 2 and >= 1 pulse\n"); return 1; }
        const double sec_read = since(t_read);
        const auto t_load = clk::now();
        bsar_ctx* ctx = fa::context();
        bsar_dataset* ds = upload_proc_dataset(ctx, in, fc, fs, c, factor);
        const double sec_load = since(t_load);
        const size_t nx = in.gx.size(), ny = in.gy.size(), nz = in.gz.size();
        std::vector<std::complex<double>> img(nx * ny * nz);
        bsar_report rep{};
        const auto t0 = clk::now();
        const int st = bsar_backproject(ctx, ds, in.gx.data(), (int)nx, in.gy.data(), (int)ny, in.gz.data(), (int)nz,
                                        precision, reinterpret_cast<double*>(img.data()), &rep);
        const double wall = since(t0);
        bsar_dataset_destroy(ds);
        fa::check(st, ctx, "backproject");
        std::printf("[PERFORMANCE] back-projection: %.3f s device, %.3f s wall, %.3e voxel-pulse/s (%s)\n",
                    rep.seconds, wall, rep.voxel_pulses / (rep.seconds > 0 ? rep.seconds : 1),
                    precision == BSAR_MIXED ? "mixed" : "fp64");
        const auto t_write = clk::now();
        const std::string out = dir + "/output_3D_matrix.bin";
        fa::check(bsar_write_image(out.c_str(), reinterpret_cast<const double*>(img.data()), (int)nx, (int)ny, (int)nz,
                                   in.gx.data(), in.gy.data(), in.gz.data()),
                  nullptr, "export");
        std::printf("wrote %s\n", out.c_str());
        std::printf("[PERFORMANCE] stages: read inputs %.3f s (rootData %.1f MB), upload + x%d spline %.3f s, "
                    "back-projection %.3f s (image to host included), write image %.3f s (%.1f MB)\n",
                    sec_read, in.S * in.Nd * 16 / 1e6, factor, sec_load, wall, since(t_write), nx * ny * nz * 16 / 1e6);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[ERROR] %s\n", e.what());
        return 1;
    }
    return 0;
}
