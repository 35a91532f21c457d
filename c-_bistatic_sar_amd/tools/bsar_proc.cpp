// bsar_proc -- drop-in replacement for the reference's proc_test_script
// (proc_test_script.cpp:184-904): reads the stage-boundary file set
// build/output/proc/ (rootData_complex.bin, grid_{x,y,z}Axis.csv, DEM_X_vec.csv,
// DEM_Y_vec.csv, DEM_Z_matrix.csv, Tx_positions.csv, Rx_positions.csv,
// environment.txt), back-projects every voxel on the GPU and writes
// output_3D_matrix.bin plus output_3D_matrix_grid_{x,y,z}.csv in the same
// formats.  Radar constants default to the reference's compiled-in values
// (radar_init.cpp:9,17,44,74): fc = 10 GHz, fs = 60 MHz, c = 3e8, factor 4
// (proc_test_script.cpp:220).  Exit status 1 on any load error
// (proc_test_script.cpp:898-901).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "bsar.h"
#include "bsar_facade.hpp"

namespace fa = bsar::facade;

struct Dem {
    std::vector<double> X_vec, Y_vec;
    std::vector<std::vector<double>> Z_DEM;
};

static void usage()
{
    std::printf(
        "usage: bsar_proc [--dir DIR] [--fc HZ] [--fs HZ] [--c M_S] [--factor K] [--precision fp64|mixed]\n"
        "                 [--device N] [--check]\n"
        "  DIR defaults to build/output/proc (relative to the working directory, as the reference).\n"
        "  --check parses the inputs and prints their shapes without using the GPU.\n");
}

// initializeEnvironment, raw_data_processing.cpp:14-46: n1=..., n2=... lines,
// defaults 1.0 / 4.0.
static void read_environment(const std::string& path, double& n1, double& n2)
{
    n1 = 1.0;
    n2 = 4.0;
    std::ifstream f(path);
    if (!f.is_open()) return;
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        std::istringstream iss(line);
        std::string key, value;
        if (std::getline(iss, key, '=') && std::getline(iss, value)) {
            if (key == "n1") n1 = std::stod(value);
            else if (key == "n2") n2 = std::stod(value);
        }
    }
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
        else if (a == "--fc") fc = std::stod(next("--fc"));
        else if (a == "--fs") fs = std::stod(next("--fs"));
        else if (a == "--c") c = std::stod(next("--c"));
        else if (a == "--factor") factor = std::stoi(next("--factor"));
        else if (a == "--precision") precision = next("--precision") == "mixed" ? BSAR_MIXED : BSAR_FP64;
        else if (a == "--device") setenv("BSAR_DEVICE", next("--device").c_str(), 1);
        else if (a == "--check") check_only = true;
        else if (a == "--help" || a == "-h") { usage(); return 0; }
        else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); usage(); return 2; }
    }
    try {
        using clk = std::chrono::steady_clock;
        const auto since = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };
        const auto t_read = clk::now();
        double n1, n2;
        read_environment(dir + "/environment.txt", n1, n2);
        fa::MatrixInfo info;
        std::vector<std::complex<double>> root = fa::readBinary<double>(dir + "/rootData_complex.bin", info);
        if (info.dimensions.size() != 2) {
            std::fprintf(stderr, "[ERROR] rootData must be a 2-D matrix\n");
            return 1;
        }
        const size_t S = info.dimensions[0], Nd = info.dimensions[1];
        std::vector<double> gx = fa::readCSV(dir + "/grid_xAxis.csv");
        std::vector<double> gy = fa::readCSV(dir + "/grid_yAxis.csv");
        std::vector<double> gz = fa::readCSV(dir + "/grid_zAxis.csv");
        Dem dem;
        dem.X_vec = fa::readCSV(dir + "/DEM_X_vec.csv");
        dem.Y_vec = fa::readCSV(dir + "/DEM_Y_vec.csv");
        dem.Z_DEM = fa::readCSVMatrix(dir + "/DEM_Z_matrix.csv");
        std::vector<std::vector<double>> Tx = fa::readCSVMatrix(dir + "/Tx_positions.csv");
        std::vector<std::vector<double>> Rx = fa::readCSVMatrix(dir + "/Rx_positions.csv");
        for (const auto* t : {&Tx, &Rx})
            for (const auto& r : *t)
                if (r.size() < 3) { std::fprintf(stderr, "[ERROR] antenna rows need 3 values\n"); return 1; }
        if (Tx.size() != Rx.size()) { std::fprintf(stderr, "[ERROR] Tx and Rx counts differ\n"); return 1; }
        std::printf("inputs: rootData %zu x %zu, grid %zu x %zu x %zu, DEM %zu x %zu, %zu positions, n1=%g n2=%g\n",
                    S, Nd, gx.size(), gy.size(), gz.size(), dem.X_vec.size(), dem.Y_vec.size(), Tx.size(), n1, n2);
        if (check_only) return 0;
        if (S < 2 || Nd < 1) { std::fprintf(stderr, "[ERROR] rootData needs S >= 2 and >= 1 pulse\n"); return 1; }
        const double sec_read = since(t_read);
        const auto t_load = clk::now();

        const size_t n = Tx.size();
        std::vector<double> tx(3 * n), rx(3 * n);
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) { tx[3 * i + k] = Tx[i][k]; rx[3 * i + k] = Rx[i][k]; }
        fa::FlatDem fd(dem);
        bsar_ctx* ctx = fa::context();
        bsar_params p{fc, c, fs, n1, n2, factor, 0};
        bsar_dataset* ds = nullptr;
        fa::check(bsar_dataset_create(ctx, &p, &fd.d, tx.data(), rx.data(), (int)n, &ds), ctx, "dataset");
        // rootData flat index r + p*S is already pulse-major, range fastest
        int st = bsar_dataset_load_rootdata(ctx, ds, reinterpret_cast<const double*>(root.data()), (int)S, (int)Nd);
        if (st != BSAR_OK) { bsar_dataset_destroy(ds); fa::check(st, ctx, "rootData upload"); }
        fa::check(bsar_ctx_synchronize(ctx), ctx, "rootData upload");
        const double sec_load = since(t_load);
        const size_t nx = gx.size(), ny = gy.size(), nz = gz.size();
        std::vector<std::complex<double>> img(nx * ny * nz);
        bsar_report rep{};
        const auto t0 = clk::now();
        st = bsar_backproject(ctx, ds, gx.data(), (int)nx, gy.data(), (int)ny, gz.data(), (int)nz, precision,
                              reinterpret_cast<double*>(img.data()), &rep);
        const double wall = since(t0);
        bsar_dataset_destroy(ds);
        fa::check(st, ctx, "backproject");
        std::printf("[PERFORMANCE] back-projection: %.3f s device, %.3f s wall, %.3e voxel-pulse/s (%s)\n",
                    rep.seconds, wall, rep.voxel_pulses / (rep.seconds > 0 ? rep.seconds : 1),
                    precision == BSAR_MIXED ? "mixed" : "fp64");
        const auto t_write = clk::now();
        const std::string out = dir + "/output_3D_matrix.bin";
        st = bsar_write_image(out.c_str(), reinterpret_cast<const double*>(img.data()), (int)nx, (int)ny, (int)nz,
                              gx.data(), gy.data(), gz.data());
        fa::check(st, nullptr, "export");
        std::printf("wrote %s\n", out.c_str());
        std::printf("[PERFORMANCE] stages: read inputs %.3f s (rootData %.1f MB), upload + x%d spline %.3f s, "
                    "back-projection %.3f s (image to host included), write image %.3f s (%.1f MB)\n",
                    sec_read, S * Nd * 16 / 1e6, factor, sec_load, wall, since(t_write), nx * ny * nz * 16 / 1e6);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[ERROR] %s\n", e.what());
        return 1;
    }
    return 0;
}
