// proc_inputs.hpp -- the stage-boundary file set that the reference's
// proc_test_script reads (proc_test_script.cpp:184-440), for the host tools:
// build/output/proc/ holds rootData_complex.bin, grid_{x,y,z}Axis.csv,
// DEM_X_vec.csv, DEM_Y_vec.csv, DEM_Z_matrix.csv, Tx_positions.csv,
// Rx_positions.csv and environment.txt.  The readers throw
// std::runtime_error on a missing or malformed file, as the reference's do.
#pragma once
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "bsar.h"
#include "bsar_facade.hpp"

namespace bsar_tools {

namespace fa = bsar::facade;

// Command-line values: the whole string must be a number (an integer for
// option_integer); otherwise the tool prints the option and exits with 2.
inline double option_number(const std::string& v, const char* what)
{
    size_t used = 0;
    double x = 0.0;
    try {
        x = std::stod(v, &used);
    } catch (const std::exception&) {
        used = 0;
    }
    if (used == 0 || used != v.size()) {
        std::fprintf(stderr, "bad value '%s' for %s\n", v.c_str(), what);
        std::exit(2);
    }
    return x;
}

inline int option_integer(const std::string& v, const char* what)
{
    const double x = option_number(v, what);
    if (!(x >= -2147483648.0 && x <= 2147483647.0) || x != (double)(long long)x) {
        std::fprintf(stderr, "bad value '%s' for %s\n", v.c_str(), what);
        std::exit(2);
    }
    return (int)x;
}

struct Dem {
    std::vector<double> X_vec, Y_vec;
    std::vector<std::vector<double>> Z_DEM;
};

// initializeEnvironment, raw_data_processing.cpp:14-46: n1=..., n2=... lines,
// defaults 1.0 / 4.0.
inline void read_environment(const std::string& path, double& n1, double& n2)
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

struct ProcInputs {
    double n1 = 1.0, n2 = 4.0;
    std::vector<std::complex<double>> root;   // flat r + p * S (pulse-major, range fastest)
    size_t S = 0, Nd = 0;                     // header dims: samples, data pulses
    std::vector<double> gx, gy, gz;
    Dem dem;
    std::vector<std::vector<double>> Tx, Rx;
};

// Reads the file set in dir.  Returns "" when it is usable, else what is wrong.
inline std::string read_proc_inputs(const std::string& dir, ProcInputs& in)
{
    read_environment(dir + "/environment.txt", in.n1, in.n2);
    fa::MatrixInfo info;
    in.root = fa::readBinary<double>(dir + "/rootData_complex.bin", info);
    if (info.dimensions.size() != 2) return "rootData must be a 2-D matrix";
    in.S = info.dimensions[0];
    in.Nd = info.dimensions[1];
    in.gx = fa::readCSV(dir + "/grid_xAxis.csv");
    in.gy = fa::readCSV(dir + "/grid_yAxis.csv");
    in.gz = fa::readCSV(dir + "/grid_zAxis.csv");
    in.dem.X_vec = fa::readCSV(dir + "/DEM_X_vec.csv");
    in.dem.Y_vec = fa::readCSV(dir + "/DEM_Y_vec.csv");
    in.dem.Z_DEM = fa::readCSVMatrix(dir + "/DEM_Z_matrix.csv");
    in.Tx = fa::readCSVMatrix(dir + "/Tx_positions.csv");
    in.Rx = fa::readCSVMatrix(dir + "/Rx_positions.csv");
    for (const auto* t : {&in.Tx, &in.Rx})
        for (const auto& r : *t)
            if (r.size() < 3) return "antenna rows need 3 values";
    if (in.Tx.size() != in.Rx.size()) return "Tx and Rx counts differ";
    return "";
}

inline void print_proc_inputs(const ProcInputs& in)
{
    std::printf("inputs: rootData %zu x %zu, grid %zu x %zu x %zu, DEM %zu x %zu, %zu positions, n1=%g n2=%g\n",
                in.S, in.Nd, in.gx.size(), in.gy.size(), in.gz.size(), in.dem.X_vec.size(), in.dem.Y_vec.size(),
                in.Tx.size(), in.n1, in.n2);
}

// The dataset of the file set on ctx's device: DEM, antennas, radar constants
// and the range data upsampled x factor.  Throws on failure.
inline bsar_dataset* upload_proc_dataset(bsar_ctx* ctx, const ProcInputs& in, double fc, double fs, double c,
                                         int factor)
{
    const size_t n = in.Tx.size();
    std::vector<double> tx(3 * n), rx(3 * n);
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            tx[3 * i + k] = in.Tx[i][k];
            rx[3 * i + k] = in.Rx[i][k];
        }
    fa::FlatDem fd(in.dem);
    bsar_params p{fc, c, fs, in.n1, in.n2, factor, 0};
    bsar_dataset* ds = nullptr;
    fa::check(bsar_dataset_create(ctx, &p, &fd.d, tx.data(), rx.data(), (int)n, &ds), ctx, "dataset");
    // rootData flat index r + p*S is already pulse-major, range fastest
    const int st = bsar_dataset_load_rootdata(ctx, ds, reinterpret_cast<const double*>(in.root.data()), (int)in.S,
                                              (int)in.Nd);
    if (st != BSAR_OK) {
        bsar_dataset_destroy(ds);
        fa::check(st, ctx, "rootData upload");
    }
    fa::check(bsar_ctx_synchronize(ctx), ctx, "rootData upload");
    return ds;
}

}  // namespace bsar_tools
