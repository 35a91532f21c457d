// Test program: the reference's scenario set-up (gs_test_script.cpp:22-94)
// written against bsar_facade.hpp, host only (no GPU).  Structs declared
// here with the reference's member names stand in for the caller's types.
// Prints the Tx/Rx trajectories, grid axes and DEM as "%.17g" text.
#include <cstdio>
#include <string>
#include <vector>

#include "bsar_facade.hpp"

struct TrajectoryPoints { std::vector<double> Px, Py, Pz; };
struct GridToProc { std::vector<double> xAxis, yAxis, zAxis; };
struct Grid3D {
    std::vector<std::vector<std::vector<double>>> X, Y, Z;
    GridToProc gridToProc;
    int nx, ny, nz;
};
struct Target { double pos[3]; double rcs; };
struct RadarTx { double NumVoltasEsp, RaioMenorEsp, RaioMaiorEsp, AltMaiorEsp, AltMenorEsp, Vt, PRF; };
struct ProcessedDEM { std::vector<double> X_vec, Y_vec; std::vector<std::vector<double>> Z_DEM; };

using namespace bsar::facade;

static void vec(const char* tag, const std::vector<double>& v)
{
    std::printf("%s %zu", tag, v.size());
    for (double x : v) std::printf(" %.17g", x);
    std::printf("\n");
}

// argv[1]: DEM file name passed to loadDEM (argv[2] == "geotiff": loadDEMGeoTIFF)
int main(int argc, char** argv)
{
    RadarTx tx{2, 172.5, 147.5, 120.0, 80.0, 120.0, 400.0};
    TrajectoryPoints a = generateTransmitterTrajectory(tx);
    TrajectoryPoints b = generateReceiverTrajectory(tx);
    TrajectoryPoints c = funcao_espiral(3, 20.0, 30.0, 15.0, 10.0, 25.0, 200.0);
    vec("txx", a.Px); vec("txy", a.Py); vec("txz", a.Pz);
    vec("rxx", b.Px); vec("rxy", b.Py); vec("rxz", b.Pz);
    vec("spx", c.Px);
    Target t{{20.0, 0.0, -5.0}, 1.0};
    Grid3D g = crearGrid(t, 0.04, 0.8, 0.08, 0.8);
    vec("gx", g.gridToProc.xAxis); vec("gy", g.gridToProc.yAxis); vec("gz", g.gridToProc.zAxis);
    std::printf("mesh %d %d %d %.17g %.17g %.17g\n", g.nx, g.ny, g.nz, g.X[g.ny / 2][g.nx / 2][g.nz / 2],
                g.Y[g.ny / 2][g.nx / 2][g.nz / 2], g.Z[g.ny / 2][g.nx / 2][g.nz / 2]);
    if (argc > 1) {
        try {
            ProcessedDEM d = (argc > 2 && std::string(argv[2]) == "geotiff") ? (ProcessedDEM)loadDEMGeoTIFF(argv[1])
                                                                               : (ProcessedDEM)loadDEM(argv[1]);
            vec("demx", d.X_vec); vec("demy", d.Y_vec);
            std::vector<double> z;
            for (const auto& r : d.Z_DEM) z.insert(z.end(), r.begin(), r.end());
            std::printf("demrows %zu\n", d.Z_DEM.size());
            vec("demz", z);
        } catch (const std::exception& e) {
            std::printf("error %s\n", e.what());
            return 3;
        }
    }
    return 0;
}
