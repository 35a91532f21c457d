// Test program: reference-style call sites compiled against bsar_facade.hpp.
// Structs declared here with the reference's member names stand in for the
// caller's own types.
#include <cstdio>
#include <string>
#include <vector>

#include "bsar_facade.hpp"

struct ProcessedDEM { std::vector<double> X_vec, Y_vec; std::vector<std::vector<double>> Z_DEM; };
struct TrajectoryMatrix { std::vector<std::vector<double>> Tx, Rx; std::vector<double> P; };
struct Environment { double n1, n2; };
struct RefractionPoints { std::vector<std::vector<double>> P_refrac_ida, P_refrac_volta; };
struct ReflectionPoints { std::vector<std::vector<double>> P_reflection; };
// the caller's own range type (raw_data_processing.h:18-23 under another name):
// the facade's result converts to it
struct CallerRangeData { std::vector<double> r1Tx, r2Tx, r1Rx, r2Rx; };

using namespace bsar::facade;

static bool rd(double& v) { return std::scanf("%lf", &v) == 1; }

// stdin: nx ny, X, Y, Z (ny rows), n, Tx rows, Rx rows, P, n1 n2 (any malformed
// or oversized input exits with status 2 before the GPU is used).
// `facade_demo reflect` (host only): nx ny, X, Y, Z rows with zc values each
// (first read zc), m, m pairs (tx ty tz rx ry rz) -> calculateReflectionPoint.
int main(int argc, char** argv)
{
    ProcessedDEM dem;
    int nx, ny, n;
    if (std::scanf("%d %d", &nx, &ny) != 2 || nx < 0 || ny < 0 || nx > 4096 || ny > 4096) return 2;
    const bool reflect = argc > 1 && std::string(argv[1]) == "reflect";
    int zr = ny, zc = nx;
    if (reflect && std::scanf("%d %d", &zr, &zc) != 2) return 2;
    if (!reflect && (nx < 1 || ny < 1)) return 2;
    if (zr < 0 || zc < 0 || zr > 4096 || zc > 4096) return 2;
    dem.X_vec.resize(nx); dem.Y_vec.resize(ny); dem.Z_DEM.assign(zr, std::vector<double>(zc));
    for (auto& v : dem.X_vec) if (!rd(v)) return 2;
    for (auto& v : dem.Y_vec) if (!rd(v)) return 2;
    for (auto& r : dem.Z_DEM) for (auto& v : r) if (!rd(v)) return 2;
    if (reflect) {
        int m;
        if (std::scanf("%d", &m) != 1 || m < 0 || m > 1000000) return 2;
        for (int k = 0; k < m; ++k) {
            double a[6];
            for (double& v : a) if (!rd(v)) return 2;
            std::vector<double> q = calculateReflectionPoint(a[0], a[1], a[2], a[3], a[4], a[5], dem);
            std::printf("%a %a %a\n", q[0], q[1], q[2]);
        }
        return 0;
    }
    TrajectoryMatrix t;
    if (std::scanf("%d", &n) != 1 || n < 1 || n > 100000) return 2;
    t.Tx.assign(n, std::vector<double>(3)); t.Rx.assign(n, std::vector<double>(3)); t.P.resize(3);
    for (auto& r : t.Tx) for (auto& v : r) if (!rd(v)) return 2;
    for (auto& r : t.Rx) for (auto& v : r) if (!rd(v)) return 2;
    for (auto& v : t.P) if (!rd(v)) return 2;
    Environment env;
    if (!rd(env.n1) || !rd(env.n2)) return 2;
    RefractionPoints refr;
    ReflectionPoints refl;
    CallerRangeData r = calculaSlantRangeFermat(dem, t, env, refl, refr, false);
    for (int i = 0; i < n; ++i)
        std::printf("%a %a %a %a %a %a %a %a %a %a %a %a %a\n", refr.P_refrac_ida[i][0], refr.P_refrac_ida[i][1],
                    refr.P_refrac_ida[i][2], refr.P_refrac_volta[i][0], refr.P_refrac_volta[i][1],
                    refr.P_refrac_volta[i][2], r.r1Tx[i], r.r2Tx[i], r.r1Rx[i], r.r2Rx[i],
                    refl.P_reflection[i][0], refl.P_reflection[i][1], refl.P_reflection[i][2]);
    std::printf("%a\n", interpolateDEMElevation(t.P[0], t.P[1], dem));
    return 0;
}
