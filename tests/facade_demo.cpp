// Test program: reference-style call sites compiled against bsar_facade.hpp.
// Structs declared here with the reference's member names stand in for the
// caller's own types.
#include <cstdio>
#include <vector>

#include "bsar_facade.hpp"

struct ProcessedDEM { std::vector<double> X_vec, Y_vec; std::vector<std::vector<double>> Z_DEM; };
struct TrajectoryMatrix { std::vector<std::vector<double>> Tx, Rx; std::vector<double> P; };
struct Environment { double n1, n2; };
struct RefractionPoints { std::vector<std::vector<double>> P_refrac_ida, P_refrac_volta; };
struct ReflectionPoints { std::vector<std::vector<double>> P_reflection; };

using namespace bsar::facade;

static bool rd(double& v) { return std::scanf("%lf", &v) == 1; }

// stdin: nx ny, X, Y, Z (ny rows), n, Tx rows, Rx rows, P, n1 n2 (any malformed
// or oversized input exits with status 2 before the GPU is used)
int main()
{
    ProcessedDEM dem;
    int nx, ny, n;
    if (std::scanf("%d %d", &nx, &ny) != 2 || nx < 1 || ny < 1 || nx > 4096 || ny > 4096) return 2;
    dem.X_vec.resize(nx); dem.Y_vec.resize(ny); dem.Z_DEM.assign(ny, std::vector<double>(nx));
    for (auto& v : dem.X_vec) if (!rd(v)) return 2;
    for (auto& v : dem.Y_vec) if (!rd(v)) return 2;
    for (auto& r : dem.Z_DEM) for (auto& v : r) if (!rd(v)) return 2;
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
    RangeData r = calculaSlantRangeFermat(dem, t, env, refl, refr, false);
    for (int i = 0; i < n; ++i)
        std::printf("%a %a %a %a %a %a %a %a %a %a\n", refr.P_refrac_ida[i][0], refr.P_refrac_ida[i][1],
                    refr.P_refrac_ida[i][2], refr.P_refrac_volta[i][0], refr.P_refrac_volta[i][1],
                    refr.P_refrac_volta[i][2], r.r1Tx[i], r.r2Tx[i], r.r1Rx[i], r.r2Rx[i]);
    std::printf("%a\n", interpolateDEMElevation(t.P[0], t.P[1], dem));
    return 0;
}
