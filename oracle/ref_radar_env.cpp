// ref_radar_env.cpp -- radar constants for the reference's own
// proc_test_script, read from the environment so that the unmodified driver
// images a workload at that workload's carrier, sampling rate and speed of
// light (the reference compiles them in: radar_init.cpp:5-89).
//
// TEST INFRASTRUCTURE ONLY (the CPU baseline and the reference-parity check
// of the benchmark).  Linked in place of src/radar_init.cpp into
// _ref/proc_test_script_env; every other source of that executable is the
// reference's own, compiled where it lies.  The driver reads, of these
// structs (radar_params.h:7-66), RadarTx::FreqPortadora (proc_test_script.
// cpp:213), RadarRx::fs (:220-223), System::VelocidadeLuz (:213, 561) and the
// target position set by updatePointToProc (:456-461); the other members are
// zero.  Unset variables give the reference's compiled-in values.
//
//   BSAR_REF_FC  carrier [Hz]              (reference: 10e9)
//   BSAR_REF_FS  receiver ADC rate [Hz]    (reference: 60e6, before the x4)
//   BSAR_REF_C   propagation speed [m/s]   (reference: 3e8)
#include <cstdlib>

#include "radar_init.h"

namespace {

double env_or(const char* name, double fallback)
{
    const char* s = std::getenv(name);
    if (!s || !*s) return fallback;
    char* end = nullptr;
    const double v = std::strtod(s, &end);
    return (end && *end == '\0') ? v : fallback;
}

}  // namespace

RadarTx initializeRadarTx()
{
    RadarTx t{};
    t.FreqPortadora = env_or("BSAR_REF_FC", 10e9);
    t.fs = env_or("BSAR_REF_FS", 60e6);
    return t;
}

RadarRx initializeRadarRx()
{
    RadarRx r{};
    r.fs = env_or("BSAR_REF_FS", 60e6);
    return r;
}

Target initializeTarget()
{
    Target g{};
    g.rcs = 1.0;
    return g;
}

System initializeSystem()
{
    System s{};
    s.VelocidadeLuz = env_or("BSAR_REF_C", 3e8);
    return s;
}

void updatePointToProc(Target& target, double x, double y, double z)
{
    target.pos[0] = x;
    target.pos[1] = y;
    target.pos[2] = z;
}
