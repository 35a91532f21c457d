// bsar_sim.h -- forward model (processRawData) launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bsar_kernels.h"

namespace bsar {

struct SimArgs {
    double n1, n2;
    double c;       // System.VelocidadeLuz
    double fs;      // RadarRx.fs (delay bins)
    double lamb;    // RadarTx.lamb = c / fc
    int S;          // System.IndiceMaximo (range samples)
};

hipError_t launch_sim_terms(const DemDev& dem, const Traj& tr, const double* scat, const double2* rcs, int ns,
                            const SimArgs& sa, int* bins, double2* vals, hipStream_t s);
hipError_t launch_sim_place(const int* bins, const double2* vals, int ns, int n, int S, double2* aux, hipStream_t s);

}  // namespace bsar
