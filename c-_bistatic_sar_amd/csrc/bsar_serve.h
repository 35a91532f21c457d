// bsar_serve.h -- device-side job table and work list of the serving loop.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bsar_kernels.h"

namespace bsar {

// One admitted job: its image's argument blocks (first K pulses of its
// dataset, its grid, its accumulators).
struct SrvJobDev {
    BpArgs a;        // FP64 form accumulates in a.acc
    MixedArgs m;     // mixed form accumulates in m.acc
};

// Work item: thread positions [vox0, vox0 + nvox) (nvox <= 256, mapped to
// voxels by band_voxel) of job `slot`, pulses [p0, p1) of its pulse group g.
struct SrvItem {
    int slot, g;
    int64_t vox0;
    int nvox, p0, p1, pad;
};

// Kernel variants: index = 2 * mixed + lds.  Variant k pulls items
// [begin[k], end[k]) through the counter next[k] (= begin[k] at the start of
// a round).
constexpr int kSrvVariants = 4;
struct SrvList {
    unsigned next[kSrvVariants];
    unsigned begin[kSrvVariants];
    unsigned end[kSrvVariants];
    unsigned pad[kSrvVariants];
    SrvItem items[1];
};

constexpr int kSrvListCap = 32768;   // items per round

// DEM slots (8 bytes) staged in LDS by the serving kernel.  Its dynamic LDS
// is fixed at capture time, so it is sized for four resident workgroups per
// CU (4 x 36 KiB of the 160 KiB); larger DEMs take the global-memory variant.
constexpr int kSrvLdsDem = 4608;
inline size_t srv_list_bytes() { return sizeof(SrvList) + (kSrvListCap - 1) * sizeof(SrvItem); }

const void* serve_kernel(int variant);
hipError_t launch_serve(int variant, int blocks, const SrvJobDev* jobs, SrvList* wl, hipStream_t s);

}  // namespace bsar
