// Host stand-in for the few HIP names c-_bistatic_sar_amd/csrc/bsar_device.h
// uses, so its refraction search compiles with g++ for the CPU test
// tests/test_search_host.py (test infrastructure, never in the library).
// The AMDGPU builtins map to the host's correctly rounded FP32 / FP64
// operations: the search's bounds and estimates carry error margins larger
// than the device instructions' (1 ulp), and its exact evaluations use only
// + - * / sqrt, the same on both.
#pragma once
#include <math.h>
#include <stdint.h>

#include <algorithm>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline

struct float2 { float x, y; };
struct double2 { double x, y; };
inline float2 make_float2(float a, float b) { return {a, b}; }
struct bsar_host_dim3 { unsigned x = 0; };
static bsar_host_dim3 threadIdx, blockDim{1};
inline void __syncthreads() {}
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }

inline float bsar_host_sqrtf(float a) { return sqrtf(a); }
// The test's planted mistake scales the FP32 reciprocal root (used by the
// search's lower bounds only) when set away from 1.
extern float bsar_host_rsqf_scale;
inline float bsar_host_rsqf(float a) { return bsar_host_rsqf_scale / sqrtf(a); }
inline float bsar_host_rcpf(float a) { return 1.0f / a; }
inline double bsar_host_rsq(double a) { return 1.0 / sqrt(a); }
#define __builtin_amdgcn_sqrtf bsar_host_sqrtf
#define __builtin_amdgcn_rsqf bsar_host_rsqf
#define __builtin_amdgcn_rcpf bsar_host_rcpf
#define __builtin_amdgcn_rsq bsar_host_rsq
