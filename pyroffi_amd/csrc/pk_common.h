// Host-side plumbing shared by the C-ABI entry points: status codes, the
// thread-local error message, argument checks and launch-error capture.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "../../include/pyroffi_amd.h"

namespace pk {

// Upper bounds the kernels size their LDS tables and register arrays for.
// n_joints <= 64 matches the reference FK limit (_fk_cuda_kernel.cu:26); the
// actuated/end-effector limits match _ik_cuda_helpers.cuh:26-36, which the
// reference never checks at run time.
constexpr int kMaxJoints = PK_MAX_JOINTS;
constexpr int kMaxAct = PK_MAX_ACT;
constexpr int kMaxEE = PK_MAX_EE;
constexpr int kMaxGrids = PK_MAX_GRIDS;

// Lanes per wavefront on CDNA; every kernel here is written for wave64.
constexpr int kWave = 64;

int set_error(int code, const char* fmt, ...);

inline int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(PK_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  return PK_OK;
}

}  // namespace pk

#define PK_REQUIRE(cond, ...)                                \
  do {                                                       \
    if (!(cond)) return ::pk::set_error(PK_ERR_INVALID, __VA_ARGS__); \
  } while (0)
