/* Host stand-in for <hip/hip_runtime.h>, used only by tests/harness/field_host.cpp so
 * that the device field arithmetic (icp-gpu_amd/csrc/field.hpp) can be compiled by g++
 * and checked on the CPU.  It defines the function qualifiers as plain C++. */
#pragma once
#define __device__
#define __host__
#define __forceinline__ inline __attribute__((always_inline))
#define __noinline__ __attribute__((noinline))
/* one lane per "wave" on the host: reading the first active lane is the identity */
#define __builtin_amdgcn_readfirstlane(x) (x)
