// CPU stand-in for <hip/hip_runtime_api.h>: just enough of the HIP host API
// for kernels.hip / kernels.hpp to compile as host C++ in the kernel
// emulation tests (tests/emu).  Test infrastructure only.
#pragma once
#include <stddef.h>
#include <stdint.h>

typedef enum { hipSuccess = 0, hipErrorLaunchFailure = 719 } hipError_t;
typedef void *hipStream_t;

struct dim3 {
    uint32_t x, y, z;
    dim3(uint32_t x_ = 1, uint32_t y_ = 1, uint32_t z_ = 1) : x(x_), y(y_), z(z_) {}
};

inline thread_local hipError_t emu_last_error = hipSuccess;
inline hipError_t hipGetLastError() {
    hipError_t e = emu_last_error;
    emu_last_error = hipSuccess;
    return e;
}
