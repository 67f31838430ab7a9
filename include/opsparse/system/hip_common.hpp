// Type aliases and HIP error handling of the drop-in C++ API.
// Counterpart of include/opsparse/system/cuda_common.hpp (mint/mdouble
// :13-14, check_cuda :19-25, memcpy helpers :35-51, warm-up :53-59), on the
// HIP runtime.  Errors throw std::runtime_error after printing, as the
// reference's check_cuda does (it throws std::exception).
#ifndef OPSPARSE_SYSTEM_HIP_COMMON_HPP_
#define OPSPARSE_SYSTEM_HIP_COMMON_HPP_

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace opsparse {

using mint = int;
using mdouble = double;

inline void check_hip(hipError_t err, const char *file, int line)
{
    if (err != hipSuccess) {
        std::printf("%s in %s at line %d\n", hipGetErrorString(err), file, line);
        throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(err));
    }
}

template <typename T>
inline void d2h(T *dst, const T *src, std::size_t bytes)
{
    check_hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), __FILE__, __LINE__);
}

template <typename T>
inline void h2d(T *dst, const T *src, std::size_t bytes)
{
    check_hip(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), __FILE__, __LINE__);
}

template <typename T>
inline void d2d(T *dst, const T *src, std::size_t bytes)
{
    check_hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), __FILE__, __LINE__);
}

// Forces HIP runtime initialisation before timed work.
inline void hip_runtime_warmup()
{
    void *d = nullptr;
    check_hip(hipMalloc(&d, 4), __FILE__, __LINE__);
    check_hip(hipFree(d), __FILE__, __LINE__);
    check_hip(hipDeviceSynchronize(), __FILE__, __LINE__);
}

}  // namespace opsparse

#define OPSPARSE_CHECK_HIP(err) (::opsparse::check_hip((err), __FILE__, __LINE__))

#endif  // OPSPARSE_SYSTEM_HIP_COMMON_HPP_
