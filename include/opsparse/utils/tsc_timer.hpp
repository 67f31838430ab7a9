// Host wall clock used by every timing in the library and its benchmark:
// CLOCK_MONOTONIC_RAW in seconds, the convention of the reference's
// fast_clock_time (include/opsparse/utils/tsc_timer.hpp:10-19).
#ifndef OPSPARSE_UTILS_TSC_TIMER_HPP_
#define OPSPARSE_UTILS_TSC_TIMER_HPP_

#include <time.h>

namespace opsparse {

inline double fast_clock_time()
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC_RAW, &ts);
    return static_cast<double>(ts.tv_sec) + static_cast<double>(ts.tv_nsec) * 1e-9;
}

}  // namespace opsparse

#endif  // OPSPARSE_UTILS_TSC_TIMER_HPP_
