// Sweep driver: C = A * A for one MatrixMarket file through the safe entry
// point, reporting the outcome as tagged lines an outer sweep parses
// (opsparse_amd/sweep.py).  Counterpart of the reference's
// benchmarks/opsparse_reg.cu:46-117, with the same lines and fields:
//   STATUS,<ok|nonsquare|int32_overflow|oom|cuda_error|exception>
//   PERF,<peak_used_bytes>,<free_before>,<free_after>,<total_bytes>
//   META,<M>,<N>,<A_nnz>,<C_nnz>,<total_flop>
//   TIMING_CSV,<name>,<total_ms>,<gflops>
//   ERROR,<message>                 (when there is one)
// The exit code is 0 for every classified outcome (overflow and
// out-of-memory included) and nonzero only for bad usage or a crash.
//
// usage: opsparse_reg <matrix.mtx>
#include <cstdio>
#include <exception>
#include <string>

#include "opsparse/core_algo/spgemm/spgemm.hpp"
#include "opsparse/profile/perf.hpp"
#include "opsparse/system/hip_common.hpp"

namespace {

// File name without directories and without ".mtx".
std::string matrix_name(const std::string &path)
{
    const std::string::size_type slash = path.rfind('/');
    std::string f = slash == std::string::npos ? path : path.substr(slash + 1);
    if (f.size() > 4 && f.compare(f.size() - 4, 4, ".mtx") == 0) f.resize(f.size() - 4);
    return f;
}

void emit(const char *fmt_status, const std::string &error)
{
    std::printf("STATUS,%s\n", fmt_status);
    if (!error.empty()) std::printf("ERROR,%s\n", error.c_str());
    std::fflush(stdout);
}

}  // namespace

int main(int argc, char **argv)
{
    using namespace opsparse;
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <matrix.mtx>\n", argv[0]);
        return 2;
    }
    const std::string path = argv[1];
    CSR A;
    try {
        A.construct(path);
    } catch (const std::exception &e) {
        emit("exception", std::string("load:") + e.what());
        return 0;
    }
    if (A.M != A.N) {
        std::printf("STATUS,nonsquare\n");
        std::printf("META,%d,%d,%d,-1,0\n", A.M, A.N, A.nnz);
        std::fflush(stdout);
        return 0;
    }
    try {
        A.H2D();
    } catch (const std::exception &e) {
        const hipError_t err = hipGetLastError();
        emit(err == hipErrorOutOfMemory ? "oom" : "cuda_error", std::string("h2d:") + e.what());
        return 0;
    }
    hip_runtime_warmup();
    CSR C;
    SpgemmMeta meta;
    SpgemmTimings timing;
    SpgemmPerf perf;
    // One call, no warm-up: a sweep records the first call of each matrix.
    const SpgemmStatus st = opsparse_spgemm_safe(A, A, C, meta, timing, perf);
    std::printf("STATUS,%s\n", to_cstr(st));
    std::printf("PERF,%zu,%zu,%zu,%zu\n", perf.peak_used_bytes, perf.free_before_bytes, perf.free_after_bytes,
                perf.device_total_bytes);
    std::printf("META,%d,%d,%d,%lld,%lld\n", A.M, A.N, A.nnz, perf.c_nnz, perf.total_flop);
    if (!perf.error_msg.empty()) std::printf("ERROR,%s\n", perf.error_msg.c_str());
    if (st == SpgemmStatus::ok && timing.total > 0) {
        const double gflops = 2.0 * static_cast<double>(perf.total_flop) / 1e9 / timing.total;
        std::printf("TIMING_CSV,%s,%.3f,%.3f\n", matrix_name(path).c_str(), timing.total * 1e3, gflops);
    }
    try {
        C.release();
        A.release();
    } catch (...) {
    }
    std::fflush(stdout);
    return 0;
}
