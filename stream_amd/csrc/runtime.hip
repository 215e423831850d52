// Process context and error reporting for the C-ABI.
#include <mutex>
#include <string>

#include "runtime.hpp"

namespace stream_amd {

namespace {
std::mutex g_err_mu;
std::string g_err;
}  // namespace

void set_last_error(const std::string& msg) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = msg;
    std::fprintf(stderr, "[stream_amd] %s\n", msg.c_str());
}

const char* last_error() {
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_err.c_str();
}

Context& ctx() {
    static Context c;
    static std::once_flag once;
    std::call_once(once, [] {
        SA_HIP(hipGetDevice(&c.device));
        SA_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        SA_HIP(hipHostMalloc(&c.pinned, 4096, hipHostMallocDefault));
        SA_HIP(hipMalloc((void**)&c.d_flags, 64 * sizeof(uint32_t)));
        SA_HIP(hipMemsetAsync(c.d_flags, 0, 64 * sizeof(uint32_t), c.stream));
        SA_HIP(hipStreamSynchronize(c.stream));
    });
    return c;
}

}  // namespace stream_amd

extern "C" {

// Last error message recorded by any entry point ("" when none).
const char* Stream_last_error(void) { return stream_amd::last_error(); }

// Number of visible HIP devices (0 without a GPU); never throws.
int32_t Stream_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// Block until all queued device work is done.
int32_t Stream_synchronize(void) {
    return stream_amd::guarded([] { SA_HIP(hipStreamSynchronize(stream_amd::ctx().stream)); });
}

}  // extern "C"
