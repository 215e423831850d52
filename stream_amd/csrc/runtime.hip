// Process context and error reporting for the C-ABI.
#include <atomic>
#include <mutex>
#include <string>

#include "runtime.hpp"

namespace stream_amd {

namespace {
std::mutex g_err_mu;
std::string g_err;
std::atomic<int64_t> g_bytes{0};
std::atomic<int64_t> g_peak{0};
}  // namespace

void note_device_bytes(int64_t delta) {
    const int64_t now = g_bytes.fetch_add(delta) + delta;
    int64_t pk = g_peak.load();
    while (now > pk && !g_peak.compare_exchange_weak(pk, now)) {
    }
}

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

namespace {
struct Sections {
    bool armed = false;
    bool created = false;
    hipEvent_t ev[kMaxMarks];
    bool recorded[kMaxMarks] = {};
};
Sections& sections() {
    static Sections t;
    return t;
}
}  // namespace

void sections_arm(bool on) {
    Sections& t = sections();
    if (on && !t.created) {
        for (auto& e : t.ev) SA_HIP(hipEventCreate(&e));
        t.created = true;
    }
    for (bool& r : t.recorded) r = false;
    t.armed = on;
}

void section_mark(hipStream_t s, int id) {
    Sections& t = sections();
    if (!t.armed || id < 0 || id >= kMaxMarks) return;
    SA_HIP(hipEventRecord(t.ev[id], s));
    t.recorded[id] = true;
}

float sections_elapsed(int a, int b) {
    Sections& t = sections();
    if (a < 0 || b < 0 || a >= kMaxMarks || b >= kMaxMarks || !t.created || !t.recorded[a] || !t.recorded[b])
        throw std::logic_error("section marks not recorded");
    SA_HIP(hipEventSynchronize(t.ev[b]));
    float ms = 0.f;
    SA_HIP(hipEventElapsedTime(&ms, t.ev[a], t.ev[b]));
    return ms;
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

// Device bytes held by the library's buffers now and at the high-water mark
// (reset_peak != 0 restarts the mark at the current value), and the device
// free / total memory as the runtime reports it.  Any pointer may be null.
int32_t Stream_device_bytes(uint64_t* current, uint64_t* peak, int32_t reset_peak, uint64_t* dev_free,
                            uint64_t* dev_total) {
    return stream_amd::guarded([&] {
        if (reset_peak) stream_amd::g_peak.store(stream_amd::g_bytes.load());
        if (current) *current = (uint64_t)stream_amd::g_bytes.load();
        if (peak) *peak = (uint64_t)stream_amd::g_peak.load();
        if (dev_free || dev_total) {
            size_t f = 0, t = 0;
            (void)stream_amd::ctx();
            SA_HIP(hipMemGetInfo(&f, &t));
            if (dev_free) *dev_free = f;
            if (dev_total) *dev_total = t;
        }
    });
}

// Points as testcases/IMR26/uniform2D.cpp:130-141 and uniform3D.cpp draw
// them: srand(seed), then (float)rand() / RAND_MAX per coordinate, point by
// point; out [n][dim].  Host only (the C library's generator).
int32_t Stream_uniform_cloud_f32(uint32_t n, uint32_t dim, uint32_t seed, float* out) {
    return stream_amd::guarded([&] {
        if (!out || dim < 1 || dim > 3) throw std::invalid_argument("Stream_uniform_cloud_f32: bad arguments");
        srand(seed);
        const uint64_t total = (uint64_t)n * dim;
        for (uint64_t k = 0; k < total; ++k) out[k] = (float)rand() / RAND_MAX;
    });
}

// Block until all queued device work is done.
int32_t Stream_synchronize(void) {
    return stream_amd::guarded([] { SA_HIP(hipStreamSynchronize(stream_amd::ctx().stream)); });
}

}  // extern "C"
