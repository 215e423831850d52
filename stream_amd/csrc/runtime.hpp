// Device runtime for the Stream MI355X path: error handling, the per-process
// stream, grow-only device buffers and pinned host scalars.
//
// Replaces the AVA runtime layer the reference builds on (AvaDeviceArray,
// deep_copy, gpu_memset, ERRCHK; src/core/ava_instanciation.cpp:20).  One
// HIP stream per process at default priority carries every kernel; host
// synchronisation happens only where a caller asks for a host-side value.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

namespace stream_amd {

// Status codes returned by the *_status entry points (0 = ok).
enum Status : int32_t {
    kOk = 0,
    kErrHip = 1,           // a HIP runtime call failed
    kErrArgument = 2,      // invalid argument (size, null pointer, state)
    kErrCapacity = 3,      // a fixed-capacity device structure overflowed
    kErrNotConverged = 4,  // iterative solver hit its iteration cap
    kErrSpin = 5,          // a bounded device wait ran out of spins
};

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const char* what) : std::runtime_error(what), code(c) {}
};

inline void hip_check(hipError_t e, const char* expr, const char* file, int line) {
    if (e != hipSuccess) {
        char buf[512];
        std::snprintf(buf, sizeof(buf), "HIP error %d (%s) at %s:%d: %s", (int)e, hipGetErrorString(e), file, line,
                      expr);
        throw HipError(e, buf);
    }
}
#define SA_HIP(expr) ::stream_amd::hip_check((expr), #expr, __FILE__, __LINE__)

// Last error message, readable through Stream_last_error().
void set_last_error(const std::string& msg);
const char* last_error();

// The process-wide execution context: one non-blocking stream (default
// priority) plus a small pinned host area for scalar read-backs.
struct Context {
    hipStream_t stream = nullptr;
    int device = 0;
    void* pinned = nullptr;  // 4 KiB pinned host staging
    uint32_t* d_flags = nullptr;  // device error/overflow flags (64 words)
};
Context& ctx();

// Grow-only device buffer.  resize() never shrinks; contents are not
// preserved across a growing resize.
// Bytes held by DBuf buffers of this process, now and at the high-water
// mark since the last reset (Stream_device_bytes).
void note_device_bytes(int64_t delta);

template <class T> struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    size_t cap = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) {
            (void)hipFree(p);
            note_device_bytes(-(int64_t)(cap * sizeof(T)));
        }
        p = nullptr;
        n = cap = 0;
    }
    void resize(size_t m) {
        if (m > cap) {
            if (p) {
                SA_HIP(hipFree(p));
                note_device_bytes(-(int64_t)(cap * sizeof(T)));
            }
            p = nullptr;
            cap = 0;
            size_t c = m < 64 ? 64 : m;
            SA_HIP(hipMalloc((void**)&p, c * sizeof(T)));
            cap = c;
            note_device_bytes((int64_t)(c * sizeof(T)));
        }
        n = m;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Blocking copies on the context stream.
template <class T> void h2d(T* d, const T* h, size_t n) {
    if (!n) return;
    SA_HIP(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, ctx().stream));
    SA_HIP(hipStreamSynchronize(ctx().stream));
}
template <class T> void d2h(T* h, const T* d, size_t n) {
    if (!n) return;
    SA_HIP(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, ctx().stream));
    SA_HIP(hipStreamSynchronize(ctx().stream));
}
template <class T> void d2d(T* dst, const T* src, size_t n) {
    if (!n) return;
    SA_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, ctx().stream));
}
// Read one device scalar through the pinned area (synchronises the stream).
template <class T> T read_scalar(const T* d) {
    Context& c = ctx();
    SA_HIP(hipMemcpyAsync(c.pinned, d, sizeof(T), hipMemcpyDeviceToHost, c.stream));
    SA_HIP(hipStreamSynchronize(c.stream));
    return *(const T*)c.pinned;
}

// Section timer for per-kernel figures: while armed, section_mark(s, id)
// records event `id` (0 .. kMaxMarks - 1) on stream s; unarmed it does
// nothing.  sections_elapsed(a, b) is the time in ms from mark a to mark b
// of the last armed run (both marks must have been recorded; synchronises).
constexpr int kMaxMarks = 16;
// Marks of the mesh-and-assemble step, in launch order.
enum SectionMark : int {
    kMarkBuild0 = 0,    // LBVH build starts
    kMarkMorton = 1,    // global box + Morton codes + digit histograms done
    kMarkSort = 2,      // radix sort passes done
    kMarkGather = 3,    // Morton-order gather done
    kMarkKarras = 4,    // hierarchy done
    kMarkRefit = 5,     // refit done
    kMarkCompute0 = 6,  // alpha-shape compute starts
    kMarkFill = 7,      // candidate search done
    kMarkDelaunay = 8,  // local Delaunay + filter (+ exact redo) done
    kMarkScans = 9,     // element / edge scans done
    kMarkCompress = 10, // compress done
    kMarkAsm0 = 11,     // P1 assembly starts
    kMarkAsm = 12,      // P1 assembly kernel done
    kMarkCompress0 = 13, // compress starts (after any host read of the counts)
    kMarkDelaunayMain = 14,  // local Delaunay of every node done (exact redo next)
};
void sections_arm(bool on);
void section_mark(hipStream_t s, int id);
float sections_elapsed(int a, int b);

inline unsigned grid_for(size_t n, unsigned block, unsigned max_blocks = 65535u * 4u) {
    size_t g = (n + block - 1) / block;
    if (g == 0) g = 1;
    if (g > max_blocks) g = max_blocks;
    return (unsigned)g;
}

// Run a C-ABI body, converting exceptions into a status code.
template <class F> int32_t guarded(F&& f) {
    try {
        f();
        return kOk;
    } catch (const HipError& e) {
        set_last_error(e.what());
        return kErrHip;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return kErrArgument;
    }
}

}  // namespace stream_amd
