// runtime.cpp — library state: one-time initialisation, the HIP stream, the
// device workspace arena, host-pointer staging and error reporting.
//
// BLIS's runtime services are CPU-specific (pack-buffer pools, thread
// decorators; frame/base/bli_pba.c, frame/3/bli_l3_decor.c). Their GPU
// counterparts are kept here:
//   bli_init_once (frame/base/bli_init.c:67-97)   -> init_once(): creates the
//                                                   stream, arena and tables;
//   pack-buffer pools (bli_pba.c:93-250)          -> workspace(): grow-only
//                                                   device slabs per slot;
//   bli_check_error_code (bli_check.c:40-58)      -> abort_with(): message and
//                                                   abort, as BLIS does in
//                                                   FULL error-checking mode.
#include "internal.hpp"

#include <algorithm>
#include <iterator>

#include <atomic>
#include <cstring>
#include <string>

namespace blis_amd {

namespace {

struct State {
    std::once_flag once;
    int device = 0;
    hipStream_t stream = nullptr;
    bool user_stream = false;
    bool sync = true;
    std::mutex ws_mutex;
    void* ws[WS_SLOTS] = {};
    size_t ws_bytes[WS_SLOTS] = {};
    std::mutex api;
};

State& S() {
    static State s;
    return s;
}

} // namespace

std::mutex& api_mutex() { return S().api; }

void abort_with(err_t code, const char* where) {
    std::fprintf(stderr, "libblis_amd: %s: %s (error code %d)\n", where,
                 bli_error_string_for_code(code), (int)code);
    std::fflush(stderr);
    std::abort();
}

void check_hip(hipError_t e, const char* what, const char* file, int line) {
    if (e == hipSuccess) return;
    std::fprintf(stderr, "libblis_amd: HIP error '%s' (%d) from %s at %s:%d\n",
                 hipGetErrorString(e), (int)e, what, file, line);
    std::fflush(stderr);
    std::abort();
}

void init_once() {
    State& s = S();
    std::call_once(s.once, [&s]() {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n == 0) {
            std::fprintf(stderr, "libblis_amd: no HIP device available (%s)\n",
                         hipGetErrorString(e));
            std::abort();
        }
        HIP_CHECK(hipGetDevice(&s.device));
        // A blocking stream: it orders with the legacy default stream, so a
        // call sees all work the caller queued there before it (BLIS calls
        // are synchronous and observe all prior writes).
        HIP_CHECK(hipStreamCreateWithFlags(&s.stream, hipStreamDefault));
    });
}

hipStream_t stream() {
    init_once();
    return S().stream;
}

int device_id() {
    init_once();
    return S().device;
}

bool synchronous() {
    return S().sync != 0;
}

void sync_if_needed() {
    if (S().sync) HIP_CHECK(hipStreamSynchronize(S().stream));
}

void* workspace(size_t bytes, int slot) {
    State& s = S();
    if (slot < 0 || slot >= WS_SLOTS) abort_with(BLIS_INVALID_UKR_ID, "workspace slot");
    std::lock_guard<std::mutex> g(s.ws_mutex);
    if (bytes == 0) bytes = 256;
    if (s.ws_bytes[slot] < bytes) {
        if (s.ws[slot]) {
            // Previous users are ordered on the library stream.
            HIP_CHECK(hipStreamSynchronize(s.stream));
            HIP_CHECK(hipFree(s.ws[slot]));
            s.ws[slot] = nullptr;
            s.ws_bytes[slot] = 0;
        }
        size_t want = bytes + bytes / 4;
        want = (want + 4095) & ~(size_t)4095;
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            abort_with(BLIS_MALLOC_RETURNED_NULL, "workspace allocation");
        }
        s.ws[slot] = p;
        s.ws_bytes[slot] = want;
    }
    return s.ws[slot];
}

bool is_device_pointer(const void* p) {
    if (p == nullptr) return true;
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof(attr));
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;  // unregistered host memory
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

Staged::~Staged() {
    if (buf) {
        if (writeback && host) {
            HIP_CHECK(hipMemcpyAsync((char*)host + lo, buf, bytes, hipMemcpyDeviceToHost,
                                     S().stream));
            HIP_CHECK(hipStreamSynchronize(S().stream));
        }
        HIP_CHECK(hipFree(buf));
    }
}

// Element offsets touched by an m x n view with strides rs, cs.
static void span_of(int64_t m, int64_t n, int64_t rs, int64_t cs, int64_t& lo, int64_t& hi) {
    lo = 0; hi = 0;
    if (m <= 0 || n <= 0) return;
    const int64_t r = (m - 1) * rs, c = (n - 1) * cs;
    lo = (r < 0 ? r : 0) + (c < 0 ? c : 0);
    hi = (r > 0 ? r : 0) + (c > 0 ? c : 0);
}

bool DeviceRanges::is_device(const void* p) {
    if (p == nullptr) return true;
    const uintptr_t u = (uintptr_t)p;
    auto it = std::upper_bound(r.begin(), r.end(), u,
                               [](uintptr_t v, const std::pair<uintptr_t, uintptr_t>& e) {
                                   return v < e.first;
                               });
    if (it != r.begin() && u < std::prev(it)->second) return true;
    if (!is_device_pointer(p)) return false;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) == hipSuccess && size > 0) {
        const std::pair<uintptr_t, uintptr_t> e{(uintptr_t)base, (uintptr_t)base + size};
        r.insert(std::upper_bound(r.begin(), r.end(), e), e);
    } else {
        (void)hipGetLastError();
    }
    return true;
}

static void stage_span(Staged& s, const void* p, Dt dt, int64_t lo, int64_t hi, bool is_output,
                       DeviceRanges* ranges = nullptr) {
    s.dev = const_cast<void*>(p);
    if (p == nullptr || (ranges ? ranges->is_device(p) : is_device_pointer(p))) return;
    const size_t es = dt_size(dt);
    s.host = const_cast<void*>(p);
    s.lo = lo * (int64_t)es;
    s.bytes = (size_t)(hi - lo + 1) * es;
    s.writeback = is_output;
    HIP_CHECK(hipMalloc(&s.buf, s.bytes));
    HIP_CHECK(hipMemcpyAsync(s.buf, (const char*)p + s.lo, s.bytes, hipMemcpyHostToDevice,
                             S().stream));
    s.dev = (char*)s.buf - s.lo;
}

void stage(Staged& s, const void* p, Dt dt, int64_t m, int64_t n, int64_t rs, int64_t cs,
           bool is_output, DeviceRanges* ranges) {
    int64_t lo, hi;
    span_of(m, n, rs, cs, lo, hi);
    if (m <= 0 || n <= 0) { s.dev = const_cast<void*>(p); return; }
    stage_span(s, p, dt, lo, hi, is_output, ranges);
}

void stage_complex(Staged& s, const void* p, Dt dt, int64_t m, int64_t n, int64_t rs,
                   int64_t cs, bool is_output) {
    if (m <= 0 || n <= 0) { s.dev = const_cast<void*>(p); return; }
    int64_t lo, hi;
    span_of(m, n, rs, cs, lo, hi);
    // In reals: element e spans [2e, 2e + 1].
    stage_span(s, p, dt, 2 * lo, 2 * hi + 1, is_output);
}

void stage_vec(Staged& s, const void* p, Dt dt, int64_t n, int64_t inc, bool is_output) {
    if (n <= 0) { s.dev = const_cast<void*>(p); return; }
    int64_t lo, hi;
    span_of(n, 1, inc, 1, lo, hi);
    stage_span(s, p, dt, lo, hi, is_output);
}

} // namespace blis_amd

using namespace blis_amd;

extern "C" {

// Library extension: run subsequent calls on a caller-provided HIP stream
// (nullptr restores the library's own stream). When sync is 0 calls return
// without waiting for the device; results are then ordered on that stream.
__attribute__((visibility("default")))
void bli_amd_set_stream(void* hip_stream, int sync) {
    init_once();
    State& s = S();
    // Calls in flight hold the API mutex for as long as they use the stream.
    std::lock_guard<std::mutex> lock(s.api);
    static hipStream_t own = nullptr;
    if (!s.user_stream) own = s.stream;
    if (hip_stream) { s.stream = (hipStream_t)hip_stream; s.user_stream = true; }
    else { s.stream = own; s.user_stream = false; }
    s.sync = sync != 0;
}

__attribute__((visibility("default")))
void* bli_amd_get_stream(void) { return (void*)stream(); }

__attribute__((visibility("default")))
void bli_amd_synchronize(void) { HIP_CHECK(hipStreamSynchronize(stream())); }

} // extern "C"
