// Host-memory implementation of the HIP runtime subset the device context
// uses (see hip/hip_runtime.h in this directory).  Every device allocation
// is a heap block in a registry; copies, fills and the checked kernel
// launches verify their ranges against the live blocks, so a stale pointer
// (a buffer freed while a captured graph still names it) is reported by
// name, and AddressSanitizer sees every host-side access.

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "hip/hip_runtime.h"

struct mock_graph {
    std::vector<std::pair<const char*, std::function<void()>>> nodes;
};
struct mock_graph_exec {
    std::vector<std::pair<const char*, std::function<void()>>> nodes;
};
struct mock_stream {
    bool capturing = false;
    mock_graph* graph = nullptr;
};
struct mock_event {
    int recorded = 0;
};

namespace {
std::map<uintptr_t, size_t>& registry() {
    static std::map<uintptr_t, size_t> r;
    return r;
}
std::map<uintptr_t, size_t>& symbols() {
    static std::map<uintptr_t, size_t> r;
    return r;
}
long long g_fail_after = -1;
long long g_malloc_calls = 0;
int g_capturing = 0;

// the live block (heap allocation or device variable) holding address a, or end
bool find_block(uintptr_t a, uintptr_t& base, size_t& size) {
    for (auto* r : {&registry(), &symbols()}) {
        auto it = r->upper_bound(a);
        if (it == r->begin()) continue;
        --it;
        if (a >= it->first && a <= it->first + it->second) {
            base = it->first;
            size = it->second;
            return true;
        }
    }
    return false;
}

[[noreturn]] void die(const char* what, const char* msg) {
    std::fprintf(stderr, "MOCK-HIP ERROR in %s: %s\n", what, msg);
    std::fflush(stderr);
    std::abort();
}

// A synchronous runtime call while a stream is being captured would end
// the capture with an error on the real runtime.
void no_capture(const char* what) {
    if (g_capturing) die(what, "synchronous runtime call during stream capture");
}
}  // namespace

namespace mock {
void check_dev(const void* p, size_t bytes, const char* what, bool allow_null) {
    if (!p) {
        if (allow_null) return;
        die(what, "null device pointer");
    }
    const uintptr_t a = (uintptr_t)p;
    uintptr_t base = 0;
    size_t size = 0;
    if (!find_block(a, base, size)) die(what, "pointer is not inside a live device allocation");
    if (a + bytes > base + size) {
        char m[256];
        std::snprintf(m, sizeof(m), "range [%p, +%zu) outside the live allocation [%p, +%zu)", p, bytes,
                      (void*)base, size);
        die(what, m);
    }
}

const void* symbol(const void* p, size_t bytes) {
    symbols()[(uintptr_t)p] = bytes;
    return p;
}

void launch(hipStream_t s, const char* name, std::function<void()> body) {
    if (!s) die(name, "launch on a null stream");
    if (s->capturing)
        s->graph->nodes.emplace_back(name, std::move(body));
    else
        body();
}

void fail_malloc_after(long long n) { g_fail_after = n; }
long long live_allocations() { return (long long)registry().size(); }
long long malloc_calls() { return g_malloc_calls; }
}  // namespace mock

hipError_t hipMalloc(void** p, size_t bytes) {
    no_capture("hipMalloc");
    ++g_malloc_calls;
    *p = nullptr;
    if (g_fail_after == 0) {
        g_fail_after = -1;
        return hipErrorOutOfMemory;
    }
    if (g_fail_after > 0) --g_fail_after;
    void* q = std::malloc(bytes ? bytes : 1);
    if (!q) return hipErrorOutOfMemory;
    std::memset(q, 0xA5, bytes);  // device memory is not zeroed
    registry()[(uintptr_t)q] = bytes;
    *p = q;
    return hipSuccess;
}

hipError_t hipFree(void* p) {
    no_capture("hipFree");
    if (!p) return hipSuccess;
    auto& r = registry();
    auto it = r.find((uintptr_t)p);
    if (it == r.end()) die("hipFree", "pointer is not the base of a live device allocation");
    r.erase(it);
    std::free(p);
    return hipSuccess;
}

hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    no_capture("hipMemcpy");
    if (bytes == 0) return hipSuccess;
    if (kind == hipMemcpyHostToDevice || kind == hipMemcpyDeviceToDevice) mock::check_dev(dst, bytes, "hipMemcpy dst");
    if (kind == hipMemcpyDeviceToHost || kind == hipMemcpyDeviceToDevice) mock::check_dev(src, bytes, "hipMemcpy src");
    std::memcpy(dst, src, bytes);
    return hipSuccess;
}

hipError_t hipMemset(void* p, int value, size_t bytes) {
    no_capture("hipMemset");
    if (bytes == 0) return hipSuccess;
    mock::check_dev(p, bytes, "hipMemset");
    std::memset(p, value, bytes);
    return hipSuccess;
}

hipError_t hipMemsetAsync(void* p, int value, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    mock::check_dev(p, bytes, "hipMemsetAsync");
    mock::launch(s, "hipMemsetAsync", [=] {
        mock::check_dev(p, bytes, "hipMemsetAsync (replay)");
        std::memset(p, value, bytes);
    });
    return hipSuccess;
}

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    *s = new mock_stream();
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    if (s && s->capturing) die("hipStreamDestroy", "stream destroyed during capture");
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
    if (s && s->capturing) die("hipStreamSynchronize", "synchronize of a capturing stream");
    no_capture("hipStreamSynchronize");
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    if (!s) return hipErrorStreamCaptureUnsupported;
    if (s->capturing) die("hipStreamBeginCapture", "stream already capturing");
    s->capturing = true;
    s->graph = new mock_graph();
    ++g_capturing;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
    if (!s || !s->capturing) die("hipStreamEndCapture", "stream not capturing");
    s->capturing = false;
    --g_capturing;
    *g = s->graph;
    s->graph = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
    auto* x = new mock_graph_exec();
    x->nodes = g->nodes;
    *e = x;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    delete g;
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
    if (!e) die("hipGraphLaunch", "null graph");
    if (!s || s->capturing) die("hipGraphLaunch", "launch on a null or capturing stream");
    for (auto& n : e->nodes) n.second();
    return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e) {
    *e = new mock_event();
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    if (!e) die("hipEventRecord", "null event");
    if (s && s->capturing) die("hipEventRecord", "event recorded during capture");
    e->recorded = 1;
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
    if (!e || !e->recorded) die("hipEventSynchronize", "event never recorded");
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    if (!a || !b || !a->recorded || !b->recorded) die("hipEventElapsedTime", "event never recorded");
    *ms = 0.001f;
    return hipSuccess;
}
hipError_t hipGetSymbolAddress(void** p, const void* symbol) {
    if (!symbols().count((uintptr_t)symbol)) die("hipGetSymbolAddress", "not a registered device variable");
    *p = const_cast<void*>(symbol);
    return hipSuccess;
}
hipError_t hipGetDeviceCount(int* n) {
    *n = 1;
    return hipSuccess;
}
hipError_t hipGetDevice(int* d) {
    *d = 0;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) {
    const char* e = std::getenv("MOCK_NUM_CUS");
    *v = e ? std::atoi(e) : 256;
    return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
    std::memset(p, 0, sizeof(*p));
    std::strcpy(p->name, "mock");
    std::strcpy(p->gcnArchName, "mock");
    p->multiProcessorCount = 256;
    return hipSuccess;
}
hipError_t hipDeviceSynchronize() {
    no_capture("hipDeviceSynchronize");
    return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t e) {
    switch (e) {
        case hipSuccess: return "no error";
        case hipErrorOutOfMemory: return "out of memory (injected)";
        case hipErrorInvalidValue: return "invalid value";
        default: return "mock error";
    }
}
