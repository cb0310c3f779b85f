// runtime.cpp -- error text, device selection, pointer kinds, kernel profiler.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "plain.h"
#include "profiler.h"
#include "runtime.h"
#include "tuning.h"
#include "sedi_internal.h"

namespace sedi {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

const char* last_error() { return g_err.c_str(); }

int ensure_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        set_error("no HIP device available (%s)", hipGetErrorString(e));
        return SEDI_ERR_HIP;
    }
    return SEDI_OK;
}

bool is_device_pointer(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

// ---- profiler --------------------------------------------------------------
namespace {
struct Rec {
    int64_t launches = 0;
    double ms = 0.0;
};
struct Pending {
    std::string name;
    hipEvent_t a, b;
};
std::mutex g_mu;
std::string g_target;  // "" = off, "*" = all
std::map<std::string, Rec> g_recs;
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_pool;

hipEvent_t take_event() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

bool profile_any() {
    std::lock_guard<std::mutex> lk(g_mu);
    return !g_target.empty();
}

static std::atomic<bool> g_graphs{true};
bool graphs_enabled() { return g_graphs.load(); }
void set_graphs_enabled(bool on) { g_graphs.store(on); }

static thread_local bool g_suppress = false;
ProfileSuppress::ProfileSuppress() : prev(g_suppress) { g_suppress = true; }
ProfileSuppress::~ProfileSuppress() { g_suppress = prev; }

bool profile_wants(const char* name) {
    if (g_suppress || g_target.empty()) return false;
    return g_target == "*" || g_target == name;
}

int profile_begin(const char* name, hipStream_t stream) {
    if (!profile_wants(name)) return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    Pending p;
    p.name = name;
    p.a = take_event();
    p.b = take_event();
    if (!p.a || !p.b) return -1;
    (void)hipEventRecord(p.a, stream);
    g_pending.push_back(p);
    return (int)g_pending.size() - 1;
}

void profile_end(int token, hipStream_t stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (token < 0 || token >= (int)g_pending.size()) return;
    (void)hipEventRecord(g_pending[token].b, stream);
}

void profile_flush() {
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& p : g_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            Rec& r = g_recs[p.name];
            r.launches += 1;
            r.ms += ms;
        }
        g_pool.push_back(p.a);
        g_pool.push_back(p.b);
    }
    g_pending.clear();
}

}  // namespace sedi

extern "C" {

const char* sedi_last_error(void) { return sedi::last_error(); }

int sedi_version(void) { return 1; }

int sedi_init(int device) {
    int rc = sedi::ensure_device();
    if (rc) return rc;
    SEDI_HIP(hipSetDevice(device));
    SEDI_HIP(hipFree(nullptr));
    // this device's address of the allocation-free solve's working array
    return sedi::plain_noalloc_setup();
}

int sedi_graphs_enable(int on) {
    sedi::set_graphs_enabled(on != 0);
    return SEDI_OK;
}

int sedi_tuning_lookup(const char* key, const sedi_grid* grid, int fallback) {
    if (!key || !grid) return fallback;
    const sedi::Shape sh{(int)grid->n1, (int)grid->n2, (int)grid->n3};
    const bool nearest = std::strncmp(key, "k:", 2) == 0;
    return sedi::tuned_value(key, sh, fallback, nearest);
}

int sedi_profile_enable(const char* kernel_name) {
    sedi::profile_flush();
    std::lock_guard<std::mutex> lk(sedi::g_mu);
    sedi::g_target = kernel_name ? kernel_name : "";
    return SEDI_OK;
}

int sedi_profile_get(const char* kernel_name, int64_t* launches, double* total_ms) {
    sedi::profile_flush();
    std::lock_guard<std::mutex> lk(sedi::g_mu);
    auto it = sedi::g_recs.find(kernel_name ? kernel_name : "");
    if (it == sedi::g_recs.end()) {
        sedi::set_error("kernel '%s' was not timed", kernel_name ? kernel_name : "(null)");
        return SEDI_ERR_ARG;
    }
    if (launches) *launches = it->second.launches;
    if (total_ms) *total_ms = it->second.ms;
    return SEDI_OK;
}

int sedi_profile_reset(void) {
    sedi::profile_flush();
    std::lock_guard<std::mutex> lk(sedi::g_mu);
    sedi::g_recs.clear();
    return SEDI_OK;
}

}  // extern "C"
