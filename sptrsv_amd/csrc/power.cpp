// power.cpp -- board power sampling through amdsmi, replacing the NVML
// polling thread of the reference (nvmlPower.cpp:22-158, nvmlPower.hpp:17-18).
//
// Differences by design: amdsmi is loaded with dlopen when sampling starts
// (the solver library does not depend on it); the thread samples on a timer
// instead of busy-polling; the GPU is matched by PCI address to the HIP
// device; errors are returned instead of exit(0) (nvmlPower.cpp:71-137).
// The reading is amdsmi_power_info_t::current_socket_power (W), the MI300+
// field, else average_socket_power; amdsmi sets fields it does not support
// to UINT32_MAX, and such a field (or 0) is no reading -- a sample without a
// reading is not counted.  The running mean is returned in mW like
// nvmlAPIEnd.  Like nvmlAPIRun / nvmlAPIEnd (nvmlPower.cpp:68, 150), each
// run initialises amdsmi and each end shuts it down; a run still going when
// the library is unloaded is stopped and shut down then.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <atomic>
#include <chrono>
#include <mutex>
#include <string>
#include <thread>

#include "common.hpp"

namespace {

// Subset of amdsmi.h used here (types spelled as in the public header).
typedef void* smi_handle;
struct smi_power_info {
    uint64_t socket_power;
    uint32_t current_socket_power;
    uint32_t average_socket_power;
    uint64_t gfx_voltage;
    uint64_t soc_voltage;
    uint64_t mem_voltage;
    uint32_t power_limit;
    uint64_t reserved[18];
};
typedef int (*fn_init)(uint64_t);
typedef int (*fn_shut)(void);
typedef int (*fn_sockets)(uint32_t*, smi_handle*);
typedef int (*fn_procs)(smi_handle, uint32_t*, smi_handle*);
typedef int (*fn_power)(smi_handle, smi_power_info*);
typedef int (*fn_bdf)(smi_handle, uint64_t*);
constexpr uint64_t kInitGpus = 1ull << 1;
constexpr uint32_t kUnsupported = 0xFFFFFFFFu;  // amdsmi: field not supported

// Socket power in W from one amdsmi reading, or a negative value when the
// reading holds none (fields unsupported or zero).
double watts_of(const smi_power_info& q) {
    if (q.current_socket_power != 0 && q.current_socket_power != kUnsupported) return q.current_socket_power;
    if (q.average_socket_power != 0 && q.average_socket_power != kUnsupported) return q.average_socket_power;
    return -1.0;
}

struct Sampler {
    std::mutex mu;
    void* lib = nullptr;
    fn_init init = nullptr;
    fn_shut shut = nullptr;
    fn_sockets sockets = nullptr;
    fn_procs procs = nullptr;
    fn_power power = nullptr;
    fn_bdf bdf = nullptr;
    smi_handle gpu = nullptr;
    std::thread th;
    std::atomic<bool> run{false};
    double mean_w = 0.0;
    int samples = 0;
    bool inited = false;
};
Sampler g_s;

bool load(Sampler& s) {
    if (s.lib) return true;
    const char* names[] = {"libamd_smi.so", "libamd_smi.so.26", "/opt/rocm/lib/libamd_smi.so"};
    for (const char* n : names) {
        s.lib = dlopen(n, RTLD_NOW | RTLD_LOCAL);
        if (s.lib) break;
    }
    if (!s.lib) return false;
    s.init = (fn_init)dlsym(s.lib, "amdsmi_init");
    s.shut = (fn_shut)dlsym(s.lib, "amdsmi_shut_down");
    s.sockets = (fn_sockets)dlsym(s.lib, "amdsmi_get_socket_handles");
    s.procs = (fn_procs)dlsym(s.lib, "amdsmi_get_processor_handles");
    s.power = (fn_power)dlsym(s.lib, "amdsmi_get_power_info");
    s.bdf = (fn_bdf)dlsym(s.lib, "amdsmi_get_gpu_device_bdf");
    return s.init && s.shut && s.sockets && s.procs && s.power;
}

// amdsmi_bdf_t packs function:3, device:5, bus:8, domain:48 from bit 0.
uint64_t pack_bdf(int domain, int bus, int dev, int fn) {
    return (uint64_t)(fn & 7) | ((uint64_t)(dev & 31) << 3) | ((uint64_t)(bus & 255) << 8) |
           ((uint64_t)domain << 16);
}

smi_handle find_gpu(Sampler& s, int device) {
    uint32_t ns = 0;
    if (s.sockets(&ns, nullptr) != 0 || ns == 0) return nullptr;
    smi_handle socks[64];
    if (ns > 64) ns = 64;
    if (s.sockets(&ns, socks) != 0) return nullptr;
    hipDeviceProp_t p;
    bool have_bdf = hipGetDeviceProperties(&p, device) == hipSuccess;
    uint64_t want = have_bdf ? pack_bdf(p.pciDomainID, p.pciBusID, p.pciDeviceID, 0) : 0;
    smi_handle first = nullptr;
    int idx = 0;
    for (uint32_t i = 0; i < ns; ++i) {
        uint32_t np = 0;
        if (s.procs(socks[i], &np, nullptr) != 0 || np == 0) continue;
        smi_handle procs[64];
        if (np > 64) np = 64;
        if (s.procs(socks[i], &np, procs) != 0) continue;
        for (uint32_t j = 0; j < np; ++j, ++idx) {
            if (!first) first = procs[j];
            uint64_t b = 0;
            if (have_bdf && s.bdf && s.bdf(procs[j], &b) == 0 && (b & ~7ull) == (want & ~7ull))
                return procs[j];
        }
    }
    return first;  // single-GPU fallback
}

// Stops the sampling thread and shuts amdsmi down (caller holds s.mu).
void stop_locked(Sampler& s) {
    s.run = false;
    if (s.th.joinable()) s.th.join();
    if (s.inited) {
        (void)s.shut();
        s.inited = false;
    }
    s.gpu = nullptr;
}

// Library teardown: a sampler left running would otherwise hold a joinable
// thread and an initialised amdsmi into process exit.
__attribute__((destructor)) void power_teardown() {
    std::lock_guard<std::mutex> g(g_s.mu);
    stop_locked(g_s);
}

}  // namespace

using namespace sptrsv;

extern "C" {

int amdsmiAPIRun(const char* label, int device, int period_us) {
    (void)label;
    Sampler& s = g_s;
    std::lock_guard<std::mutex> g(s.mu);
    if (s.run.load()) return fail(SPTRSV_E_INVALID_ARG, "power sampling already running");
    if (!load(s)) return fail(SPTRSV_E_UNAVAILABLE, "amdsmi library not loadable");
    if (s.init(kInitGpus) != 0) return fail(SPTRSV_E_UNAVAILABLE, "amdsmi_init failed");
    s.inited = true;
    s.gpu = find_gpu(s, device);
    smi_power_info pi{};
    if (!s.gpu || s.power(s.gpu, &pi) != 0 || watts_of(pi) < 0) {
        const char* why = !s.gpu ? "no amdsmi GPU handle" : "amdsmi reports no socket power for this GPU";
        stop_locked(s);
        return fail(SPTRSV_E_UNAVAILABLE, why);
    }
    s.mean_w = 0.0;
    s.samples = 0;
    s.run = true;
    int period = period_us > 0 ? period_us : 1000;
    s.th = std::thread([&s, period]() {
        while (s.run.load()) {
            smi_power_info q{};
            if (s.power(s.gpu, &q) == 0) {
                const double w = watts_of(q);
                if (w >= 0) {
                    ++s.samples;
                    s.mean_w += (w - s.mean_w) / s.samples;  // running mean, nvmlPower.cpp:52
                }
            }
            std::this_thread::sleep_for(std::chrono::microseconds(period));
        }
    });
    return SPTRSV_OK;
}

float amdsmiAPIEnd(void) {
    Sampler& s = g_s;
    std::lock_guard<std::mutex> g(s.mu);
    if (!s.run.load()) return -1.0f;
    stop_locked(s);
    return s.samples ? (float)(s.mean_w * 1000.0) : -1.0f;
}

int amdsmiAPISamples(void) { return g_s.samples; }

}  // extern "C"
