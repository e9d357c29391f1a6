// gu_util.cpp -- power sampling and trace ranges (util_device.cu:79-130).
//
// The reference samples board power through NVML on an std::async thread
// between start_power_sample() and end_power_sample() and returns the mean
// in watts, or -1 on boards NVML cannot read.  Here the sampler reads the
// socket power of the chosen device through AMD SMI (libamd_smi), and also
// reads the device's energy accumulator at both ends, so the interval's
// average power is known even when it is shorter than one sampling period.
// AMD SMI is loaded with dlopen when a sampler is first started: the
// library itself never depends on it, and a host without it gets
// GU_ERR_UNSUPPORTED (the reference's "-1, power can only be measured on
// Tesla boards").
//
// Trace ranges: gu_trace_push / gu_trace_pop forward to roctxRangePushA /
// roctxRangePop of librocprofiler-sdk-roctx (dlopen'ed on first use), so a
// `rocprofv3 --marker-trace` run attributes kernels to the engine's phases;
// without the library they do nothing.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <climits>
#include <cstdint>
#include <mutex>
#include <new>
#include <thread>

#include "../../include/graph_utils_amd.h"

namespace {

// ---- AMD SMI, resolved at run time -----------------------------------------
// Only the handful of types and entry points used here are declared; their
// layout follows amd_smi/amdsmi.h of ROCm 7.
typedef void *smi_handle;
typedef struct {
  uint64_t socket_power;          // W
  uint32_t current_socket_power;  // W (MI300 and later)
  uint32_t average_socket_power;  // W (earlier boards)
  uint64_t gfx_voltage, soc_voltage, mem_voltage;
  uint32_t power_limit;
  uint64_t reserved[18];
} smi_power_info;
constexpr uint64_t kSmiInitGpus = 1ull << 1;  // AMDSMI_INIT_AMD_GPUS

struct Smi {
  void *lib = nullptr;
  int (*init)(uint64_t) = nullptr;
  int (*shut_down)() = nullptr;
  int (*sockets)(uint32_t *, smi_handle *) = nullptr;
  int (*processors)(smi_handle, uint32_t *, smi_handle *) = nullptr;
  int (*from_bdf)(uint64_t, smi_handle *) = nullptr;
  int (*power)(smi_handle, smi_power_info *) = nullptr;
  int (*energy)(smi_handle, uint64_t *, float *, uint64_t *) = nullptr;
  int users = 0;  // samplers alive (amdsmi_init / amdsmi_shut_down pair)
};
std::mutex g_smi_mu;
Smi g_smi;

bool smi_load_locked() {
  if (g_smi.lib) return true;
  void *h = dlopen("libamd_smi.so", RTLD_NOW | RTLD_LOCAL);
  if (!h) h = dlopen("/opt/rocm/lib/libamd_smi.so", RTLD_NOW | RTLD_LOCAL);
  if (!h) return false;
  Smi s;
  s.lib = h;
  s.init = (int (*)(uint64_t))dlsym(h, "amdsmi_init");
  s.shut_down = (int (*)())dlsym(h, "amdsmi_shut_down");
  s.sockets = (int (*)(uint32_t *, smi_handle *))dlsym(h, "amdsmi_get_socket_handles");
  s.processors = (int (*)(smi_handle, uint32_t *, smi_handle *))dlsym(h, "amdsmi_get_processor_handles");
  s.from_bdf = (int (*)(uint64_t, smi_handle *))dlsym(h, "amdsmi_get_processor_handle_from_bdf");
  s.power = (int (*)(smi_handle, smi_power_info *))dlsym(h, "amdsmi_get_power_info");
  s.energy = (int (*)(smi_handle, uint64_t *, float *, uint64_t *))dlsym(h, "amdsmi_get_energy_count");
  if (!s.init || !s.shut_down || !s.sockets || !s.processors || !s.power) {
    dlclose(h);
    return false;
  }
  g_smi = s;
  return true;
}

// Watts from one power reading: the current socket power on MI300 and
// later, else the socket or averaged power; -1 when none is reported.
double smi_watts(const smi_power_info &p) {
  const uint32_t none32 = UINT32_MAX;
  if (p.current_socket_power != none32 && p.current_socket_power != 0) return p.current_socket_power;
  if (p.socket_power != (uint64_t)none32 && p.socket_power != UINT64_MAX && p.socket_power != 0)
    return (double)p.socket_power;
  if (p.average_socket_power != none32 && p.average_socket_power != 0) return p.average_socket_power;
  return -1.0;
}

// The AMD SMI processor of HIP device `dev`: by PCI address, else the only
// GPU processor when exactly one is visible.
bool smi_processor(int dev, smi_handle *out) {
  int bus = 0, devno = 0, dom = 0;
  if (g_smi.from_bdf && hipDeviceGetAttribute(&bus, hipDeviceAttributePciBusId, dev) == hipSuccess &&
      hipDeviceGetAttribute(&devno, hipDeviceAttributePciDeviceId, dev) == hipSuccess &&
      hipDeviceGetAttribute(&dom, hipDeviceAttributePciDomainId, dev) == hipSuccess) {
    // amdsmi_bdf_t: function (3 bits), device (5), bus (8), domain (48)
    const uint64_t bdf = ((uint64_t)(devno & 31) << 3) | ((uint64_t)(bus & 255) << 8) | ((uint64_t)dom << 16);
    if (g_smi.from_bdf(bdf, out) == 0) return true;
  }
  uint32_t ns = 0;
  if (g_smi.sockets(&ns, nullptr) != 0 || ns == 0 || ns > 64) return false;
  smi_handle socks[64];
  if (g_smi.sockets(&ns, socks) != 0) return false;
  smi_handle found = nullptr;
  int count = 0;
  for (uint32_t i = 0; i < ns; i++) {
    uint32_t np = 0;
    if (g_smi.processors(socks[i], &np, nullptr) != 0 || np == 0 || np > 64) continue;
    smi_handle procs[64];
    if (g_smi.processors(socks[i], &np, procs) != 0) continue;
    for (uint32_t j = 0; j < np; j++) {
      found = procs[j];
      count++;
    }
  }
  if (count != 1) return false;
  *out = found;
  return true;
}

}  // namespace

struct gu_power_s {
  int device = 0;
  int period_ms = 100;
  smi_handle proc = nullptr;
  std::atomic<bool> run{false};
  std::thread th;
  double sum = 0.0, maxw = 0.0;
  int samples = 0;
  bool have_energy = false;
  uint64_t e0 = 0, t0 = 0;
  float res_uj = 0.0f;
  std::chrono::steady_clock::time_point w0;
};

extern "C" int gu_power_start(int32_t device, int32_t period_ms, gu_power_t *out) {
  if (!out || period_ms <= 0) return GU_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return GU_ERR_NODEVICE;
  if (device < 0 || device >= count) return GU_ERR_INVALID;
  gu_power_s *p = new (std::nothrow) gu_power_s();
  if (!p) return GU_ERR_NOMEM;
  p->device = device;
  p->period_ms = period_ms;
  {
    std::lock_guard<std::mutex> lk(g_smi_mu);
    if (!smi_load_locked()) {
      delete p;
      return GU_ERR_UNSUPPORTED;
    }
    if (g_smi.users == 0 && g_smi.init(kSmiInitGpus) != 0) {
      delete p;
      return GU_ERR_UNSUPPORTED;
    }
    g_smi.users++;
    smi_power_info pi{};
    if (!smi_processor(device, &p->proc) || g_smi.power(p->proc, &pi) != 0 || smi_watts(pi) < 0) {
      if (--g_smi.users == 0) g_smi.shut_down();
      delete p;
      return GU_ERR_UNSUPPORTED;
    }
    if (g_smi.energy) p->have_energy = g_smi.energy(p->proc, &p->e0, &p->res_uj, &p->t0) == 0 && p->res_uj > 0;
  }
  p->w0 = std::chrono::steady_clock::now();
  p->run = true;
  try {
    p->th = std::thread([p]() {
    // one reading, then one per period until stopped (power_sample's loop)
    do {
      smi_power_info pi{};
      if (g_smi.power(p->proc, &pi) == 0) {
        const double w = smi_watts(pi);
        if (w >= 0) {
          p->sum += w;
          p->maxw = p->samples ? (w > p->maxw ? w : p->maxw) : w;
          p->samples++;
        }
      }
      for (int slept = 0; slept < p->period_ms && p->run.load(); slept += 5)
        std::this_thread::sleep_for(std::chrono::milliseconds(p->period_ms - slept < 5 ? p->period_ms - slept : 5));
    } while (p->run.load());
    });
  } catch (...) {  // no thread: release AMD SMI and report it
    std::lock_guard<std::mutex> lk(g_smi_mu);
    if (--g_smi.users == 0) g_smi.shut_down();
    delete p;
    return GU_ERR_NOMEM;
  }
  *out = p;
  return GU_OK;
}

extern "C" int gu_power_stop(gu_power_t p, gu_power_result *r) {
  if (!p) return GU_ERR_INVALID;
  // end_power_sample synchronizes the device before stopping the sampler
  int prev = 0;
  hipError_t e = hipGetDevice(&prev);
  if (e == hipSuccess) e = hipSetDevice(p->device);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  (void)hipSetDevice(prev);
  p->run = false;
  if (p->th.joinable()) p->th.join();
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - p->w0).count();
  gu_power_result res{};
  res.samples = p->samples;
  res.avg_w = p->samples ? p->sum / p->samples : -1.0;
  res.max_w = p->samples ? p->maxw : -1.0;
  res.seconds = secs;
  res.energy_avg_w = -1.0;
  res.energy_j = -1.0;
  {
    std::lock_guard<std::mutex> lk(g_smi_mu);
    uint64_t e1 = 0, t1 = 0;
    float r1 = 0.0f;
    // an accumulator that did not move (an interval shorter than its
    // update period) reports nothing rather than 0 W
    if (p->have_energy && g_smi.energy(p->proc, &e1, &r1, &t1) == 0 && e1 > p->e0) {
      res.energy_j = (double)(e1 - p->e0) * (double)p->res_uj * 1e-6;
      // the counter's own timestamps (ns) when it reports them, else the wall clock
      const double dt = t1 > p->t0 ? (double)(t1 - p->t0) * 1e-9 : secs;
      if (dt > 0) res.energy_avg_w = res.energy_j / dt;
    }
    if (--g_smi.users == 0) g_smi.shut_down();
  }
  delete p;
  if (r) *r = res;
  return e == hipSuccess ? GU_OK : GU_ERR_HIP;
}

// ---- trace ranges -----------------------------------------------------------
namespace {
std::once_flag g_tx_once;
int (*g_tx_push)(const char *) = nullptr;
int (*g_tx_pop)() = nullptr;
void tx_load() {
  void *h = dlopen("librocprofiler-sdk-roctx.so.1", RTLD_NOW | RTLD_LOCAL);
  if (!h) h = dlopen("/opt/rocm/lib/librocprofiler-sdk-roctx.so.1", RTLD_NOW | RTLD_LOCAL);
  if (!h) return;
  g_tx_push = (int (*)(const char *))dlsym(h, "roctxRangePushA");
  g_tx_pop = (int (*)())dlsym(h, "roctxRangePop");
  if (!g_tx_push || !g_tx_pop) g_tx_push = nullptr, g_tx_pop = nullptr;
}
}  // namespace

extern "C" int gu_trace_push(const char *msg) {
  if (!msg) return GU_ERR_INVALID;
  std::call_once(g_tx_once, tx_load);
  if (!g_tx_push) return GU_ERR_UNSUPPORTED;
  g_tx_push(msg);
  return GU_OK;
}

extern "C" int gu_trace_pop(void) {
  std::call_once(g_tx_once, tx_load);
  if (!g_tx_pop) return GU_ERR_UNSUPPORTED;
  g_tx_pop();
  return GU_OK;
}
