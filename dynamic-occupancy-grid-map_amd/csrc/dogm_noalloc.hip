// dogm_update_grid_noalloc: the full updateGrid step of a plain handle with its plan-sized step
// buffers (work tables, scan tile totals, segmented-reduction carries, sort tile histograms) taken
// from a static device array instead of device allocations, for callers that cannot allocate
// (code under stream capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The array `scratch` below has a fixed size (kScratchBytes); each device the library runs on
// holds its own copy until the process ends.  Its device address is looked up once per device
// with hipGetSymbolAddress and only that address is handed to kernels.  Every call on a device
// shares that device's array: calls on one device that overlap in time (from different host
// threads) race, so callers serialise them.  A call leaves the array holding its step buffers;
// the next call rebuilds every table it reads (nothing in the array needs to start at zero).
#include "dogm_internal.h"

#include <mutex>

using namespace dogm_hip;

namespace
{

constexpr size_t kScratchBytes = size_t(16) << 20;  // 16 MiB: steps of up to about 1.3e7 particles
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace
{

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, looked up once per device with that device current.
int scratch_address(int device, void** out, std::string* why)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices)
    {
        *why = "device " + std::to_string(device) + " is outside the supported range";
        return DOGM_ERR_INVALID_ARGUMENT;
    }
    if (!g_scratch_ptr[device])
    {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess)
        {
            *why = "hipGetDevice failed";
            return DOGM_ERR_HIP;
        }
        if (cur != device && hipSetDevice(device) != hipSuccess)
        {
            *why = "hipSetDevice failed";
            return DOGM_ERR_HIP;
        }
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device)
            (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
        {
            *why = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return DOGM_ERR_HIP;
        }
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return DOGM_OK;
}

int noalloc_fail(Engine* E, int code, const std::string& msg)
{
    E->last_error = std::string("updateGridNoAlloc: ") + msg;
    return code;
}

}  // namespace

extern "C" int dogm_update_grid_noalloc(dogm_handle* h, const dogm_measurement_cell* meas, float new_x, float new_y,
                                        float new_yaw, float dt, int is_device)
{
    if (!h)
        return DOGM_ERR_INVALID_ARGUMENT;
    Engine* E = &h->eng;
    Vehicle* v = &h->veh;
    if (!meas)
        return noalloc_fail(E, DOGM_ERR_INVALID_ARGUMENT, "measurement_grid is null");
    if (E->m_cap < 1)
        return noalloc_fail(E, DOGM_ERR_INVALID_ARGUMENT, "the handle's step arrays are not reserved");
    Vehicle* list[1] = {v};
    std::string why;
    size_t need = 0;
    int st = engine_step_scratch_bytes(list, 1, &need, &why);
    if (st != DOGM_OK)
        return noalloc_fail(E, st, why);
    if (need > kScratchBytes)
        return noalloc_fail(E, DOGM_ERR_INVALID_ARGUMENT,
                            "this grid's step needs " + std::to_string(need) + " bytes of step scratch; the static " +
                                "scratch holds " + std::to_string(kScratchBytes) +
                                " (use dogm_update_grid for larger grids)");
    void* base = nullptr;
    if ((st = scratch_address(E->device, &base, &why)) != DOGM_OK)
        return noalloc_fail(E, st, why);

    // dogm_update_grid's sequence (dogm.cu:115-131) on the handle's engine, with the step buffers
    // bound to the static array, no step graph and no event timing (both would create objects).
    if ((st = vehicle_take_measurement(E, v, meas, is_device)) != DOGM_OK)
        return st;
    vehicle_update_pose(E, v, new_x, new_y, new_yaw, true);
    v->dt = dt;
    if ((st = engine_bind_scratch(E, base, kScratchBytes, list, 1, &why)) != DOGM_OK)
        return noalloc_fail(E, st, why);
    const bool graph = E->use_graph;
    bool timing[K_COUNT];
    for (int k = 0; k < K_COUNT; ++k)
    {
        timing[k] = E->timing_on[k];
        E->timing_on[k] = false;
    }
    E->use_graph = false;
    st = engine_run(E, list, 1, ST_ALL);
    const hipError_t e = hipStreamSynchronize(E->stream);  // the kernels use the bound buffers
    const int dc = engine_device_check(E);
    engine_unbind_scratch(E);
    E->use_graph = graph;
    for (int k = 0; k < K_COUNT; ++k)
        E->timing_on[k] = timing[k];
    if (st != DOGM_OK)
        return st;
    if (e != hipSuccess)
        return noalloc_fail(E, DOGM_ERR_HIP, hipGetErrorString(e));
    return dc;
}
