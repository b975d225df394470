// C-ABI of libdogm_hip.so: handle lifecycle, the per-frame stage sequence and read-backs.
// The stage order is that of dogm::DOGM::updateGrid (dogm/src/dogm.cu:115-131).
#include "dogm_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

using namespace dogm_hip;

namespace dogm_hip
{
const char* const kKernelNames[K_COUNT] = {
    "predict",        "move_particles",  "shift_field",     "sort_hist",       "sort_upsweep",
    "sort_scan",      "sort_downsweep",  "sort_gather",     "reset_bounds",    "seg_occ",
    "seg_occ_fix",    "mass_update",     "normalize",       "seg_pers",        "seg_pers_fix",
    "post_weights",   "empty_cells",     "scan_tiles",      "scan_top",        "birth_cells",
    "birth_particles", "init_cells",     "init_particles",  "copy_masses",     "seg_moments",
    "seg_moments_fix", "cdf_write",      "exp_tiles",       "rs_merge",        "fill",            "pack_grid",
    "unpack_grid",
};

static hipEvent_t take_event(dogm_handle* h)
{
    if (!h->event_pool.empty())
    {
        hipEvent_t e = h->event_pool.back();
        h->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

void launch_begin(dogm_handle* h, int kid)
{
    if (!h->timing_on[kid])
        return;
    TimedLaunch t{kid, take_event(h), take_event(h)};
    if (t.a)
        (void)hipEventRecord(t.a, h->stream);
    h->timed_pending.push_back(t);
}

void launch_end(dogm_handle* h, int kid)
{
    if (!h->timing_on[kid] || h->timed_pending.empty())
        return;
    TimedLaunch& t = h->timed_pending.back();
    if (t.b)
        (void)hipEventRecord(t.b, h->stream);
}

static void collect_timing(dogm_handle* h)
{
    for (auto& t : h->timed_pending)
    {
        float ms = 0.0f;
        if (t.a && t.b && hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess)
        {
            h->timing_ms[t.kid] += ms;
            h->timing_n[t.kid] += 1;
        }
        if (t.a)
            h->event_pool.push_back(t.a);
        if (t.b)
            h->event_pool.push_back(t.b);
    }
    h->timed_pending.clear();
}

}  // namespace dogm_hip

namespace
{

int fail(dogm_handle* h, int code, const std::string& msg)
{
    if (h)
        h->last_error = msg;
    return code;
}

#define HIP_TRY(h, expr)                                                                                  \
    do                                                                                                    \
    {                                                                                                     \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return fail((h), DOGM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));            \
    } while (0)

// Checks for launch errors queued so far.
int check_launch(dogm_handle* h, const char* stage)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(h, DOGM_ERR_HIP, std::string(stage) + ": " + hipGetErrorString(e));
    return DOGM_OK;
}

// Waits for the stream, then reports device-side check failures.
int finish(dogm_handle* h, const char* stage)
{
    int st = check_launch(h, stage);
    if (st != DOGM_OK)
        return st;
    const hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess)
        return fail(h, DOGM_ERR_HIP, std::string(stage) + ": " + hipGetErrorString(e));
    collect_timing(h);
    int err = 0;
    HIP_TRY(h, hipMemcpy(&err, &h->scalars->err, sizeof(int), hipMemcpyDeviceToHost));
    if (err != 0)
    {
        const int zero = 0;
        HIP_TRY(h, hipMemcpy(&h->scalars->err, &zero, sizeof(int), hipMemcpyHostToDevice));
        if (err & DEV_ERR_SPIN_LIMIT)
            return fail(h, DOGM_ERR_DEVICE_CHECK,
                        std::string(stage) + ": a workgroup gave up waiting for its predecessor");
        return fail(h, DOGM_ERR_DEVICE_CHECK, std::string(stage) + ": device check failed");
    }
    return DOGM_OK;
}

int bits_for(long long x)  // number of bits needed to represent x >= 0
{
    int b = 0;
    while (b < 62 && (1ll << b) <= x)
        ++b;
    return b;
}

template <typename T>
int dalloc(dogm_handle* h, T** p, size_t count, const char* what)
{
    *p = nullptr;
    if (count == 0)
        return DOGM_OK;
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
    if (e != hipSuccess)
        return fail(h, e == hipErrorOutOfMemory ? DOGM_ERR_OUT_OF_MEMORY : DOGM_ERR_HIP,
                    std::string("allocating ") + what + ": " + hipGetErrorString(e));
    return DOGM_OK;
}

void free_particles(DevParticles& p)
{
    (void)hipFree(p.state);
    (void)hipFree(p.idx);
    (void)hipFree(p.weight);
    (void)hipFree(p.assoc);
    p = DevParticles{};
}

int alloc_particles(dogm_handle* h, DevParticles& p, int n)
{
    int st;
    if ((st = dalloc(h, &p.state, n, "particle state")) || (st = dalloc(h, &p.idx, n, "particle index")) ||
        (st = dalloc(h, &p.weight, n, "particle weight")) || (st = dalloc(h, &p.assoc, n, "particle flags")))
        return st;
    if (n > 0)
    {
        HIP_TRY(h, hipMemset(p.state, 0, sizeof(float4) * n));
        HIP_TRY(h, hipMemset(p.idx, 0, sizeof(int) * n));
        HIP_TRY(h, hipMemset(p.weight, 0, sizeof(float) * n));
        HIP_TRY(h, hipMemset(p.assoc, 0, n));
    }
    return DOGM_OK;
}

void release(dogm_handle* h)
{
    if (!h)
        return;
    (void)hipDeviceSynchronize();
    free_particles(h->part[0]);
    free_particles(h->part[1]);
    free_particles(h->birth);
    (void)hipFree(h->grid.start);
    (void)hipFree(h->grid.end);
    for (int f = 0; f < GF_COUNT; ++f)
        (void)hipFree(h->grid.f[f]);
    void* bufs[] = {h->free_alt,       h->field_tmp,     h->meas,          h->weight_array, h->born,
                    h->m_raw,          h->pend,          h->ancestors,     h->sort_keys[0], h->sort_keys[1],
                    h->sort_vals[0],   h->sort_vals[1],  h->sort_ghist,    h->sort_thist,   h->seg.flags,
                    h->seg.head_partial, h->seg.tail_partial, h->scan_tile_sum, h->scan_tile_pre, h->cdf,
                    h->scalars,        h->aos_tmp,       h->cdf_tile_incl};
    for (void* b : bufs)
        (void)hipFree(b);
    for (auto& t : h->timed_pending)
    {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (auto e : h->event_pool)
        (void)hipEventDestroy(e);
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
}

int validate(const dogm_params* p, std::string* why)
{
    if (!p)
        return (*why = "params is null", DOGM_ERR_INVALID_ARGUMENT);
    if (!(p->size > 0.0f) || !(p->resolution > 0.0f) || !std::isfinite(p->size) || !std::isfinite(p->resolution))
        return (*why = "size and resolution must be positive and finite", DOGM_ERR_INVALID_ARGUMENT);
    const double gsd = (double)(p->size / p->resolution);
    if (!(gsd >= 1.0) || gsd > 46340.0)
        return (*why = "grid_size = int(size / resolution) must be in [1, 46340]", DOGM_ERR_INVALID_ARGUMENT);
    if (p->particle_count < 1 || p->new_born_particle_count < 0)
        return (*why = "particle_count must be >= 1 and new_born_particle_count >= 0", DOGM_ERR_INVALID_ARGUMENT);
    if ((long long)p->particle_count + p->new_born_particle_count > (1ll << 30))
        return (*why = "particle_count + new_born_particle_count must not exceed 2^30", DOGM_ERR_INVALID_ARGUMENT);
    return DOGM_OK;
}

// ---- stages ------------------------------------------------------------------------------------

int stage_measurement(dogm_handle* h, const dogm_measurement_cell* m, int is_device)
{
    if (!m)
        return fail(h, DOGM_ERR_INVALID_ARGUMENT, "measurement_grid is null");
    HIP_TRY(h, hipMemcpyAsync(h->meas, m, sizeof(dogm_measurement_cell) * (size_t)h->C,
                              is_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, h->stream));
    if (!h->first_meas)
    {
        run_init_particles(h);
        h->first_meas = true;
    }
    return check_launch(h, "updateMeasurementGrid");
}

void shift_grid_eager(dogm_handle* h, int ix, int iy)
{
    auto shift_ptr = [&](void** field) {
        run_shift_field(h, *field, h->field_tmp, ix, iy);
        void* old = *field;
        *field = h->field_tmp;
        h->field_tmp = static_cast<float*>(old);
    };
    shift_ptr(reinterpret_cast<void**>(&h->grid.start));
    shift_ptr(reinterpret_cast<void**>(&h->grid.end));
    for (int f = 0; f < GF_COUNT; ++f)
        shift_ptr(reinterpret_cast<void**>(&h->grid.f[f]));
}

// dogm.cu:161-203.  lazy: the shift is applied inside the following predict (particles) and
// occupancy update (free mass, the only grid field read before being rewritten in a frame).
int stage_pose(dogm_handle* h, float new_x, float new_y, float new_yaw, bool lazy)
{
    if (!h->first_pose)
    {
        h->pos_x = new_x;
        h->pos_y = new_y;
        h->yaw = new_yaw;
        h->first_pose = true;
        return DOGM_OK;
    }
    const float x_diff = new_x - h->pos_x;
    const float y_diff = new_y - h->pos_y;
    const float res = h->params.resolution;
    if (std::fabs(x_diff) > res || std::fabs(y_diff) > res)
    {
        const int x_move = -static_cast<int>(x_diff / res);
        const int y_move = -static_cast<int>(y_diff / res);
        if (lazy)
        {
            h->pend_part_shift = true;
            h->pend_px = (float)x_move;
            h->pend_py = (float)y_move;
            h->pend_grid_shift = true;
            h->pend_gx = x_move;
            h->pend_gy = y_move;
        }
        else
        {
            run_move_particles(h, (float)x_move, (float)y_move);
            shift_grid_eager(h, x_move, y_move);
        }
        h->pos_x = new_x;
        h->pos_y = new_y;
        h->yaw = new_yaw;
    }
    return check_launch(h, "updatePose");
}

int stage_predict(dogm_handle* h, float dt)
{
    run_predict(h, dt, h->pend_part_shift, h->pend_px, h->pend_py);
    h->pend_part_shift = false;
    return check_launch(h, "particlePrediction");
}

int stage_assignment(dogm_handle* h)
{
    run_reset_bounds(h);
    const int *skeys = nullptr, *perm = nullptr;
    run_sort(h, &skeys, &perm);
    run_sort_gather(h, skeys, perm);
    return check_launch(h, "particleAssignment");
}

int stage_occupancy(dogm_handle* h, float dt)
{
    // freespace_discount^dt with std::pow(float, float) as in dogm.cu:291.
    const float alpha = std::pow(h->params.freespace_discount, dt);
    run_occupancy(h, alpha, h->pend_grid_shift, h->pend_gx, h->pend_gy);
    h->pend_grid_shift = false;
    return check_launch(h, "gridCellOccupancyUpdate");
}

int stage_persistent(dogm_handle* h)
{
    run_persistent(h);
    return check_launch(h, "updatePersistentParticles");
}

int stage_births(dogm_handle* h)
{
    run_births(h);
    return check_launch(h, "initializeNewParticles");
}

int stage_moments(dogm_handle* h)
{
    run_moments(h);
    return check_launch(h, "statisticalMoments");
}

int stage_resampling(dogm_handle* h)
{
    run_resampling(h);
    return check_launch(h, "resampling");
}

int copy_out(dogm_handle* h, void* dst, const void* src, size_t bytes)
{
    if (bytes == 0)
        return DOGM_OK;
    if (!dst)
        return fail(h, DOGM_ERR_INVALID_ARGUMENT, "output buffer is null");
    HIP_TRY(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return DOGM_OK;
}

int copy_in(dogm_handle* h, void* dst, const void* src, size_t bytes)
{
    if (bytes == 0)
        return DOGM_OK;
    if (!src)
        return fail(h, DOGM_ERR_INVALID_ARGUMENT, "input buffer is null");
    HIP_TRY(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return DOGM_OK;
}

}  // namespace

// ================================================================================================
extern "C" {

const char* dogm_status_string(int status)
{
    switch (status)
    {
    case DOGM_OK: return "ok";
    case DOGM_ERR_INVALID_ARGUMENT: return "invalid argument";
    case DOGM_ERR_OUT_OF_MEMORY: return "out of device memory";
    case DOGM_ERR_HIP: return "HIP runtime error";
    case DOGM_ERR_DEVICE_CHECK: return "device-side check failed";
    case DOGM_ERR_NO_DEVICE: return "no HIP device";
    default: return "unknown status";
    }
}

const char* dogm_last_error(const dogm_handle* h)
{
    return h ? h->last_error.c_str() : "null handle";
}

int dogm_create_ex(const dogm_params* params, uint64_t seed, uint32_t instance, dogm_handle** out)
{
    if (!out)
        return DOGM_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    std::string why;
    int st = validate(params, &why);
    if (st != DOGM_OK)
        return st;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return DOGM_ERR_NO_DEVICE;

    dogm_handle* h = new (std::nothrow) dogm_handle();
    if (!h)
        return DOGM_ERR_OUT_OF_MEMORY;
    h->params = *params;
    h->gs = static_cast<int>(params->size / params->resolution);  // dogm.cu:34
    h->C = h->gs * h->gs;
    h->N = params->particle_count;
    h->B = params->new_born_particle_count;
    h->seed = seed;
    h->instance = instance;
    const int C = h->C, N = h->N, B = h->B;

    auto bail = [&](int code) {
        release(h);
        return code;
    };
    if (hipGetDevice(&h->device) != hipSuccess)
        return bail(DOGM_ERR_NO_DEVICE);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(DOGM_ERR_HIP);

    h->sort_passes = std::max(1, (bits_for(C - 1) + 7) / 8);
    h->sort_tiles = div_up(N, 256 * 8);
    h->seg_tiles_max = div_up(N, 256 * 4);
    h->scan_tiles_max = div_up(std::max((long long)C, (long long)N + B + 1), 1024);

    if ((st = alloc_particles(h, h->part[0], N)) || (st = alloc_particles(h, h->part[1], N)) ||
        (st = alloc_particles(h, h->birth, B)))
        return bail(st);
    if ((st = dalloc(h, &h->grid.start, C, "grid")) || (st = dalloc(h, &h->grid.end, C, "grid")))
        return bail(st);
    for (int f = 0; f < GF_COUNT; ++f)
        if ((st = dalloc(h, &h->grid.f[f], C, "grid")))
            return bail(st);
    if ((st = dalloc(h, &h->free_alt, C, "grid")) || (st = dalloc(h, &h->field_tmp, C, "grid")) ||
        (st = dalloc(h, &h->meas, C, "measurement grid")) || (st = dalloc(h, &h->weight_array, N, "weights")) ||
        (st = dalloc(h, &h->born, C, "born masses")) || (st = dalloc(h, &h->m_raw, C, "cell masses")) ||
        (st = dalloc(h, &h->pend, C, "particle orders")) || (st = dalloc(h, &h->ancestors, N, "ancestors")) ||
        (st = dalloc(h, &h->sort_keys[0], N, "sort")) || (st = dalloc(h, &h->sort_keys[1], N, "sort")) ||
        (st = dalloc(h, &h->sort_vals[0], N, "sort")) || (st = dalloc(h, &h->sort_vals[1], N, "sort")) ||
        (st = dalloc(h, &h->sort_ghist, (size_t)4 * 256, "sort")) ||
        (st = dalloc(h, &h->sort_thist, (size_t)256 * h->sort_tiles, "sort")) ||
        (st = dalloc(h, &h->seg.flags, h->seg_tiles_max, "segments")) ||
        (st = dalloc(h, &h->seg.head_partial, (size_t)h->seg_tiles_max * 5, "segments")) ||
        (st = dalloc(h, &h->seg.tail_partial, (size_t)h->seg_tiles_max * 5, "segments")) ||
        (st = dalloc(h, &h->scan_tile_sum, h->scan_tiles_max, "scan")) ||
        (st = dalloc(h, &h->scan_tile_pre, h->scan_tiles_max, "scan")) ||
        (st = dalloc(h, &h->cdf, (size_t)N + B, "cdf")) || (st = dalloc(h, &h->scalars, 1, "scalars")) ||
        (st = dalloc(h, &h->cdf_tile_incl, h->scan_tiles_max, "cdf")) ||
        (st = dalloc(h, &h->aos_tmp, C, "grid staging")))
        return bail(st);

    // initGridCellsKernel (init.cu:69-84): start = end = -1, masses 0, measurement (0, 0, 1, 1);
    // the remaining grid fields and buffers start at zero.
    if (hipMemset(h->grid.start, 0xff, sizeof(int) * C) != hipSuccess ||
        hipMemset(h->grid.end, 0xff, sizeof(int) * C) != hipSuccess)
        return bail(DOGM_ERR_HIP);
    for (int f = 0; f < GF_COUNT; ++f)
        if (hipMemset(h->grid.f[f], 0, sizeof(float) * C) != hipSuccess)
            return bail(DOGM_ERR_HIP);
    float* zbufs[] = {h->free_alt, h->weight_array, h->born, h->m_raw};
    const size_t zcnt[] = {(size_t)C, (size_t)N, (size_t)C, (size_t)C};
    for (int i = 0; i < 4; ++i)
        if (zcnt[i] && hipMemset(zbufs[i], 0, sizeof(float) * zcnt[i]) != hipSuccess)
            return bail(DOGM_ERR_HIP);
    if (hipMemset(h->ancestors, 0, sizeof(int) * N) != hipSuccess ||
        hipMemset(h->scalars, 0, sizeof(DevScalars)) != hipSuccess)
        return bail(DOGM_ERR_HIP);
    {
        std::vector<dogm_measurement_cell> m0((size_t)C, dogm_measurement_cell{0.0f, 0.0f, 1.0f, 1.0f});
        if (hipMemcpy(h->meas, m0.data(), sizeof(dogm_measurement_cell) * C, hipMemcpyHostToDevice) != hipSuccess)
            return bail(DOGM_ERR_HIP);
    }
    if (hipDeviceSynchronize() != hipSuccess)
        return bail(DOGM_ERR_HIP);
    *out = h;
    return DOGM_OK;
}

int dogm_create(const dogm_params* params, dogm_handle** out)
{
    return dogm_create_ex(params, DOGM_DEFAULT_SEED, 0u, out);
}

void dogm_destroy(dogm_handle* h)
{
    release(h);
}

#define NEED_H(h)                              \
    do                                         \
    {                                          \
        if (!(h))                              \
            return DOGM_ERR_INVALID_ARGUMENT;  \
    } while (0)

int dogm_update_grid(dogm_handle* h, const dogm_measurement_cell* meas, float new_x, float new_y, float new_yaw,
                     float dt, int is_device)
{
    NEED_H(h);
    int st;
    if ((st = stage_measurement(h, meas, is_device)) || (st = stage_pose(h, new_x, new_y, new_yaw, true)) ||
        (st = stage_predict(h, dt)) || (st = stage_assignment(h)) || (st = stage_occupancy(h, dt)) ||
        (st = stage_persistent(h)) || (st = stage_births(h)) || (st = stage_moments(h)) ||
        (st = stage_resampling(h)))
    {
        h->pend_part_shift = h->pend_grid_shift = false;
        (void)hipStreamSynchronize(h->stream);
        return st;
    }
    return finish(h, "updateGrid");
}

int dogm_update_measurement_grid(dogm_handle* h, const dogm_measurement_cell* meas, int is_device)
{
    NEED_H(h);
    int st = stage_measurement(h, meas, is_device);
    return st ? st : finish(h, "updateMeasurementGrid");
}

int dogm_update_pose(dogm_handle* h, float new_x, float new_y, float new_yaw)
{
    NEED_H(h);
    int st = stage_pose(h, new_x, new_y, new_yaw, false);
    return st ? st : finish(h, "updatePose");
}

int dogm_initialize_particles(dogm_handle* h)
{
    NEED_H(h);
    run_init_particles(h);
    return finish(h, "initializeParticles");
}

int dogm_particle_prediction(dogm_handle* h, float dt)
{
    NEED_H(h);
    int st = stage_predict(h, dt);
    return st ? st : finish(h, "particlePrediction");
}

int dogm_particle_assignment(dogm_handle* h)
{
    NEED_H(h);
    int st = stage_assignment(h);
    return st ? st : finish(h, "particleAssignment");
}

int dogm_grid_cell_occupancy_update(dogm_handle* h, float dt)
{
    NEED_H(h);
    int st = stage_occupancy(h, dt);
    return st ? st : finish(h, "gridCellOccupancyUpdate");
}

int dogm_update_persistent_particles(dogm_handle* h)
{
    NEED_H(h);
    int st = stage_persistent(h);
    return st ? st : finish(h, "updatePersistentParticles");
}

int dogm_initialize_new_particles(dogm_handle* h)
{
    NEED_H(h);
    int st = stage_births(h);
    return st ? st : finish(h, "initializeNewParticles");
}

int dogm_statistical_moments(dogm_handle* h)
{
    NEED_H(h);
    int st = stage_moments(h);
    return st ? st : finish(h, "statisticalMoments");
}

int dogm_resampling(dogm_handle* h)
{
    NEED_H(h);
    int st = stage_resampling(h);
    return st ? st : finish(h, "resampling");
}

// ---- read-backs --------------------------------------------------------------------------------

int dogm_get_grid_cells(dogm_handle* h, dogm_grid_cell* out)
{
    NEED_H(h);
    run_pack_grid(h);
    int st = check_launch(h, "getGridCells");
    return st ? st : copy_out(h, out, h->aos_tmp, sizeof(dogm_grid_cell) * (size_t)h->C);
}

int dogm_get_measurement_cells(dogm_handle* h, dogm_measurement_cell* out)
{
    NEED_H(h);
    return copy_out(h, out, h->meas, sizeof(dogm_measurement_cell) * (size_t)h->C);
}

static int get_set(dogm_handle* h, const DevParticles& p, int n, dogm_vec4* state, int* idx, float* weight,
                   uint8_t* assoc)
{
    int st;
    if ((state && (st = copy_out(h, state, p.state, sizeof(float4) * (size_t)n))) ||
        (idx && (st = copy_out(h, idx, p.idx, sizeof(int) * (size_t)n))) ||
        (weight && (st = copy_out(h, weight, p.weight, sizeof(float) * (size_t)n))) ||
        (assoc && (st = copy_out(h, assoc, p.assoc, (size_t)n))))
        return st;
    return DOGM_OK;
}

static int put_set(dogm_handle* h, const DevParticles& p, int n, const dogm_vec4* state, const int* idx,
                   const float* weight, const uint8_t* assoc)
{
    if (idx)
        for (int i = 0; i < n; ++i)
            if (idx[i] < 0 || idx[i] >= h->C)
                return fail(h, DOGM_ERR_INVALID_ARGUMENT,
                            "grid_cell_idx[" + std::to_string(i) + "] = " + std::to_string(idx[i]) +
                                " is outside [0, grid_size^2)");
    int st;
    if ((state && (st = copy_in(h, p.state, state, sizeof(float4) * (size_t)n))) ||
        (idx && (st = copy_in(h, p.idx, idx, sizeof(int) * (size_t)n))) ||
        (weight && (st = copy_in(h, p.weight, weight, sizeof(float) * (size_t)n))) ||
        (assoc && (st = copy_in(h, p.assoc, assoc, (size_t)n))))
        return st;
    return DOGM_OK;
}

int dogm_get_particles(dogm_handle* h, dogm_vec4* state, int* grid_cell_idx, float* weight, uint8_t* associated)
{
    NEED_H(h);
    return get_set(h, h->part[h->cur], h->N, state, grid_cell_idx, weight, associated);
}

int dogm_get_birth_particles(dogm_handle* h, dogm_vec4* state, int* grid_cell_idx, float* weight,
                             uint8_t* associated)
{
    NEED_H(h);
    return get_set(h, h->birth, h->B, state, grid_cell_idx, weight, associated);
}

int dogm_set_particles(dogm_handle* h, const dogm_vec4* state, const int* grid_cell_idx, const float* weight,
                       const uint8_t* associated)
{
    NEED_H(h);
    return put_set(h, h->part[h->cur], h->N, state, grid_cell_idx, weight, associated);
}

int dogm_set_birth_particles(dogm_handle* h, const dogm_vec4* state, const int* grid_cell_idx, const float* weight,
                             const uint8_t* associated)
{
    NEED_H(h);
    return put_set(h, h->birth, h->B, state, grid_cell_idx, weight, associated);
}

int dogm_set_grid_cells(dogm_handle* h, const dogm_grid_cell* cells)
{
    NEED_H(h);
    int st = copy_in(h, h->aos_tmp, cells, sizeof(dogm_grid_cell) * (size_t)h->C);
    if (st)
        return st;
    run_unpack_grid(h);
    return finish(h, "setGridCells");
}

int dogm_get_weight_array(dogm_handle* h, float* out)
{
    NEED_H(h);
    return copy_out(h, out, h->weight_array, sizeof(float) * (size_t)h->N);
}

int dogm_set_weight_array(dogm_handle* h, const float* in)
{
    NEED_H(h);
    return copy_in(h, h->weight_array, in, sizeof(float) * (size_t)h->N);
}

int dogm_get_born_masses(dogm_handle* h, float* out)
{
    NEED_H(h);
    return copy_out(h, out, h->born, sizeof(float) * (size_t)h->C);
}

int dogm_set_born_masses(dogm_handle* h, const float* in)
{
    NEED_H(h);
    return copy_in(h, h->born, in, sizeof(float) * (size_t)h->C);
}

int dogm_get_resampled_indices(dogm_handle* h, int* out)
{
    NEED_H(h);
    return copy_out(h, out, h->ancestors, sizeof(int) * (size_t)h->N);
}

int dogm_get_grid_size(const dogm_handle* h) { return h ? h->gs : 0; }
int dogm_get_particle_count(const dogm_handle* h) { return h ? h->N : 0; }
int dogm_get_new_born_particle_count(const dogm_handle* h) { return h ? h->B : 0; }
float dogm_get_resolution(const dogm_handle* h) { return h ? h->params.resolution : 0.0f; }
float dogm_get_position_x(const dogm_handle* h) { return h ? h->pos_x : 0.0f; }
float dogm_get_position_y(const dogm_handle* h) { return h ? h->pos_y : 0.0f; }
float dogm_get_yaw(const dogm_handle* h) { return h ? h->yaw : 0.0f; }

// ---- device memory helpers ---------------------------------------------------------------------

int dogm_device_malloc(void** ptr, size_t bytes)
{
    if (!ptr)
        return DOGM_ERR_INVALID_ARGUMENT;
    const hipError_t e = hipMalloc(ptr, bytes);
    if (e == hipErrorOutOfMemory)
        return DOGM_ERR_OUT_OF_MEMORY;
    return e == hipSuccess ? DOGM_OK : DOGM_ERR_HIP;
}

int dogm_device_free(void* ptr)
{
    return hipFree(ptr) == hipSuccess ? DOGM_OK : DOGM_ERR_HIP;
}

int dogm_memcpy(void* dst, const void* src, size_t bytes, int kind)
{
    hipMemcpyKind k;
    switch (kind)
    {
    case 0: k = hipMemcpyHostToDevice; break;
    case 1: k = hipMemcpyDeviceToHost; break;
    case 2: k = hipMemcpyDeviceToDevice; break;
    default: return DOGM_ERR_INVALID_ARGUMENT;
    }
    return hipMemcpy(dst, src, bytes, k) == hipSuccess ? DOGM_OK : DOGM_ERR_HIP;
}

int dogm_device_synchronize(void)
{
    return hipDeviceSynchronize() == hipSuccess ? DOGM_OK : DOGM_ERR_HIP;
}

// ---- kernel timing -----------------------------------------------------------------------------

int dogm_kernel_count(void) { return K_COUNT; }

const char* dogm_kernel_name(int kernel_id)
{
    return (kernel_id >= 0 && kernel_id < K_COUNT) ? kKernelNames[kernel_id] : nullptr;
}

int dogm_set_kernel_timing(dogm_handle* h, int kernel_id, int enable)
{
    NEED_H(h);
    if (kernel_id < 0 || kernel_id >= K_COUNT)
        return fail(h, DOGM_ERR_INVALID_ARGUMENT, "kernel id out of range");
    h->timing_on[kernel_id] = enable != 0;
    return DOGM_OK;
}

int dogm_get_kernel_timing(dogm_handle* h, int kernel_id, double* total_ms, long long* launches)
{
    NEED_H(h);
    if (kernel_id < 0 || kernel_id >= K_COUNT)
        return fail(h, DOGM_ERR_INVALID_ARGUMENT, "kernel id out of range");
    if (total_ms)
        *total_ms = h->timing_ms[kernel_id];
    if (launches)
        *launches = h->timing_n[kernel_id];
    return DOGM_OK;
}

int dogm_reset_kernel_timing(dogm_handle* h)
{
    NEED_H(h);
    for (int k = 0; k < K_COUNT; ++k)
    {
        h->timing_ms[k] = 0.0;
        h->timing_n[k] = 0;
    }
    return DOGM_OK;
}

}  // extern "C"
