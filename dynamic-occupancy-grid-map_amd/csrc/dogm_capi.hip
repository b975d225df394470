// C-ABI of libdogm_hip.so: the plain handle (one vehicle), the batch of vehicles, the per-frame
// stage sequence and read-backs.  The stage order is that of dogm::DOGM::updateGrid
// (dogm/src/dogm.cu:115-131).
#include "dogm_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <new>

using namespace dogm_hip;

// A batch: one engine and the vehicles it owns, addressed by id.
struct dogm_batch
{
    Engine eng;
    Engine aux;             // second lane of steps over two or more vehicles (created on first use)
    bool aux_ready = false;
    int lanes = 0;  // 0: as the tuning of each step gives
    std::map<int, std::unique_ptr<Vehicle>> veh;
    int next_id = 1;
    uint64_t bytes = 0;
    ExtractCfg ext;  // extraction run inside every step (what = 0: none)
};

namespace
{

int fail(Engine* E, int code, const std::string& msg)
{
    if (E)
        E->last_error = msg;
    return code;
}

#define HIP_TRY(E, expr)                                                                                  \
    do                                                                                                    \
    {                                                                                                     \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return fail((E), DOGM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));            \
    } while (0)

int check_launch(Engine* E, const char* stage)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(E, DOGM_ERR_HIP, std::string(stage) + ": " + hipGetErrorString(e));
    return DOGM_OK;
}

// Waits for the stream and collects kernel timings.
int finish(Engine* E, const char* stage)
{
    int st = check_launch(E, stage);
    if (st != DOGM_OK)
        return st;
    const hipError_t e = hipStreamSynchronize(E->stream);
    const int dc = engine_device_check(E);
    if (e != hipSuccess)
        return fail(E, DOGM_ERR_HIP, std::string(stage) + ": " + hipGetErrorString(e));
    engine_collect_timing(E);
    return dc;
}

// Runs `stages` on one vehicle and waits.
int run_one(Engine* E, Vehicle* v, int stages, const char* stage)
{
    Vehicle* list[1] = {v};
    int st = engine_run(E, list, 1, stages);
    if (st != DOGM_OK)
    {
        (void)hipStreamSynchronize(E->stream);
        return st;
    }
    return finish(E, stage);
}

int copy_out(Engine* E, void* dst, const void* src, size_t bytes)
{
    if (bytes == 0)
        return DOGM_OK;
    if (!dst)
        return fail(E, DOGM_ERR_INVALID_ARGUMENT, "output buffer is null");
    HIP_TRY(E, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, E->stream));
    HIP_TRY(E, hipStreamSynchronize(E->stream));
    return DOGM_OK;
}

int copy_in(Engine* E, void* dst, const void* src, size_t bytes)
{
    if (bytes == 0)
        return DOGM_OK;
    if (!src)
        return fail(E, DOGM_ERR_INVALID_ARGUMENT, "input buffer is null");
    HIP_TRY(E, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, E->stream));
    HIP_TRY(E, hipStreamSynchronize(E->stream));
    return DOGM_OK;
}

int ensure_aos(Engine* E, Vehicle* v)
{
    if (v->aos_tmp)
        return DOGM_OK;
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(&v->aos_tmp), sizeof(dogm_grid_cell) * (size_t)v->C);
    if (e != hipSuccess)
    {
        v->aos_tmp = nullptr;
        return fail(E, e == hipErrorOutOfMemory ? DOGM_ERR_OUT_OF_MEMORY : DOGM_ERR_HIP,
                    std::string("allocating the grid packing buffer: ") + hipGetErrorString(e));
    }
    return DOGM_OK;
}

int get_grid_cells(Engine* E, Vehicle* v, dogm_grid_cell* out)
{
    int st = ensure_aos(E, v);
    if (st)
        return st;
    run_pack_grid(E, v);
    st = check_launch(E, "getGridCells");
    return st ? st : copy_out(E, out, v->aos_tmp, sizeof(dogm_grid_cell) * (size_t)v->C);
}

// Device buffer of n (x, y, vx, vy) records for packing: the second particle buffer, which no step
// reads before writing it, or a lazy buffer for a birth set larger than N.
int state_staging(Engine* E, Vehicle* v, int n, float4** out)
{
    if (n <= v->N)
    {
        *out = reinterpret_cast<float4*>(v->part[1].pos);
        return DOGM_OK;
    }
    if (!v->state_tmp)
    {
        const hipError_t e = hipMalloc(reinterpret_cast<void**>(&v->state_tmp), sizeof(float4) * (size_t)v->B);
        if (e != hipSuccess)
        {
            v->state_tmp = nullptr;
            return fail(E, e == hipErrorOutOfMemory ? DOGM_ERR_OUT_OF_MEMORY : DOGM_ERR_HIP,
                        std::string("allocating the particle packing buffer: ") + hipGetErrorString(e));
        }
    }
    *out = v->state_tmp;
    return DOGM_OK;
}

}  // namespace

int dogm_hip::read_particles(Engine* E, Vehicle* v, const DevParticles& p, int n, dogm_vec4* state, int* idx,
                             float* weight, uint8_t* assoc)
{
    // A deferred resampled set is written out first (the state staging below reuses part[1]).
    vehicle_materialize(E, v);
    int st;
    if (state)
    {
        float4* tmp = nullptr;
        if ((st = state_staging(E, v, n, &tmp)))
            return st;
        run_pack_state(E, p, n, tmp);
        if ((st = check_launch(E, "packing particle state")) || (st = copy_out(E, state, tmp, sizeof(float4) * (size_t)n)))
            return st;
    }
    if ((idx && (st = copy_out(E, idx, p.idx, sizeof(int) * (size_t)n))) ||
        (weight && (st = copy_out(E, weight, p.weight, sizeof(float) * (size_t)n))) ||
        (assoc && (st = copy_out(E, assoc, p.assoc, (size_t)n))))
        return st;
    return DOGM_OK;
}

int dogm_hip::write_particles(Engine* E, Vehicle* v, const DevParticles& p, int n, const dogm_vec4* state,
                              const int* idx, const float* weight, const uint8_t* assoc)
{
    if (idx)
        for (int i = 0; i < n; ++i)
            if (idx[i] < 0 || idx[i] >= v->C)
                return fail(E, DOGM_ERR_INVALID_ARGUMENT,
                            "grid_cell_idx[" + std::to_string(i) + "] = " + std::to_string(idx[i]) +
                                " is outside [0, grid_size^2)");
    vehicle_materialize(E, v);  // the staging reuses part[1], and a deferred set reads the birth set
    int st;
    if (state)
    {
        float4* tmp = nullptr;
        if ((st = state_staging(E, v, n, &tmp)) || (st = copy_in(E, tmp, state, sizeof(float4) * (size_t)n)))
            return st;
        run_unpack_state(E, p, n, tmp);
        if ((st = check_launch(E, "unpacking particle state")))
            return st;
    }
    if (weight && p.weight == v->part[0].weight)
        v->w_uniform = false;  // the next prediction reads every weight
    if ((idx && (st = copy_in(E, p.idx, idx, sizeof(int) * (size_t)n))) ||
        (weight && (st = copy_in(E, p.weight, weight, sizeof(float) * (size_t)n))) ||
        (assoc && (st = copy_in(E, p.assoc, assoc, (size_t)n))))
        return st;
    return DOGM_OK;
}

namespace
{

void release(dogm_handle* h)
{
    if (!h)
        return;
    if (h->eng.stream)
        (void)hipStreamSynchronize(h->eng.stream);
    vehicle_free(&h->veh);
    engine_release(&h->eng);
    delete h;
}

int set_timing(Engine* E, int kernel_id, int enable)
{
    if (kernel_id < 0 || kernel_id >= K_COUNT)
        return fail(E, DOGM_ERR_INVALID_ARGUMENT, "kernel id out of range");
    E->timing_on[kernel_id] = enable != 0;
    return DOGM_OK;
}

int get_timing(Engine* E, int kernel_id, double* total_ms, long long* launches)
{
    if (kernel_id < 0 || kernel_id >= K_COUNT)
        return fail(E, DOGM_ERR_INVALID_ARGUMENT, "kernel id out of range");
    if (total_ms)
        *total_ms = E->timing_ms[kernel_id];
    if (launches)
        *launches = E->timing_n[kernel_id];
    return DOGM_OK;
}

void reset_timing(Engine* E)
{
    for (int k = 0; k < K_COUNT; ++k)
    {
        E->timing_ms[k] = 0.0;
        E->timing_n[k] = 0;
    }
}

// Switching modes keeps the grid where it is: the metric frame starts at the current pose.
int set_ego_mode(Engine* E, Vehicle* v, int mode)
{
    if (mode != DOGM_EGO_REFERENCE && mode != DOGM_EGO_METRIC)
        return fail(E, DOGM_ERR_INVALID_ARGUMENT, "unknown ego mode " + std::to_string(mode));
    if (mode == DOGM_EGO_METRIC && v->ego_mode != DOGM_EGO_METRIC && v->first_pose)
    {
        v->ref_x = v->pos_x;
        v->ref_y = v->pos_y;
        v->ref_yaw = v->yaw;
    }
    v->ego_mode = mode;
    return DOGM_OK;
}

Vehicle* find_vehicle(dogm_batch* b, int id)
{
    auto it = b->veh.find(id);
    return it == b->veh.end() ? nullptr : it->second.get();
}

}  // namespace

// ---- result extraction getters -------------------------------------------------------------------
namespace dogm_hip
{

int extract_ensure(Engine* E, Vehicle* v, const ExtractCfg& want, bool reuse)
{
    const bool vel = (want.what & DOGM_EXTRACT_VELOCITY) != 0;
    if (reuse && (v->ext_what & want.what) == want.what && v->ext_source == want.source &&
        (!vel || (v->ext_min_occ == want.min_occ && v->ext_min_vel == want.min_vel)))
        return DOGM_OK;
    E->ext = want;
    return run_one(E, v, ST_EXTRACT, "extraction");
}

int extract_read_occupancy(Engine* E, Vehicle* v, float* out)
{
    return copy_out(E, out, v->x_occ, sizeof(float) * (size_t)v->C);
}

int extract_read_image(Engine* E, Vehicle* v, uint8_t* out)
{
    return copy_out(E, out, v->x_img, (size_t)v->C);
}

int extract_read_velocity(Engine* E, Vehicle* v, dogm_velocity_cell* out, int capacity, int* count)
{
    if (!count)
        return fail(E, DOGM_ERR_INVALID_ARGUMENT, "count is null");
    if (capacity < 0 || (capacity > 0 && !out))
        return fail(E, DOGM_ERR_INVALID_ARGUMENT, "capacity must be >= 0, with an output buffer when > 0");
    int total = 0;
    int st = copy_out(E, &total, v->x_off + units_for(v->C), sizeof(int));
    if (st)
        return st;
    if (total < 0 || total > v->C)
        return fail(E, DOGM_ERR_DEVICE_CHECK, "velocity cell count out of range");
    *count = total;
    const int n = std::min(total, capacity);
    for (int r0 = 0; r0 < n; r0 += (int)v->x_rec_cap)
    {
        const int k = std::min(n - r0, (int)v->x_rec_cap);
        Vehicle chunk = {};
        chunk.grid = v->grid;
        chunk.gs = v->gs;
        chunk.x_idx = v->x_idx + r0;
        chunk.x_rec = v->x_rec;
        run_extract_gather(E, &chunk, k);
        if ((st = check_launch(E, "cells with velocity")) ||
            (st = copy_out(E, out + r0, v->x_rec, sizeof(dogm_velocity_cell) * (size_t)k)))
            return st;
    }
    return DOGM_OK;
}

}  // namespace dogm_hip

// Single-dispatch device-to-device copy (16-byte vectors, grid-stride), on the null stream.
__global__ __launch_bounds__(BLOCK) void k_copy_bytes(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                      size_t n16, const uint8_t* __restrict__ src_tail,
                                                      uint8_t* __restrict__ dst_tail, int tail)
{
    const size_t stride = (size_t)gridDim.x * BLOCK;
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
    if (blockIdx.x == 0 && (int)threadIdx.x < tail)
        dst_tail[threadIdx.x] = src_tail[threadIdx.x];
}

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
    return h ? h->eng.last_error.c_str() : "null handle";
}

int dogm_create_ex(const dogm_params* params, uint64_t seed, uint32_t instance, dogm_handle** out)
{
    if (!out)
        return DOGM_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    std::string why;
    int st = vehicle_validate(params, &why);
    if (st != DOGM_OK)
        return st;
    dogm_handle* h = new (std::nothrow) dogm_handle();
    if (!h)
        return DOGM_ERR_OUT_OF_MEMORY;
    // The per-vehicle step arrays are reserved here, so a step on this handle allocates only its
    // plan-sized buffers (and dogm_update_grid_noalloc none at all).
    if ((st = engine_init(&h->eng, &why)) != DOGM_OK ||
        (st = vehicle_alloc(&h->veh, params, seed, instance, &h->eng.last_error)) ||
        (st = engine_reserve(&h->eng, 1, &h->eng.last_error)))
    {
        release(h);
        return st;
    }
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
    Engine* E = &h->eng;
    Vehicle* v = &h->veh;
    if (!meas)
        return fail(E, DOGM_ERR_INVALID_ARGUMENT, "measurement_grid is null");
    int st = vehicle_take_measurement(E, v, meas, is_device);
    if (st)
        return st;
    vehicle_update_pose(E, v, new_x, new_y, new_yaw, true);
    v->dt = dt;
    return run_one(E, v, ST_ALL, "updateGrid");
}

int dogm_update_measurement_grid(dogm_handle* h, const dogm_measurement_cell* meas, int is_device)
{
    NEED_H(h);
    if (!meas)
        return fail(&h->eng, DOGM_ERR_INVALID_ARGUMENT, "measurement_grid is null");
    int st = vehicle_set_measurement(&h->eng, &h->veh, meas, is_device);
    if (st)
        return st;
    if (h->veh.init_pending)
        return run_one(&h->eng, &h->veh, ST_INIT, "updateMeasurementGrid");
    return finish(&h->eng, "updateMeasurementGrid");
}

int dogm_update_pose(dogm_handle* h, float new_x, float new_y, float new_yaw)
{
    NEED_H(h);
    vehicle_update_pose(&h->eng, &h->veh, new_x, new_y, new_yaw, false);
    return finish(&h->eng, "updatePose");
}

int dogm_initialize_particles(dogm_handle* h)
{
    NEED_H(h);
    h->veh.init_pending = true;
    return run_one(&h->eng, &h->veh, ST_INIT, "initializeParticles");
}

int dogm_particle_prediction(dogm_handle* h, float dt)
{
    NEED_H(h);
    h->veh.dt = dt;
    return run_one(&h->eng, &h->veh, ST_PREDICT, "particlePrediction");
}

int dogm_particle_assignment(dogm_handle* h)
{
    NEED_H(h);
    return run_one(&h->eng, &h->veh, ST_ASSIGN, "particleAssignment");
}

int dogm_grid_cell_occupancy_update(dogm_handle* h, float dt)
{
    NEED_H(h);
    h->veh.dt = dt;
    return run_one(&h->eng, &h->veh, ST_OCC, "gridCellOccupancyUpdate");
}

int dogm_update_persistent_particles(dogm_handle* h)
{
    NEED_H(h);
    return run_one(&h->eng, &h->veh, ST_PERS, "updatePersistentParticles");
}

int dogm_initialize_new_particles(dogm_handle* h)
{
    NEED_H(h);
    return run_one(&h->eng, &h->veh, ST_BIRTH, "initializeNewParticles");
}

int dogm_statistical_moments(dogm_handle* h)
{
    NEED_H(h);
    return run_one(&h->eng, &h->veh, ST_MOMENTS, "statisticalMoments");
}

int dogm_resampling(dogm_handle* h)
{
    NEED_H(h);
    return run_one(&h->eng, &h->veh, ST_RESAMPLE, "resampling");
}

// ---- read-backs --------------------------------------------------------------------------------

int dogm_get_grid_cells(dogm_handle* h, dogm_grid_cell* out)
{
    NEED_H(h);
    return get_grid_cells(&h->eng, &h->veh, out);
}

int dogm_get_measurement_cells(dogm_handle* h, dogm_measurement_cell* out)
{
    NEED_H(h);
    return copy_out(&h->eng, out, h->veh.meas, sizeof(dogm_measurement_cell) * (size_t)h->veh.C);
}

int dogm_get_particles(dogm_handle* h, dogm_vec4* state, int* grid_cell_idx, float* weight, uint8_t* associated)
{
    NEED_H(h);
    return read_particles(&h->eng, &h->veh, h->veh.part[0], h->veh.N, state, grid_cell_idx, weight, associated);
}

int dogm_get_birth_particles(dogm_handle* h, dogm_vec4* state, int* grid_cell_idx, float* weight,
                             uint8_t* associated)
{
    NEED_H(h);
    return read_particles(&h->eng, &h->veh, h->veh.birth, h->veh.B, state, grid_cell_idx, weight, associated);
}

int dogm_set_particles(dogm_handle* h, const dogm_vec4* state, const int* grid_cell_idx, const float* weight,
                       const uint8_t* associated)
{
    NEED_H(h);
    return write_particles(&h->eng, &h->veh, h->veh.part[0], h->veh.N, state, grid_cell_idx, weight, associated);
}

int dogm_set_birth_particles(dogm_handle* h, const dogm_vec4* state, const int* grid_cell_idx, const float* weight,
                             const uint8_t* associated)
{
    NEED_H(h);
    return write_particles(&h->eng, &h->veh, h->veh.birth, h->veh.B, state, grid_cell_idx, weight, associated);
}

int dogm_set_grid_cells(dogm_handle* h, const dogm_grid_cell* cells)
{
    NEED_H(h);
    int st = ensure_aos(&h->eng, &h->veh);
    if (st)
        return st;
    st = copy_in(&h->eng, h->veh.aos_tmp, cells, sizeof(dogm_grid_cell) * (size_t)h->veh.C);
    if (st)
        return st;
    run_unpack_grid(&h->eng, &h->veh);
    h->veh.nz_valid = false;
    return finish(&h->eng, "setGridCells");
}

int dogm_get_occupancy(dogm_handle* h, int source, float* out)
{
    NEED_H(h);
    int st = extract_ensure(&h->eng, &h->veh, ExtractCfg{DOGM_EXTRACT_OCCUPANCY, source, 0.0f, 0.0f}, false);
    return st ? st : extract_read_occupancy(&h->eng, &h->veh, out);
}

int dogm_get_occupancy_image(dogm_handle* h, int source, uint8_t* out)
{
    NEED_H(h);
    int st = extract_ensure(&h->eng, &h->veh, ExtractCfg{DOGM_EXTRACT_IMAGE, source, 0.0f, 0.0f}, false);
    return st ? st : extract_read_image(&h->eng, &h->veh, out);
}

int dogm_get_cells_with_velocity(dogm_handle* h, float min_occupancy, float min_velocity, dogm_velocity_cell* out,
                                 int capacity, int* count)
{
    NEED_H(h);
    int st = extract_ensure(&h->eng, &h->veh,
                            ExtractCfg{DOGM_EXTRACT_VELOCITY, DOGM_SOURCE_GRID, min_occupancy, min_velocity}, false);
    return st ? st : extract_read_velocity(&h->eng, &h->veh, out, capacity, count);
}

int dogm_set_graph(dogm_handle* h, int enable)
{
    NEED_H(h);
    h->eng.use_graph = enable != 0;
    return DOGM_OK;
}

int dogm_set_resampling_mode(dogm_handle* h, int mode)
{
    NEED_H(h);
    if (mode != DOGM_RESAMPLE_MULTINOMIAL && mode != DOGM_RESAMPLE_SYSTEMATIC)
        return fail(&h->eng, DOGM_ERR_INVALID_ARGUMENT, "unknown resampling mode " + std::to_string(mode));
    h->veh.draw_mode = mode;
    return DOGM_OK;
}

int dogm_get_resampling_mode(const dogm_handle* h)
{
    return h ? h->veh.draw_mode : -1;
}

int dogm_set_ego_mode(dogm_handle* h, int mode)
{
    NEED_H(h);
    return set_ego_mode(&h->eng, &h->veh, mode);
}

int dogm_get_ego_mode(const dogm_handle* h)
{
    return h ? h->veh.ego_mode : -1;
}

int dogm_get_weight_array(dogm_handle* h, float* out)
{
    NEED_H(h);
    return copy_out(&h->eng, out, h->veh.weight_array, sizeof(float) * (size_t)h->veh.N);
}

int dogm_set_weight_array(dogm_handle* h, const float* in)
{
    NEED_H(h);
    return copy_in(&h->eng, h->veh.weight_array, in, sizeof(float) * (size_t)h->veh.N);
}

int dogm_get_born_masses(dogm_handle* h, float* out)
{
    NEED_H(h);
    return copy_out(&h->eng, out, h->veh.born, sizeof(float) * (size_t)h->veh.C);
}

int dogm_set_born_masses(dogm_handle* h, const float* in)
{
    NEED_H(h);
    return copy_in(&h->eng, h->veh.born, in, sizeof(float) * (size_t)h->veh.C);
}

int dogm_get_resampled_indices(dogm_handle* h, int* out)
{
    NEED_H(h);
    return copy_out(&h->eng, out, h->veh.ancestors, sizeof(int) * (size_t)h->veh.N);
}

int dogm_set_cdf_mode(dogm_handle* h, int mode)
{
    NEED_H(h);
    if (mode < DOGM_CDF_EXACT || mode > DOGM_CDF_FP32C)
        return fail(&h->eng, DOGM_ERR_INVALID_ARGUMENT, "unknown CDF mode " + std::to_string(mode));
    h->veh.cdf_mode = mode;
    return DOGM_OK;
}

int dogm_get_cdf_mode(const dogm_handle* h) { return h ? h->veh.cdf_mode : -1; }
uint32_t dogm_get_epoch(const dogm_handle* h) { return h ? h->veh.epoch : 0u; }

int dogm_get_grid_size(const dogm_handle* h) { return h ? h->veh.gs : 0; }
int dogm_get_particle_count(const dogm_handle* h) { return h ? h->veh.N : 0; }
int dogm_get_new_born_particle_count(const dogm_handle* h) { return h ? h->veh.B : 0; }
float dogm_get_resolution(const dogm_handle* h) { return h ? h->veh.params.resolution : 0.0f; }
float dogm_get_position_x(const dogm_handle* h) { return h ? h->veh.pos_x : 0.0f; }
float dogm_get_position_y(const dogm_handle* h) { return h ? h->veh.pos_y : 0.0f; }
float dogm_get_yaw(const dogm_handle* h) { return h ? h->veh.yaw : 0.0f; }

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
    if (hipMemcpy(dst, src, bytes, k) != hipSuccess)
        return DOGM_ERR_HIP;
    // The copy runs on the null stream, which the library's non-blocking streams do not wait for:
    // finish it before a step may read the destination.
    return hipDeviceSynchronize() == hipSuccess ? DOGM_OK : DOGM_ERR_HIP;
}

int dogm_device_copy(void* dst, const void* src, size_t bytes)
{
    if ((!dst || !src) && bytes)
        return DOGM_ERR_INVALID_ARGUMENT;
    if (((uintptr_t)dst | (uintptr_t)src) & 15u)
        return DOGM_ERR_INVALID_ARGUMENT;
    const size_t n16 = bytes / 16;
    const int tail = (int)(bytes % 16);
    const int blocks = (int)std::min<size_t>(std::max<size_t>(1, (n16 + BLOCK - 1) / BLOCK), 16384);
    hipLaunchKernelGGL(k_copy_bytes, dim3(blocks), dim3(BLOCK), 0, nullptr, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), n16, static_cast<const uint8_t*>(src) + n16 * 16,
                       static_cast<uint8_t*>(dst) + n16 * 16, tail);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return DOGM_ERR_HIP;
    return DOGM_OK;
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
    return set_timing(&h->eng, kernel_id, enable);
}

int dogm_get_kernel_timing(dogm_handle* h, int kernel_id, double* total_ms, long long* launches)
{
    NEED_H(h);
    return get_timing(&h->eng, kernel_id, total_ms, launches);
}

int dogm_reset_kernel_timing(dogm_handle* h)
{
    NEED_H(h);
    reset_timing(&h->eng);
    return DOGM_OK;
}

// ---- batches -----------------------------------------------------------------------------------

int dogm_batch_create(dogm_batch** out)
{
    if (!out)
        return DOGM_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    dogm_batch* b = new (std::nothrow) dogm_batch();
    if (!b)
        return DOGM_ERR_OUT_OF_MEMORY;
    std::string why;
    int st = engine_init(&b->eng, &why);
    if (st != DOGM_OK)
    {
        engine_release(&b->eng);
        delete b;
        return st;
    }
    *out = b;
    return DOGM_OK;
}

void dogm_batch_destroy(dogm_batch* b)
{
    if (!b)
        return;
    if (b->eng.stream)
        (void)hipStreamSynchronize(b->eng.stream);
    for (auto& kv : b->veh)
        vehicle_free(kv.second.get());
    b->veh.clear();
    if (b->aux_ready)
        engine_release(&b->aux);
    engine_release(&b->eng);
    delete b;
}

const char* dogm_batch_last_error(const dogm_batch* b)
{
    return b ? b->eng.last_error.c_str() : "null batch";
}

int dogm_vehicle_footprint(const dogm_params* params, uint64_t* bytes)
{
    std::string why;
    if (!bytes)
        return DOGM_ERR_INVALID_ARGUMENT;
    int st = vehicle_validate(params, &why);
    if (st != DOGM_OK)
        return st;
    *bytes = (uint64_t)vehicle_footprint(params);
    return DOGM_OK;
}

int dogm_batch_add_vehicle(dogm_batch* b, const dogm_params* params, uint64_t seed, uint32_t instance,
                           int* vehicle_id)
{
    if (!b || !vehicle_id)
        return DOGM_ERR_INVALID_ARGUMENT;
    *vehicle_id = 0;
    std::unique_ptr<Vehicle> v(new (std::nothrow) Vehicle());
    if (!v)
        return DOGM_ERR_OUT_OF_MEMORY;
    int st = vehicle_alloc(v.get(), params, seed, instance, &b->eng.last_error);
    if (st != DOGM_OK)
    {
        vehicle_free(v.get());
        return st;
    }
    const int id = b->next_id++;
    b->bytes += v->bytes;
    b->veh[id] = std::move(v);
    *vehicle_id = id;
    return DOGM_OK;
}

int dogm_batch_remove_vehicle(dogm_batch* b, int vehicle_id)
{
    if (!b)
        return DOGM_ERR_INVALID_ARGUMENT;
    Vehicle* v = find_vehicle(b, vehicle_id);
    if (!v)
        return fail(&b->eng, DOGM_ERR_INVALID_ARGUMENT, "unknown vehicle id " + std::to_string(vehicle_id));
    (void)hipStreamSynchronize(b->eng.stream);
    b->bytes -= v->bytes;
    vehicle_free(v);
    b->veh.erase(vehicle_id);
    return DOGM_OK;
}

int dogm_batch_vehicle_count(const dogm_batch* b)
{
    return b ? (int)b->veh.size() : 0;
}

uint64_t dogm_batch_device_bytes(const dogm_batch* b)
{
    if (!b)
        return 0;
    uint64_t x = 0;
    for (const auto& kv : b->veh)
        x += kv.second->x_bytes;
    return b->bytes + x;
}

int dogm_batch_measurement_buffer(dogm_batch* b, int vehicle_id, dogm_measurement_cell** device_ptr)
{
    if (!b || !device_ptr)
        return DOGM_ERR_INVALID_ARGUMENT;
    Vehicle* v = find_vehicle(b, vehicle_id);
    if (!v)
        return fail(&b->eng, DOGM_ERR_INVALID_ARGUMENT, "unknown vehicle id " + std::to_string(vehicle_id));
    *device_ptr = v->meas;
    return DOGM_OK;
}

int dogm_batch_update(dogm_batch* b, int count, const int* vehicle_ids, const dogm_measurement_cell* const* meas,
                      const float* x, const float* y, const float* yaw, const float* dt, int is_device)
{
    if (!b)
        return DOGM_ERR_INVALID_ARGUMENT;
    Engine* E = &b->eng;
    if (count < 1 || !vehicle_ids || !x || !y || !yaw || !dt)
        return fail(E, DOGM_ERR_INVALID_ARGUMENT, "count must be >= 1 and ids, poses and dt non-null");
    std::vector<Vehicle*> list((size_t)count);
    for (int i = 0; i < count; ++i)
    {
        Vehicle* v = find_vehicle(b, vehicle_ids[i]);
        if (!v)
            return fail(E, DOGM_ERR_INVALID_ARGUMENT, "unknown vehicle id " + std::to_string(vehicle_ids[i]));
        if (v->uid == 0)
            return fail(E, DOGM_ERR_INVALID_ARGUMENT, "vehicle is not allocated");
        list[i] = v;
    }
    for (int i = 0; i < count; ++i)
        for (int j = i + 1; j < count; ++j)
            if (list[i] == list[j])
                return fail(E, DOGM_ERR_INVALID_ARGUMENT, "a vehicle appears twice in one step");
    std::vector<FrameIn> in((size_t)count);
    for (int i = 0; i < count; ++i)
        in[i] = FrameIn{meas ? meas[i] : nullptr, is_device, x[i], y[i], yaw[i], dt[i]};
    E->ext = b->ext;
    if (count >= 2 && b->lanes != 1 && !b->aux_ready)
    {
        std::string why;
        const int st = engine_init(&b->aux, &why);
        if (st != DOGM_OK)
        {
            engine_release(&b->aux);
            b->aux = Engine{};
            return fail(E, st, why);
        }
        for (int k = 0; k < K_COUNT; ++k)
            b->aux.timing_on[k] = E->timing_on[k];
        b->aux.use_graph = E->use_graph;
        b->aux_ready = true;
    }
    Engine* lanes[2] = {E, &b->aux};
    return engine_frame_lanes(lanes, b->aux_ready ? b->lanes : 1, list.data(), in.data(), count,
                              b->ext.what != 0);
}

int dogm_batch_get_grid_cells(dogm_batch* b, int vehicle_id, dogm_grid_cell* out)
{
    if (!b)
        return DOGM_ERR_INVALID_ARGUMENT;
    Vehicle* v = find_vehicle(b, vehicle_id);
    if (!v)
        return fail(&b->eng, DOGM_ERR_INVALID_ARGUMENT, "unknown vehicle id " + std::to_string(vehicle_id));
    return get_grid_cells(&b->eng, v, out);
}

int dogm_batch_get_particles(dogm_batch* b, int vehicle_id, dogm_vec4* state, int* grid_cell_idx, float* weight,
                             uint8_t* associated)
{
    if (!b)
        return DOGM_ERR_INVALID_ARGUMENT;
    Vehicle* v = find_vehicle(b, vehicle_id);
    if (!v)
        return fail(&b->eng, DOGM_ERR_INVALID_ARGUMENT, "unknown vehicle id " + std::to_string(vehicle_id));
    return read_particles(&b->eng, v, v->part[0], v->N, state, grid_cell_idx, weight, associated);
}

int dogm_batch_get_resampled_indices(dogm_batch* b, int vehicle_id, int* out)
{
    if (!b)
        return DOGM_ERR_INVALID_ARGUMENT;
    Vehicle* v = find_vehicle(b, vehicle_id);
    if (!v)
        return fail(&b->eng, DOGM_ERR_INVALID_ARGUMENT, "unknown vehicle id " + std::to_string(vehicle_id));
    return copy_out(&b->eng, out, v->ancestors, sizeof(int) * (size_t)v->N);
}

int dogm_batch_set_cdf_mode(dogm_batch* b, int vehicle_id, int mode)
{
    if (!b)
        return DOGM_ERR_INVALID_ARGUMENT;
    Vehicle* v = find_vehicle(b, vehicle_id);
    if (!v)
        return fail(&b->eng, DOGM_ERR_INVALID_ARGUMENT, "unknown vehicle id " + std::to_string(vehicle_id));
    if (mode < DOGM_CDF_EXACT || mode > DOGM_CDF_FP32C)
        return fail(&b->eng, DOGM_ERR_INVALID_ARGUMENT, "unknown CDF mode " + std::to_string(mode));
    v->cdf_mode = mode;
    return DOGM_OK;
}

int dogm_batch_set_resampling_mode(dogm_batch* b, int vehicle_id, int mode)
{
    if (!b)
        return DOGM_ERR_INVALID_ARGUMENT;
    Vehicle* v = find_vehicle(b, vehicle_id);
    if (!v)
        return fail(&b->eng, DOGM_ERR_INVALID_ARGUMENT, "unknown vehicle id " + std::to_string(vehicle_id));
    if (mode != DOGM_RESAMPLE_MULTINOMIAL && mode != DOGM_RESAMPLE_SYSTEMATIC)
        return fail(&b->eng, DOGM_ERR_INVALID_ARGUMENT, "unknown resampling mode " + std::to_string(mode));
    v->draw_mode = mode;
    return DOGM_OK;
}

int dogm_batch_set_ego_mode(dogm_batch* b, int vehicle_id, int mode)
{
    if (!b)
        return DOGM_ERR_INVALID_ARGUMENT;
    Vehicle* v = find_vehicle(b, vehicle_id);
    if (!v)
        return fail(&b->eng, DOGM_ERR_INVALID_ARGUMENT, "unknown vehicle id " + std::to_string(vehicle_id));
    return set_ego_mode(&b->eng, v, mode);
}

int dogm_batch_set_graph(dogm_batch* b, int enable)
{
    if (!b)
        return DOGM_ERR_INVALID_ARGUMENT;
    b->eng.use_graph = enable != 0;
    b->aux.use_graph = enable != 0;
    return DOGM_OK;
}

int dogm_batch_set_lanes(dogm_batch* b, int lanes)
{
    if (!b)
        return DOGM_ERR_INVALID_ARGUMENT;
    if (lanes < 0 || lanes > 2)
        return fail(&b->eng, DOGM_ERR_INVALID_ARGUMENT, "lanes must be 1, 2 or 0 (as the tuning gives)");
    b->lanes = lanes;
    return DOGM_OK;
}

int dogm_batch_set_extraction(dogm_batch* b, int what, float min_occupancy, float min_velocity)
{
    if (!b)
        return DOGM_ERR_INVALID_ARGUMENT;
    if (what & ~(DOGM_EXTRACT_OCCUPANCY | DOGM_EXTRACT_IMAGE | DOGM_EXTRACT_VELOCITY))
        return fail(&b->eng, DOGM_ERR_INVALID_ARGUMENT, "unknown extraction flags");
    b->ext = ExtractCfg{what, DOGM_SOURCE_GRID, min_occupancy, min_velocity};
    return DOGM_OK;
}

}  // extern "C"

namespace
{

int batch_extract(dogm_batch* b, int vehicle_id, int what, Vehicle** out)
{
    if (!b)
        return DOGM_ERR_INVALID_ARGUMENT;
    Vehicle* v = find_vehicle(b, vehicle_id);
    if (!v)
        return fail(&b->eng, DOGM_ERR_INVALID_ARGUMENT, "unknown vehicle id " + std::to_string(vehicle_id));
    *out = v;
    return extract_ensure(&b->eng, v, ExtractCfg{what, DOGM_SOURCE_GRID, b->ext.min_occ, b->ext.min_vel}, true);
}

}  // namespace

extern "C" {

int dogm_batch_get_occupancy(dogm_batch* b, int vehicle_id, float* out)
{
    Vehicle* v = nullptr;
    const int st = batch_extract(b, vehicle_id, DOGM_EXTRACT_OCCUPANCY, &v);
    return st ? st : extract_read_occupancy(&b->eng, v, out);
}

int dogm_batch_get_occupancy_image(dogm_batch* b, int vehicle_id, uint8_t* out)
{
    Vehicle* v = nullptr;
    const int st = batch_extract(b, vehicle_id, DOGM_EXTRACT_IMAGE, &v);
    return st ? st : extract_read_image(&b->eng, v, out);
}

int dogm_batch_get_cells_with_velocity(dogm_batch* b, int vehicle_id, dogm_velocity_cell* out, int capacity,
                                       int* count)
{
    Vehicle* v = nullptr;
    const int st = batch_extract(b, vehicle_id, DOGM_EXTRACT_VELOCITY, &v);
    return st ? st : extract_read_velocity(&b->eng, v, out, capacity, count);
}

int dogm_batch_set_kernel_timing(dogm_batch* b, int kernel_id, int enable)
{
    if (!b)
        return DOGM_ERR_INVALID_ARGUMENT;
    const int st = set_timing(&b->eng, kernel_id, enable);
    if (st == DOGM_OK && b->aux_ready)
        (void)set_timing(&b->aux, kernel_id, enable);
    return st;
}

int dogm_batch_get_kernel_timing(dogm_batch* b, int kernel_id, double* total_ms, long long* launches)
{
    if (!b)
        return DOGM_ERR_INVALID_ARGUMENT;
    double ms = 0.0, ms2 = 0.0;
    long long n = 0, n2 = 0;
    int st = get_timing(&b->eng, kernel_id, &ms, &n);
    if (st == DOGM_OK && b->aux_ready)
        st = get_timing(&b->aux, kernel_id, &ms2, &n2);
    if (st != DOGM_OK)
        return st;
    if (total_ms)
        *total_ms = ms + ms2;
    if (launches)
        *launches = n + n2;
    return DOGM_OK;
}

int dogm_batch_reset_kernel_timing(dogm_batch* b)
{
    if (!b)
        return DOGM_ERR_INVALID_ARGUMENT;
    reset_timing(&b->eng);
    if (b->aux_ready)
        reset_timing(&b->aux);
    return DOGM_OK;
}

}  // extern "C"
