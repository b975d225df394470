// Internal declarations of libdogm_hip.so: device buffers, handle state, kernel launchers.
#pragma once

#include "../../include/dogm/dogm_c.h"
#include "dogm_numerics.h"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace dogm_hip
{

using dogm_num::fx_t;
using dogm_num::ufx_t;

constexpr int BLOCK = 256;  // threads per workgroup (4 wave64)
constexpr int WAVE = 64;

// Grid cell fields stored structure-of-arrays (one f32 array per field), in dogm_grid_cell order
// after start_idx / end_idx.
enum GridField
{
    GF_NEW_BORN,
    GF_PERS,
    GF_FREE,
    GF_OCC,
    GF_PRED,
    GF_MU_A,
    GF_MU_UA,
    GF_W_A,
    GF_W_UA,
    GF_MEAN_X,
    GF_MEAN_Y,
    GF_VAR_X,
    GF_VAR_Y,
    GF_COVAR,
    GF_COUNT
};

struct DevParticles
{
    float4* state;  // (x, y, vx, vy), cells and cells/s
    int* idx;       // grid cell index
    float* weight;
    uint8_t* assoc;
};

struct DevGrid
{
    int* start;
    int* end;
    float* f[GF_COUNT];
};

// Scalars produced and consumed on the device within one frame (no host round trips).
struct DevScalars
{
    fx_t total;        // last exact scan total (particle orders or resampling CDF)
    fx_t exp_total;    // resampling: sum of the N + 1 exponentials
    dogm_num::Spacing spacing;  // resampling: draw mapping derived from exp_total and total
    double scale;      // particle-order scale v / total
    float joint_max;   // resampling: f32 of the CDF total
    float new_weight;  // resampling: joint_max / N
    int err;           // non-zero: a device-side check failed (see DevErr)
    int pad[3];
};

enum DevErr
{
    DEV_ERR_SPIN_LIMIT = 1,  // a bounded wait on another workgroup gave up
};

// Per-tile carry of the exact segmented reductions.
constexpr int SEG_TAIL_CONT = 1;  // tile's last segment continues into the next tile
constexpr int SEG_NOHEAD = 2;     // no segment starts inside the tile

struct SegCarry
{
    uint8_t* flags;      // [tiles]
    fx_t* head_partial;  // [tiles * NCH]
    fx_t* tail_partial;  // [tiles * NCH]
};

// Kernel ids for optional per-kernel event timing.
enum KernelId
{
    K_PREDICT,
    K_MOVE_PARTICLES,
    K_SHIFT_FIELD,
    K_SORT_HIST,
    K_SORT_UPSWEEP,
    K_SORT_SCAN,
    K_SORT_DOWNSWEEP,
    K_SORT_GATHER,
    K_RESET_BOUNDS,
    K_SEG_OCC,
    K_SEG_OCC_FIX,
    K_MASS_UPDATE,
    K_NORMALIZE,
    K_SEG_PERS,
    K_SEG_PERS_FIX,
    K_POST_WEIGHTS,
    K_EMPTY_CELLS,
    K_SCAN_TILES,
    K_SCAN_TOP,
    K_BIRTH_CELLS,
    K_BIRTH_PARTICLES,
    K_INIT_CELLS,
    K_INIT_PARTICLES,
    K_COPY_MASSES,
    K_SEG_MOM,
    K_SEG_MOM_FIX,
    K_CDF_WRITE,
    K_EXP_TILES,
    K_RS_MERGE,
    K_FILL,
    K_PACK_GRID,
    K_UNPACK_GRID,
    K_COUNT
};

extern const char* const kKernelNames[K_COUNT];

struct TimedLaunch
{
    int kid;
    hipEvent_t a, b;
};

}  // namespace dogm_hip

// Handle state.  All device buffers are allocated once at creation.
struct dogm_handle
{
    dogm_params params{};
    int gs = 0, C = 0, N = 0, B = 0;
    uint64_t seed = 0;
    uint32_t instance = 0;
    uint32_t epoch = 0;  // Philox epoch: incremented by every random-consuming stage

    int device = 0;
    hipStream_t stream = nullptr;

    dogm_hip::DevParticles part[2]{};  // part[cur] is the current particle array
    int cur = 0;
    dogm_hip::DevParticles birth{};
    dogm_hip::DevGrid grid{};
    float* free_alt = nullptr;    // second buffer for free_mass (occupancy reads one, writes the other)
    float* field_tmp = nullptr;   // scratch for eager grid shifts
    dogm_measurement_cell* meas = nullptr;
    float* weight_array = nullptr;
    float* born = nullptr;
    float* m_raw = nullptr;  // per-cell predicted occupancy before the >1 clamp
    int* pend = nullptr;     // particle-order ends P_j (birth / initial placement)
    int* ancestors = nullptr;

    // radix sort scratch
    int sort_passes = 0;
    int sort_tiles = 0;
    int* sort_keys[2] = {nullptr, nullptr};
    int* sort_vals[2] = {nullptr, nullptr};
    unsigned* sort_ghist = nullptr;  // [passes * 256]
    unsigned* sort_thist = nullptr;  // [256 * tiles]

    // exact segmented reductions
    int seg_tiles_max = 0;
    dogm_hip::SegCarry seg{};

    // exact scans
    int scan_tiles_max = 0;
    dogm_hip::fx_t* scan_tile_sum = nullptr;
    dogm_hip::fx_t* scan_tile_pre = nullptr;
    dogm_hip::fx_t* cdf = nullptr;
    dogm_hip::fx_t* cdf_tile_incl = nullptr;  // inclusive CDF value at the end of every scan tile

    dogm_hip::DevScalars* scalars = nullptr;
    dogm_grid_cell* aos_tmp = nullptr;  // packing buffer for grid read-backs / writes

    // pose (dogm/src/dogm.cu:161-203)
    bool first_pose = false;
    bool first_meas = false;
    float pos_x = 0.0f, pos_y = 0.0f, yaw = 0.0f;
    // Shift decided by the pose update inside dogm_update_grid, applied inside predict (particles)
    // and the occupancy update (free mass) of the same frame.
    bool pend_part_shift = false;
    float pend_px = 0.0f, pend_py = 0.0f;
    bool pend_grid_shift = false;
    int pend_gx = 0, pend_gy = 0;

    // kernel timing
    bool timing_on[dogm_hip::K_COUNT] = {};
    double timing_ms[dogm_hip::K_COUNT] = {};
    long long timing_n[dogm_hip::K_COUNT] = {};
    std::vector<dogm_hip::TimedLaunch> timed_pending;
    std::vector<hipEvent_t> event_pool;

    std::string last_error;
};

namespace dogm_hip
{

// Launch helpers (dogm_capi.hip): open / close an event bracket for kernel `kid` when timed.
void launch_begin(dogm_handle* h, int kid);
void launch_end(dogm_handle* h, int kid);

inline int div_up(long long a, long long b)
{
    return (int)((a + b - 1) / b);
}

// ---- kernel launchers (dogm_kernels.hip) ----
void run_fill_u32(dogm_handle* h, unsigned* p, unsigned v, long long n);
void run_fill_i32(dogm_handle* h, int* p, int v, long long n);
void run_move_particles(dogm_handle* h, float mx, float my);
void run_shift_field(dogm_handle* h, const void* src, void* dst, int ix, int iy);
void run_predict(dogm_handle* h, float dt, bool move, float mx, float my);
void run_reset_bounds(dogm_handle* h);
void run_sort_gather(dogm_handle* h, const int* skeys, const int* perm);
void run_occupancy(dogm_handle* h, float alpha, bool shifted, int ix, int iy);
void run_persistent(dogm_handle* h);
void run_births(dogm_handle* h);
void run_init_particles(dogm_handle* h);
void run_moments(dogm_handle* h);
void run_resampling(dogm_handle* h);
void run_pack_grid(dogm_handle* h);
void run_unpack_grid(dogm_handle* h);

// ---- radix sort (dogm_sort.hip) ----
// Stable ascending sort of h->part[cur].idx; returns device pointers to the sorted keys and the
// permutation (source slot of every sorted position).
void run_sort(dogm_handle* h, const int** skeys, const int** perm);

}  // namespace dogm_hip
