// Internal declarations of libdogm_hip.so.
//
// Execution model.  Every update runs as a *step* over a list of vehicles (independent grids,
// each with its own particles, cells, parameters and Philox instance).  The plain dogm::DOGM
// handle is a step over one vehicle; the batched and serving entry points put many vehicles of
// uneven size into the same kernel launches.  A step uploads one VehDesc per vehicle (sizes,
// parameters, RNG epochs, buffer pointers) and a work table per element space: entry u of a
// space's table names the vehicle and the local 256-element unit that workgroup-unit u covers.
// Every vehicle's units are padded to a multiple of 8, so a workgroup tile of up to 2048
// elements never straddles two vehicles, and all indices inside a vehicle are its own (cell
// bounds, ancestors, particle slots) exactly as in a single-grid run.
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

// pos and vel share one allocation of 16 bytes per particle (pos first); read-backs pack them into
// the reference's (x, y, vx, vy) records.
struct DevParticles
{
    float2* pos;    // (x, y), cells
    float2* vel;    // (vx, vy), cells/s
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

// ---------------------------------------------------------------------------------------------
// Work tables
// ---------------------------------------------------------------------------------------------
enum Space
{
    SP_P,  // particles (N per vehicle); also the N resampling draws
    SP_B,  // birth particles (B)
    SP_C,  // grid cells (C)
    SP_J,  // joint resampling weights [persistent | birth] (N + B)
    SP_E,  // exponentials of the resampling spacings (N + 1)
    SP_COUNT
};

// The resampling CDF index has one entry per CDF segment: 256 consecutive joint weights, one
// wave's share (64 lanes x 4) of a 1024-element scan tile.
constexpr int CDF_SEG = 256;
constexpr int CDF_SEGS_PER_TILE = 4;

constexpr int UNIT = 256;          // elements per table entry
constexpr int UNIT_ALIGN = 8;      // units per vehicle are padded to a multiple of this
constexpr int TAB_VEH_SHIFT = 20;  // entry = vehicle << 20 | local unit
constexpr uint32_t TAB_UNIT_MASK = (1u << TAB_VEH_SHIFT) - 1u;
constexpr int MAX_STEP_VEHICLES = 1 << (32 - TAB_VEH_SHIFT);
constexpr long long MAX_SPACE_LEN = (long long)UNIT << TAB_VEH_SHIFT;  // elements per vehicle and space

// What the kernels read about one vehicle of a step.  The current particle buffer is part[0]
// at the start of the step and the current free-mass buffer is free[0]; stages that write out of
// place flip to index 1 and the host swaps the vehicle's buffers after the step.
// Metric ego compensation: rotation by -dpsi about the grid centre (c) plus a whole-cell translation.
struct EgoXform
{
    float c;       // grid centre in cells (grid_size / 2)
    float cs, sn;  // cos / sin of the yaw change
    float tx, ty;  // translation in cells
};

struct VehDesc
{
    int gs, C, N, B;
    uint64_t seed;
    uint32_t inst;
    uint32_t ep_init, ep_pred, ep_birth, ep_rs;  // Philox epochs of this step's random stages
    float ps, sp, sv, pB, sd_birth, vmax;        // dogm.h:29-59 (cell units, as passed)
    float alpha;                                 // freespace_discount^dt (host std::pow)
    float dt;
    int init;                // first measurement: particles are initialised in this step
    int wuni;                // every weight predict reads equals weight[0] (resampled or initialised)
    int rs_gather;           // part[0] is a deferred resampled set: predict reads it through ancestors
    int rs_defer;            // the resampling writes ancestors and weight[0] only (a deferred set)
    int move;                // pending particle shift (fused into predict)
    float mx, my;
    int shifted;             // pending grid shift (fused into the occupancy update)
    int gx, gy;
    int xform_p;             // pending metric transform of the particles (fused into predict)
    int xform_g;             // ... and of the free mass (fused into the occupancy update)
    EgoXform xt;
    int cdf_mode;            // resampling CDF arithmetic (dogm_num::CdfMode)
    int draw_mode;           // resampling draws (dogm_num::DrawMode)
    int u0[SP_COUNT];        // first table unit of this vehicle per space
    int nu[SP_COUNT];        // its units per space (multiple of UNIT_ALIGN)
    DevParticles part[2];
    DevParticles birth;
    int* start;
    int* end;
    float* f[GF_COUNT];      // f[GF_FREE] is not used by step kernels: see free[]
    float* free[2];
    dogm_measurement_cell* meas;
    // A device measurement grid of this step not yet copied into meas: the occupancy update reads it
    // from here and writes the copy (dogm.cu:205-215) as it goes, instead of a separate copy.
    const dogm_measurement_cell* meas_in;
    float* weight_array;
    float4* cellrec;  // Vehicle::cellrec
    float2* cellin;   // Vehicle::cellin
    float* born;
    float* m_raw;
    int* pend;
    int* ancestors;
    int* sort_keys[2];
    int* sort_vals[2];
    fx_t* cdf;
    // result extraction (allocated on first use)
    float* x_occ;                 // [C] pignistic occupancy
    uint8_t* x_img;               // [C] 8-bit occupancy image
    unsigned long long* x_mask;   // [4 per unit] velocity-test ballots of the unit's waves
    int* x_cnt;                   // [units] passing cells per unit
    int* x_off;                   // [units + 1] their exclusive prefix; x_off[units] = total
    int* x_idx;                   // [C] passing cells, row-major
    // Occupancy bits of the last full update (bit c % 64 of nz[c / 64]: cell c had particles) and
    // whether they describe the grid in memory (see Vehicle::nz_valid).
    unsigned long long* nz;
    int nz_valid;
};

// Scalars produced and consumed on the device within one step, one per vehicle.
struct DevScalars
{
    fx_t total;        // last exact scan total (particle orders or resampling CDF)
    fx_t exp_total;    // resampling: sum of the N + 1 exponentials
    dogm_num::Spacing spacing;  // resampling: draw mapping derived from exp_total and total
    double scale;      // particle-order scale v / total
    double ftotal;     // resampling, float CDF modes: the CDF total (envelope maximum)
    float joint_max;   // resampling: f32 of the CDF total
    float new_weight;  // resampling: joint_max / N
    int pad[2];
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
    K_BUILD_TABLES,
    K_EXTRACT_MARK,
    K_EXTRACT_SCAN,
    K_EXTRACT_COMPACT,
    K_EXTRACT_GATHER,
    K_EGO_PARTICLES,
    K_EGO_FIELD,
    K_PACK_STATE,
    K_UNPACK_STATE,
    K_MATERIALIZE,
    K_COUNT
};

extern const char* const kKernelNames[K_COUNT];

struct TimedLaunch
{
    int kid;
    hipEvent_t a, b;
};

// ---------------------------------------------------------------------------------------------
// Host-side objects
// ---------------------------------------------------------------------------------------------

// Device buffers and host state of one vehicle.
struct Vehicle
{
    dogm_params params{};
    int gs = 0, C = 0, N = 0, B = 0;
    uint64_t seed = 0;
    uint32_t instance = 0;
    uint32_t epoch = 0;  // Philox epoch: incremented by every random-consuming stage
    uint64_t uid = 0;    // unique per allocation (work tables are reused while the list is unchanged)
    bool init_pending = false;  // initialise particles in the next step that includes ST_INIT
    bool w_uniform = false;     // every weight of part[0] equals part[0].weight[0]
    // part[0] is the resampled set by reference only: particle k is particle ancestors[k] of part[1]
    // (ancestors[k] < N) or of the birth set (ancestors[k] - N), with weight part[0].weight[0].  The
    // next step's predict reads it that way; anything else that reads or writes part[0], part[1] or
    // the birth set first writes it out (vehicle_materialize).
    bool rs_pending = false;
    int cdf_mode = 0;           // resampling CDF arithmetic (dogm_num::CdfMode)
    int draw_mode = 0;          // resampling draws (dogm_num::DrawMode)
    float dt = 0.0f;            // time step of the next step

    DevParticles part[2]{};  // part[0] is current
    DevParticles birth{};
    DevGrid grid{};          // grid.f[GF_FREE] is the current free mass
    float* free_alt = nullptr;
    float* field_tmp = nullptr;  // scratch for eager grid shifts
    // Per cell (p_A mu_A, (1 - p_A) mu_UA, rho_p, 0): written by the persistent update for cells with
    // particles when the moments of the same step form the posterior weights.
    float4* cellrec = nullptr;
    // Per cell with particles (likelihood, Σw before clamping): written by the occupancy update for the
    // persistent update of the same step (one 8-byte gather per particle instead of two).
    float2* cellin = nullptr;
    dogm_measurement_cell* meas = nullptr;
    // Device measurement grid of the next full update step, copied into meas by that step's
    // occupancy update (nullptr: meas already holds the grid).
    const dogm_measurement_cell* meas_src = nullptr;
    float* weight_array = nullptr;
    float* born = nullptr;
    float* m_raw = nullptr;  // per-cell predicted occupancy before the >1 clamp
    int* pend = nullptr;     // particle-order ends P_j (birth / initial placement)
    int* ancestors = nullptr;
    int* sort_keys[2] = {nullptr, nullptr};
    int* sort_vals[2] = {nullptr, nullptr};
    fx_t* cdf = nullptr;
    dogm_grid_cell* aos_tmp = nullptr;  // packing buffer for grid read-backs / writes (lazy)
    // Packing buffer of (x, y, vx, vy) records for particle read-backs / writes: part[1] (free between
    // steps), or this lazy buffer for a birth set larger than N.
    float4* state_tmp = nullptr;
    // Occupancy bits written by the occupancy update of every full update step.  While nz_valid
    // holds, a cell whose bit is clear has start = end = -1 and zero mu_A, mu_UA and velocity
    // moments in memory, so the next full update skips rewriting them.  Anything else that writes
    // those fields (stage-by-stage steps, eager shifts and rotations, grid-cell injection) clears
    // nz_valid, and the next full update rewrites every cell.
    unsigned long long* nz = nullptr;
    bool nz_valid = false;
    size_t bytes = 0;                   // device bytes held

    // Result extraction buffers (allocated on first use, not part of the footprint) and what the
    // vehicle's last step extracted (ext_what = 0 once a later step changed the state).
    float* x_occ = nullptr;
    uint8_t* x_img = nullptr;
    unsigned long long* x_mask = nullptr;
    int* x_cnt = nullptr;
    int* x_off = nullptr;
    int* x_idx = nullptr;
    dogm_velocity_cell* x_rec = nullptr;
    long long x_rec_cap = 0;
    size_t x_bytes = 0;
    int ext_what = 0;
    int ext_source = 0;
    float ext_min_occ = 0.0f, ext_min_vel = 0.0f;

    // pose (dogm/src/dogm.cu:161-203)
    int ego_mode = 0;  // DOGM_EGO_REFERENCE / DOGM_EGO_METRIC
    double ref_x = 0.0, ref_y = 0.0, ref_yaw = 0.0;  // metric mode: pose the grid frame sits at
    bool first_pose = false;
    bool first_meas = false;
    float pos_x = 0.0f, pos_y = 0.0f, yaw = 0.0f;
    // Shift decided by a pose update inside an update step, applied inside predict (particles) and
    // the occupancy update (free mass) of the same step.
    bool pend_part_shift = false;
    float pend_px = 0.0f, pend_py = 0.0f;
    bool pend_grid_shift = false;
    int pend_gx = 0, pend_gy = 0;
    // Metric ego mode: the transform decided by a pose update inside an update step, applied inside
    // predict (particles) and the occupancy update (free mass) of the same step.
    bool pend_xp = false, pend_xg = false;
    EgoXform pend_xt{};
};

// Stage selection of one step.
enum StageBits
{
    ST_ALL = 0xff,         // the eight stages of updateGrid
    ST_EXTRACT = 1 << 8,   // result extraction after the stages (Engine::ext), not part of ST_ALL
    ST_INIT = 1 << 0,
    ST_PREDICT = 1 << 1,
    ST_ASSIGN = 1 << 2,
    ST_OCC = 1 << 3,
    ST_PERS = 1 << 4,
    ST_BIRTH = 1 << 5,
    ST_MOMENTS = 1 << 6,
    ST_RESAMPLE = 1 << 7,
};

// What a step with ST_EXTRACT extracts (dogm/demo/utils/image_creation.cpp:19-66).
struct ExtractCfg
{
    int what = 0;    // DOGM_EXTRACT_* bits
    int source = 0;  // DOGM_SOURCE_GRID / DOGM_SOURCE_MEASUREMENT
    float min_occ = 0.0f, min_vel = 0.0f;
};

// Radix-sort pass forms (dogm_sort.hip).
enum SortForm
{
    SORT_REDUCE_SCAN = 0,  // upsweep, digit scan, downsweep per pass
    SORT_ONESWEEP = 1      // one launch per pass, decoupled lookback
};

// Launch choices of a step that do not change its results (dogm_tuning.hip); the defaults are the
// built-in choices.
struct Tuning
{
    int sort = SORT_REDUCE_SCAN;  // SortForm of the radix passes
    int merge_slice = 0;          // exact-CDF merge: CDF tiles staged per workgroup (1, 2; 0: by size)
    int pred_block = 512;         // predict: 512 threads x 4 particles or 256 x 8 per workgroup
    int seg_items = 8;            // occupancy and persistent reductions: particles per thread (4, 8)
    int lanes = 2;                // engines a batched step of two or more vehicles is split over
};

// Distinct for distinct tunings (part of a step graph's launch signature).
inline uint64_t tuning_signature(const Tuning& t)
{
    return (uint64_t)(unsigned)t.sort | (uint64_t)(unsigned)t.merge_slice << 4 |
           (uint64_t)(unsigned)t.pred_block << 8 | (uint64_t)(unsigned)t.seg_items << 20 |
           (uint64_t)(unsigned)t.lanes << 28;
}

bool tuning_valid(const Tuning& t, std::string* why);
// The table's choice for a step of `vehicles` vehicles and `particles` particles in all (or the
// override, or the built-in choices).
Tuning tuning_for(int vehicles, long long particles);

// Execution context on one device and stream: step descriptors, work tables and the shared
// scratch of the scans, segmented reductions and the sort.
struct Engine
{
    int device = 0;
    hipStream_t stream = nullptr;

    // step plan
    int m = 0;                       // vehicles of the current plan
    std::vector<uint64_t> plan_veh;  // uids of the vehicles the tables were built for
    long long units[SP_COUNT] = {};
    long long tab_off[SP_COUNT] = {};
    int passes = 1;
    int m_cap = 0;
    VehDesc* h_desc = nullptr;  // pinned staging
    VehDesc* d_desc = nullptr;
    int* h_upre = nullptr;      // pinned: per space, m + 1 unit prefixes
    int* d_upre = nullptr;
    long long tab_cap = 0;
    uint32_t* d_tab = nullptr;

    // scratch sized for the largest plan so far
    long long scan_tiles_cap = 0;
    fx_t* scan_tile_sum = nullptr;
    fx_t* scan_tile_pre = nullptr;
    fx_t* exp_tile_sum = nullptr;  // resampling: the exponentials' tile totals (scanned beside the CDF's)
    // Resampling CDF index: exact mode, the inclusive end of every CDF segment (CDF_SEGS_PER_TILE
    // per scan tile); float modes, two doubles per scan tile (envelope at the end and before it).
    fx_t* cdf_tile_incl = nullptr;
    long long seg_tiles_cap = 0;
    SegCarry seg{};
    long long sort_tiles_cap = 0;
    unsigned* sort_status = nullptr;  // [sort tiles * 256] decoupled-lookback words of the radix passes
    long long status_plan = -1;       // plan whose tiles sort_status holds lookback words for
    int lookback_spin = 1 << 20;      // polls before a lookback wait gives up (DOGM_ERR_DEVICE_CHECK)
    unsigned* sort_tickets = nullptr; // [4] tile tickets of the passes (after sort_ghist's histograms)
    // sort_ghist is all zero: every sort clears the histograms it consumed in its last pass, so the
    // next step's predict can count into them without a fill (false after a step that stopped early).
    bool ghist_clean = false;
    int ghist_cap = 0;
    unsigned* sort_ghist = nullptr;  // [m * 4 * 256]
    int sc_cap = 0;
    DevScalars* scalars = nullptr;   // [m]
    // Device-side check failures (a lookback wait that ran out): set by kernels in host-mapped
    // memory, read and cleared by the host after every step.
    unsigned* h_err = nullptr;
    unsigned* d_err = nullptr;

    // kernel timing
    bool timing_on[K_COUNT] = {};
    double timing_ms[K_COUNT] = {};
    long long timing_n[K_COUNT] = {};
    std::vector<TimedLaunch> timed_pending;
    std::vector<hipEvent_t> event_pool;

    ExtractCfg ext;  // read by steps with ST_EXTRACT
    Tuning tune;     // launch choices of the current step
    bool tune_locked = false;  // set by a caller that chose tune for a whole multi-lane step

    // Step graphs: with use_graph, a step's kernel sequence is captured once into a HIP graph and
    // replayed while the plan (vehicle list, tables, scratch) and the launch signature are unchanged.
    bool use_graph = false;
    long long plan_epoch = 0;  // incremented whenever the tables / scratch are rebuilt
    hipGraphExec_t graph_exec = nullptr;
    long long graph_plan = -1;
    uint64_t graph_sig[3] = {0, 0, 0};

    // Step buffers bound to caller memory (engine_bind_scratch): the engine's own buffers and their
    // capacities wait here until engine_unbind_scratch puts them back.
    struct OwnStepBuffers
    {
        uint32_t* d_tab;
        long long tab_cap;
        fx_t *scan_tile_sum, *scan_tile_pre, *exp_tile_sum, *cdf_tile_incl;
        long long scan_tiles_cap;
        SegCarry seg;
        long long seg_tiles_cap;
        unsigned* sort_status;
        long long sort_tiles_cap;
    } own{};
    bool bound = false;

    std::string last_error;

    const uint32_t* tab(int s) const { return d_tab + tab_off[s]; }
    int tiles(int s, int q) const { return (int)(units[s] / q); }
};

// ---- launch helpers (dogm_engine.hip) ----
void launch_begin(Engine* E, int kid);
void launch_end(Engine* E, int kid);

inline int div_up(long long a, long long b)
{
    return (int)((a + b - 1) / b);
}

// ---- engine (dogm_engine.hip) ----
long long units_for(long long len);  // table units of a space of `len` elements (padded)

int engine_init(Engine* E, std::string* why);
void engine_release(Engine* E);
// Reserves the per-vehicle step arrays (descriptors, unit prefixes, digit histograms, scalars) for
// steps of up to m vehicles, so that later steps of that size allocate only plan-sized buffers.
int engine_reserve(Engine* E, int m, std::string* why);
// Bytes of the plan-sized step buffers (work tables, scan tile totals, segmented-reduction carries,
// sort tile histograms) of a step over `veh`, each buffer starting on a 256-byte boundary.
int engine_step_scratch_bytes(Vehicle* const* veh, int m, size_t* bytes, std::string* why);
// Points those buffers into caller memory [base, base + bytes) for the following steps over `veh`
// (the engine then allocates none of them), and forces the next step to rebuild its tables there.
int engine_bind_scratch(Engine* E, void* base, size_t bytes, Vehicle* const* veh, int m, std::string* why);
// Gives the engine its own buffers back (after the bound steps have completed on its stream).
void engine_unbind_scratch(Engine* E);
// Builds descriptors and tables for `veh` (reusing the tables while the vehicle list is
// unchanged) and uploads the descriptors; the stage set fixes the per-step fields (epochs, shifts).
int engine_plan(Engine* E, Vehicle* const* veh, int m, int stages);
// Runs the selected stages of one step on `veh` (asynchronous on E->stream) and updates the
// vehicles' host state (epochs, buffer swaps, pending shifts).
int engine_run(Engine* E, Vehicle* const* veh, int m, int stages);
void engine_collect_timing(Engine* E);
// After a step's stream has been waited for: DOGM_ERR_DEVICE_CHECK (and the message) when a kernel
// of the step flagged a failed device-side check, else DOGM_OK.
int engine_device_check(Engine* E);
int vehicle_validate(const dogm_params* p, std::string* why);
size_t vehicle_footprint(const dogm_params* p);  // device bytes vehicle_alloc takes
int vehicle_alloc(Vehicle* v, const dogm_params* p, uint64_t seed, uint32_t instance, std::string* why);
void vehicle_free(Vehicle* v);
// Queues the copy of a frame's measurement grid (dogm.cu:205-215); the first one arms the particle
// initialisation of the next step.
int vehicle_set_measurement(Engine* E, Vehicle* v, const dogm_measurement_cell* m, int is_device);
// The same for a frame whose full update step follows at once: a device grid after the first is
// copied by that step's occupancy update (Vehicle::meas_src) instead of a separate copy.
int vehicle_take_measurement(Engine* E, Vehicle* v, const dogm_measurement_cell* m, int is_device);
// dogm.cu:161-203.  lazy: the shift is applied inside the next step's predict (particles) and
// occupancy update (free mass); otherwise at once.
void vehicle_update_pose(Engine* E, Vehicle* v, float new_x, float new_y, float new_yaw, bool lazy);
// Writes out v's deferred resampled set, if any (Vehicle::rs_pending), on E's stream.
void vehicle_materialize(Engine* E, Vehicle* v);

// One updateGrid frame for each of `m` distinct vehicles in one step.
struct FrameIn
{
    const dogm_measurement_cell* meas;  // nullptr: the grid already sits in the vehicle's buffer
    int is_device;
    float x, y, yaw, dt;
};
// Measurement copies, lazy pose updates, the full stage sequence (plus ST_EXTRACT with
// E->ext when `extract`), and a wait for the stream.
int engine_frame(Engine* E, Vehicle* const* veh, const FrameIn* in, int m, bool extract = false);
// The same step split over two engines ("lanes", each with its own stream, tables and scratch):
// every vehicle runs all its stages on one lane, and the two lanes' launches overlap on the device
// (one lane's small, latency-bound launches and tails run beside the other's bulk work).  Results
// equal those of engine_frame.  nl = 0 takes the lane count from the tuning of the whole step (both
// lanes then use that step's tuning); lanes[1] must be initialised unless nl == 1.  Errors are
// reported on lanes[0].
int engine_frame_lanes(Engine* const* lanes, int nl, Vehicle* const* veh, const FrameIn* in, int m,
                       bool extract = false);
// Allocates the extraction buffers `what` needs on a vehicle.
int vehicle_ext_alloc(Vehicle* v, int what, std::string* why);
// Lidar generator internals used by the server (dogm_lidar.hip).
int lidar_grid_size(const dogm_lidar* l);
int lidar_beam_count(const dogm_lidar* l);

// Velocity records are packed and read back in chunks of at most this many.
constexpr int EXTRACT_REC_CHUNK = 1 << 16;
// Device bytes the extraction buffers for `what` take on a vehicle of C cells (record staging included).
size_t extract_footprint(int C, int what);

// ---- result getters shared by the plain, batch and server entry points (dogm_capi.hip) ----
// Runs the extraction `want` on v and waits, unless `reuse` and v's last step already extracted it
// (same source and, for the velocity test, the same thresholds).
int extract_ensure(Engine* E, Vehicle* v, const ExtractCfg& want, bool reuse);
int extract_read_occupancy(Engine* E, Vehicle* v, float* out);
int extract_read_image(Engine* E, Vehicle* v, uint8_t* out);
int extract_read_velocity(Engine* E, Vehicle* v, dogm_velocity_cell* out, int capacity, int* count);

// ---- stage launchers (dogm_kernels.hip, dogm_sort.hip, dogm_resample.hip) ----
// `pph` is the index (0 / 1) of the current particle buffer of the step's descriptors.
void run_fill_u32(Engine* E, unsigned* p, unsigned v, long long n);
void run_build_tables(Engine* E);
// Stages that run in the same step share work: predict counts the sort's digit histogram (hist),
// the occupancy update zeroes the persistent / moment fields of empty cells (zero_empty bits 0 / 1),
// the weight normalisation moves into the persistent update's load and the posterior weights into
// the moments' load.  Each stage run alone keeps the reference's separate passes.
void run_init_particles(Engine* E);                // pph 0
// aos: the state goes to the CDF buffer as 16-byte records for the gathering sort pass of the same
// step (run_sort with hist_done and gather), which reads it there; for vehicles whose weights are
// uniform (VehDesc::wuni) predict then leaves the weights in place and the gathering pass forms
// each predicted weight from weight[0] and the predicted position, with predict's arithmetic.
void run_predict(Engine* E, int pph, bool hist, bool aos);
// defer_bounds: the last sort pass writes the particles directly and the occupancy reduction records
// the cell bounds (bounds = true there; the reduction then reads the sorted particles' weights).
// write_wa: that pass also writes the weight copy into weight_array (not needed when the persistent
// update of the same step overwrites weight_array).
void run_assignment(Engine* E, int pph, bool hist_done, bool defer_bounds, bool write_wa);  // reads pph, writes pph ^ 1
// born_sums: the occupancy update leaves the born masses' exact sum per cell unit for the births of the
// same step (born_unit_sums), whose scan then skips its tile pass.
// exp_sums: the same launch computes the resampling exponentials' tile totals (exp_tile_sum) for the
// resampling of the same step (exp_done there).
void run_occupancy(Engine* E, int pph, int zero_empty, bool defer_normalize, bool bounds, bool born_sums,
                   bool exp_sums);  // free[0] -> free[1]
void run_persistent(Engine* E, int pph, bool fused_normalize, bool empty_done, bool defer_post);
void run_births(Engine* E, bool born_sums);
// Unit sums of the born masses between the occupancy update and the births of a step: the CDF index
// buffer (SCAN_Q units per scan tile fit its CDF_SEGS_PER_TILE entries per tile), unused until the
// resampling.
inline fx_t* born_unit_sums(Engine* E) { return E->cdf_tile_incl; }
void run_moments(Engine* E, int pph, bool fused_post, bool empty_done);
// modes: bit m set when a vehicle of the step resamples with CDF mode m (dogm_num::CdfMode).
void run_resampling(Engine* E, int pph, int modes, bool exp_done);  // reads pph, writes pph ^ 1
// Flags of the gathering sort pass.
constexpr int GATHER_AOS = 1;  // the state is in the CDF buffer as 16-byte records (predict, same step)
constexpr int GATHER_WA = 2;   // also write the weight copy into weight_array
// gather: the last pass writes the particles (and, with write_wa, the weight copy into weight_array).
void run_sort(Engine* E, int pph, bool hist_done, bool gather, bool write_wa);
// Single-vehicle kernels (no tables): eager ego shift and grid packing.
void run_move_particles(Engine* E, const Vehicle* v, float mx, float my);
// Writes out a deferred resampled set (Vehicle::rs_pending) into part[0].
void run_materialize(Engine* E, const Vehicle* v);
void run_ego_particles(Engine* E, const Vehicle* v, const EgoXform& t);
void run_ego_field(Engine* E, const Vehicle* v, const void* src, void* dst, const EgoXform& t);
void run_shift_field(Engine* E, const Vehicle* v, const void* src, void* dst, int ix, int iy);
void run_pack_grid(Engine* E, const Vehicle* v);
void run_unpack_grid(Engine* E, const Vehicle* v);
// (x, y, vx, vy) records <-> the pos / vel arrays of n particles.
void run_pack_state(Engine* E, const DevParticles& p, int n, float4* out);
void run_unpack_state(Engine* E, const DevParticles& p, int n, const float4* in);
// Particle read-back / write of one of v's sets (part[0] or birth) in the reference's layout:
// (x, y, vx, vy) records, cell index, weight, association; null outputs are skipped.
int read_particles(Engine* E, Vehicle* v, const DevParticles& p, int n, dogm_vec4* state, int* idx, float* weight,
                   uint8_t* assoc);
int write_particles(Engine* E, Vehicle* v, const DevParticles& p, int n, const dogm_vec4* state, const int* idx,
                    const float* weight, const uint8_t* assoc);
// Result extraction over the step's vehicles (dogm_extract.hip); fsel selects the current free mass
// buffer of the step's descriptors (1 after an occupancy update in the same step).
void run_extract(Engine* E, int fsel);
// Packs records of the first n cells of v->x_idx into v->x_rec.
void run_extract_gather(Engine* E, const Vehicle* v, int n);

}  // namespace dogm_hip

// The plain handle: one engine, one vehicle.
struct dogm_handle
{
    dogm_hip::Engine eng;
    dogm_hip::Vehicle veh;
};
