/*
 * C-ABI of the MI355X (gfx950) DOGM particle-filter update, exported by libdogm_hip.so.
 *
 * This is the drop-in boundary for the path behind dogm::DOGM::updateGrid of the reference
 * (dogm/include/dogm/dogm.h:20-199, dogm/src/dogm.cu:115-131).  The inline C++ class in
 * dogm/dogm.h wraps it with the reference's surface; other hosts (ctypes, cgo, JNI) bind these
 * functions directly.  Plain pointers and sizes only: device memory lives behind the opaque handle.
 *
 * Conventions
 *  - Every function returning int returns DOGM_OK (0) or a DOGM_ERR_* code; the message of the last
 *    failure on a handle is available from dogm_last_error().  The reference only prints CUDA errors
 *    and continues (dogm/include/dogm/cuda_utils.h:13-24).
 *  - All calls on a handle are blocking and use the handle's own HIP stream (the reference's
 *    updateGrid ends in cudaDeviceSynchronize, dogm/src/dogm.cu:130).  A handle is not thread-safe;
 *    use one handle per host thread.
 *  - Particle state is in grid cells and cells/s, as in the reference (predict.cu:27-33).
 */
#ifndef DOGM_C_H
#define DOGM_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same fields, order and types as dogm::DOGM::Params (dogm/include/dogm/dogm.h:26-60). */
typedef struct dogm_params
{
    float size;                          /* grid side [m] */
    float resolution;                    /* [m/cell]; grid_size = (int)(size / resolution) */
    int particle_count;                  /* persistent particles N */
    int new_born_particle_count;         /* birth particles B */
    float persistence_prob;              /* p_S */
    float stddev_process_noise_position; /* [cells] */
    float stddev_process_noise_velocity; /* [cells/s] */
    float birth_prob;                    /* p_B */
    float stddev_velocity;               /* birth velocity sigma [cells/s] */
    float init_max_velocity;             /* first-frame velocity U(-v, v) [cells/s] */
    float freespace_discount;            /* alpha = freespace_discount^dt */
} dogm_params;

/* 64-byte cell, same layout as dogm::GridCell (dogm/include/dogm/dogm_types.h:13-33). */
typedef struct dogm_grid_cell
{
    int start_idx;
    int end_idx;
    float new_born_occ_mass;
    float pers_occ_mass;
    float free_mass;
    float occ_mass;
    float pred_occ_mass;
    float mu_A;
    float mu_UA;
    float w_A;
    float w_UA;
    float mean_x_vel;
    float mean_y_vel;
    float var_x_vel;
    float var_y_vel;
    float covar_xy_vel;
} dogm_grid_cell;

/* 16-byte cell, same layout as dogm::MeasurementCell (dogm_types.h:35-41). */
typedef struct dogm_measurement_cell
{
    float free_mass;
    float occ_mass;
    float likelihood;
    float p_A;
} dogm_measurement_cell;

/* Particle state (x, y, vx, vy); same 16-byte layout as the reference's glm::vec4. */
typedef struct dogm_vec4
{
    float x, y, z, w;
} dogm_vec4;

typedef struct dogm_handle dogm_handle;

enum
{
    DOGM_OK = 0,
    DOGM_ERR_INVALID_ARGUMENT = 1,
    DOGM_ERR_OUT_OF_MEMORY = 2,
    DOGM_ERR_HIP = 3,
    DOGM_ERR_DEVICE_CHECK = 4, /* a kernel-side check failed (capacity or consistency) */
    DOGM_ERR_NO_DEVICE = 5
};

/* Default seed: the reference seeds every RNG state with 123456 (dogm/src/dogm.cu:101). */
#define DOGM_DEFAULT_SEED 123456ull

/* ----------------------------------------------------------------------------------------------
 * Construction (replaces DOGM::DOGM / ~DOGM / initialize, dogm/src/dogm.cu:33-113)
 * ---------------------------------------------------------------------------------------------- */
int dogm_create(const dogm_params* params, dogm_handle** out);
/* Extension: explicit Philox seed and instance id (the instance id selects an independent stream,
 * so several filters can share a seed). */
int dogm_create_ex(const dogm_params* params, uint64_t seed, uint32_t instance, dogm_handle** out);
void dogm_destroy(dogm_handle* h);
const char* dogm_last_error(const dogm_handle* h);
const char* dogm_status_string(int status);

/* ----------------------------------------------------------------------------------------------
 * Per-frame update (replaces DOGM::updateGrid, dogm/include/dogm/dogm.h:82-83, dogm/src/dogm.cu:115-131)
 * measurement_grid: grid_size^2 cells, row-major (index = x + grid_size * y), device memory when
 * is_device != 0, host memory otherwise.  Borrowed: copied before the call returns.
 * ---------------------------------------------------------------------------------------------- */
int dogm_update_grid(dogm_handle* h, const dogm_measurement_cell* measurement_grid, float new_x, float new_y,
                     float new_yaw, float dt, int is_device);

/* No counterpart in the reference: the same update as dogm_update_grid, for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The step's plan-sized buffers (work tables, scan tile totals, segmented-reduction
 * carries, sort tile histograms) come from a static 16 MiB device array per GPU instead of
 * allocations, and the call allocates nothing (no step graph, no kernel timing).  A grid whose step
 * needs more than 16 MiB (about 1.3e7 particles) is rejected with DOGM_ERR_INVALID_ARGUMENT.  All
 * calls on one GPU share that array: calls that overlap in time on one GPU (from several host
 * threads) race and must be serialised by the caller.  Results equal dogm_update_grid's. */
int dogm_update_grid_noalloc(dogm_handle* h, const dogm_measurement_cell* measurement_grid, float new_x,
                             float new_y, float new_yaw, float dt, int is_device);

/* Private steps of updateGrid in the reference, exported so tests can drive them. */
int dogm_update_measurement_grid(dogm_handle* h, const dogm_measurement_cell* measurement_grid,
                                 int is_device);                                    /* dogm.cu:205-215 */
int dogm_update_pose(dogm_handle* h, float new_x, float new_y, float new_yaw);  /* dogm.cu:161-203 */

/* Public stage methods (dogm/include/dogm/dogm.h:148-156). */
int dogm_initialize_particles(dogm_handle* h);                /* dogm.cu:217-241 */
int dogm_particle_prediction(dogm_handle* h, float dt);       /* dogm.cu:243-260 */
int dogm_particle_assignment(dogm_handle* h);                 /* dogm.cu:262-281 */
int dogm_grid_cell_occupancy_update(dogm_handle* h, float dt); /* dogm.cu:283-298 */
int dogm_update_persistent_particles(dogm_handle* h);         /* dogm.cu:300-321 */
int dogm_initialize_new_particles(dogm_handle* h);            /* dogm.cu:323-348 */
int dogm_statistical_moments(dogm_handle* h);                 /* dogm.cu:350-384 */
int dogm_resampling(dogm_handle* h);                          /* dogm.cu:386-423 */

/* ----------------------------------------------------------------------------------------------
 * Read-backs (replace getGridCells / getMeasurementCells / getParticles, dogm/src/dogm.cu:133-159,
 * and the scalar getters, dogm/include/dogm/dogm.h:90-139).  Output buffers are host memory owned by
 * the caller and sized for grid_size^2 cells or particle_count particles.
 * ---------------------------------------------------------------------------------------------- */
int dogm_get_grid_cells(dogm_handle* h, dogm_grid_cell* out);
int dogm_get_measurement_cells(dogm_handle* h, dogm_measurement_cell* out);
int dogm_get_particles(dogm_handle* h, dogm_vec4* state, int* grid_cell_idx, float* weight, uint8_t* associated);
int dogm_get_grid_size(const dogm_handle* h);
int dogm_get_particle_count(const dogm_handle* h);
int dogm_get_new_born_particle_count(const dogm_handle* h);
float dogm_get_resolution(const dogm_handle* h);
float dogm_get_position_x(const dogm_handle* h);
float dogm_get_position_y(const dogm_handle* h);
float dogm_get_yaw(const dogm_handle* h);

/* ----------------------------------------------------------------------------------------------
 * Inspection and state injection.  The reference exposes all of its device buffers as public data
 * members (dogm/include/dogm/dogm.h:158-191: particle_array, birth_particle_array, grid_cell_array,
 * weight_array, born_masses_array ...); these calls replace that direct access.
 * ---------------------------------------------------------------------------------------------- */
int dogm_get_birth_particles(dogm_handle* h, dogm_vec4* state, int* grid_cell_idx, float* weight,
                             uint8_t* associated);
int dogm_get_weight_array(dogm_handle* h, float* out);      /* N floats */
int dogm_get_born_masses(dogm_handle* h, float* out);       /* grid_size^2 floats */
int dogm_get_resampled_indices(dogm_handle* h, int* out);   /* N ancestor indices of the last resampling */
int dogm_set_particles(dogm_handle* h, const dogm_vec4* state, const int* grid_cell_idx, const float* weight,
                       const uint8_t* associated);
int dogm_set_birth_particles(dogm_handle* h, const dogm_vec4* state, const int* grid_cell_idx,
                             const float* weight, const uint8_t* associated);
int dogm_set_grid_cells(dogm_handle* h, const dogm_grid_cell* cells);
int dogm_set_weight_array(dogm_handle* h, const float* in);
int dogm_set_born_masses(dogm_handle* h, const float* in);

/* ----------------------------------------------------------------------------------------------
 * Result extraction on the device.  Replaces the demo's per-frame post-processing, which reads the
 * whole grid back (getGridCells, 64 bytes per cell) and loops over it on the host
 * (dogm/demo/utils/image_creation.cpp:19-66 computeCellsWithVelocity, :70-87
 * compute_measurement_grid_image, :112-128 compute_dogm_image).  Outputs are compact: 4 or 1 bytes
 * per cell, or one record per cell that passes the velocity test.
 *
 *  pignistic occupancy  p = occ + 0.5 * (1 - occ - free)   (f32, image_creation.cpp:19-22)
 *  image                255 - floor(255 * p), clamped to [0, 255] (image_creation.cpp:80-83, 123-125)
 *  velocity test        p >= min_occupancy and v^T S^-1 v >= min_velocity, v = (mean_x_vel,
 *                       mean_y_vel), S = [[var_x_vel, covar_xy_vel], [covar_xy_vel, var_y_vel]]
 *                       (image_creation.cpp:35-50).  S^-1 is OpenCV's closed 2x2 inverse: determinant
 *                       in f64, elements f32(element * (1 / det)), all zero when det == 0; both
 *                       products are 2-term dot products summed in f64 and rounded to f32.
 * ---------------------------------------------------------------------------------------------- */
enum
{
    DOGM_SOURCE_GRID = 0,        /* the filter's grid cells (occ_mass, free_mass) */
    DOGM_SOURCE_MEASUREMENT = 1  /* the current measurement grid (occupancy and image only) */
};
enum
{
    DOGM_EXTRACT_OCCUPANCY = 1, /* f32 pignistic occupancy per cell */
    DOGM_EXTRACT_IMAGE = 2,     /* 8-bit occupancy image per cell */
    DOGM_EXTRACT_VELOCITY = 4   /* cells that pass the velocity test */
};

/* A cell that passed the velocity test: the demo's Point<dogm::GridCell> (dbscan.h:10-17) without
 * cluster_id.  x, y = column and row of the cell (image_creation.cpp:53-56). */
typedef struct dogm_velocity_cell
{
    float x, y;
    dogm_grid_cell data;
} dogm_velocity_cell;

int dogm_get_occupancy(dogm_handle* h, int source, float* out);         /* grid_size^2 floats */
int dogm_get_occupancy_image(dogm_handle* h, int source, uint8_t* out); /* grid_size^2 bytes */
/* Passing cells in row-major order (the demo's loop order).  Writes min(*count, capacity) records;
 * *count = number of passing cells.  out may be NULL when capacity is 0. */
int dogm_get_cells_with_velocity(dogm_handle* h, float min_occupancy, float min_velocity, dogm_velocity_cell* out,
                                 int capacity, int* count);

/* ----------------------------------------------------------------------------------------------
 * Accuracy harness (host code): DBSCAN over the cells that pass the velocity test, cluster means and
 * MAE / RMSE against ground-truth objects -- the reference demo's evaluation
 * (dogm/demo/utils/dbscan.cpp:24-114, precision_evaluator.cpp:37-122, metrics.cpp:7-53).
 * DBSCAN is the textbook algorithm with the reference's conventions (a neighbourhood includes its
 * point; core points have >= min_cells neighbours within eps, distance sqrtf of the cell-coordinate
 * differences; clusters numbered by first core point in input order; breadth-first expansion in
 * input order; a border point keeps the first cluster that reaches it).  The reference's
 * expandCluster labels the first |seeds| entries of its point array instead of the seeds
 * (dbscan.cpp:62-65); that defect is not reproduced.
 * ---------------------------------------------------------------------------------------------- */
/* Cluster mean: position in cell units (dogm_cluster_means) or metres (evaluator), velocity in
 * cells/s or m/s. */
typedef struct dogm_cluster
{
    float x, y, vx, vy;
    int cells;
} dogm_cluster;

/* A ground-truth object (or an error vector): position [m] in the grid frame, velocity [m/s]. */
typedef struct dogm_object
{
    float x, y, vx, vy;
} dogm_object;

/* cluster_id[i]: cluster of cell i, -1 for noise; *clusters = number of clusters. */
int dogm_dbscan(const dogm_velocity_cell* cells, int n, float eps, int min_cells, int* cluster_id, int* clusters);
/* Mean cell coordinates and mean velocities per cluster, f32 sums in input order
 * (precision_evaluator.cpp:98-110, before its metre conversion). */
int dogm_cluster_means(const dogm_velocity_cell* cells, int n, const int* cluster_id, int clusters,
                       dogm_cluster* out);

typedef struct dogm_evaluator dogm_evaluator;
/* The demo uses resolution = the grid's, max_assignment_distance 5 m, eps 3 cells, min_cells 5
 * (precision_evaluator.cpp:15-17). */
int dogm_evaluator_create(float resolution, float max_assignment_distance, float eps, int min_cells,
                          dogm_evaluator** out);
void dogm_evaluator_destroy(dogm_evaluator* e);
/* One frame: clusters the cells, converts the means to metres (x resolution), assigns each cluster
 * to the closest object closer than max_assignment_distance and accumulates its error; a cluster
 * without one counts as unassigned.  Cluster means (metres) are written to clusters_out (up to
 * clusters_cap) and their number to *clusters.  Frames without cells or objects add nothing. */
int dogm_evaluator_add_step(dogm_evaluator* e, const dogm_velocity_cell* cells, int n, const dogm_object* truth,
                            int objects, dogm_cluster* clusters_out, int clusters_cap, int* clusters);
/* MAE and RMSE of (x, y, vx, vy) over every assigned cluster so far (sums in f64). */
int dogm_evaluator_summary(const dogm_evaluator* e, dogm_object* mae, dogm_object* rmse, long long* detections,
                           long long* unassigned);

/* ----------------------------------------------------------------------------------------------
 * Measurement-grid generation on the device (the step before updateGrid).  Replaces the reference's
 * LaserMeasurementGrid::generateGrid (dogm/demo/simulator/mapping/laser_to_meas_grid.cu:25-70), a polar
 * inverse-sensor texture rendered to a Cartesian grid with OpenGL, by a direct ray cast with a
 * two-level inverse sensor model:
 *   free_mass on every cell a beam samples before its return (samples every resolution / 2 metres
 *   from the sensor, distance < min(range, max_range)); occ_mass and free 0 on the cell of a return
 *   (range <= max_range); unknown (0, 0) elsewhere; likelihood = p_A = 1 (measurement_grid.cu:129-130).
 * The sensor sits at the grid centre, 0.5 * grid_size * resolution metres along both axes; a
 * sample at (x, y) metres lies in cell floor(x / resolution) + grid_size * floor(y / resolution).
 * Positions are f64 with that operation order, so grids are reproducible bit for bit.
 * ---------------------------------------------------------------------------------------------- */
typedef struct dogm_lidar_params
{
    int grid_size;     /* cells per side of the generated grids */
    int beam_count;
    double resolution; /* [m/cell] */
    double max_range;  /* [m] */
    float occ_mass;    /* occupied mass at a return (0.7 in the benchmark scenes) */
    float free_mass;   /* free mass along a beam (0.4) */
} dogm_lidar_params;

typedef struct dogm_lidar dogm_lidar;

/* beam_cos / beam_sin: unit direction of each beam (both NULL: 360 degrees, beam b at
 * 2 pi b / beam_count, directions from the host's cos / sin in double). */
int dogm_lidar_create(const dogm_lidar_params* params, const double* beam_cos, const double* beam_sin,
                      dogm_lidar** out);
void dogm_lidar_destroy(dogm_lidar* l);
const char* dogm_lidar_last_error(const dogm_lidar* l);
/* ranges: beam_count distances [m] in host memory (inf or NaN: no return).  device_out: grid_size^2
 * cells of device memory, e.g. a vehicle's measurement buffer (dogm_batch_measurement_buffer).
 * Blocking. */
int dogm_lidar_generate(dogm_lidar* l, const double* ranges, dogm_measurement_cell* device_out);
/* count grids in the same launches: ranges holds count * beam_count distances, grid i from
 * ranges + i * beam_count into device_out[i]. */
int dogm_lidar_generate_many(dogm_lidar* l, int count, const double* ranges,
                             dogm_measurement_cell* const* device_out);

/* ----------------------------------------------------------------------------------------------
 * Device-memory helpers for callers that pass device measurement grids (the reference's callers use
 * cudaMalloc / cudaMemcpy for this, e.g. dogm/demo/simulator/mapping/laser_to_meas_grid.cu:15).
 * kind: 0 host->device, 1 device->host, 2 device->device.
 * ---------------------------------------------------------------------------------------------- */
int dogm_device_malloc(void** ptr, size_t bytes);
int dogm_device_free(void* ptr);
int dogm_memcpy(void* dst, const void* src, size_t bytes, int kind);
/* Device-to-device copy in one kernel dispatch (16-byte aligned pointers); blocking.  Also the
 * reference transfer of known size for scaling memory-traffic counters. */
int dogm_device_copy(void* dst, const void* src, size_t bytes);
int dogm_device_synchronize(void);

/* ----------------------------------------------------------------------------------------------
 * Kernel timing.  When enabled, the handle records a HIP event pair around every launch of the
 * selected kernel (on the handle's stream) and accumulates the durations.
 * ---------------------------------------------------------------------------------------------- */
int dogm_kernel_count(void);
const char* dogm_kernel_name(int kernel_id);
int dogm_set_kernel_timing(dogm_handle* h, int kernel_id, int enable);
/* Total milliseconds and launch count accumulated for kernel_id since the last reset. */
int dogm_get_kernel_timing(dogm_handle* h, int kernel_id, double* total_ms, long long* launches);
int dogm_reset_kernel_timing(dogm_handle* h);

/* ----------------------------------------------------------------------------------------------
 * Tuning (no counterpart in the reference, whose launch geometry is fixed: dogm/src/dogm.cu:45-47).
 * Launch choices that change no result, per step class: the radix sort's pass form
 * (DOGM_SORT_REDUCE_SCAN: upsweep, digit scan and scatter per pass; DOGM_SORT_ONESWEEP: one launch
 * per pass with a decoupled lookback), the exact-CDF merge's staged CDF tiles per workgroup
 * (merge_slice 1 or 2; 0 = by grid size), predict's workgroup (pred_block 512 threads x 4 particles
 * or 256 x 8), the occupancy and persistent reductions' particles per thread (seg_items 4 or 8) and
 * the lanes of batched steps (1 or 2).  A step of m vehicles and P particles in all takes the first
 * entry of the process-wide table whose vehicle and particle ranges hold (m, P).  The table is the
 * JSON file written by the autotuner (dynamic-occupancy-grid-map_amd/autotune.py), loaded on first
 * use from tuning_table.json beside the library's directory or from the file DOGM_TUNING_TABLE
 * names ("none": no table); steps without a matching entry use the built-in choices.
 * ---------------------------------------------------------------------------------------------- */
enum dogm_sort_form
{
    DOGM_SORT_REDUCE_SCAN = 0,
    DOGM_SORT_ONESWEEP = 1
};

typedef struct dogm_tuning
{
    int sort;        /* dogm_sort_form */
    int merge_slice; /* 0, 1 or 2 */
    int pred_block;  /* 256 or 512 */
    int seg_items;   /* 4 or 8 */
    int lanes;       /* 1 or 2 */
} dogm_tuning;

/* Replaces the table with the one in `path` (NULL: no table). */
int dogm_tuning_load(const char* path);
/* Forces `t` for every following step in the process (NULL: back to the table). */
int dogm_tuning_override(const dogm_tuning* t);
/* The choice a step of `vehicles` vehicles and `particles` particles in all gets. */
int dogm_tuning_lookup(int vehicles, long long particles, dogm_tuning* out);
/* Entries of the loaded table. */
int dogm_tuning_entries(void);

/* ----------------------------------------------------------------------------------------------
 * Precision of the resampling CDF (the cumulative-weight scan behind resampling.cu:34-47 and the
 * new weight total / N, dogm.cu:386-423).  Particle state and weights stay f32 in every mode.
 *  DOGM_CDF_EXACT (default): exact 128-bit fixed-point prefix sums; the draws and the ancestors
 *    do not depend on any summation order.
 *  DOGM_CDF_FP64 / DOGM_CDF_FP32 / DOGM_CDF_FP32C: an f64, f32 or compensated-f32 (double-float
 *    pairs: TwoSum of the high parts, low parts added with the error, renormalised) scan with one
 *    fixed association (4-element
 *    sequential folds, 64-lane wave scans, sequential carries across waves and 1024-entry tiles),
 *    ancestor = first entry of its running maximum >= the draw.  FP32 is the reference's class of
 *    scan (thrust::inclusive_scan in f32, dogm.cu:405-407).
 * ---------------------------------------------------------------------------------------------- */
enum
{
    DOGM_CDF_EXACT = 0,
    DOGM_CDF_FP64 = 1,
    DOGM_CDF_FP32 = 2,
    DOGM_CDF_FP32C = 3
};
int dogm_set_cdf_mode(dogm_handle* h, int mode);
int dogm_get_cdf_mode(const dogm_handle* h);
/* Step graphs (opt-in): the kernel sequence of an update step is captured once into a HIP graph
 * and replayed while the vehicle list, its tables and the step's launch signature (stage set,
 * first-frame initialisation, CDF modes, extraction) are unchanged; results are identical to eager
 * launches.  Disabled while any kernel timing is on. */
int dogm_set_graph(dogm_handle* h, int enable);

/* ----------------------------------------------------------------------------------------------
 * Resampling draws.  DOGM_RESAMPLE_MULTINOMIAL (default) is the reference's: N sorted i.i.d. uniform
 * draws over the CDF (resampling.cu:17-47), generated here as exponential spacings.
 * DOGM_RESAMPLE_SYSTEMATIC (opt-in): draw k at (k + u) / N of the total with one uniform u per
 * resampling, so a particle of weight w gets floor(N w / W) or ceil(N w / W) copies; no
 * exponentials are generated.  Works with every CDF mode.
 * ---------------------------------------------------------------------------------------------- */
enum
{
    DOGM_RESAMPLE_MULTINOMIAL = 0,
    DOGM_RESAMPLE_SYSTEMATIC = 1
};
int dogm_set_resampling_mode(dogm_handle* h, int mode);
int dogm_get_resampling_mode(const dogm_handle* h);

/* ----------------------------------------------------------------------------------------------
 * Ego-motion compensation mode.  DOGM_EGO_REFERENCE (default) is the reference's updatePose
 * (dogm.cu:161-203): a move of more than one resolution shifts particles and grid by the truncated
 * cell offset with the sign convention of ego_motion_compensation.cu:16-43, the sub-cell remainder is
 * dropped and yaw is stored but never applied.
 * DOGM_EGO_METRIC (opt-in, no counterpart in the reference): the grid frame is the sensor frame (axes
 * along the vehicle at its yaw, the sensor at the grid centre); a world point w sits at cell
 * c + R(-yaw) (w - ref) / resolution.  Every pose change applies the whole-cell part of the translation
 * (the remainder carries into later frames: `ref` advances by exactly the applied distance) and the
 * yaw change as a rotation about the grid centre -- particle positions and velocities exactly, grid
 * fields by nearest-cell resampling; cells whose source lies outside the grid become 0.  For a sensor
 * whose grid does not turn with the vehicle, pass yaw 0.  Switching to it starts the metric frame at
 * the current pose.
 * ---------------------------------------------------------------------------------------------- */
enum
{
    DOGM_EGO_REFERENCE = 0,
    DOGM_EGO_METRIC = 1
};
int dogm_set_ego_mode(dogm_handle* h, int mode);
int dogm_get_ego_mode(const dogm_handle* h);

/* Philox epoch of the handle's next random stage (the draws of a resampling called now use it);
 * lets a host reference rebuild the same uniforms. */
uint32_t dogm_get_epoch(const dogm_handle* h);

/* ----------------------------------------------------------------------------------------------
 * Batched update of many independent grids ("vehicles").  No counterpart in the reference, which
 * runs one dogm::DOGM per grid, one frame per updateGrid call (dogm/src/dogm.cu:115-131).
 *
 * A batch owns vehicles of any size (each with its own dogm_params, Philox seed and instance) and
 * one HIP stream.  dogm_batch_update runs one full updateGrid frame for a list of its vehicles in
 * the same kernel launches: every stage launches once for the whole list, with a work table that
 * maps each workgroup to a vehicle and a tile of that vehicle.  The result for a vehicle is
 * bit-identical to the plain handle created with the same params, seed and instance and fed the
 * same frames (dogm_create_ex + dogm_update_grid), whatever other vehicles share its steps.
 * Vehicles keep all their state in device memory between steps.
 * ---------------------------------------------------------------------------------------------- */
typedef struct dogm_batch dogm_batch;

int dogm_batch_create(dogm_batch** out);
void dogm_batch_destroy(dogm_batch* b);
const char* dogm_batch_last_error(const dogm_batch* b);
/* Device bytes one vehicle with these params holds (for admission against a memory budget). */
int dogm_vehicle_footprint(const dogm_params* params, uint64_t* bytes);
int dogm_batch_add_vehicle(dogm_batch* b, const dogm_params* params, uint64_t seed, uint32_t instance,
                           int* vehicle_id);
int dogm_batch_remove_vehicle(dogm_batch* b, int vehicle_id);
int dogm_batch_vehicle_count(const dogm_batch* b);
uint64_t dogm_batch_device_bytes(const dogm_batch* b);
/* Device address of a vehicle's measurement grid (grid_size^2 cells): a producer on the same
 * device may write the next frame there and pass a NULL measurement pointer to dogm_batch_update. */
int dogm_batch_measurement_buffer(dogm_batch* b, int vehicle_id, dogm_measurement_cell** device_ptr);
/* One frame for `count` distinct vehicles.  Per vehicle i: measurement grid meas[i] (device memory
 * when is_device != 0; NULL, or meas itself NULL, means the grid already sits in the vehicle's
 * measurement buffer), ego pose (x[i], y[i], yaw[i]) and time step dt[i].  Blocking. */
int dogm_batch_update(dogm_batch* b, int count, const int* vehicle_ids, const dogm_measurement_cell* const* meas,
                      const float* x, const float* y, const float* yaw, const float* dt, int is_device);
int dogm_batch_get_grid_cells(dogm_batch* b, int vehicle_id, dogm_grid_cell* out);
int dogm_batch_get_particles(dogm_batch* b, int vehicle_id, dogm_vec4* state, int* grid_cell_idx, float* weight,
                             uint8_t* associated);
int dogm_batch_get_resampled_indices(dogm_batch* b, int vehicle_id, int* out);
int dogm_batch_set_cdf_mode(dogm_batch* b, int vehicle_id, int mode);
int dogm_batch_set_ego_mode(dogm_batch* b, int vehicle_id, int mode);
int dogm_batch_set_resampling_mode(dogm_batch* b, int vehicle_id, int mode);
/* Lanes of a step (1, 2, or 0 = as the tuning of each step gives: default).  With 2, a step over
 * two or more vehicles splits them over two engines with their own HIP streams, balanced by
 * per-frame traffic; both lanes' launches run concurrently on the device, which fills one lane's
 * small launches and tails with the other's work.  Results do not depend on the lane count. */
int dogm_batch_set_lanes(dogm_batch* b, int lanes);
/* Step graphs for the batch (both lanes), as dogm_set_graph. */
int dogm_batch_set_graph(dogm_batch* b, int enable);
/* Result extraction of the grid source inside every dogm_batch_update step, for all vehicles of
 * the step in the same launches; what = OR of DOGM_EXTRACT_* bits (0 = off).  The getters below
 * return the results of the vehicle's last step, or extract them on demand (with these
 * thresholds) when that step did not. */
int dogm_batch_set_extraction(dogm_batch* b, int what, float min_occupancy, float min_velocity);
int dogm_batch_get_occupancy(dogm_batch* b, int vehicle_id, float* out);
int dogm_batch_get_occupancy_image(dogm_batch* b, int vehicle_id, uint8_t* out);
int dogm_batch_get_cells_with_velocity(dogm_batch* b, int vehicle_id, dogm_velocity_cell* out, int capacity,
                                       int* count);
int dogm_batch_set_kernel_timing(dogm_batch* b, int kernel_id, int enable);
int dogm_batch_get_kernel_timing(dogm_batch* b, int kernel_id, double* total_ms, long long* launches);
int dogm_batch_reset_kernel_timing(dogm_batch* b);

/* ----------------------------------------------------------------------------------------------
 * Serving: continuous batching of frames from many vehicles.  No counterpart in the reference.
 *
 * Vehicles are admitted against a device-memory budget (their footprint, dogm_vehicle_footprint,
 * is reserved until release) and keep all filter state resident.  Frames are queued with their
 * arrival time; a step runs the oldest queued frame of every vehicle that has one (at most
 * max_batch frames, 0 = no limit) in one batched update, so frames that arrive while a step is on
 * the GPU join the next step.  A vehicle's frames run in submission order, and every result is
 * bit-identical to the vehicle's frames run alone through dogm_update_grid.
 * The measurement grid passed to dogm_server_submit must stay valid until the frame's result is
 * reported.  Drive the queue either with dogm_server_step (caller's thread and clock) or with the
 * worker thread (dogm_server_start / dogm_server_stop, stamped with dogm_server_clock).
 * ---------------------------------------------------------------------------------------------- */
typedef struct dogm_server dogm_server;

typedef struct dogm_server_config
{
    uint64_t hbm_budget_bytes; /* admission budget for resident vehicle state */
    int max_batch;             /* frames per step, 0 = every vehicle with a queued frame */
} dogm_server_config;

typedef struct dogm_frame_result
{
    uint64_t ticket;
    int vehicle_id;
    int status;     /* DOGM_OK or the step's error */
    int batch_size; /* frames in the step that ran this one */
    int pad;
    double t_arrival; /* as submitted */
    double t_start;   /* step start */
    double t_done;    /* step end (result ready) */
} dogm_frame_result;

int dogm_server_create(const dogm_server_config* cfg, dogm_server** out);
void dogm_server_destroy(dogm_server* s);
const char* dogm_server_last_error(const dogm_server* s);
double dogm_server_clock(const dogm_server* s); /* seconds since creation (steady clock) */
/* DOGM_ERR_OUT_OF_MEMORY when the vehicle's footprint does not fit the remaining budget. */
int dogm_server_admit(dogm_server* s, const dogm_params* params, uint64_t seed, uint32_t instance, int* vehicle_id);
/* Waits for the vehicle's in-flight frames when the worker runs; fails if frames are queued otherwise. */
int dogm_server_release(dogm_server* s, int vehicle_id);
uint64_t dogm_server_resident_bytes(const dogm_server* s);
int dogm_server_submit(dogm_server* s, int vehicle_id, const dogm_measurement_cell* meas, int is_device, float x,
                       float y, float yaw, float dt, double t_arrival, uint64_t* ticket);
/* A frame given as a lidar scan: at the start of its step the server generates the vehicle's
 * measurement grid from `ranges` (beam_count distances, copied at submission) with `lidar`
 * (dogm_lidar_generate_many over every scan frame of the step), then runs the update.  The
 * generator's grid_size must equal the vehicle's; the server uses it during steps, so do not call
 * it elsewhere while the server may step. */
int dogm_server_submit_scan(dogm_server* s, int vehicle_id, dogm_lidar* lidar, const double* ranges, float x,
                            float y, float yaw, float dt, double t_arrival, uint64_t* ticket);
int dogm_server_queued(const dogm_server* s);
/* One step on the caller's thread: start = now, done = now + the step's measured duration. */
int dogm_server_step(dogm_server* s, double now, int* frames_run, double* step_seconds);
int dogm_server_start(dogm_server* s);
int dogm_server_stop(dogm_server* s); /* runs the queue empty, then joins the worker */
int dogm_server_wait_idle(dogm_server* s);
int dogm_server_poll(dogm_server* s, dogm_frame_result* out, int max_results, int* count);
int dogm_server_get_grid_cells(dogm_server* s, int vehicle_id, dogm_grid_cell* out);
int dogm_server_get_particles(dogm_server* s, int vehicle_id, dogm_vec4* state, int* grid_cell_idx, float* weight,
                              uint8_t* associated);
int dogm_server_stats(const dogm_server* s, long long* steps, long long* frames);
/* Extraction inside every step (as dogm_batch_set_extraction); the getters read the vehicle's
 * latest results, compact instead of the 64-byte-per-cell dogm_server_get_grid_cells. */
int dogm_server_set_extraction(dogm_server* s, int what, float min_occupancy, float min_velocity);
int dogm_server_get_occupancy(dogm_server* s, int vehicle_id, float* out);
int dogm_server_get_occupancy_image(dogm_server* s, int vehicle_id, uint8_t* out);
int dogm_server_get_cells_with_velocity(dogm_server* s, int vehicle_id, dogm_velocity_cell* out, int capacity,
                                        int* count);

#ifdef __cplusplus
}
#endif

#endif /* DOGM_C_H */
