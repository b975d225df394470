// dogm::DOGM, drop-in for dogm/include/dogm/dogm.h of the reference (class at dogm.h:20-199).
//
// An inline C++ wrapper over the C-ABI of libdogm_hip.so (dogm_c.h); link with -ldogm_hip.  The
// public surface is the reference's: Params (same fields and order, now with defaults), the
// constructor, updateGrid, getGridCells / getMeasurementCells / getParticles, the scalar getters,
// the public stage methods and the public counts.  Differences: device buffers sit behind an
// opaque handle instead of public pointer members, errors throw std::runtime_error instead of being
// printed and ignored (cuda_utils.h:13-24), and the class is move-only.
#pragma once

#include "dogm_c.h"
#include "dogm_types.h"

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace dogm
{

class DOGM
{
public:
    struct Params
    {
        // Grid size [m]
        float size = 50.0f;
        // Grid cell size [m/cell]
        float resolution = 0.2f;
        // Number of persistent particles
        int particle_count = 300000;
        // Number of birth particles
        int new_born_particle_count = 30000;
        // Probability of persistence
        float persistence_prob = 0.99f;
        // Process noise position
        float stddev_process_noise_position = 0.1f;
        // Process noise velocity
        float stddev_process_noise_velocity = 2.0f;
        // Probability of birth
        float birth_prob = 0.02f;
        // Velocity to sample birth particles from (normal distribution) [m/s]
        float stddev_velocity = 30.0f;
        // Velocity to sample the initial particles from (uniform distribution) [m/s]
        float init_max_velocity = 15.0f;
        // Freespace discount is calculated as freespace_discount^dT
        float freespace_discount = 0.01f;
    };

    explicit DOGM(const Params& p, uint64_t seed = DOGM_DEFAULT_SEED, uint32_t instance = 0) : params(p)
    {
        const dogm_params cp = toC(p);
        check(dogm_create_ex(&cp, seed, instance, &h_), "DOGM");
        grid_size = dogm_get_grid_size(h_);
        grid_cell_count = grid_size * grid_size;
        particle_count = dogm_get_particle_count(h_);
        new_born_particle_count = dogm_get_new_born_particle_count(h_);
    }

    ~DOGM() { dogm_destroy(h_); }

    DOGM(const DOGM&) = delete;
    DOGM& operator=(const DOGM&) = delete;
    DOGM(DOGM&& o) noexcept { *this = std::move(o); }
    DOGM& operator=(DOGM&& o) noexcept
    {
        if (this != &o)
        {
            dogm_destroy(h_);
            h_ = o.h_;
            o.h_ = nullptr;
            params = o.params;
            grid_size = o.grid_size;
            grid_cell_count = o.grid_cell_count;
            particle_count = o.particle_count;
            new_born_particle_count = o.new_born_particle_count;
        }
        return *this;
    }

    /**
     * Updates the grid map and particle filter to the new timestep (dogm.h:75-83).
     * @param device whether the measurement grid resides in GPU memory (default: true).
     */
    void updateGrid(MeasurementCell* measurement_grid, float new_x, float new_y, float new_yaw, float dt,
                    bool device = true)
    {
        check(dogm_update_grid(h_, reinterpret_cast<const dogm_measurement_cell*>(measurement_grid), new_x, new_y,
                               new_yaw, dt, device ? 1 : 0),
              "updateGrid");
    }

    std::vector<GridCell> getGridCells() const
    {
        std::vector<GridCell> cells(static_cast<size_t>(grid_cell_count));
        check(dogm_get_grid_cells(h_, reinterpret_cast<dogm_grid_cell*>(cells.data())), "getGridCells");
        return cells;
    }

    std::vector<MeasurementCell> getMeasurementCells() const
    {
        std::vector<MeasurementCell> cells(static_cast<size_t>(grid_cell_count));
        check(dogm_get_measurement_cells(h_, reinterpret_cast<dogm_measurement_cell*>(cells.data())),
              "getMeasurementCells");
        return cells;
    }

    /** Host copy of the persistent particles; the caller must call .free() (dogm.cu:153-159). */
    ParticlesSoA getParticles() const
    {
        ParticlesSoA particles(particle_count, false);
        check(dogm_get_particles(h_, reinterpret_cast<dogm_vec4*>(particles.state), particles.grid_cell_idx,
                                 particles.weight, reinterpret_cast<uint8_t*>(particles.associated)),
              "getParticles");
        return particles;
    }

    int getGridSize() const { return grid_size; }
    float getResolution() const { return params.resolution; }
    float getPositionX() const { return dogm_get_position_x(h_); }
    float getYaw() const { return dogm_get_yaw(h_); }
    float getPositionY() const { return dogm_get_position_y(h_); }

    // Public stage methods (dogm.h:148-156).
    void initializeParticles() { check(dogm_initialize_particles(h_), "initializeParticles"); }
    void particlePrediction(float dt) { check(dogm_particle_prediction(h_, dt), "particlePrediction"); }
    void particleAssignment() { check(dogm_particle_assignment(h_), "particleAssignment"); }
    void gridCellOccupancyUpdate(float dt) { check(dogm_grid_cell_occupancy_update(h_, dt), "gridCellOccupancyUpdate"); }
    void updatePersistentParticles() { check(dogm_update_persistent_particles(h_), "updatePersistentParticles"); }
    void initializeNewParticles() { check(dogm_initialize_new_particles(h_), "initializeNewParticles"); }
    void statisticalMoments() { check(dogm_statistical_moments(h_), "statisticalMoments"); }
    void resampling() { check(dogm_resampling(h_), "resampling"); }

    /** Replaces the persistent particles (the reference's callers write its public device arrays). */
    void setParticles(const ParticlesSoA& p)
    {
        if (p.size != particle_count)
            throw std::invalid_argument("setParticles: size mismatch");
        check(dogm_set_particles(h_, reinterpret_cast<const dogm_vec4*>(p.state), p.grid_cell_idx, p.weight,
                                 reinterpret_cast<const uint8_t*>(p.associated)),
              "setParticles");
    }

    /** The underlying C handle, for calls that have no C++ counterpart. */
    dogm_handle* handle() const { return h_; }

public:
    Params params;
    int grid_size = 0;
    int grid_cell_count = 0;
    int particle_count = 0;
    int new_born_particle_count = 0;

private:
    static dogm_params toC(const Params& p)
    {
        return dogm_params{p.size,
                           p.resolution,
                           p.particle_count,
                           p.new_born_particle_count,
                           p.persistence_prob,
                           p.stddev_process_noise_position,
                           p.stddev_process_noise_velocity,
                           p.birth_prob,
                           p.stddev_velocity,
                           p.init_max_velocity,
                           p.freespace_discount};
    }

    void check(int status, const char* what) const
    {
        if (status != DOGM_OK)
            throw std::runtime_error(std::string(what) + ": " + dogm_status_string(status) + ": " +
                                     (h_ ? dogm_last_error(h_) : ""));
    }

    dogm_handle* h_ = nullptr;
};

}  // namespace dogm
