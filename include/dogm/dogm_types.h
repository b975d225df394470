// Data types of the DOGM filter, drop-in for dogm/include/dogm/dogm_types.h of the reference.
//
// Same layouts as the reference: GridCell (64 B, dogm_types.h:13-33), MeasurementCell (16 B, :35-41),
// Particle (:43-49) and ParticlesSoA (:51-144, one block holding state | grid_cell_idx | weight |
// associated).  glm::vec4 is replaced by an own 16-byte vec4 with the operations callers use
// (operator[], ==, +, *); no CUDA, HIP or GLM headers are needed.
#pragma once

#include "dogm_c.h"

#include <cassert>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace dogm
{

struct vec4
{
    float x, y, z, w;

    vec4() : x(0.0f), y(0.0f), z(0.0f), w(0.0f) {}
    vec4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
    explicit vec4(float s) : x(s), y(s), z(s), w(s) {}

    float& operator[](int i) { return (&x)[i]; }
    const float& operator[](int i) const { return (&x)[i]; }

    bool operator==(const vec4& o) const { return x == o.x && y == o.y && z == o.z && w == o.w; }
    bool operator!=(const vec4& o) const { return !(*this == o); }
    vec4 operator+(const vec4& o) const { return vec4(x + o.x, y + o.y, z + o.z, w + o.w); }
    vec4 operator-(const vec4& o) const { return vec4(x - o.x, y - o.y, z - o.z, w - o.w); }
    vec4 operator*(float s) const { return vec4(x * s, y * s, z * s, w * s); }
    vec4& operator+=(const vec4& o) { return *this = *this + o; }
};

inline vec4 operator*(float s, const vec4& v)
{
    return vec4(s * v.x, s * v.y, s * v.z, s * v.w);
}

struct GridCell
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
};

struct MeasurementCell
{
    float free_mass;
    float occ_mass;
    float likelihood;
    float p_A;
};

struct Particle
{
    int grid_cell_idx;
    float weight;
    bool associated;
    vec4 state;
};

static_assert(sizeof(vec4) == sizeof(dogm_vec4), "vec4 layout");
static_assert(sizeof(GridCell) == sizeof(dogm_grid_cell), "GridCell layout");
static_assert(sizeof(MeasurementCell) == sizeof(dogm_measurement_cell), "MeasurementCell layout");
static_assert(sizeof(bool) == 1, "associated flags are bytes");

// Particle arrays in one block: state | grid_cell_idx | weight | associated (dogm_types.h:137-143).
// Host blocks use malloc/free; device blocks use the library's device allocator.  As in the reference
// there is no destructor: the owner calls free().
struct ParticlesSoA
{
    vec4* state = nullptr;
    int* grid_cell_idx = nullptr;
    float* weight = nullptr;
    bool* associated = nullptr;

    void* memory_block = nullptr;
    int size = 0;
    bool device = true;

    ParticlesSoA() = default;
    ParticlesSoA(int new_size, bool is_device) { init(new_size, is_device); }
    // Shallow copy of the pointers, as the reference's implicit copy constructor.
    ParticlesSoA(const ParticlesSoA&) = default;

    void init(int new_size, bool is_device)
    {
        size = new_size;
        device = is_device;
        const size_t num_bytes = (size_t)size * sizeof(Particle);
        if (device)
        {
            if (dogm_device_malloc(&memory_block, num_bytes) != DOGM_OK)
                throw std::runtime_error("ParticlesSoA: device allocation failed");
        }
        else
        {
            memory_block = std::malloc(num_bytes);
            if (!memory_block && num_bytes)
                throw std::bad_alloc();
        }
        assignPointers();
    }

    void free()
    {
        assert(size);
        if (device)
            dogm_device_free(memory_block);
        else
            std::free(memory_block);
        memory_block = nullptr;
        size = 0;
    }

    void copy(const ParticlesSoA& other)
    {
        assert(size && size == other.size);
        int kind = 2;
        if (device != other.device)
            kind = device ? 0 : 1;
        const size_t num_bytes = (size_t)size * sizeof(Particle);
        if (!device && !other.device)
            std::memcpy(memory_block, other.memory_block, num_bytes);
        else if (dogm_memcpy(memory_block, other.memory_block, num_bytes, kind) != DOGM_OK)
            throw std::runtime_error("ParticlesSoA: copy failed");
        assignPointers();
    }

    ParticlesSoA& operator=(const ParticlesSoA& other)
    {
        if (this != &other)
            copy(other);
        return *this;
    }

    void assignPointers()
    {
        state = reinterpret_cast<vec4*>(memory_block);
        grid_cell_idx = reinterpret_cast<int*>(state + size);
        weight = reinterpret_cast<float*>(grid_cell_idx + size);
        associated = reinterpret_cast<bool*>(weight + size);
    }
};

}  // namespace dogm
