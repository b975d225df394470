// Measurement-grid generation on gfx950: the step before updateGrid.
//
// The reference turns a lidar scan into a measurement grid on the GPU through a polar inverse-sensor
// texture rendered into a Cartesian framebuffer with OpenGL (dogm/demo/simulator/mapping/
// laser_to_meas_grid.cu:25-70, kernel/measurement_grid.cu:46-132).  This generator ray-casts each
// beam straight into the Cartesian grid with a two-level inverse sensor model:
//   - every cell a beam samples (half-cell steps from the sensor, strictly before the return or
//     max_range) gets free mass `free_mass`;
//   - the cell of each return gets occupied mass `occ_mass` and no free mass;
//   - all other cells are unknown (0, 0); likelihood = p_A = 1 (measurement_grid.cu:129-130).
// The sensor sits at the grid centre; index = x + grid_size * y.  Positions are computed in f64 with
// one fixed operation order (centre + dir * dist, then / resolution, then floor), so the grid is a
// pure function of the ranges and the beam directions.
//
// Three launches per call, each covering every grid of the call (blockIdx.y = grid):
//   clear  16-byte stores of (0, 0, 1, 1) over the cells;
//   free   one thread per (beam, step) sample: an idempotent 4-byte store of free_mass;
//   occ    one thread per beam: occ_mass and free 0 at the return's cell (after every free store).
#include "dogm_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

struct dogm_lidar
{
    dogm_lidar_params p{};
    int steps = 0;          // samples per beam
    double step = 0.0;      // sample spacing [m]
    double centre = 0.0;    // sensor position in grid metres (both axes)
    hipStream_t stream = nullptr;
    double* d_dir = nullptr;       // [2 * beams]: cos, then sin
    double* d_ranges = nullptr;    // [cap * beams]
    double* h_ranges = nullptr;    // pinned staging
    dogm_measurement_cell** d_out = nullptr;  // [cap]
    dogm_measurement_cell** h_out = nullptr;  // pinned staging
    int cap = 0;
    std::string last_error;
};

namespace dogm_hip
{

struct LidarArgs
{
    int gs;
    int beams;
    int steps;
    double res;
    double step;
    double centre;
    double max_range;
    float occ_mass;
    float free_mass;
};

__global__ __launch_bounds__(BLOCK) void k_lidar_clear(dogm_measurement_cell* const* __restrict__ outs, int cells)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= cells)
        return;
    *reinterpret_cast<float4*>(outs[blockIdx.y] + c) = make_float4(0.0f, 0.0f, 1.0f, 1.0f);
}

// End of the free samples: min(range, max_range); a NaN range has no free samples.
__device__ __forceinline__ double lidar_limit(double r, double max_range)
{
    return (r < max_range || r != r) ? r : max_range;
}

__global__ __launch_bounds__(BLOCK) void k_lidar_free(dogm_measurement_cell* const* __restrict__ outs,
                                                      const double* __restrict__ ranges,
                                                      const double* __restrict__ dir, LidarArgs a)
{
    const long long t = (long long)blockIdx.x * BLOCK + threadIdx.x;
    const long long n = (long long)a.beams * a.steps;
    if (t >= n)
        return;
    const int b = (int)(t / a.steps);
    const int k = (int)(t - (long long)b * a.steps);
    const double lim = lidar_limit(ranges[(size_t)blockIdx.y * a.beams + b], a.max_range);
    const double dist = (double)k * a.step;
    if (!(dist < lim))
        return;
    const double xs = a.centre + dir[b] * dist;
    const double ys = a.centre + dir[a.beams + b] * dist;
    const double fx = floor(xs / a.res), fy = floor(ys / a.res);
    if (!(fx >= 0.0 && fx < (double)a.gs && fy >= 0.0 && fy < (double)a.gs))
        return;
    const int c = (int)fx + a.gs * (int)fy;
    outs[blockIdx.y][c].free_mass = a.free_mass;
}

__global__ __launch_bounds__(BLOCK) void k_lidar_occ(dogm_measurement_cell* const* __restrict__ outs,
                                                     const double* __restrict__ ranges,
                                                     const double* __restrict__ dir, LidarArgs a)
{
    const int b = blockIdx.x * BLOCK + threadIdx.x;
    if (b >= a.beams)
        return;
    const double r = ranges[(size_t)blockIdx.y * a.beams + b];
    if (!(r <= a.max_range))
        return;  // no return within range (inf, NaN or beyond max_range)
    const double fx = floor((a.centre + dir[b] * r) / a.res);
    const double fy = floor((a.centre + dir[a.beams + b] * r) / a.res);
    if (!(fx >= 0.0 && fx < (double)a.gs && fy >= 0.0 && fy < (double)a.gs))
        return;
    const int c = (int)fx + a.gs * (int)fy;
    *reinterpret_cast<float2*>(&outs[blockIdx.y][c].free_mass) = make_float2(0.0f, a.occ_mass);
}

int lidar_grid_size(const dogm_lidar* l)
{
    return l->p.grid_size;
}

int lidar_beam_count(const dogm_lidar* l)
{
    return l->p.beam_count;
}

}  // namespace dogm_hip

using namespace dogm_hip;

namespace
{

int lfail(dogm_lidar* l, int code, const std::string& msg)
{
    if (l)
        l->last_error = msg;
    return code;
}

void lidar_release(dogm_lidar* l)
{
    if (l->stream)
        (void)hipStreamSynchronize(l->stream);
    (void)hipFree(l->d_dir);
    (void)hipFree(l->d_ranges);
    (void)hipFree(l->d_out);
    if (l->h_ranges)
        (void)hipHostFree(l->h_ranges);
    if (l->h_out)
        (void)hipHostFree(l->h_out);
    if (l->stream)
        (void)hipStreamDestroy(l->stream);
}

int lidar_reserve(dogm_lidar* l, int count)
{
    if (count <= l->cap)
        return DOGM_OK;
    const int cap = std::max(count, 2 * l->cap);
    (void)hipStreamSynchronize(l->stream);
    (void)hipFree(l->d_ranges);
    (void)hipFree(l->d_out);
    if (l->h_ranges)
        (void)hipHostFree(l->h_ranges);
    if (l->h_out)
        (void)hipHostFree(l->h_out);
    l->d_ranges = nullptr;
    l->d_out = nullptr;
    l->h_ranges = nullptr;
    l->h_out = nullptr;
    l->cap = 0;
    const size_t nr = (size_t)cap * (size_t)l->p.beam_count;
    if (hipMalloc(reinterpret_cast<void**>(&l->d_ranges), sizeof(double) * nr) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&l->d_out), sizeof(void*) * (size_t)cap) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&l->h_ranges), sizeof(double) * nr) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&l->h_out), sizeof(void*) * (size_t)cap) != hipSuccess)
        return lfail(l, DOGM_ERR_OUT_OF_MEMORY, "allocating lidar staging buffers");
    l->cap = cap;
    return DOGM_OK;
}

}  // namespace

extern "C" {

int dogm_lidar_create(const dogm_lidar_params* p, const double* beam_cos, const double* beam_sin, dogm_lidar** out)
{
    if (!p || !out)
        return DOGM_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (p->grid_size < 1 || p->grid_size > 46340 || p->beam_count < 1 || !(p->resolution > 0.0) ||
        !(p->max_range > 0.0) || !std::isfinite(p->max_range) || !std::isfinite(p->resolution))
        return DOGM_ERR_INVALID_ARGUMENT;
    if ((beam_cos == nullptr) != (beam_sin == nullptr))
        return DOGM_ERR_INVALID_ARGUMENT;
    dogm_lidar* l = new (std::nothrow) dogm_lidar();
    if (!l)
        return DOGM_ERR_OUT_OF_MEMORY;
    l->p = *p;
    l->step = 0.5 * p->resolution;
    l->steps = (int)std::ceil(p->max_range / l->step) + 1;
    l->centre = 0.5 * p->grid_size * p->resolution;
    if ((long long)l->steps * p->beam_count > (long long)INT32_MAX * 64)
    {
        delete l;
        return DOGM_ERR_INVALID_ARGUMENT;
    }
    std::vector<double> dir(2 * (size_t)p->beam_count);
    for (int b = 0; b < p->beam_count; ++b)
    {
        if (beam_cos)
        {
            dir[b] = beam_cos[b];
            dir[p->beam_count + b] = beam_sin[b];
        }
        else
        {
            // 360 degrees, beam b at 2 pi b / beam_count.
            const double ang = 2.0 * M_PI * (double)b / (double)p->beam_count;
            dir[b] = std::cos(ang);
            dir[p->beam_count + b] = std::sin(ang);
        }
    }
    int st = DOGM_OK;
    if (hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking) != hipSuccess)
        st = DOGM_ERR_HIP;
    else if (hipMalloc(reinterpret_cast<void**>(&l->d_dir), sizeof(double) * dir.size()) != hipSuccess)
        st = DOGM_ERR_OUT_OF_MEMORY;
    else if (hipMemcpy(l->d_dir, dir.data(), sizeof(double) * dir.size(), hipMemcpyHostToDevice) != hipSuccess)
        st = DOGM_ERR_HIP;
    else
        st = lidar_reserve(l, 1);
    if (st != DOGM_OK)
    {
        lidar_release(l);
        delete l;
        return st;
    }
    *out = l;
    return DOGM_OK;
}

void dogm_lidar_destroy(dogm_lidar* l)
{
    if (!l)
        return;
    lidar_release(l);
    delete l;
}

const char* dogm_lidar_last_error(const dogm_lidar* l)
{
    return l ? l->last_error.c_str() : "null lidar";
}

int dogm_lidar_generate_many(dogm_lidar* l, int count, const double* ranges, dogm_measurement_cell* const* device_out)
{
    if (!l)
        return DOGM_ERR_INVALID_ARGUMENT;
    if (count < 1 || count > 65535 || !ranges || !device_out)
        return lfail(l, DOGM_ERR_INVALID_ARGUMENT, "count must be in [1, 65535] with ranges and outputs");
    for (int i = 0; i < count; ++i)
        if (!device_out[i])
            return lfail(l, DOGM_ERR_INVALID_ARGUMENT, "output grid " + std::to_string(i) + " is null");
    int st = lidar_reserve(l, count);
    if (st != DOGM_OK)
        return st;
    const int beams = l->p.beam_count;
    const int gs = l->p.grid_size;
    std::memcpy(l->h_ranges, ranges, sizeof(double) * (size_t)count * beams);
    std::memcpy(l->h_out, device_out, sizeof(void*) * (size_t)count);
    if (hipMemcpyAsync(l->d_ranges, l->h_ranges, sizeof(double) * (size_t)count * beams, hipMemcpyHostToDevice,
                       l->stream) != hipSuccess ||
        hipMemcpyAsync(l->d_out, l->h_out, sizeof(void*) * (size_t)count, hipMemcpyHostToDevice, l->stream) !=
            hipSuccess)
        return lfail(l, DOGM_ERR_HIP, "uploading ranges");
    const LidarArgs a{gs, beams, l->steps, l->p.resolution, l->step, l->centre, l->p.max_range, l->p.occ_mass,
                      l->p.free_mass};
    const int cells = gs * gs;
    const long long samples = (long long)beams * l->steps;
    hipLaunchKernelGGL(k_lidar_clear, dim3(div_up(cells, BLOCK), count), dim3(BLOCK), 0, l->stream, l->d_out, cells);
    hipLaunchKernelGGL(k_lidar_free, dim3(div_up(samples, BLOCK), count), dim3(BLOCK), 0, l->stream, l->d_out,
                       l->d_ranges, l->d_dir, a);
    hipLaunchKernelGGL(k_lidar_occ, dim3(div_up(beams, BLOCK), count), dim3(BLOCK), 0, l->stream, l->d_out,
                       l->d_ranges, l->d_dir, a);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipStreamSynchronize(l->stream);
    if (e != hipSuccess)
        return lfail(l, DOGM_ERR_HIP, std::string("generating measurement grids: ") + hipGetErrorString(e));
    return DOGM_OK;
}

int dogm_lidar_generate(dogm_lidar* l, const double* ranges, dogm_measurement_cell* device_out)
{
    dogm_measurement_cell* outs[1] = {device_out};
    return dogm_lidar_generate_many(l, 1, ranges, outs);
}

}  // extern "C"
