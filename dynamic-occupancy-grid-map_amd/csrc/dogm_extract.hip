// Result extraction on gfx950: the per-frame post-processing the reference's demo runs on the host
// over a full grid read-back (dogm/demo/utils/image_creation.cpp:19-66, 70-87, 112-128), done on the
// device for every vehicle of a step so that a caller reads back 4 or 1 bytes per cell, or only the
// cells that pass the velocity test.
//
// Three launches over the step's cell work table:
//   mark     one thread per cell: pignistic occupancy (f32 and / or 8-bit image) and the velocity
//            test; each wave stores its 64-lane ballot of passing cells and each 256-cell unit its
//            passing count.  The five moment fields are read only for cells whose occupancy
//            reaches the threshold, so an unoccupied cell costs the 8 bytes of occ / free.
//   scan     one workgroup per vehicle: exclusive prefix of the unit counts.
//   compact  units with passing cells write their cell indices at the unit's offset plus the
//            cell's rank among the unit's ballots, so the list is in row-major order.
// A getter then gathers the records of the listed cells (64-byte cell plus column / row).
#include "dogm_device.h"

namespace dogm_hip
{

// image_creation.cpp:19-22, f32 with the reference's association.
__device__ __forceinline__ float pignistic(float free_mass, float occ_mass)
{
    return occ_mass + 0.5f * (1.0f - occ_mass - free_mass);
}

// 255 - floor(255 p) as uint8 (image_creation.cpp:123-125; :82 truncates, the same for p in
// [0, 1]).  The reference's cast is undefined outside [0, 255]; here the floor is clamped (NaN -> 0).
__device__ __forceinline__ uint8_t occupancy_gray(float p)
{
    float g = floorf(p * 255.0f);
    g = fminf(fmaxf(g, 0.0f), 255.0f);
    return (uint8_t)(255 - (int)g);
}

// v^T S^-1 v (image_creation.cpp:35-50) as OpenCV evaluates it for CV_32F: cv::Mat::inv() of a 2x2
// matrix by its closed form (f64 determinant, f32(element * (1 / det)), zero matrix when det == 0),
// then two matrix products whose 2-term sums are accumulated in f64 and stored as f32.
__device__ __forceinline__ float mahalanobis2(float vx, float vy, float sxx, float sxy, float syy)
{
    const double det = (double)sxx * syy - (double)sxy * sxy;
    float i00 = 0.0f, i01 = 0.0f, i10 = 0.0f, i11 = 0.0f;
    if (det != 0.0)
    {
        const double r = 1.0 / det;
        i00 = (float)(syy * r);
        i11 = (float)(sxx * r);
        i01 = (float)(-sxy * r);
        i10 = (float)(-sxy * r);
    }
    const float t0 = (float)((double)vx * i00 + (double)vy * i10);
    const float t1 = (float)((double)vx * i01 + (double)vy * i11);
    return (float)((double)t0 * vx + (double)t1 * vy);
}

struct ExtractArgs
{
    int what;
    int source;
    int fsel;
    float min_occ;
    float min_vel;
};

__global__ __launch_bounds__(BLOCK) void k_extract_mark(const VehDesc* __restrict__ descs,
                                                        const uint32_t* __restrict__ tab, ExtractArgs a)
{
    __shared__ int s_cnt[BLOCK / WAVE];
    const TileRef tr = tile_ref(tab, 1, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    const int c = tr.base + threadIdx.x;
    const bool in = c < d.C;
    float occ = 0.0f, fr = 0.0f;
    if (in)
    {
        if (a.source == DOGM_SOURCE_GRID)
        {
            occ = d.f[GF_OCC][c];
            fr = d.free[a.fsel][c];
        }
        else
        {
            const float2 m = *reinterpret_cast<const float2*>(&d.meas[c].free_mass);
            fr = m.x;
            occ = m.y;
        }
    }
    const float p = pignistic(fr, occ);
    if (in && (a.what & DOGM_EXTRACT_OCCUPANCY))
        d.x_occ[c] = p;
    if (in && (a.what & DOGM_EXTRACT_IMAGE))
        d.x_img[c] = occupancy_gray(p);
    if (!(a.what & DOGM_EXTRACT_VELOCITY))
        return;
    bool pass = false;
    if (in && p >= a.min_occ)
    {
        const float m = mahalanobis2(d.f[GF_MEAN_X][c], d.f[GF_MEAN_Y][c], d.f[GF_VAR_X][c], d.f[GF_COVAR][c],
                                     d.f[GF_VAR_Y][c]);
        pass = m >= a.min_vel;
    }
    const uint64_t bal = __ballot(pass);
    const int w = threadIdx.x / WAVE;
    if ((threadIdx.x & (WAVE - 1)) == 0)
    {
        d.x_mask[(tr.base / WAVE) + w] = bal;
        s_cnt[w] = __popcll(bal);
    }
    __syncthreads();
    if (threadIdx.x == 0)
        d.x_cnt[tr.base / UNIT] = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
}

// One workgroup per vehicle: x_off = exclusive prefix of x_cnt over the vehicle's units,
// x_off[units] = total.
__global__ __launch_bounds__(BLOCK) void k_extract_scan(const VehDesc* __restrict__ descs)
{
    __shared__ int s_w[BLOCK / WAVE];
    const VehDesc& d = descs[blockIdx.x];
    const int n = d.nu[SP_C];
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = threadIdx.x / WAVE;
    int carry = 0;
    for (int k0 = 0; k0 < n; k0 += BLOCK)
    {
        const int k = k0 + threadIdx.x;
        const int v = k < n ? d.x_cnt[k] : 0;
        int x = v;
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1)
        {
            const int y = __shfl_up(x, o, WAVE);
            if (lane >= o)
                x += y;
        }
        if (lane == WAVE - 1)
            s_w[w] = x;
        __syncthreads();
        int pre = carry;
        for (int j = 0; j < w; ++j)
            pre += s_w[j];
        const int tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        if (k < n)
            d.x_off[k] = pre + x - v;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        d.x_off[n] = carry;
}

__global__ __launch_bounds__(BLOCK) void k_extract_compact(const VehDesc* __restrict__ descs,
                                                           const uint32_t* __restrict__ tab)
{
    const TileRef tr = tile_ref(tab, 1, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    const int u = tr.base / UNIT;
    if (d.x_cnt[u] == 0)
        return;
    const int w = threadIdx.x / WAVE;
    const unsigned long long* masks = d.x_mask + (size_t)u * (UNIT / WAVE);
    int pre = d.x_off[u];
    for (int j = 0; j < w; ++j)
        pre += __popcll(masks[j]);
    const uint64_t mine = masks[w];
    const int lane = threadIdx.x & (WAVE - 1);
    if ((mine >> lane) & 1ull)
        d.x_idx[pre + __popcll(mine & lanemask_lt())] = tr.base + threadIdx.x;
}

void run_extract(Engine* E, int fsel)
{
    const ExtractArgs a{E->ext.what, E->ext.source, fsel, E->ext.min_occ, E->ext.min_vel};
    launch_begin(E, K_EXTRACT_MARK);
    hipLaunchKernelGGL(k_extract_mark, dim3(E->tiles(SP_C, 1)), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_C),
                       a);
    launch_end(E, K_EXTRACT_MARK);
    if (!(a.what & DOGM_EXTRACT_VELOCITY))
        return;
    launch_begin(E, K_EXTRACT_SCAN);
    hipLaunchKernelGGL(k_extract_scan, dim3(E->m), dim3(BLOCK), 0, E->stream, E->d_desc);
    launch_end(E, K_EXTRACT_SCAN);
    launch_begin(E, K_EXTRACT_COMPACT);
    hipLaunchKernelGGL(k_extract_compact, dim3(E->tiles(SP_C, 1)), dim3(BLOCK), 0, E->stream, E->d_desc,
                       E->tab(SP_C));
    launch_end(E, K_EXTRACT_COMPACT);
}

// Records of the first n listed cells: the cell as dogm_grid_cell plus its column / row.
__global__ __launch_bounds__(BLOCK) void k_extract_gather(DevGrid G, const int* __restrict__ idx, int n, int gs,
                                                          dogm_velocity_cell* __restrict__ out)
{
    const int r = blockIdx.x * BLOCK + threadIdx.x;
    if (r >= n)
        return;
    const int c = idx[r];
    dogm_velocity_cell o;
    o.x = (float)(c % gs);
    o.y = (float)(c / gs);
    dogm_grid_cell& g = o.data;
    g.start_idx = G.start[c];
    g.end_idx = G.end[c];
    g.new_born_occ_mass = G.f[GF_NEW_BORN][c];
    g.pers_occ_mass = G.f[GF_PERS][c];
    g.free_mass = G.f[GF_FREE][c];
    g.occ_mass = G.f[GF_OCC][c];
    g.pred_occ_mass = G.f[GF_PRED][c];
    g.mu_A = G.f[GF_MU_A][c];
    g.mu_UA = G.f[GF_MU_UA][c];
    g.w_A = G.f[GF_W_A][c];
    g.w_UA = G.f[GF_W_UA][c];
    g.mean_x_vel = G.f[GF_MEAN_X][c];
    g.mean_y_vel = G.f[GF_MEAN_Y][c];
    g.var_x_vel = G.f[GF_VAR_X][c];
    g.var_y_vel = G.f[GF_VAR_Y][c];
    g.covar_xy_vel = G.f[GF_COVAR][c];
    out[r] = o;
}

void run_extract_gather(Engine* E, const Vehicle* v, int n)
{
    if (n <= 0)
        return;
    launch_begin(E, K_EXTRACT_GATHER);
    hipLaunchKernelGGL(k_extract_gather, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, E->stream, v->grid, v->x_idx, n,
                       v->gs, v->x_rec);
    launch_end(E, K_EXTRACT_GATHER);
}

}  // namespace dogm_hip
