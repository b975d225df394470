// Result extraction on gfx950: the per-frame post-processing the reference's demo runs on the host
// over a full grid read-back (dogm/demo/utils/image_creation.cpp:19-66, 70-87, 112-128), done on the
// device for every vehicle of a step so that a caller reads back 4 or 1 bytes per cell, or only the
// cells that pass the velocity test.
//
// Three launches over the step's cell work table:
//   mark     one wave per 256-cell unit, four consecutive cells per lane (16-byte loads of occ /
//            free, 16-byte occupancy and 4-byte image stores): pignistic occupancy and the velocity
//            test; each unit stores its four 64-lane ballots of passing cells and its passing
//            count.  The five moment fields are read only for cells whose occupancy reaches the
//            threshold, so an unoccupied cell costs the 8 bytes of occ / free.
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

constexpr int XCELLS = 4;                         // consecutive cells per lane
constexpr int XUNITS = BLOCK / WAVE;              // units per mark workgroup (one wave per unit)
static_assert(WAVE * XCELLS == UNIT, "one wave covers one unit");

// Mark: wave w of the workgroup covers unit (tile base / UNIT + w), lane l its cells 4l .. 4l + 3.
// The unit's four ballots are stored as masks j = 0..3 (bit l of mask j <-> cell 4l + j) with the
// unit's passing count.
__global__ __launch_bounds__(BLOCK) void k_extract_mark(const VehDesc* __restrict__ descs,
                                                        const uint32_t* __restrict__ tab, ExtractArgs a)
{
    const TileRef tr = tile_ref(tab, XUNITS, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    const int w = threadIdx.x / WAVE;
    const int lane = threadIdx.x & (WAVE - 1);
    const int u = tr.base / UNIT + w;
    const int c0 = u * UNIT + lane * XCELLS;
    float occ[XCELLS], fr[XCELLS];
    if (a.source == DOGM_SOURCE_GRID)
    {
        if (c0 + XCELLS <= d.C)
        {
            const float4 o = *reinterpret_cast<const float4*>(d.f[GF_OCC] + c0);
            const float4 f = *reinterpret_cast<const float4*>(d.free[a.fsel] + c0);
            occ[0] = o.x, occ[1] = o.y, occ[2] = o.z, occ[3] = o.w;
            fr[0] = f.x, fr[1] = f.y, fr[2] = f.z, fr[3] = f.w;
        }
        else
        {
#pragma unroll
            for (int j = 0; j < XCELLS; ++j)
            {
                const bool in = c0 + j < d.C;
                occ[j] = in ? d.f[GF_OCC][c0 + j] : 0.0f;
                fr[j] = in ? d.free[a.fsel][c0 + j] : 0.0f;
            }
        }
    }
    else
    {
#pragma unroll
        for (int j = 0; j < XCELLS; ++j)
        {
            float2 m = make_float2(0.0f, 0.0f);
            if (c0 + j < d.C)
                m = *reinterpret_cast<const float2*>(&d.meas[c0 + j].free_mass);
            fr[j] = m.x;
            occ[j] = m.y;
        }
    }
    float p[XCELLS];
#pragma unroll
    for (int j = 0; j < XCELLS; ++j)
        p[j] = pignistic(fr[j], occ[j]);
    const bool full = c0 + XCELLS <= d.C;
    if (a.what & DOGM_EXTRACT_OCCUPANCY)
    {
        if (full)
            *reinterpret_cast<float4*>(d.x_occ + c0) = make_float4(p[0], p[1], p[2], p[3]);
        else
            for (int j = 0; j < XCELLS; ++j)
                if (c0 + j < d.C)
                    d.x_occ[c0 + j] = p[j];
    }
    if (a.what & DOGM_EXTRACT_IMAGE)
    {
        if (full)
        {
            const uint32_t g = (uint32_t)occupancy_gray(p[0]) | ((uint32_t)occupancy_gray(p[1]) << 8) |
                               ((uint32_t)occupancy_gray(p[2]) << 16) | ((uint32_t)occupancy_gray(p[3]) << 24);
            *reinterpret_cast<uint32_t*>(d.x_img + c0) = g;
        }
        else
        {
            for (int j = 0; j < XCELLS; ++j)
                if (c0 + j < d.C)
                    d.x_img[c0 + j] = occupancy_gray(p[j]);
        }
    }
    if (!(a.what & DOGM_EXTRACT_VELOCITY))
        return;
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < XCELLS; ++j)
    {
        const int c = c0 + j;
        bool pass = false;
        if (c < d.C && p[j] >= a.min_occ)
        {
            const float m = mahalanobis2(d.f[GF_MEAN_X][c], d.f[GF_MEAN_Y][c], d.f[GF_VAR_X][c], d.f[GF_COVAR][c],
                                         d.f[GF_VAR_Y][c]);
            pass = m >= a.min_vel;
        }
        const uint64_t bal = __ballot(pass);
        cnt += __popcll(bal);
        if (lane == 0)
            d.x_mask[(size_t)u * XCELLS + j] = bal;
    }
    if (lane == 0)
        d.x_cnt[u] = cnt;
}

constexpr int XSCAN_BLOCK = 1024;
constexpr int XSCAN_ITEMS = UNIT_ALIGN;  // consecutive units per thread (a vehicle's units come in 8s)

// One workgroup per vehicle: x_off = exclusive prefix of x_cnt over the vehicle's units,
// x_off[units] = total.  Each pass covers XSCAN_BLOCK * 8 units, eight consecutive per thread.
__global__ __launch_bounds__(XSCAN_BLOCK) void k_extract_scan(const VehDesc* __restrict__ descs)
{
    __shared__ int s_w[XSCAN_BLOCK / WAVE];
    const VehDesc& d = descs[blockIdx.x];
    const int n = d.nu[SP_C];
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = threadIdx.x / WAVE;
    int carry = 0;
    for (int k0 = 0; k0 < n; k0 += XSCAN_BLOCK * XSCAN_ITEMS)
    {
        const int k = k0 + threadIdx.x * XSCAN_ITEMS;
        int v[XSCAN_ITEMS] = {};
        if (k < n)
        {
            const int4 a = *reinterpret_cast<const int4*>(d.x_cnt + k);
            const int4 b = *reinterpret_cast<const int4*>(d.x_cnt + k + 4);
            v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
        }
        int t = 0;
#pragma unroll
        for (int j = 0; j < XSCAN_ITEMS; ++j)
            t += v[j];
        int x = t;
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
        int pre = carry, tot = 0;
        for (int j = 0; j < XSCAN_BLOCK / WAVE; ++j)
        {
            const int s_j = s_w[j];
            pre += j < w ? s_j : 0;
            tot += s_j;
        }
        if (k < n)
        {
            int e[XSCAN_ITEMS];
            int run = pre + x - t;
#pragma unroll
            for (int j = 0; j < XSCAN_ITEMS; ++j)
            {
                e[j] = run;
                run += v[j];
            }
            *reinterpret_cast<int4*>(d.x_off + k) = make_int4(e[0], e[1], e[2], e[3]);
            *reinterpret_cast<int4*>(d.x_off + k + 4) = make_int4(e[4], e[5], e[6], e[7]);
        }
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        d.x_off[n] = carry;
}

// Compact: one wave per unit (two units per wave, eight per workgroup); units without passing
// cells are skipped.  Cell 4l + j of the unit has rank
//   sum_j' popc(mask_j' & lanes below l) + sum_{j' < j} bit l of mask_j'.
__global__ __launch_bounds__(BLOCK) void k_extract_compact(const VehDesc* __restrict__ descs,
                                                           const uint32_t* __restrict__ tab)
{
    const TileRef tr = tile_ref(tab, UNIT_ALIGN, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    const int w = threadIdx.x / WAVE;
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t below = lanemask_lt();
    for (int k = 0; k < UNIT_ALIGN / XUNITS; ++k)
    {
        const int u = tr.base / UNIT + k * XUNITS + w;
        if (d.x_cnt[u] == 0)
            continue;
        uint64_t m[XCELLS];
        int r = d.x_off[u];
#pragma unroll
        for (int j = 0; j < XCELLS; ++j)
        {
            m[j] = d.x_mask[(size_t)u * XCELLS + j];
            r += __popcll(m[j] & below);
        }
#pragma unroll
        for (int j = 0; j < XCELLS; ++j)
        {
            if ((m[j] >> lane) & 1ull)
                d.x_idx[r++] = u * UNIT + lane * XCELLS + j;
        }
    }
}

void run_extract(Engine* E, int fsel)
{
    const ExtractArgs a{E->ext.what, E->ext.source, fsel, E->ext.min_occ, E->ext.min_vel};
    launch_begin(E, K_EXTRACT_MARK);
    hipLaunchKernelGGL(k_extract_mark, dim3(E->tiles(SP_C, XUNITS)), dim3(BLOCK), 0, E->stream, E->d_desc,
                       E->tab(SP_C), a);
    launch_end(E, K_EXTRACT_MARK);
    if (!(a.what & DOGM_EXTRACT_VELOCITY))
        return;
    launch_begin(E, K_EXTRACT_SCAN);
    hipLaunchKernelGGL(k_extract_scan, dim3(E->m), dim3(XSCAN_BLOCK), 0, E->stream, E->d_desc);
    launch_end(E, K_EXTRACT_SCAN);
    launch_begin(E, K_EXTRACT_COMPACT);
    hipLaunchKernelGGL(k_extract_compact, dim3(E->tiles(SP_C, UNIT_ALIGN)), dim3(BLOCK), 0, E->stream, E->d_desc,
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
