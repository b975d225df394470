// Stage kernels of the DOGM particle-filter update on gfx950.
//
// Layout in HBM, per vehicle: particles structure-of-arrays (float4 state, int cell index, float
// weight, uint8 association), grid cells structure-of-arrays (one array per field), measurement
// cells as given by the caller (16-byte AoS).  Every per-cell sum is an exact 128-bit fixed-point
// segmented reduction over the cell-sorted particles (dogm_device.h).  Step kernels find their
// vehicle and local tile through the step's work tables (dogm_internal.h).
#include "dogm_device.h"

namespace dogm_hip
{

// ---------------------------------------------------------------------------------------------
// Utility kernels
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_fill_u32(unsigned* __restrict__ p, unsigned v, long long n)
{
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n)
        p[i] = v;
}

void run_fill_u32(Engine* E, unsigned* p, unsigned v, long long n)
{
    if (n <= 0)
        return;
    launch_begin(E, K_FILL);
    hipLaunchKernelGGL(k_fill_u32, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, E->stream, p, v, n);
    launch_end(E, K_FILL);
}

struct TabArgs
{
    long long units[SP_COUNT];
};

// Expands the per-space unit prefixes of the step's vehicles (upre[s * (m + 1) + v]) into the
// work tables: entry i of space s = vehicle << 20 | local unit.
__global__ __launch_bounds__(BLOCK) void k_build_tables(const int* __restrict__ upre, int m, TabArgs a,
                                                        uint32_t* __restrict__ tab, long long total)
{
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= total)
        return;
    int s = 0;
    long long u = i;
    while (s < SP_COUNT - 1 && u >= a.units[s])
    {
        u -= a.units[s];
        ++s;
    }
    const int* pre = upre + s * (m + 1);
    // last v in [0, m) with pre[v] <= u
    int lo = 0, count = m;
    while (count > 0)
    {
        const int step = count / 2;
        if (pre[lo + step + 1] <= u)
        {
            lo += step + 1;
            count -= step + 1;
        }
        else
        {
            count = step;
        }
    }
    tab[i] = ((uint32_t)lo << TAB_VEH_SHIFT) | (uint32_t)(u - pre[lo]);
}

void run_build_tables(Engine* E)
{
    TabArgs a{};
    long long total = 0;
    for (int s = 0; s < SP_COUNT; ++s)
    {
        a.units[s] = E->units[s];
        total += E->units[s];
    }
    if (total <= 0)
        return;
    launch_begin(E, K_BUILD_TABLES);
    hipLaunchKernelGGL(k_build_tables, dim3(div_up(total, BLOCK)), dim3(BLOCK), 0, E->stream, E->d_upre, E->m, a,
                       E->d_tab, total);
    launch_end(E, K_BUILD_TABLES);
}

// ---------------------------------------------------------------------------------------------
// Ego-motion compensation (dogm/src/kernel/ego_motion_compensation.cu:16-43), eager form for the
// single-vehicle pose update.  Inside an update step the shift is fused into predict and the
// occupancy update.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_move_particles(DevParticles P, int n, float mx, float my)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n)
        return;
    float2 s = P.pos[i];
    s.x = s.x - mx;
    s.y = s.y - my;
    P.pos[i] = s;
}

void run_move_particles(Engine* E, const Vehicle* v, float mx, float my)
{
    launch_begin(E, K_MOVE_PARTICLES);
    hipLaunchKernelGGL(k_move_particles, dim3(div_up(v->N, BLOCK)), dim3(BLOCK), 0, E->stream, v->part[0], v->N, mx,
                       my);
    launch_end(E, K_MOVE_PARTICLES);
}

// Metric ego compensation of one particle: position rotated by -dpsi about the grid centre and
// translated by whole cells, velocity rotated (a pure translation keeps positions exact).
__device__ __forceinline__ float4 ego_particle(float4 s, const EgoXform& t)
{
    if (t.sn == 0.0f && t.cs == 1.0f)
        return make_float4(s.x + t.tx, s.y + t.ty, s.z, s.w);
    const float dx = s.x - t.c, dy = s.y - t.c;
    const float rx = t.cs * dx + t.sn * dy;
    const float ry = t.cs * dy - t.sn * dx;
    const float vx = t.cs * s.z + t.sn * s.w;
    const float vy = t.cs * s.w - t.sn * s.z;
    return make_float4((t.c + rx) + t.tx, (t.c + ry) + t.ty, vx, vy);
}

// Source cell of destination cell c under the transform, by nearest cell: the cell containing
// c + R(dpsi) ((x + 0.5, y + 0.5) - c - t); -1 outside the grid.
__device__ __forceinline__ int ego_source_cell(int c, int gs, const EgoXform& t)
{
    const float qx = ((float)(c % gs) + 0.5f - t.c) - t.tx;
    const float qy = ((float)(c / gs) + 0.5f - t.c) - t.ty;
    const float sx = floorf(t.c + (t.cs * qx - t.sn * qy));
    const float sy = floorf(t.c + (t.sn * qx + t.cs * qy));
    const bool in = sx >= 0.0f && sx < (float)gs && sy >= 0.0f && sy < (float)gs;
    return in ? (int)sx + gs * (int)sy : -1;
}

__global__ __launch_bounds__(BLOCK) void k_ego_particles(DevParticles P, int n, EgoXform t)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n)
        return;
    const float2 p = P.pos[i], v = P.vel[i];
    const float4 s = ego_particle(make_float4(p.x, p.y, v.x, v.y), t);
    P.pos[i] = make_float2(s.x, s.y);
    P.vel[i] = make_float2(s.z, s.w);
}

void run_ego_particles(Engine* E, const Vehicle* v, const EgoXform& t)
{
    if (v->N <= 0)
        return;
    launch_begin(E, K_EGO_PARTICLES);
    hipLaunchKernelGGL(k_ego_particles, dim3(div_up(v->N, BLOCK)), dim3(BLOCK), 0, E->stream, v->part[0], v->N, t);
    launch_end(E, K_EGO_PARTICLES);
}

// The same transform on a grid field (nearest source cell; sources outside the grid give 0).
__global__ __launch_bounds__(BLOCK) void k_ego_field(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                     int gs, EgoXform t)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= gs * gs)
        return;
    const int s = ego_source_cell(c, gs, t);
    dst[c] = s >= 0 ? src[s] : 0u;
}

void run_ego_field(Engine* E, const Vehicle* v, const void* src, void* dst, const EgoXform& t)
{
    launch_begin(E, K_EGO_FIELD);
    hipLaunchKernelGGL(k_ego_field, dim3(div_up(v->C, BLOCK)), dim3(BLOCK), 0, E->stream,
                       static_cast<const uint32_t*>(src), static_cast<uint32_t*>(dst), v->gs, t);
    launch_end(E, K_EGO_FIELD);
}

// new[x, y] = old[x + ix, y + iy] when 0 < x + ix < gs and 0 < y + iy < gs (strict lower bound,
// ego_motion_compensation.cu:38), else 0 (the reference memsets the grid first, dogm.cu:185).
__global__ __launch_bounds__(BLOCK) void k_shift_field(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                       int gs, int ix, int iy)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= gs * gs)
        return;
    const int x = c % gs, y = c / gs;
    const int sx = x + ix, sy = y + iy;
    const bool in = sx > 0 && sx < gs && sy > 0 && sy < gs;
    dst[c] = in ? src[sx + gs * sy] : 0u;
}

void run_shift_field(Engine* E, const Vehicle* v, const void* src, void* dst, int ix, int iy)
{
    launch_begin(E, K_SHIFT_FIELD);
    hipLaunchKernelGGL(k_shift_field, dim3(div_up(v->C, BLOCK)), dim3(BLOCK), 0, E->stream,
                       static_cast<const uint32_t*>(src), static_cast<uint32_t*>(dst), v->gs, ix, iy);
    launch_end(E, K_SHIFT_FIELD);
}

// ---------------------------------------------------------------------------------------------
// Prediction (dogm/src/kernel/predict.cu:16-52), with the pending ego shift of the step fused in
// (moveParticlesKernel runs right before it in the reference).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ int clamp_cell(float x, int gs)
{
    if (x >= (float)(gs - 1))
        return gs - 1;
    if (x > 0.0f)
        return (int)x;
    return 0;
}

// A workgroup covers one sort tile: PB threads of SORT_TILE / PB particles each (strided by PB).
// 512 x 4 (the built-in choice) gives a single 2e6-particle grid about 8 waves per SIMD to hide the
// Philox / Box-Muller latency; 256 x 8 is the other choice of the tuning.

// With passes > 0 the workgroup also counts the radix digits of the new cell indices for every
// sort pass into ghist[(vehicle * 4 + pass) * 256 + digit] (the sort's one-time histogram) and
// clears the sort passes' tile tickets (the passes run after predict on the same stream); with
// thist set (reduce-then-scan sort) it stores the first pass's digit counts of its tile as
// thist[digit * tiles + tile]: a workgroup covers exactly one sort tile, so the sort skips that
// upsweep.  With aos (the sort's gathering pass follows in the same step) the new state goes to the
// vehicle's CDF buffer, free until the resampling, as 16-byte (x, y, vx, vy) records, so that the
// gather reads one 16-byte record per particle at random instead of two 8-byte ones from two arrays;
// nothing else reads the positions or velocities in between.
// start = end = -1 for the cells of work-table group `group` (THREADS x 4 cells), four cells per
// thread with 16-byte stores.  With valid occupancy bits only groups of four containing a cell that
// had particles are written (the others still hold -1, and rewriting -1 is harmless).
template <int THREADS>
__device__ __forceinline__ void reset_bounds_cells(const VehDesc* __restrict__ descs, const uint32_t* __restrict__ tab,
                                                   int group)
{
    const TileRef tr = tile_ref(tab, THREADS * 4 / UNIT, group);
    const VehDesc& d = descs[tr.v];
    const int c = tr.base + 4 * threadIdx.x;
    if (c >= d.C)
        return;
    if (d.nz_valid && ((d.nz[c >> 6] >> (c & 63)) & 0xfull) == 0)
        return;
    if (c + 4 <= d.C)
    {
        const int4 m1 = make_int4(-1, -1, -1, -1);
        *reinterpret_cast<int4*>(d.start + c) = m1;
        *reinterpret_cast<int4*>(d.end + c) = m1;
    }
    else
    {
        for (int k = c; k < d.C; ++k)
        {
            d.start[k] = -1;
            d.end[k] = -1;
        }
    }
}

template <int PRED_BLOCK>
__global__ __launch_bounds__(PRED_BLOCK) void k_predict(const VehDesc* __restrict__ descs, const uint32_t* __restrict__ tab,
                                                   int pph, int passes, unsigned* __restrict__ ghist,
                                                   unsigned* __restrict__ tickets, unsigned* __restrict__ thist,
                                                   int tiles, int aos, const uint32_t* __restrict__ tab_c)
{
    constexpr int PRED_ITEMS = SORT_TILE / PRED_BLOCK;
    static_assert(PRED_ITEMS * PRED_BLOCK == SORT_TILE, "predict workgroups are the sort's tiles");
    if ((int)blockIdx.x >= tiles)
    {
        // The assignment's bounds reset, in workgroups after the particle tiles (memory work beside
        // the ALU-bound prediction).
        reset_bounds_cells<PRED_BLOCK>(descs, tab_c, (int)blockIdx.x - tiles);
        return;
    }
    __shared__ unsigned s_h[4][RADIX];
    if (passes > 0 && blockIdx.x == 0 && threadIdx.x < 4)
        tickets[threadIdx.x] = 0u;
    const TileRef tr = tile_ref(tab, SORT_Q, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    const int n = d.N;
    if (tr.base >= n)
    {
        if (passes > 0 && thist)
            for (int i = threadIdx.x; i < RADIX; i += PRED_BLOCK)
                thist[(size_t)i * tiles + blockIdx.x] = 0u;
        return;
    }
    if (passes > 0)
    {
        for (int i = threadIdx.x; i < 4 * RADIX; i += PRED_BLOCK)
            (&s_h[0][0])[i] = 0u;
        __syncthreads();
    }
    const DevParticles P = d.part[pph];
    const DevParticles Q = d.part[pph ^ 1];  // sources of a deferred resampled set
    const DevParticles Bq = d.birth;
    const int gs = d.gs;
    const float dt = d.dt;
    const float zero = 0.0f, one = 1.0f;
    for (int it = 0; it < PRED_ITEMS; ++it)
    {
        const int i = tr.base + it * PRED_BLOCK + threadIdx.x;
        const bool valid = i < n;
        int cell = 0;
        if (valid)
        {
            float2 ps, vs;
            if (d.rs_gather)
            {
                // A deferred resampled set (see merge_write): the state of the particle's ancestor
                // (part[pph ^ 1] below N, the birth set above), and its association flag written here
                // for the sort.
                const int a = d.ancestors[i];
                const bool from_p = a < n;
                const int o = from_p ? a : a - n;
                ps = (from_p ? Q.pos : Bq.pos)[o];
                vs = (from_p ? Q.vel : Bq.vel)[o];
                P.assoc[i] = (from_p ? Q.assoc : Bq.assoc)[o];
            }
            else
            {
                ps = P.pos[i];
                vs = P.vel[i];
            }
            float4 s = make_float4(ps.x, ps.y, vs.x, vs.y);
            // The gathering pass forms the weight (always so for a deferred set: the host reads one
            // through its ancestors only in full steps, whose weights are uniform).
            const bool keep_w = aos && d.wuni;
            const float w_in = keep_w ? 0.0f : P.weight[i];
            if (d.move)
            {
                s.x = s.x - d.mx;
                s.y = s.y - d.my;
            }
            else if (d.xform_p)
            {
                s = ego_particle(s, d.xt);
            }
            const u32x4 r = draw(d.seed, d.inst, TAG_PREDICT, d.ep_pred, (uint32_t)i);
            float n0, n1, n2, n3;
            box_muller(r.x, r.y, &n0, &n1);
            box_muller(r.z, r.w, &n2, &n3);
            const float npx = normal_scale(n0, zero, d.sp), npy = normal_scale(n1, zero, d.sp);
            const float nvx = normal_scale(n2, zero, d.sv), nvy = normal_scale(n3, zero, d.sv);
            // glm mat4 * vec4 association ((c0 x + c1 y) + (c2 vx + c3 vy)) of the transposed
            // constant velocity matrix (dogm.cu:246-253).
            const float sx = (one * s.x + zero * s.y) + (dt * s.z + zero * s.w);
            const float sy = (zero * s.x + one * s.y) + (zero * s.z + dt * s.w);
            const float svx = (zero * s.x + zero * s.y) + (one * s.z + zero * s.w);
            const float svy = (zero * s.x + zero * s.y) + (zero * s.z + one * s.w);
            float4 o;
            o.x = sx + npx;
            o.y = sy + npy;
            o.z = svx + nvx;
            o.w = svy + nvy;
            const float w = predicted_weight(d, w_in, o);
            cell = clamp_cell(o.x, gs) + gs * clamp_cell(o.y, gs);
            if (aos)
            {
                stream_store(reinterpret_cast<float4*>(d.cdf) + i, o);
            }
            else
            {
                P.pos[i] = make_float2(o.x, o.y);
                P.vel[i] = make_float2(o.z, o.w);
            }
            if (!keep_w)
                P.weight[i] = w;
            stream_store(&P.idx[i], cell);
        }
        for (int ps = 0; ps < passes; ++ps)
            wave_count(s_h[ps], ((unsigned)cell >> (ps * RADIX_BITS)) & (RADIX - 1), valid);
    }
    if (passes > 0)
    {
        __syncthreads();
        unsigned* gh = ghist + (size_t)tr.v * 4 * RADIX;
        for (int i = threadIdx.x; i < passes * RADIX; i += PRED_BLOCK)
        {
            const unsigned v = (&s_h[0][0])[i];
            if (v)
                atomicAdd(&gh[i], v);
        }
        if (thist)
            for (int i = threadIdx.x; i < RADIX; i += PRED_BLOCK)
                thist[(size_t)i * tiles + blockIdx.x] = s_h[0][i];
    }
}

void run_predict(Engine* E, int pph, bool hist, bool aos)
{
    const int blocks = E->tiles(SP_P, SORT_Q);
    if (blocks <= 0)
        return;
    if (hist && !E->ghist_clean)
        run_fill_u32(E, E->sort_ghist, 0u, (long long)E->m * 4 * RADIX);
    E->ghist_clean = false;  // until the sort of this step has cleared them again
    unsigned* thist = E->tune.sort == SORT_REDUCE_SCAN ? E->sort_status : nullptr;
    const int passes = hist ? E->passes : 0;
    // With the assignment in the same step (hist), its bounds reset runs in extra workgroups.
    const int pb = E->tune.pred_block == 256 ? 256 : 512;
    const int resets = hist ? E->tiles(SP_C, pb * 4 / UNIT) : 0;
    launch_begin(E, K_PREDICT);
    if (pb == 256)
        hipLaunchKernelGGL(k_predict<256>, dim3(blocks + resets), dim3(256), 0, E->stream, E->d_desc, E->tab(SP_P),
                           pph, passes, E->sort_ghist, E->sort_tickets, thist, blocks, aos ? 1 : 0, E->tab(SP_C));
    else
        hipLaunchKernelGGL(k_predict<512>, dim3(blocks + resets), dim3(512), 0, E->stream, E->d_desc, E->tab(SP_P),
                           pph, passes, E->sort_ghist, E->sort_tickets, thist, blocks, aos ? 1 : 0, E->tab(SP_C));
    launch_end(E, K_PREDICT);
}

// ---------------------------------------------------------------------------------------------
// Assignment (dogm/src/dogm.cu:262-281, init.cu:86-93, particle_to_grid.cu:16-44)
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_reset_bounds(const VehDesc* __restrict__ descs,
                                                        const uint32_t* __restrict__ tab)
{
    reset_bounds_cells<BLOCK>(descs, tab, blockIdx.x);
}

// Gathers the payload into sorted order and records first/last occurrence of every cell.
__global__ __launch_bounds__(BLOCK) void k_sort_gather(const VehDesc* __restrict__ descs,
                                                       const uint32_t* __restrict__ tab, int pph, int fin)
{
    const TileRef tr = tile_ref(tab, 1, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    const int k = tr.base + threadIdx.x;
    const int n = d.N;
    if (k >= n)
        return;
    const DevParticles src = d.part[pph];
    const DevParticles dst = d.part[pph ^ 1];
    const int* skeys = d.sort_keys[fin];
    const int sv = d.sort_vals[fin][k];
    const int o = sort_value_index(sv);
    const int key = skeys[k];
    const float w = src.weight[o];
    dst.pos[k] = src.pos[o];
    dst.vel[k] = src.vel[o];
    dst.weight[k] = w;
    dst.assoc[k] = sort_value_assoc(sv);
    dst.idx[k] = key;
    d.weight_array[k] = w;
    if (k == 0 || skeys[k - 1] != key)
        d.start[key] = k;
    if (k == n - 1 || skeys[k + 1] != key)
        d.end[key] = k;
}

void run_assignment(Engine* E, int pph, bool hist_done, bool defer_bounds, bool write_wa)
{
    if (!hist_done)  // else predict's launch reset the bounds
    {
        launch_begin(E, K_RESET_BOUNDS);
        hipLaunchKernelGGL(k_reset_bounds, dim3(E->tiles(SP_C, 4)), dim3(BLOCK), 0, E->stream, E->d_desc,
                           E->tab(SP_C));
        launch_end(E, K_RESET_BOUNDS);
    }
    run_sort(E, pph, hist_done, defer_bounds, write_wa || !defer_bounds);
    if (defer_bounds)
        return;  // the last sort pass wrote the particles; the occupancy reduction records the bounds
    launch_begin(E, K_SORT_GATHER);
    hipLaunchKernelGGL(k_sort_gather, dim3(E->tiles(SP_P, 1)), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_P),
                       pph, (E->passes - 1) & 1);
    launch_end(E, K_SORT_GATHER);
}

// ---------------------------------------------------------------------------------------------
// Occupancy / mass update (dogm/src/dogm.cu:283-298, mass_update.cu:16-93)
// ---------------------------------------------------------------------------------------------
// IT particles per thread (the tuning's seg_items); register budgets of six waves per SIMD for 8
// (no spills) and eight for 4.
template <int IT>
struct OccOp
{
    static constexpr int NCH = 1;
    static constexpr int ITEMS = IT;
    static constexpr int WAVES = IT == 8 ? 6 : 8;
    int pph;
    int bounds;  // record start_idx / end_idx (dogm.cu:262-281) when the assignment deferred them
    __device__ const int* keys(const VehDesc& d) const { return d.part[pph].idx; }
    __device__ void head(const VehDesc& d, int key, int p) const
    {
        if (bounds)
            d.start[key] = p;
    }
    __device__ void end(const VehDesc& d, int key, int p) const
    {
        if (bounds)
            d.end[key] = p;
    }
    __device__ void load(const VehDesc& d, int p0, int cnt, const int (&)[ITEMS], float (&v)[ITEMS][NCH]) const
    {
        // After the gathering sort pass of the same step the particles' weights equal weight_array
        // (which that pass may leave unwritten); otherwise weight_array is what the step starts from.
        float w[ITEMS];
        load_run<ITEMS>(bounds ? d.part[pph].weight : d.weight_array, p0, cnt, w);
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
            v[i][0] = w[i];
    }
    __device__ void finalize(const VehDesc& d, int key, const fx_t (&s)[1]) const { d.m_raw[key] = fx_to_f32(s[0]); }
    // A cell crossing tiles is marked with NaN (an exact sum is never NaN); the occupancy update
    // sums its tile partials.
    __device__ void cont(const VehDesc& d, int key) const { d.m_raw[key] = __builtin_nanf(""); }
};

// zero_empty: bit 0 writes mu_A = mu_UA = 0, bit 1 the five velocity moments = 0 for cells without
// particles (what the persistent-update and moments stages write there), when those stages run
// in the same step.  With both (a full update) the kernel also writes the occupancy bits, and with
// valid bits from the previous full update it skips the zeros of cells that were empty then (their
// fields are still zero in memory).
// Cells whose particles cross reduction tiles get their weight sum here from the reduction's tile
// partials (the first tile's tail, whole middle tiles, the last tile's head), as the fix-up pass
// would, and m_raw is written for them.
// One cell; returns its born mass rho_b.
__device__ __forceinline__ float mass_update_cell(const VehDesc& d, int c, int zero_empty, const SegCarry& carry,
                                                  int occ_tile)
{
    const int gs = d.gs;
    const float* free_src = d.free[0];
    const float one = 1.0f;
    float m_occ_pred = 0.0f;
    const bool occupied = d.start[c] != -1;
    bool zero_fields = !occupied && zero_empty;
    if (zero_empty == 3)
    {
        // Work-table units are 256 cells, so each wave covers one aligned 64-cell word.
        const unsigned long long was = d.nz_valid ? d.nz[c >> 6] : ~0ull;
        zero_fields = zero_fields && ((was >> (c & 63)) & 1ull);
        const unsigned long long now = __ballot(occupied);
        if ((c & 63) == 0)
            d.nz[c >> 6] = now;
    }
    if (occupied)
    {
        const int OT = occ_tile;  // particles per reduction tile
        m_occ_pred = d.m_raw[c];
        if (m_occ_pred != m_occ_pred)  // marked: the cell crosses reduction tiles
        {
            const int ts = d.start[c] / OT, te = d.end[c] / OT;
            const int t0 = d.u0[SP_P] / (OT / UNIT);
            fx_t sum = carry.tail_partial[t0 + ts];
            for (int t = ts + 1; t < te; ++t)
                sum += carry.tail_partial[t0 + t];
            sum += carry.head_partial[t0 + te];
            m_occ_pred = fx_to_f32(sum);
            d.m_raw[c] = m_occ_pred;
        }
    }
    else if (zero_fields)
    {
        if (zero_empty & 1)
        {
            d.f[GF_MU_A][c] = 0.0f;
            d.f[GF_MU_UA][c] = 0.0f;
        }
        if (zero_empty & 2)
            for (int f = GF_MEAN_X; f <= GF_COVAR; ++f)
                d.f[f][c] = 0.0f;
    }
    const float m_raw_c = m_occ_pred;  // Σw of the cell before the clamp
    if (m_occ_pred > one)
        m_occ_pred = one;
    float fprev;
    if (d.shifted)
    {
        const int x = c % gs, y = c / gs;
        const int sx = x + d.gx, sy = y + d.gy;
        fprev = (sx > 0 && sx < gs && sy > 0 && sy < gs) ? free_src[sx + gs * sy] : 0.0f;
    }
    else if (d.xform_g)
    {
        const int sc = ego_source_cell(c, gs, d.xt);
        fprev = sc >= 0 ? free_src[sc] : 0.0f;
    }
    else
    {
        fprev = free_src[c];
    }
    // The step's measurement grid, copied into the vehicle's buffer on the way when it comes from
    // the caller's device buffer (dogm.cu:205-215).
    const dogm_measurement_cell mc = d.meas_in ? d.meas_in[c] : d.meas[c];
    if (d.meas_in)
        stream_store(&d.meas[c], mc);
    if ((zero_empty & 1) && occupied)
        stream_store(&d.cellin[c], make_float2(mc.likelihood, m_raw_c));
    const float pB = d.pB;
    const float m_free_pred = fminf(d.alpha * fprev, one - m_occ_pred);
    const float unknown_pred = one - m_occ_pred - m_free_pred;
    const float meas_unknown = one - mc.free_mass - mc.occ_mass;
    const float K = m_free_pred * mc.occ_mass + m_occ_pred * mc.free_mass;
    const float occ_up =
        (m_occ_pred * meas_unknown + unknown_pred * mc.occ_mass + m_occ_pred * mc.occ_mass) / (one - K);
    const float free_up =
        (m_free_pred * meas_unknown + unknown_pred * mc.free_mass + m_free_pred * mc.free_mass) / (one - K);
    const float rho_b = (occ_up * pB * (one - m_occ_pred)) / (m_occ_pred + pB * (one - m_occ_pred));
    const float rho_p = occ_up - rho_b;
    stream_store(&d.born[c], rho_b);
    stream_store(&d.f[GF_PERS][c], rho_p);
    stream_store(&d.f[GF_NEW_BORN][c], rho_b);
    stream_store(&d.free[1][c], free_up);
    stream_store(&d.f[GF_OCC][c], occ_up);
    stream_store(&d.f[GF_PRED][c], m_occ_pred);
    return rho_b;
}

// One workgroup per 256-cell unit.  With born_sums set (the births run in the same step), the unit's
// exact sum of born masses goes to born_sums[unit] for the births' top-level scan, which then needs no
// tile pass over the born masses.  With exp_sums set (the resampling runs in the same step), the
// workgroups after the `units` cell units compute the tile totals of the resampling's exponentials
// (ALU work beside the memory-bound cell update) into exp_sums.
__global__ __launch_bounds__(BLOCK) void k_mass_update(const VehDesc* __restrict__ descs,
                                                       const uint32_t* __restrict__ tab, int zero_empty,
                                                       SegCarry carry, int occ_tile, fx_t* __restrict__ born_sums,
                                                       int units, const uint32_t* __restrict__ tab_e,
                                                       fx_t* __restrict__ exp_sums)
{
    if ((int)blockIdx.x >= units)
    {
        scan_tile_body(SrcExp{}, descs, tab_e, (int)blockIdx.x - units, exp_sums, (fx_t*)nullptr);
        return;
    }
    const TileRef tr = tile_ref(tab, 1, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    const int c = tr.base + threadIdx.x;
    float rho_b = 0.0f;
    if (c < d.C)
        rho_b = mass_update_cell(d, c, zero_empty, carry, occ_tile);
    if (born_sums)
    {
        fx_t total;
        block_exclusive_scan_fx(c < d.C ? fx_from_f32(rho_b) : (fx_t)0, &total);
        if (threadIdx.x == 0)
            born_sums[blockIdx.x] = total;
    }
}

// normalize_weights (mass_update.cu:51-59) for cells whose predicted occupancy exceeds 1.
__global__ __launch_bounds__(BLOCK) void k_normalize(const VehDesc* __restrict__ descs,
                                                     const uint32_t* __restrict__ tab, int pph)
{
    const TileRef tr = tile_ref(tab, 1, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    const int i = tr.base + threadIdx.x;
    if (i >= d.N)
        return;
    const DevParticles P = d.part[pph];
    const float m = d.m_raw[P.idx[i]];
    if (m > 1.0f)
    {
        const float w = d.weight_array[i] / m;
        d.weight_array[i] = w;
        P.weight[i] = w;
    }
}

void run_occupancy(Engine* E, int pph, int zero_empty, bool defer_normalize, bool bounds, bool born_sums,
                   bool exp_sums)
{
    if (E->tune.seg_items == 4)
        launch_segreduce(E, K_SEG_OCC, K_SEG_OCC_FIX, OccOp<4>{pph, bounds ? 1 : 0}, false);
    else
        launch_segreduce(E, K_SEG_OCC, K_SEG_OCC_FIX, OccOp<8>{pph, bounds ? 1 : 0}, false);
    launch_begin(E, K_MASS_UPDATE);
    const int units = E->tiles(SP_C, 1);
    const int etiles = exp_sums ? E->tiles(SP_E, SCAN_Q) : 0;
    hipLaunchKernelGGL(k_mass_update, dim3(units + etiles), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_C),
                       zero_empty, E->seg, BLOCK * E->tune.seg_items, born_sums ? born_unit_sums(E) : (fx_t*)nullptr,
                       units, E->tab(SP_E), E->exp_tile_sum);
    launch_end(E, K_MASS_UPDATE);
    if (defer_normalize)
        return;
    launch_begin(E, K_NORMALIZE);
    hipLaunchKernelGGL(k_normalize, dim3(E->tiles(SP_P, 1)), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_P),
                       pph);
    launch_end(E, K_NORMALIZE);
}

// ---------------------------------------------------------------------------------------------
// Persistent particle update (dogm/src/dogm.cu:300-321, update_persistent_particles.cu:16-87)
// ---------------------------------------------------------------------------------------------
// Kernel1 (lw = likelihood * w, written to weight_array) is fused into the reduction's load.
// norm: normalize_weights (mass_update.cu:51-59) is applied here, when the occupancy update ran in the
// same step: w <- w / m_occ_pred for cells whose predicted occupancy exceeds 1 (the particle weight is
// rewritten; weight_array is overwritten by lw below either way).
template <int IT>
struct PersOp
{
    static constexpr int NCH = 1;
    static constexpr int ITEMS = IT;
    static constexpr int WAVES = IT == 8 ? 6 : 8;
    int pph;
    int norm;
    int rec;  // write the cell factors of the moments' posterior weights (cellrec)
    __device__ const int* keys(const VehDesc& d) const { return d.part[pph].idx; }
    __device__ void head(const VehDesc&, int, int) const {}
    __device__ void end(const VehDesc&, int, int) const {}
    __device__ void cont(const VehDesc&, int) const {}
    __device__ void load(const VehDesc& d, int p0, int cnt, const int (&key)[ITEMS], float (&v)[ITEMS][NCH]) const
    {
        float w[ITEMS], lw[ITEMS], lik[ITEMS];
        load_run<ITEMS>(d.part[pph].weight, p0, cnt, w);
        if (norm)
        {
            // The occupancy update of this step left (likelihood, Σw) per cell with particles; every
            // item's pair is loaded before the first store.
            float m[ITEMS];
#pragma unroll
            for (int i = 0; i < ITEMS; ++i)
            {
                const float2 ci = i < cnt ? d.cellin[key[i]] : make_float2(0.0f, 0.0f);
                lik[i] = ci.x;
                m[i] = ci.y;
            }
#pragma unroll
            for (int i = 0; i < ITEMS; ++i)
            {
                if (i < cnt && m[i] > 1.0f)
                {
                    w[i] = w[i] / m[i];
                    d.part[pph].weight[p0 + i] = w[i];
                }
            }
        }
        else
        {
#pragma unroll
            for (int i = 0; i < ITEMS; ++i)
                lik[i] = i < cnt ? d.meas[key[i]].likelihood : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
        {
            lw[i] = i < cnt ? lik[i] * w[i] : 0.0f;
            v[i][0] = lw[i];
        }
        store_run<ITEMS>(d.weight_array, p0, cnt, lw);
    }
    __device__ void finalize(const VehDesc& d, int key, const fx_t (&s)[1]) const
    {
        const float acc = fx_to_f32(s[0]);
        const float rho_p = d.f[GF_PERS][key];
        const float mu_A = acc > 0.0f ? rho_p / acc : 0.0f;
        d.f[GF_MU_A][key] = mu_A;
        const float pred = d.f[GF_PRED][key];
        const float mu_UA = pred > 0.0f ? rho_p / pred : 0.0f;
        d.f[GF_MU_UA][key] = mu_UA;
        if (rec)
        {
            // The factors of w = (p_A mu_A) w_lik + ((1 - p_A) mu_UA) w_prior in that association
            // (update_persistent_particles.cu:76-87), and rho_p for the moments.
            const float pA = d.meas[key].p_A;
            d.cellrec[key] = make_float4(pA * mu_A, (1.0f - pA) * mu_UA, rho_p, 0.0f);
        }
    }
};

// Writes 0 into fields [f_lo, f_hi] of every cell without particles.
__global__ __launch_bounds__(BLOCK) void k_empty_cells(const VehDesc* __restrict__ descs,
                                                       const uint32_t* __restrict__ tab, int f_lo, int f_hi)
{
    const TileRef tr = tile_ref(tab, 1, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    const int c = tr.base + threadIdx.x;
    if (c >= d.C || d.start[c] != -1)
        return;
    for (int f = f_lo; f <= f_hi; ++f)
        d.f[f][c] = 0.0f;
}

__global__ __launch_bounds__(BLOCK) void k_post_weights(const VehDesc* __restrict__ descs,
                                                        const uint32_t* __restrict__ tab, int pph)
{
    const TileRef tr = tile_ref(tab, 1, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    const int i = tr.base + threadIdx.x;
    if (i >= d.N)
        return;
    const DevParticles P = d.part[pph];
    const int c = P.idx[i];
    const float pA = d.meas[c].p_A;
    d.weight_array[i] = pA * d.f[GF_MU_A][c] * d.weight_array[i] + (1.0f - pA) * d.f[GF_MU_UA][c] * P.weight[i];
}

void run_persistent(Engine* E, int pph, bool fused_normalize, bool empty_done, bool defer_post)
{
    if (E->tune.seg_items == 4)
        launch_segreduce(E, K_SEG_PERS, K_SEG_PERS_FIX, PersOp<4>{pph, fused_normalize ? 1 : 0, defer_post ? 1 : 0});
    else
        launch_segreduce(E, K_SEG_PERS, K_SEG_PERS_FIX, PersOp<8>{pph, fused_normalize ? 1 : 0, defer_post ? 1 : 0});
    if (!empty_done)
    {
        launch_begin(E, K_EMPTY_CELLS);
        hipLaunchKernelGGL(k_empty_cells, dim3(E->tiles(SP_C, 1)), dim3(BLOCK), 0, E->stream, E->d_desc,
                           E->tab(SP_C), (int)GF_MU_A, (int)GF_MU_UA);
        launch_end(E, K_EMPTY_CELLS);
    }
    if (defer_post)
        return;
    launch_begin(E, K_POST_WEIGHTS);
    hipLaunchKernelGGL(k_post_weights, dim3(E->tiles(SP_P, 1)), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_P),
                       pph);
    launch_end(E, K_POST_WEIGHTS);
}

// Born masses per cell (first-frame placement when init_only, births otherwise).
struct SrcBorn
{
    static constexpr bool kNonNeg = false;
    int init_only;
    __device__ bool active(const VehDesc& d) const { return !init_only || d.init; }
    __device__ int len(const VehDesc& d) const { return d.C; }
    __device__ void load4(const VehDesc& d, int p0, int cnt, float (&v)[4]) const { load_run<4>(d.born, p0, cnt, v); }
};

__device__ __forceinline__ int order_end_at(int j, fx_t prefix, const DevScalars* sc, int C, int v)
{
    if (sc->total == 0)
        return order_end_uniform(j, C, v);
    if (j == C - 1)
        return v;
    return order_end(prefix, sc->scale, v);
}

// First-frame placement: P_j only.
struct SinkOrders
{
    __device__ void apply4(const VehDesc& d, const DevScalars* sc, int p0, int cnt, const fx_t (&pre)[4],
                           const fx_t (&)[4]) const
    {
        int e[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            e[i] = i < cnt ? order_end_at(p0 + i, pre[i], sc, d.C, d.N) : 0;
        store_run<4>(d.pend, p0, cnt, e);
    }
};

// Birth placement and weights per cell (init_new_particles.cu:126-155).
struct SinkBirth
{
    __device__ void apply4(const VehDesc& d, const DevScalars* sc, int p0, int cnt, const fx_t (&pre)[4],
                           const fx_t (&own)[4]) const
    {
        float bm[4];
        load_run<4>(d.born, p0, cnt, bm);
        const int C = d.C, v = d.B;
        int e[4];
        float a[4], u[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
            e[i] = 0;
            a[i] = 0.0f;
            u[i] = 0.0f;
            if (i < cnt)
            {
                const int j = p0 + i;
                e[i] = order_end_at(j, pre[i], sc, C, v);
                // P_{j-1}: the previous item's end, or recomputed from the exclusive prefix.
                const int s = j == 0 ? 0 : (i > 0 ? e[i - 1] : order_end_at(j - 1, pre[i] - own[i], sc, C, v));
                const int n = e[i] - s;
                if (n > 0)  // cells without birth slots keep w_A = w_UA = 0 and skip the measurement read
                {
                    const float pA = d.meas[j].p_A;
                    const int nu_A = (int)roundf((float)n * pA);
                    const int nu_UA = n - nu_A;
                    a[i] = nu_A > 0 ? (pA * bm[i]) / (float)nu_A : 0.0f;
                    u[i] = nu_UA > 0 ? ((1.0f - pA) * bm[i]) / (float)nu_UA : 0.0f;
                }
            }
        }
        store_run<4>(d.pend, p0, cnt, e);
        store_run<4>(d.f[GF_W_A], p0, cnt, a);
        store_run<4>(d.f[GF_W_UA], p0, cnt, u);
    }
};

// ---------------------------------------------------------------------------------------------
// Birth particles (dogm/src/dogm.cu:323-348, init.cu:46-67, init_new_particles.cu:157-195)
// ---------------------------------------------------------------------------------------------

// Cell of birth slot k, j = first index with pend[j] > k (pend non-decreasing, pend[C - 1] = B), for
// the 64 consecutive slots k_first .. k_first + 63 of a wave (all lanes call it; k_last is the last
// valid slot).  Lanes 0-31 and 32-63 run 32-ary searches for the cells of k_first and k_last (one
// wave-wide load per round instead of one dependent load per binary-search level); when those cells
// are fewer than 65 apart, each lane holds one pend entry of the range and finds its own cell with a
// six-step search over the lanes, otherwise it binary-searches the range.
__device__ int wave_slot_cells(const int* __restrict__ pend, int C, int k_first, int k_last, int k, int lane)
{
    const int half = lane >> 5, hl = lane & 31;
    const int target = half ? k_last : k_first;
    int lo = 0, hi = C - 1;  // the answer lies in [lo, hi]
    while (__ballot(lo < hi))
    {
        const int step = (hi - lo + 31) >> 5;
        const int pos = lo + (hl + 1) * step - 1;
        const bool le = lo < hi && pos < hi && pend[pos] <= target;
        const unsigned long long bal = __ballot(le);
        const int cnt = __popcll((bal >> (32 * half)) & 0xffffffffull);
        if (lo < hi)
        {
            const int nlo = lo + cnt * step;
            hi = min(hi, nlo + step - 1);
            lo = nlo;
        }
    }
    const int jlo = __shfl(lo, 0, WAVE);
    const int jhi = __shfl(lo, 32, WAVE);
    const int r = jhi - jlo + 1;
    if (r <= WAVE)
    {
        const int pv = jlo + lane <= jhi ? pend[jlo + lane] : 0x7fffffff;
        int cnt = 0;  // entries of the range <= k
#pragma unroll
        for (int bit = WAVE / 2; bit > 0; bit >>= 1)
        {
            const int cand = cnt + bit;
            const int v = __shfl(pv, cand - 1, WAVE);
            if (v <= k)
                cnt = cand;
        }
        return jlo + cnt;
    }
    return jlo + upper_bound_i32(pend + jlo, r, k);
}
__global__ __launch_bounds__(BLOCK) void k_birth_particles(const VehDesc* __restrict__ descs,
                                                           const uint32_t* __restrict__ tab)
{
    const TileRef tr = tile_ref(tab, 1, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    const int B = d.B;
    const int lane = threadIdx.x & (WAVE - 1);
    const int k0 = tr.base + (threadIdx.x - lane);  // the wave's first slot
    if (k0 >= B)
        return;  // wave-uniform
    const int k = k0 + lane;
    const int* pend = d.pend;
    const int j = wave_slot_cells(pend, d.C, k0, min(k0 + WAVE - 1, B - 1), k, lane);
    if (k >= B)
        return;
    const int s = j > 0 ? pend[j - 1] : 0;
    const int n = pend[j] - s;
    const int nu_A = (int)roundf((float)n * d.meas[j].p_A);
    const bool assoc = (k - s) <= nu_A;
    const u32x4 r = draw(d.seed, d.inst, TAG_BIRTH, d.ep_birth, (uint32_t)k);
    float n0, n1;
    box_muller(r.x, r.y, &n0, &n1);
    const int gs = d.gs;
    d.birth.pos[k] = make_float2((float)(j % gs) + 0.5f, (float)j / (float)gs + 0.5f);
    d.birth.vel[k] = make_float2(normal_scale(n0, 0.0f, d.sd_birth), normal_scale(n1, 0.0f, d.sd_birth));
    d.birth.idx[k] = j;
    d.birth.weight[k] = assoc ? d.f[GF_W_A][j] : d.f[GF_W_UA][j];
    d.birth.assoc[k] = assoc ? 1 : 0;
}

void run_births(Engine* E, bool born_sums)
{
    // With B = 0 the cell pass still runs (every cell gets n_j = 0 and w_A = w_UA = 0) and the
    // stage still consumes its RNG epoch, so the stream of later stages does not depend on B.
    run_exact_scan(E, SP_C, SrcBorn{0}, SinkBirth{}, SCAN_ORDERS_B, K_BIRTH_CELLS, ONLY_ALL, nullptr,
                   born_sums ? born_unit_sums(E) : nullptr);
    const int blocks = E->tiles(SP_B, 1);
    if (blocks <= 0)
        return;
    launch_begin(E, K_BIRTH_PARTICLES);
    hipLaunchKernelGGL(k_birth_particles, dim3(blocks), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_B));
    launch_end(E, K_BIRTH_PARTICLES);
}

// ---------------------------------------------------------------------------------------------
// First-measurement initialisation (dogm/src/dogm.cu:217-241, init_new_particles.cu:76-124),
// for the vehicles of the step whose descriptor has `init` set.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_copy_masses(const VehDesc* __restrict__ descs,
                                                       const uint32_t* __restrict__ tab)
{
    const TileRef tr = tile_ref(tab, 1, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    const int c = tr.base + threadIdx.x;
    if (d.init && c < d.C)
        d.born[c] = d.meas[c].occ_mass;
}

__global__ __launch_bounds__(BLOCK) void k_init_particles(const VehDesc* __restrict__ descs,
                                                          const uint32_t* __restrict__ tab)
{
    const TileRef tr = tile_ref(tab, 1, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    const int k = tr.base + threadIdx.x;
    if (!d.init || k >= d.N)
        return;
    const int gs = d.gs;
    const int j = upper_bound_i32(d.pend, d.C, k);
    const u32x4 r = draw(d.seed, d.inst, TAG_INIT_PARTICLES, d.ep_init, (uint32_t)k);
    const float v = d.vmax;
    const DevParticles P = d.part[0];
    P.pos[k] = make_float2((float)(j % gs) + 0.5f, (float)(j / gs) + 0.5f);
    P.vel[k] = make_float2(-v + (v - -v) * uniform01(r.x), -v + (v - -v) * uniform01(r.y));
    P.idx[k] = j;
    P.weight[k] = 1.0f / (float)d.N;  // dogm.cu:228
    P.assoc[k] = 0;
}

void run_init_particles(Engine* E)
{
    launch_begin(E, K_COPY_MASSES);
    hipLaunchKernelGGL(k_copy_masses, dim3(E->tiles(SP_C, 1)), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_C));
    launch_end(E, K_COPY_MASSES);
    run_exact_scan(E, SP_C, SrcBorn{1}, SinkOrders{}, SCAN_ORDERS_N, K_INIT_CELLS, ONLY_INIT);
    launch_begin(E, K_INIT_PARTICLES);
    hipLaunchKernelGGL(k_init_particles, dim3(E->tiles(SP_P, 1)), dim3(BLOCK), 0, E->stream, E->d_desc,
                       E->tab(SP_P));
    launch_end(E, K_INIT_PARTICLES);
}

// ---------------------------------------------------------------------------------------------
// Statistical moments (dogm/src/dogm.cu:350-384, statistical_moments.cu:16-106)
// ---------------------------------------------------------------------------------------------
// Per-cell sums of the five velocity channels w vx, w vy, w vx^2, w vy^2, w vx vy (f32 products
// widened to f64) in one fixed association, independent of the launch shape and of the other
// vehicles of the step: a grid's sorted particles are cut into chunks of 256 (one wave64 per chunk,
// 4 consecutive particles per lane).
//   lane:   sequential fold restarting where a cell starts (the lane's first particle always starts it);
//   wave:   segmented inclusive scan of the lanes' last folds in six DPP steps (row shifts 1, 2, 4, 8
//           inside 16-lane rows, then lane 15 of rows 0 / 2 into rows 1 / 3, lane 31 into rows 2 / 3),
//           each adding the moved value only while no cell starts in between;
//   cell:   ending in lane l: its lane-local fold if it starts in lane l, else (scan of lane l - 1) +
//           fold;
//   chunks: a cell crossing chunks is the left fold ((first piece + whole middle chunk) + ...) + last
//           piece, summed by k_moments_fix from the chunks' carries.
// The float sums are then rounded once to f32 and finalised as in statistical_moments.cu:58-106.
// With post set (the persistent update ran in the same step) the posterior weights
// (update_persistent_particles.cu:76-87) are formed in the load: w = p_A mu_A lw + (1 - p_A) mu_UA
// w_prior, written to weight_array and used for the moments.
constexpr int MOM_CH = 5;
constexpr int MOM_ITEMS = 4;                  // consecutive particles per lane
constexpr int MOM_CHUNK = WAVE * MOM_ITEMS;   // particles per wave
static_assert(MOM_CHUNK == UNIT, "a moments chunk is one work-table unit");
constexpr uint8_t MC_CONT = 1;    // the chunk's last cell continues into the next chunk
constexpr uint8_t MC_NOHEAD = 2;  // no cell starts in the chunk

// rho_p = the cell's persistent occupancy mass (d.f[GF_PERS][key]).
__device__ __forceinline__ void mom_finalize(const VehDesc& d, int key, float rho_p, const float (&s)[MOM_CH])
{
    float mx = 0.0f, my = 0.0f, vx = 0.0f, vy = 0.0f, cv = 0.0f;
    if (rho_p > 0.0f)
    {
        const float inv = 1.0f / rho_p;
        mx = inv * s[0];
        my = inv * s[1];
        vx = inv * s[2] - mx * mx;
        vy = inv * s[3] - my * my;
        cv = inv * s[4] - mx * my;
    }
    d.f[GF_MEAN_X][key] = mx;
    d.f[GF_MEAN_Y][key] = my;
    d.f[GF_VAR_X][key] = vx;
    d.f[GF_VAR_Y][key] = vy;
    d.f[GF_COVAR][key] = cv;
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_i(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, CTRL, ROW_MASK, 0xf, false);
}

// carry.flags[unit], and per unit MOM_CH doubles in carry.head_partial (piece of the chunk's first
// cell when that cell started in an earlier chunk) and carry.tail_partial (piece of its last cell).
template <int WAVES_PER_WG>
__global__ __launch_bounds__(WAVES_PER_WG* WAVE) __attribute__((amdgpu_waves_per_eu(7))) void k_moments(const VehDesc* __restrict__ descs,
                                                                const uint32_t* __restrict__ tab, int pph, int post,
                                                                SegCarry carry)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int unit = xcd_tile(blockIdx.x, gridDim.x) * WAVES_PER_WG + (int)threadIdx.x / WAVE;
    const TileRef tr = tile_ref(tab, 1, unit);
    const VehDesc& d = descs[tr.v];
    const int n = d.N;
    const int base = tr.base;
    if (base >= n)
    {
        if (lane == 0)
            carry.flags[unit] = 0;  // padding chunk past the vehicle's particles
        return;                     // wave-uniform
    }
    const int p0 = base + lane * MOM_ITEMS;
    const int cnt = max(0, min(MOM_ITEMS, n - p0));
    const int* keys = d.part[pph].idx;
    int key[MOM_ITEMS];
    load_run<MOM_ITEMS>(keys, p0, cnt, key, -1);
    int kprev = __shfl_up(key[MOM_ITEMS - 1], 1, WAVE);
    int knext = __shfl_down(key[0], 1, WAVE);
    if (lane == 0)
        kprev = base > 0 ? keys[base - 1] : -1;
    if (lane == WAVE - 1)
        knext = p0 + MOM_ITEMS < n ? keys[p0 + MOM_ITEMS] : -1;
    bool head[MOM_ITEMS], end[MOM_ITEMS], started[MOM_ITEMS];
    bool lh = false;
#pragma unroll
    for (int i = 0; i < MOM_ITEMS; ++i)
    {
        const int kp = i == 0 ? kprev : key[i - 1];
        const int kn = i == MOM_ITEMS - 1 ? knext : key[i + 1];
        head[i] = i < cnt && key[i] != kp;
        end[i] = i < cnt && key[i] != kn;
        lh = lh || head[i];
        started[i] = lh;
    }

    // The persistent masses of the cells ending here (and with post the cells' weight factors, which
    // carry them), loaded before any store so that they travel with the other loads instead of
    // waiting behind the weight_array stores.
    float rho[MOM_ITEMS];
    float4 cf[MOM_ITEMS];
#pragma unroll
    for (int i = 0; i < MOM_ITEMS; ++i)
    {
        cf[i] = post && i < cnt ? d.cellrec[key[i]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        rho[i] = post ? cf[i].z : (end[i] ? d.f[GF_PERS][key[i]] : 0.0f);
    }
    float w[MOM_ITEMS];
    load_run<MOM_ITEMS>(d.weight_array, p0, cnt, w);
    if (post)
    {
        float wp[MOM_ITEMS];
        load_run<MOM_ITEMS>(d.part[pph].weight, p0, cnt, wp);
#pragma unroll
        for (int i = 0; i < MOM_ITEMS; ++i)
            if (i < cnt)
                w[i] = cf[i].x * w[i] + cf[i].y * wp[i];  // (p_A mu_A) w + ((1 - p_A) mu_UA) w_prior
        store_run<MOM_ITEMS>(d.weight_array, p0, cnt, w);
    }
    float vx[MOM_ITEMS], vy[MOM_ITEMS];
    {
        const float2* vel = d.part[pph].vel;
        if (cnt == MOM_ITEMS)
        {
            const float4 a = *reinterpret_cast<const float4*>(vel + p0);
            const float4 b = *reinterpret_cast<const float4*>(vel + p0 + 2);
            vx[0] = a.x; vy[0] = a.y; vx[1] = a.z; vy[1] = a.w;
            vx[2] = b.x; vy[2] = b.y; vx[3] = b.z; vy[3] = b.w;
        }
        else
        {
#pragma unroll
            for (int i = 0; i < MOM_ITEMS; ++i)
            {
                const float2 v = i < cnt ? vel[p0 + i] : make_float2(0.0f, 0.0f);
                vx[i] = v.x;
                vy[i] = v.y;
            }
        }
    }

    // Combine predicates of the six scan steps (the same for every channel): the source lane is in
    // range and no cell starts between it and this lane.
    bool comb[6];
    {
        const int row = lane >> 4, rl = lane & 15;
        int f = lh ? 1 : 0;
        int g;
        g = dpp_i<0x111, 0xf>(f); comb[0] = rl >= 1 && !f; f = rl >= 1 ? (f | g) : f;
        g = dpp_i<0x112, 0xf>(f); comb[1] = rl >= 2 && !f; f = rl >= 2 ? (f | g) : f;
        g = dpp_i<0x114, 0xf>(f); comb[2] = rl >= 4 && !f; f = rl >= 4 ? (f | g) : f;
        g = dpp_i<0x118, 0xf>(f); comb[3] = rl >= 8 && !f; f = rl >= 8 ? (f | g) : f;
        const bool b15 = row == 1 || row == 3, b31 = row >= 2;
        g = dpp_i<0x142, 0xa>(f); comb[4] = b15 && !f; f = b15 ? (f | g) : f;
        comb[5] = b31 && !f;
    }
    // The chunk's first cell started in an earlier chunk (lane 0's first particle is no cell start);
    // its first end is lane first_end_lane's particle first_i.
    const bool from_before = base > 0 && !(__ballot(head[0]) & 1ull);
    const uint64_t endlanes = __ballot(end[0] || end[1] || end[2] || end[3]);
    const int first_end_lane = endlanes ? __builtin_ctzll(endlanes) : WAVE;
    const int first_i = end[0] ? 0 : end[1] ? 1 : end[2] ? 2 : 3;
    const bool nohead = __ballot(lh) == 0;
    // The chunk's last cell continues into the next chunk: lane 63 holds 4 particles and the last
    // is no cell end.
    const bool cont = (__ballot(cnt == MOM_ITEMS && !end[MOM_ITEMS - 1]) >> (WAVE - 1)) & 1ull;

    float s32[MOM_ITEMS][MOM_CH];
#pragma unroll
    for (int c = 0; c < MOM_CH; ++c)
    {
        double r[MOM_ITEMS];
#pragma unroll
        for (int i = 0; i < MOM_ITEMS; ++i)
        {
            const float wv = w[i] * vx[i], wy = w[i] * vy[i];
            const float xf = c == 0 ? wv : c == 1 ? wy : c == 2 ? wv * vx[i] : c == 3 ? wy * vy[i] : wv * vy[i];
            const double x = (double)xf;
            if (i == 0)
                r[i] = x;
            else if (i < cnt)
                r[i] = head[i] ? x : r[i - 1] + x;
            else
                r[i] = r[i - 1];
        }
        double S = r[MOM_ITEMS - 1];
        { const double y = dpp_val<0x111, 0xf>(S); S = comb[0] ? y + S : S; }
        { const double y = dpp_val<0x112, 0xf>(S); S = comb[1] ? y + S : S; }
        { const double y = dpp_val<0x114, 0xf>(S); S = comb[2] ? y + S : S; }
        { const double y = dpp_val<0x118, 0xf>(S); S = comb[3] ? y + S : S; }
        { const double y = dpp_val<0x142, 0xa>(S); S = comb[4] ? y + S : S; }
        { const double y = dpp_val<0x143, 0xc>(S); S = comb[5] ? y + S : S; }
        const double E = __shfl_up(S, 1, WAVE);
#pragma unroll
        for (int i = 0; i < MOM_ITEMS; ++i)
        {
            const double v = (started[i] || lane == 0) ? r[i] : E + r[i];
            s32[i][c] = (float)v;
            if (end[i] && from_before && lane == first_end_lane && i == first_i)
                reinterpret_cast<double*>(carry.head_partial)[(size_t)unit * MOM_CH + c] = v;
        }
        if (lane == WAVE - 1 && cont)
            reinterpret_cast<double*>(carry.tail_partial)[(size_t)unit * MOM_CH + c] = S;
    }
    if (lane == 0)
        carry.flags[unit] = (uint8_t)((cont ? MC_CONT : 0) | (nohead ? MC_NOHEAD : 0));
#pragma unroll
    for (int i = 0; i < MOM_ITEMS; ++i)
    {
        if (!end[i])
            continue;
        if (from_before && lane == first_end_lane && i == first_i)
            continue;  // the cell started in an earlier chunk: k_moments_fix
        float s[MOM_CH];
#pragma unroll
        for (int c = 0; c < MOM_CH; ++c)
            s[c] = s32[i][c];
        mom_finalize(d, key[i], rho[i], s);
    }
}

// Cells crossing chunks: one thread per chunk whose last cell starts in it and continues.
__global__ __launch_bounds__(BLOCK) void k_moments_fix(const VehDesc* __restrict__ descs,
                                                       const uint32_t* __restrict__ tab, int units, int pph,
                                                       SegCarry carry)
{
    const int u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= units)
        return;
    const uint8_t fl = carry.flags[u];
    if (!(fl & MC_CONT) || (fl & MC_NOHEAD))
        return;
    const double* tail = reinterpret_cast<const double*>(carry.tail_partial);
    const double* headp = reinterpret_cast<const double*>(carry.head_partial);
    double acc[MOM_CH];
#pragma unroll
    for (int c = 0; c < MOM_CH; ++c)
        acc[c] = tail[(size_t)u * MOM_CH + c];
    int j = u + 1;
    while ((carry.flags[j] & (MC_CONT | MC_NOHEAD)) == (MC_CONT | MC_NOHEAD))
    {
#pragma unroll
        for (int c = 0; c < MOM_CH; ++c)
            acc[c] = acc[c] + tail[(size_t)j * MOM_CH + c];
        ++j;
    }
    float s[MOM_CH];
#pragma unroll
    for (int c = 0; c < MOM_CH; ++c)
        s[c] = (float)(acc[c] + headp[(size_t)j * MOM_CH + c]);
    const TileRef tr = tile_ref(tab, 1, u);
    const VehDesc& d = descs[tr.v];
    const int key = d.part[pph].idx[tr.base + MOM_CHUNK - 1];
    mom_finalize(d, key, d.f[GF_PERS][key], s);
}

constexpr int MOM_WAVES_PER_WG = 4;

void run_moments(Engine* E, int pph, bool fused_post, bool empty_done)
{
    const int units = (int)E->units[SP_P];
    if (units > 0)
    {
        launch_begin(E, K_SEG_MOM);
        hipLaunchKernelGGL(k_moments<MOM_WAVES_PER_WG>, dim3(units / MOM_WAVES_PER_WG), dim3(MOM_WAVES_PER_WG * WAVE),
                           0, E->stream, E->d_desc, E->tab(SP_P), pph, fused_post ? 1 : 0, E->seg);
        launch_end(E, K_SEG_MOM);
        launch_begin(E, K_SEG_MOM_FIX);
        hipLaunchKernelGGL(k_moments_fix, dim3(div_up(units, BLOCK)), dim3(BLOCK), 0, E->stream, E->d_desc,
                           E->tab(SP_P), units, pph, E->seg);
        launch_end(E, K_SEG_MOM_FIX);
    }
    if (empty_done)
        return;
    launch_begin(E, K_EMPTY_CELLS);
    hipLaunchKernelGGL(k_empty_cells, dim3(E->tiles(SP_C, 1)), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_C),
                       (int)GF_MEAN_X, (int)GF_COVAR);
    launch_end(E, K_EMPTY_CELLS);
}

// ---------------------------------------------------------------------------------------------
// Grid packing for read-back / injection (dogm_grid_cell AoS <-> structure-of-arrays)
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_pack_grid(DevGrid G, dogm_grid_cell* __restrict__ out, int C)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= C)
        return;
    dogm_grid_cell g;
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
    out[c] = g;
}

__global__ __launch_bounds__(BLOCK) void k_unpack_grid(DevGrid G, const dogm_grid_cell* __restrict__ in, int C)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= C)
        return;
    const dogm_grid_cell g = in[c];
    G.start[c] = g.start_idx;
    G.end[c] = g.end_idx;
    G.f[GF_NEW_BORN][c] = g.new_born_occ_mass;
    G.f[GF_PERS][c] = g.pers_occ_mass;
    G.f[GF_FREE][c] = g.free_mass;
    G.f[GF_OCC][c] = g.occ_mass;
    G.f[GF_PRED][c] = g.pred_occ_mass;
    G.f[GF_MU_A][c] = g.mu_A;
    G.f[GF_MU_UA][c] = g.mu_UA;
    G.f[GF_W_A][c] = g.w_A;
    G.f[GF_W_UA][c] = g.w_UA;
    G.f[GF_MEAN_X][c] = g.mean_x_vel;
    G.f[GF_MEAN_Y][c] = g.mean_y_vel;
    G.f[GF_VAR_X][c] = g.var_x_vel;
    G.f[GF_VAR_Y][c] = g.var_y_vel;
    G.f[GF_COVAR][c] = g.covar_xy_vel;
}

__global__ __launch_bounds__(BLOCK) void k_pack_state(DevParticles P, int n, float4* __restrict__ out)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n)
        return;
    const float2 p = P.pos[i], v = P.vel[i];
    out[i] = make_float4(p.x, p.y, v.x, v.y);
}

__global__ __launch_bounds__(BLOCK) void k_unpack_state(DevParticles P, int n, const float4* __restrict__ in)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n)
        return;
    const float4 s = in[i];
    P.pos[i] = make_float2(s.x, s.y);
    P.vel[i] = make_float2(s.z, s.w);
}

void run_pack_state(Engine* E, const DevParticles& p, int n, float4* out)
{
    if (n <= 0)
        return;
    launch_begin(E, K_PACK_STATE);
    hipLaunchKernelGGL(k_pack_state, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, E->stream, p, n, out);
    launch_end(E, K_PACK_STATE);
}

void run_unpack_state(Engine* E, const DevParticles& p, int n, const float4* in)
{
    if (n <= 0)
        return;
    launch_begin(E, K_UNPACK_STATE);
    hipLaunchKernelGGL(k_unpack_state, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, E->stream, p, n, in);
    launch_end(E, K_UNPACK_STATE);
}

void run_pack_grid(Engine* E, const Vehicle* v)
{
    launch_begin(E, K_PACK_GRID);
    hipLaunchKernelGGL(k_pack_grid, dim3(div_up(v->C, BLOCK)), dim3(BLOCK), 0, E->stream, v->grid, v->aos_tmp, v->C);
    launch_end(E, K_PACK_GRID);
}

void run_unpack_grid(Engine* E, const Vehicle* v)
{
    launch_begin(E, K_UNPACK_GRID);
    hipLaunchKernelGGL(k_unpack_grid, dim3(div_up(v->C, BLOCK)), dim3(BLOCK), 0, E->stream, v->grid, v->aos_tmp,
                       v->C);
    launch_end(E, K_UNPACK_GRID);
}

}  // namespace dogm_hip
