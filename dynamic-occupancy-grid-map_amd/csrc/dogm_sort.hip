// Stable LSD radix sort of particle cell indices on gfx950 (replaces thrust::sort_by_key,
// dogm/src/dogm.cu:270-276).
//
// Keys are cell indices < grid_size^2, so only ceil(log2(C)) bits are sorted, 8 bits per pass.
// Per pass: tile digit counts -> per-vehicle, per-digit exclusive scan over the vehicle's tiles (plus
// the digit's base from a one-time histogram of all passes) -> ranking and scatter.  Every vehicle
// of a step is sorted independently into its own buffers; all use the step's pass count.  Ranking is stable: inside a
// tile, keys are laid out wave-striped (wave w, item j, lane l = position (w*ITEMS + j)*64 + l),
// each wave ranks its 64 keys of one item with ballot-matched peer masks against per-wave digit
// counters in LDS, and waves are combined in order.  The scatter is staged through LDS in
// digit-sorted order so that equal digits leave the tile as contiguous runs.
#include "dogm_device.h"

namespace dogm_hip
{

constexpr int SORT_WAVES = BLOCK / WAVE;

// Lanes of the wave holding the same digit (among `active` lanes).  Nearly sorted keys usually
// give one digit per wave: that case costs one ballot.
__device__ __forceinline__ uint64_t match_digit(unsigned d, uint64_t active)
{
    if (active == 0)
        return 0;
    const unsigned first = __shfl(d, (int)__builtin_ctzll(active), WAVE);
    if ((__ballot(d == first) & active) == active)
        return active;
    uint64_t peers = active;
#pragma unroll
    for (int b = 0; b < RADIX_BITS; ++b)
    {
        const bool bit = (d >> b) & 1u;
        const uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

// Histogram of every pass's digit over all keys, per vehicle: ghist[(v * 4 + pass) * 256 + digit].
__global__ __launch_bounds__(BLOCK) void k_sort_hist(const VehDesc* __restrict__ descs, const uint32_t* __restrict__ tab,
                                                     int pph, int passes, unsigned* __restrict__ ghist)
{
    __shared__ unsigned s_h[4][RADIX];
    const TileRef tr = tile_ref(tab, SORT_Q, blockIdx.x);
    const VehDesc& dv = descs[tr.v];
    const int n = dv.N;
    if (tr.base >= n)
        return;
    const int* keys = dv.part[pph].idx;
    for (int i = threadIdx.x; i < 4 * RADIX; i += BLOCK)
        (&s_h[0][0])[i] = 0u;
    __syncthreads();
    for (int i = 0; i < SORT_ITEMS; ++i)
    {
        const int p = tr.base + i * BLOCK + threadIdx.x;
        const bool valid = p < n;
        const unsigned k = valid ? (unsigned)keys[p] : 0u;
        for (int ps = 0; ps < passes; ++ps)
            wave_count(s_h[ps], (k >> (ps * RADIX_BITS)) & (RADIX - 1), valid);
    }
    __syncthreads();
    unsigned* gh = ghist + (size_t)tr.v * 4 * RADIX;
    for (int i = threadIdx.x; i < passes * RADIX; i += BLOCK)
    {
        const unsigned v = (&s_h[0][0])[i];
        if (v)
            atomicAdd(&gh[i], v);
    }
}

__device__ __forceinline__ const int* sort_keys_in(const VehDesc& d, int pph, int ps)
{
    return ps == 0 ? d.part[pph].idx : d.sort_keys[(ps - 1) & 1];
}

// Digit counts of one tile for one pass: thist[d * tiles + t].
__global__ __launch_bounds__(BLOCK) void k_sort_upsweep(const VehDesc* __restrict__ descs,
                                                        const uint32_t* __restrict__ tab, int pph, int ps, int tiles,
                                                        unsigned* __restrict__ thist)
{
    __shared__ unsigned s_h[RADIX];
    const TileRef tr = tile_ref(tab, SORT_Q, blockIdx.x);
    const VehDesc& dv = descs[tr.v];
    const int n = dv.N;
    const int shift = ps * RADIX_BITS;
    const int* keys = sort_keys_in(dv, pph, ps);
    for (int i = threadIdx.x; i < RADIX; i += BLOCK)
        s_h[i] = 0u;
    __syncthreads();
    // Every key is loaded before the first count: the compiler keeps a load behind the LDS atomics
    // of the previous item, which serialises the tile's loads.
    unsigned kd[SORT_ITEMS];
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i)
    {
        const int p = tr.base + i * BLOCK + threadIdx.x;
        kd[i] = p < n ? (unsigned)keys[p] : 0xffffffffu;
    }
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i)
        wave_count(s_h, (kd[i] >> shift) & (RADIX - 1), tr.base + i * BLOCK + (int)threadIdx.x < n);
    __syncthreads();
    for (int d = threadIdx.x; d < RADIX; d += BLOCK)
        thist[(size_t)d * tiles + blockIdx.x] = s_h[d];
}

// One workgroup per (digit, vehicle): exclusive scan over the vehicle's tiles plus the digit's
// base inside the vehicle.  Each thread scans a contiguous chunk of the row.
__global__ __launch_bounds__(BLOCK) void k_sort_scan(const VehDesc* __restrict__ descs, unsigned* __restrict__ thist,
                                                     int tiles, const unsigned* __restrict__ ghist, int ps)
{
    __shared__ unsigned s_w[SORT_WAVES];
    __shared__ unsigned s_base;
    const int d = blockIdx.x;
    const int v = blockIdx.y;
    const VehDesc& dv = descs[v];
    const unsigned* ghist_pass = ghist + ((size_t)v * 4 + ps) * RADIX;
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    if (threadIdx.x < WAVE)
    {
        unsigned b = 0;
#pragma unroll
        for (int k = 0; k < RADIX / WAVE; ++k)  // independent loads
        {
            const int e = lane + k * WAVE;
            b += e < d ? ghist_pass[e] : 0u;
        }
#pragma unroll
        for (int s = WAVE / 2; s > 0; s >>= 1)
            b += __shfl_xor(b, s, WAVE);
        if (lane == 0)
            s_base = b;
    }
    const int t0 = dv.u0[SP_P] / SORT_Q;
    const int nt = dv.nu[SP_P] / SORT_Q;
    unsigned* row = thist + (size_t)d * tiles + t0;
    const int per = (nt + BLOCK - 1) / BLOCK;
    const int b0 = min(nt, (int)threadIdx.x * per);
    const int b1 = min(nt, b0 + per);
    // Chunks of eight independent loads (a loop-carried load per tile would wait for each).
    unsigned sum = 0;
    for (int tc = b0; tc < b1; tc += 8)
    {
        unsigned x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            x[k] = tc + k < b1 ? row[tc + k] : 0u;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            sum += x[k];
    }
    unsigned incl = sum;
#pragma unroll
    for (int s = 1; s < WAVE; s <<= 1)
    {
        const unsigned u = __shfl_up(incl, s, WAVE);
        if (lane >= s)
            incl += u;
    }
    if (lane == WAVE - 1)
        s_w[wid] = incl;
    __syncthreads();
    unsigned run = s_base + incl - sum;
#pragma unroll
    for (int w = 0; w < SORT_WAVES; ++w)
        if (w < wid)
            run += s_w[w];
    for (int tc = b0; tc < b1; tc += 8)
    {
        unsigned x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            x[k] = tc + k < b1 ? row[tc + k] : 0u;
#pragma unroll
        for (int k = 0; k < 8; ++k)
        {
            if (tc + k < b1)
                row[tc + k] = run;
            run += x[k];
        }
    }
}

// Stable ranking and scatter of one tile.  Pass 0 ranks the particles' cell indices with identity
// values carrying the particle's association flag in bit 31 (read here in particle order, so the
// gather needs no random byte load); later passes read the previous pass's output.
// GATHER (last pass only): instead of the sorted keys and the permutation, write the particles in
// sorted order into part[pph ^ 1] (cell index, state, weight, association) and the weight copy into
// weight_array; the cell bounds are then recorded by the next segmented reduction.
template <bool GATHER>
__global__ __launch_bounds__(BLOCK) void k_sort_downsweep(const VehDesc* __restrict__ descs,
                                                          const uint32_t* __restrict__ tab, int pph, int ps, int tiles,
                                                          const unsigned* __restrict__ goff, int passes,
                                                          unsigned* __restrict__ ghist)
{
    __shared__ uint16_t s_wcnt[SORT_WAVES][RADIX];  // counts <= SORT_TILE; 16 bits keep 8 workgroups per CU
    __shared__ unsigned s_digoff[RADIX];
    __shared__ int s_keys[SORT_TILE];
    __shared__ int s_vals[SORT_TILE];
    __shared__ unsigned s_scan[SORT_WAVES];

    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const TileRef tr = tile_ref(tab, SORT_Q, tile);
    const VehDesc& dv = descs[tr.v];
    const int n = dv.N;
    const int tbase = tr.base;
    if (tbase >= n)
        return;
    const int shift = ps * RADIX_BITS;
    if (ps == passes - 1 && tbase == 0)
    {
        // The digit scans of every pass have run: clear this vehicle's histograms for the next sort.
        unsigned* gh = ghist + (size_t)tr.v * 4 * RADIX;
        for (int i = threadIdx.x; i < 4 * RADIX; i += BLOCK)
            gh[i] = 0u;
    }
    const int* keys_in = sort_keys_in(dv, pph, ps);
    const int* vals_in = ps == 0 ? nullptr : dv.sort_vals[(ps - 1) & 1];
    int* keys_out = dv.sort_keys[ps & 1];
    int* vals_out = dv.sort_vals[ps & 1];

    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    for (int i = threadIdx.x; i < SORT_WAVES * RADIX; i += BLOCK)
        (&s_wcnt[0][0])[i] = 0;
    __syncthreads();

    int key[SORT_ITEMS], val[SORT_ITEMS];
    unsigned rank[SORT_ITEMS];
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int j = 0; j < SORT_ITEMS; ++j)
    {
        const int p = tbase + (wid * SORT_ITEMS + j) * WAVE + lane;
        const bool valid = p < n;
        key[j] = valid ? keys_in[p] : 0;
        val[j] = valid ? (vals_in ? vals_in[p] : sort_value(p, dv.part[pph].assoc[p])) : 0;
        const unsigned d = ((unsigned)key[j] >> shift) & (RADIX - 1);
        const uint64_t peers = match_digit(d, __ballot(valid));
        unsigned r = 0;
        if (valid)
        {
            const unsigned before = s_wcnt[wid][d];
            r = before + (unsigned)__popcll(peers & lt);
            if ((peers & lt) == 0)
                s_wcnt[wid][d] = (uint16_t)(before + (unsigned)__popcll(peers));
        }
        rank[j] = r;
    }
    __syncthreads();

    // Per digit: tile total, offsets of the waves inside the digit, tile-local digit offsets.
    {
        const int d = threadIdx.x;  // BLOCK == RADIX
        unsigned run = 0;
#pragma unroll
        for (int w = 0; w < SORT_WAVES; ++w)
        {
            const unsigned c = s_wcnt[w][d];
            s_wcnt[w][d] = (uint16_t)run;
            run += c;
        }
        unsigned incl = run;
#pragma unroll
        for (int s = 1; s < WAVE; s <<= 1)
        {
            const unsigned u = __shfl_up(incl, s, WAVE);
            if (lane >= s)
                incl += u;
        }
        if (lane == WAVE - 1)
            s_scan[wid] = incl;
        __syncthreads();
        unsigned wc = 0;
#pragma unroll
        for (int w = 0; w < SORT_WAVES; ++w)
            if (w < wid)
                wc += s_scan[w];
        s_digoff[d] = wc + incl - run;
    }
    __syncthreads();

#pragma unroll
    for (int j = 0; j < SORT_ITEMS; ++j)
    {
        const int p = tbase + (wid * SORT_ITEMS + j) * WAVE + lane;
        if (p < n)
        {
            const unsigned d = ((unsigned)key[j] >> shift) & (RADIX - 1);
            const unsigned s = s_digoff[d] + s_wcnt[wid][d] + rank[j];
            s_keys[s] = key[j];
            s_vals[s] = val[j];
        }
    }
    __syncthreads();

    const int cnt = min(SORT_TILE, n - tbase);
    // Every destination offset of the thread's elements is loaded before the first store (the output
    // arrays may alias the offsets for the compiler, which would keep each load behind the previous
    // element's stores).
    constexpr int GB = SORT_TILE / BLOCK;
    int kk[GB], vv[GB];
    unsigned gg[GB];
#pragma unroll
    for (int b = 0; b < GB; ++b)
    {
        const int s = min((int)threadIdx.x + b * BLOCK, cnt - 1);
        kk[b] = s_keys[s];
        vv[b] = s_vals[s];
        const unsigned d = ((unsigned)kk[b] >> shift) & (RADIX - 1);
        gg[b] = goff[(size_t)d * tiles + tile] + ((unsigned)s - s_digoff[d]);
    }
#pragma unroll
    for (int b = 0; b < GB; ++b)
    {
        if ((int)threadIdx.x + b * BLOCK >= cnt)
            break;
        const unsigned g = gg[b];
        if constexpr (GATHER)
        {
            const DevParticles src = dv.part[pph];
            const DevParticles dst = dv.part[pph ^ 1];
            const int o = sort_value_index(vv[b]);
            const float w = src.weight[o];
            dst.pos[g] = src.pos[o];
            dst.vel[g] = src.vel[o];
            dst.weight[g] = w;
            dst.assoc[g] = sort_value_assoc(vv[b]);
            dst.idx[g] = kk[b];
            dv.weight_array[g] = w;
        }
        else
        {
            keys_out[g] = kk[b];
            vals_out[g] = vv[b];
        }
    }
}

static_assert(BLOCK == RADIX, "k_sort_downsweep assigns one digit per thread");

void run_sort(Engine* E, int pph, bool hist_done, bool gather)
{
    const int tiles = E->tiles(SP_P, SORT_Q);
    const int passes = E->passes;
    if (tiles <= 0)
        return;
    if (!hist_done)
    {
        if (!E->ghist_clean)
            run_fill_u32(E, E->sort_ghist, 0u, (long long)E->m * 4 * RADIX);
        E->ghist_clean = false;
        launch_begin(E, K_SORT_HIST);
        hipLaunchKernelGGL(k_sort_hist, dim3(tiles), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_P), pph, passes,
                           E->sort_ghist);
        launch_end(E, K_SORT_HIST);
    }

    for (int ps = 0; ps < passes; ++ps)
    {
        // With hist_done, predict has also stored the first pass's tile counts.
        if (ps > 0 || !hist_done)
        {
            launch_begin(E, K_SORT_UPSWEEP);
            hipLaunchKernelGGL(k_sort_upsweep, dim3(tiles), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_P), pph,
                               ps, tiles, E->sort_thist);
            launch_end(E, K_SORT_UPSWEEP);
        }
        launch_begin(E, K_SORT_SCAN);
        hipLaunchKernelGGL(k_sort_scan, dim3(RADIX, E->m), dim3(BLOCK), 0, E->stream, E->d_desc, E->sort_thist, tiles,
                           E->sort_ghist, ps);
        launch_end(E, K_SORT_SCAN);
        // The gathering last pass is timed as the sort's gather (the standalone gather's work).
        const bool g = gather && ps == passes - 1;
        launch_begin(E, g ? K_SORT_GATHER : K_SORT_DOWNSWEEP);
        if (g)
            hipLaunchKernelGGL(k_sort_downsweep<true>, dim3(tiles), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_P),
                               pph, ps, tiles, E->sort_thist, passes, E->sort_ghist);
        else
            hipLaunchKernelGGL(k_sort_downsweep<false>, dim3(tiles), dim3(BLOCK), 0, E->stream, E->d_desc,
                               E->tab(SP_P), pph, ps, tiles, E->sort_thist, passes, E->sort_ghist);
        launch_end(E, g ? K_SORT_GATHER : K_SORT_DOWNSWEEP);
    }
    E->ghist_clean = true;  // the last pass cleared the histograms
}

}  // namespace dogm_hip
