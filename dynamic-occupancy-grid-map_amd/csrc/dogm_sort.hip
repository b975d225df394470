// Stable LSD radix sort of particle cell indices on gfx950 (replaces thrust::sort_by_key,
// dogm/src/dogm.cu:270-276).
//
// Keys are cell indices < grid_size^2, so only ceil(log2(C)) bits are sorted, 8 bits per pass.
// Per pass: tile digit counts -> per-digit exclusive scan over tiles (plus the digit's global base
// from a one-time histogram of all passes) -> ranking and scatter.  Ranking is stable: inside a
// tile, keys are laid out wave-striped (wave w, item j, lane l = position (w*ITEMS + j)*64 + l),
// each wave ranks its 64 keys of one item with ballot-matched peer masks against per-wave digit
// counters in LDS, and waves are combined in order.  The scatter is staged through LDS in
// digit-sorted order so that equal digits leave the tile as contiguous runs.
#include "dogm_device.h"

namespace dogm_hip
{

constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;
constexpr int SORT_ITEMS = 8;
constexpr int SORT_TILE = BLOCK * SORT_ITEMS;
constexpr int SORT_WAVES = BLOCK / WAVE;

__device__ __forceinline__ uint64_t lanemask_lt()
{
    const int lane = threadIdx.x & (WAVE - 1);
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

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

// Adds one to s_h[d] for every active lane.  A wave whose keys share the digit (the usual case
// for the upper digits: resampled particles arrive in cell order) adds once; otherwise lanes add
// individually.
__device__ __forceinline__ void wave_count(unsigned* s_h, unsigned d, bool valid)
{
    const uint64_t active = __ballot(valid);
    if (active == 0)
        return;
    const unsigned first = __shfl(d, (int)__builtin_ctzll(active), WAVE);
    if ((__ballot(d == first) & active) == active)
    {
        if (valid && (active & lanemask_lt()) == 0)
            atomicAdd(&s_h[d], (unsigned)__popcll(active));
    }
    else if (valid)
    {
        atomicAdd(&s_h[d], 1u);
    }
}

// Histogram of every pass's digit over all keys.
__global__ __launch_bounds__(BLOCK) void k_sort_hist(const int* __restrict__ keys, int n, int passes,
                                                     unsigned* __restrict__ ghist)
{
    __shared__ unsigned s_h[4][RADIX];
    for (int i = threadIdx.x; i < 4 * RADIX; i += BLOCK)
        (&s_h[0][0])[i] = 0u;
    __syncthreads();
    const int base = blockIdx.x * SORT_TILE;
    for (int i = 0; i < SORT_ITEMS; ++i)
    {
        const int p = base + i * BLOCK + threadIdx.x;
        const bool valid = p < n;
        const unsigned k = valid ? (unsigned)keys[p] : 0u;
        for (int ps = 0; ps < passes; ++ps)
            wave_count(s_h[ps], (k >> (ps * RADIX_BITS)) & (RADIX - 1), valid);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * RADIX; i += BLOCK)
    {
        const unsigned v = (&s_h[0][0])[i];
        if (v)
            atomicAdd(&ghist[i], v);
    }
}

// Digit counts of one tile for one pass: thist[d * tiles + t].
__global__ __launch_bounds__(BLOCK) void k_sort_upsweep(const int* __restrict__ keys, int n, int shift, int tiles,
                                                        unsigned* __restrict__ thist)
{
    __shared__ unsigned s_h[RADIX];
    for (int i = threadIdx.x; i < RADIX; i += BLOCK)
        s_h[i] = 0u;
    __syncthreads();
    const int base = blockIdx.x * SORT_TILE;
    for (int i = 0; i < SORT_ITEMS; ++i)
    {
        const int p = base + i * BLOCK + threadIdx.x;
        const bool valid = p < n;
        wave_count(s_h, valid ? ((unsigned)keys[p] >> shift) & (RADIX - 1) : 0u, valid);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < RADIX; d += BLOCK)
        thist[(size_t)d * tiles + blockIdx.x] = s_h[d];
}

// One workgroup per digit: exclusive scan over tiles plus the digit's global base.  Each thread
// scans a contiguous chunk of the digit's row.
__global__ __launch_bounds__(BLOCK) void k_sort_scan(unsigned* __restrict__ thist, int tiles,
                                                     const unsigned* __restrict__ ghist_pass)
{
    __shared__ unsigned s_w[SORT_WAVES];
    __shared__ unsigned s_base;
    const int d = blockIdx.x;
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    if (threadIdx.x < WAVE)
    {
        unsigned b = 0;
        for (int e = lane; e < d; e += WAVE)
            b += ghist_pass[e];
#pragma unroll
        for (int s = WAVE / 2; s > 0; s >>= 1)
            b += __shfl_xor(b, s, WAVE);
        if (lane == 0)
            s_base = b;
    }
    unsigned* row = thist + (size_t)d * tiles;
    const int per = (tiles + BLOCK - 1) / BLOCK;
    const int b0 = min(tiles, (int)threadIdx.x * per);
    const int b1 = min(tiles, b0 + per);
    unsigned sum = 0;
    for (int t = b0; t < b1; ++t)
        sum += row[t];
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
    for (int t = b0; t < b1; ++t)
    {
        const unsigned x = row[t];
        row[t] = run;
        run += x;
    }
}

// Stable ranking and scatter of one tile.  vals_in == nullptr means identity values.
__global__ __launch_bounds__(BLOCK) void k_sort_downsweep(const int* __restrict__ keys_in,
                                                          const int* __restrict__ vals_in, int* __restrict__ keys_out,
                                                          int* __restrict__ vals_out, int n, int shift, int tiles,
                                                          const unsigned* __restrict__ goff)
{
    __shared__ unsigned s_wcnt[SORT_WAVES][RADIX];
    __shared__ unsigned s_digoff[RADIX];
    __shared__ int s_keys[SORT_TILE];
    __shared__ int s_vals[SORT_TILE];
    __shared__ unsigned s_scan[SORT_WAVES];

    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    const int tile = blockIdx.x;
    const int tbase = tile * SORT_TILE;
    for (int i = threadIdx.x; i < SORT_WAVES * RADIX; i += BLOCK)
        (&s_wcnt[0][0])[i] = 0u;
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
        val[j] = valid ? (vals_in ? vals_in[p] : p) : 0;
        const unsigned d = ((unsigned)key[j] >> shift) & (RADIX - 1);
        const uint64_t peers = match_digit(d, __ballot(valid));
        unsigned r = 0;
        if (valid)
        {
            const unsigned before = s_wcnt[wid][d];
            r = before + (unsigned)__popcll(peers & lt);
            if ((peers & lt) == 0)
                s_wcnt[wid][d] = before + (unsigned)__popcll(peers);
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
            s_wcnt[w][d] = run;
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
    for (int s = threadIdx.x; s < cnt; s += BLOCK)
    {
        const int k = s_keys[s];
        const unsigned d = ((unsigned)k >> shift) & (RADIX - 1);
        const unsigned g = goff[(size_t)d * tiles + tile] + ((unsigned)s - s_digoff[d]);
        keys_out[g] = k;
        vals_out[g] = s_vals[s];
    }
}

static_assert(BLOCK == RADIX, "k_sort_downsweep assigns one digit per thread");

void run_sort(dogm_handle* h, const int** skeys, const int** perm)
{
    const int n = h->N;
    const int tiles = div_up(n, SORT_TILE);
    const int passes = h->sort_passes;
    const int* kin = h->part[h->cur].idx;
    const int* vin = nullptr;

    run_fill_u32(h, h->sort_ghist, 0u, (long long)passes * RADIX);
    launch_begin(h, K_SORT_HIST);
    hipLaunchKernelGGL(k_sort_hist, dim3(tiles), dim3(BLOCK), 0, h->stream, kin, n, passes, h->sort_ghist);
    launch_end(h, K_SORT_HIST);

    for (int ps = 0; ps < passes; ++ps)
    {
        const int shift = ps * RADIX_BITS;
        int* kout = h->sort_keys[ps & 1];
        int* vout = h->sort_vals[ps & 1];
        launch_begin(h, K_SORT_UPSWEEP);
        hipLaunchKernelGGL(k_sort_upsweep, dim3(tiles), dim3(BLOCK), 0, h->stream, kin, n, shift, tiles,
                           h->sort_thist);
        launch_end(h, K_SORT_UPSWEEP);
        launch_begin(h, K_SORT_SCAN);
        hipLaunchKernelGGL(k_sort_scan, dim3(RADIX), dim3(BLOCK), 0, h->stream, h->sort_thist, tiles,
                           h->sort_ghist + ps * RADIX);
        launch_end(h, K_SORT_SCAN);
        launch_begin(h, K_SORT_DOWNSWEEP);
        hipLaunchKernelGGL(k_sort_downsweep, dim3(tiles), dim3(BLOCK), 0, h->stream, kin, vin, kout, vout, n, shift,
                           tiles, h->sort_thist);
        launch_end(h, K_SORT_DOWNSWEEP);
        kin = kout;
        vin = vout;
    }
    *skeys = kin;
    *perm = vin;
}

}  // namespace dogm_hip
