// Stable LSD radix sort of particle cell indices on gfx950 (replaces thrust::sort_by_key,
// dogm/src/dogm.cu:270-276).
//
// Keys are cell indices < grid_size^2, so only ceil(log2(C)) bits are sorted, 8 bits per pass.
// One histogram of every pass's digits is counted once (by predict when it runs in the same step).
// Two forms of a pass, chosen per step by the engine's tuning:
//  * reduce-then-scan: tile digit counts (predict stores the first pass's) -> per-vehicle,
//    per-digit exclusive scan over the vehicle's tiles plus the digit's base -> ranking and scatter,
//    tiles taken in an XCD-aware order;
//  * onesweep: a single launch per pass; a tile ranks its keys, obtains every digit's offset from
//    the earlier tiles of its vehicle by a decoupled lookback, and scatters.
// Both rank and scatter identically, so their results are identical.  Every vehicle of a step is
// sorted independently into its own buffers; all use the step's pass count.
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
    // Bit-sliced match on 32-bit halves: per bit, the lane's bit spread to 0 / all ones (e), the
    // wave's mask m of lanes with the bit set, and peers &= ~(m ^ e) (m where the bit is set, its
    // complement where not).
    uint32_t plo = (uint32_t)active, phi = (uint32_t)(active >> 32);
#pragma unroll
    for (int b = 0; b < RADIX_BITS; ++b)
    {
        const uint32_t e = (uint32_t)((int)(d << (31 - b)) >> 31);
        const uint64_t m = __ballot(e != 0u);
        plo &= ~((uint32_t)m ^ e);
        phi &= ~((uint32_t)(m >> 32) ^ e);
    }
    return ((uint64_t)phi << 32) | plo;
}

// Histogram of every pass's digit over all keys, per vehicle: ghist[(v * 4 + pass) * 256 + digit].
// Also clears the passes' tile tickets (the sort's passes run after it on the same stream).
__global__ __launch_bounds__(BLOCK) void k_sort_hist(const VehDesc* __restrict__ descs, const uint32_t* __restrict__ tab,
                                                     int pph, int passes, unsigned* __restrict__ ghist,
                                                     unsigned* __restrict__ tickets)
{
    __shared__ unsigned s_h[4][RADIX];
    if (blockIdx.x == 0 && threadIdx.x < 4)
        tickets[threadIdx.x] = 0u;
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

// ---- reduce-then-scan form: per pass an upsweep of tile digit counts, a scan per (vehicle, digit)
//      over the vehicle's tiles, and the ranking and scatter -----------------------------------

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
// sorted order into part[pph ^ 1] (cell index, state, weight, association) and, with GATHER_WA, the
// weight copy into weight_array; the cell bounds are then recorded by the next segmented reduction.
// With GATHER_AOS the state is read from the 16-byte records predict left in the CDF buffer, and
// for vehicles with uniform weights the predicted weight is formed here (predicted_weight).
template <bool GATHER>
__global__ __launch_bounds__(BLOCK) void k_sort_downsweep(const VehDesc* __restrict__ descs,
                                                          const uint32_t* __restrict__ tab, int pph, int ps, int tiles,
                                                          const unsigned* __restrict__ goff, int passes,
                                                          unsigned* __restrict__ ghist, int gflags)
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
    // The digit's destination base for this tile: one load per digit instead of one per key.
    const unsigned gbase = goff[(size_t)threadIdx.x * tiles + tile];
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
    // From here on s_digoff holds destination minus tile-local slot per digit.
    s_digoff[threadIdx.x] = gbase - s_digoff[threadIdx.x];
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
        gg[b] = s_digoff[d] + (unsigned)s;
    }
    // Uniform weights: every particle's prior weight is weight[0], which predict left in place.
    const float w_prior = GATHER && (gflags & GATHER_AOS) && dv.wuni ? dv.part[pph].weight[0] : 0.0f;
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
            float w;
            if (gflags & GATHER_AOS)
            {
                const float4 st = reinterpret_cast<const float4*>(dv.cdf)[o];
                dst.pos[g] = make_float2(st.x, st.y);
                dst.vel[g] = make_float2(st.z, st.w);
                w = dv.wuni ? predicted_weight(dv, w_prior, st) : src.weight[o];
            }
            else
            {
                w = src.weight[o];
                dst.pos[g] = src.pos[o];
                dst.vel[g] = src.vel[o];
            }
            dst.weight[g] = w;
            dst.assoc[g] = sort_value_assoc(vv[b]);
            dst.idx[g] = kk[b];
            if (gflags & GATHER_WA)
                dv.weight_array[g] = w;
        }
        else
        {
            keys_out[g] = kk[b];
            vals_out[g] = vv[b];
        }
    }
}

// ---- onesweep form: one launch per pass, digit offsets by decoupled lookback -----------------

// Decoupled-lookback status word of one (tile, digit) of a pass: epoch (2 bits, never 0 and
// different from the epoch of the pass that wrote the buffer last, so its words never read as
// published), flag (AGGREGATE: the tile's own count; PREFIX: the count of this digit in the
// vehicle's earlier tiles and this one, plus the digit's base) and a 28-bit value (counts are below
// 2^28 per vehicle).  Flag and value travel in one 32-bit atomic word, so publishing needs no fence.
constexpr unsigned LB_AGGREGATE = 1u, LB_PREFIX = 2u;
constexpr unsigned LB_VALUE_MASK = (1u << 28) - 1u;

// Epochs of the passes of a sort with `passes` > 1 passes: 1, 2, 3, 2, ..., with the last pass's
// different from both its predecessor's and 1, so every pass differs from the pass before it, also
// across sorts (the first pass of the next sort uses 1).  A fixed function of the pass index, so a
// captured step graph replays correct epochs.  One-pass sorts clear the buffer before their pass.
inline unsigned lb_epoch(int ps, int passes)
{
    if (ps == 0)
        return 1u;
    if (ps < passes - 1)
        return ps % 2 ? 2u : 3u;
    const unsigned prev = lb_epoch(ps - 1, passes);
    return prev == 2u ? 3u : 2u;
}

__device__ __forceinline__ unsigned lb_word(unsigned epoch, unsigned flag, unsigned value)
{
    return epoch << 30 | flag << 28 | value;
}

__device__ __forceinline__ void lb_publish(unsigned* w, unsigned v)
{
    __hip_atomic_store(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned lb_read(const unsigned* w)
{
    return __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Onesweep pass: ranking, digit offsets and scatter of one tile in a single launch per pass.
//
// Tiles are taken in the order workgroups start (an atomic ticket), so a tile only ever waits for
// tiles that are already running.  Per digit (one per thread) a tile publishes its own count at
// once (AGGREGATE), then walks back over the vehicle's earlier tiles adding their published values
// until it meets a PREFIX, and publishes its own PREFIX.  A vehicle's first tile publishes PREFIX
// straight away, with the digit's base inside the vehicle (the exclusive scan of the pass's
// histogram, counted by predict or k_sort_hist).  Every wait is bounded by spin_limit polls; a wait
// that runs out takes 0, sets *err (read by the host after the step: DOGM_ERR_DEVICE_CHECK) and
// the kernel still finishes (destinations only ever come out too small, so every store stays
// inside the vehicle's arrays).  spin_limit < 0 sets *err from every tile that looks back but
// waits as usual (a host-forced failure report for tests of that path).
//
// Ranking is stable: inside a tile, keys are laid out wave-striped (wave w, item j, lane l =
// position (w*ITEMS + j)*64 + l), each wave ranks its 64 keys of one item with ballot-matched peer
// masks against per-wave digit counters in LDS, and waves are combined in order.  The scatter is
// staged through LDS in digit-sorted order so that equal digits leave the tile as contiguous runs.
// Pass 0 ranks the particles' cell indices with identity values carrying the particle's association
// flag in bit 31 (read here in particle order, so the gather needs no random byte load); later
// passes read the previous pass's output.
// GATHER (last pass only): instead of the sorted keys and the permutation, write the particles in
// sorted order into part[pph ^ 1] (cell index, state, weight, association) and, with GATHER_WA, the
// weight copy into weight_array; the cell bounds are then recorded by the next segmented reduction.
// With GATHER_AOS the state is read from the 16-byte records predict left in the CDF buffer, and
// for vehicles with uniform weights the predicted weight is formed here (predicted_weight).
template <bool GATHER>
__global__ __launch_bounds__(BLOCK) void k_sort_onesweep(const VehDesc* __restrict__ descs,
                                                         const uint32_t* __restrict__ tab, int pph, int ps, int passes,
                                                         unsigned* __restrict__ ghist, unsigned* __restrict__ status,
                                                         unsigned* __restrict__ tickets, unsigned epoch,
                                                         int spin_limit, unsigned* __restrict__ err, int gflags)
{
    __shared__ uint16_t s_wcnt[SORT_WAVES][RADIX];  // counts <= SORT_TILE; 16 bits keep 8 workgroups per CU
    __shared__ unsigned s_digoff[RADIX];
    __shared__ unsigned s_gbase[RADIX];
    __shared__ int s_keys[SORT_TILE];
    __shared__ int s_vals[SORT_TILE];
    __shared__ unsigned s_scan[SORT_WAVES];
    __shared__ int s_tile;

    if (threadIdx.x == 0)
        s_tile = (int)atomicAdd(&tickets[ps], 1u);
    for (int i = threadIdx.x; i < SORT_WAVES * RADIX; i += BLOCK)
        (&s_wcnt[0][0])[i] = 0;
    __syncthreads();
    const int tile = s_tile;
    const TileRef tr = tile_ref(tab, SORT_Q, tile);
    const VehDesc& dv = descs[tr.v];
    const int n = dv.N;
    const int tbase = tr.base;
    if (tbase >= n)
        return;  // padding past the vehicle's particles: no later tile of the vehicle looks back here
    const int shift = ps * RADIX_BITS;
    const int* keys_in = sort_keys_in(dv, pph, ps);
    const int* vals_in = ps == 0 ? nullptr : dv.sort_vals[(ps - 1) & 1];
    int* keys_out = dv.sort_keys[ps & 1];
    int* vals_out = dv.sort_vals[ps & 1];

    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;

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

    const int d = threadIdx.x;  // BLOCK == RADIX: one digit per thread
    // Tile count of the digit and the offsets of the waves inside it.
    unsigned count = 0;
#pragma unroll
    for (int w = 0; w < SORT_WAVES; ++w)
    {
        const unsigned c = s_wcnt[w][d];
        s_wcnt[w][d] = (uint16_t)count;
        count += c;
    }
    unsigned* st = status + (size_t)tile * RADIX;
    unsigned excl = 0;  // digit d in the vehicle's earlier tiles, plus the digit's base
    if (tbase == 0)
    {
        // Digit base: exclusive scan of the pass's histogram over the digits (block scan).
        const unsigned h = ghist[((size_t)tr.v * 4 + ps) * RADIX + d];
        unsigned incl = h;
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
        excl = incl - h;
#pragma unroll
        for (int w = 0; w < SORT_WAVES; ++w)
            if (w < wid)
                excl += s_scan[w];
        lb_publish(&st[d], lb_word(epoch, LB_PREFIX, excl + count));
        __syncthreads();  // s_scan is reused below
        if (ps == passes - 1)
        {
            // Every pass has read its histogram: clear this vehicle's for the next sort.
            unsigned* gh = ghist + (size_t)tr.v * 4 * RADIX;
            for (int i = threadIdx.x; i < 4 * RADIX; i += BLOCK)
                gh[i] = 0u;
        }
    }
    else
    {
        lb_publish(&st[d], lb_word(epoch, LB_AGGREGATE, count));
        // Walk back in windows of LB_WINDOW tiles read at once (each read crosses to the memory
        // side, so reading one tile after the other would pay that latency per tile); the
        // vehicle's first tile always holds a PREFIX, so the walk ends there at the latest.
        constexpr int LB_WINDOW = 8;
        const int first = tile - tbase / SORT_TILE;
        // A negative spin_limit reports a give-up from every tile that looks back but still waits
        // as usual (a host-forced failure report whose results stay those of a normal pass).
        const int limit = spin_limit < 0 ? (1 << 20) : spin_limit;
        if (spin_limit < 0 && d == 0)
            *reinterpret_cast<volatile unsigned*>(err) = 1u;
        bool done = false;
        for (int j0 = tile - 1; !done; j0 -= LB_WINDOW)
        {
            unsigned x[LB_WINDOW];
#pragma unroll
            for (int k = 0; k < LB_WINDOW; ++k)
                x[k] = j0 - k >= first ? lb_read(status + (size_t)(j0 - k) * RADIX + d) : 0u;
#pragma unroll
            for (int k = 0; k < LB_WINDOW; ++k)
            {
                if (done || j0 - k < first)
                    continue;
                unsigned v = x[k];
                bool ready = (v >> 30) == epoch && ((v >> 28) & 3u) != 0u;
                for (int it = 0; !ready && it < limit; ++it)
                {
                    __builtin_amdgcn_s_sleep(1);
                    v = lb_read(status + (size_t)(j0 - k) * RADIX + d);
                    ready = (v >> 30) == epoch && ((v >> 28) & 3u) != 0u;
                }
                if (!ready)
                {
                    *reinterpret_cast<volatile unsigned*>(err) = 1u;
                    done = true;
                    continue;
                }
                excl += v & LB_VALUE_MASK;
                done = ((v >> 28) & 3u) == LB_PREFIX;
            }
            if (!done && j0 - LB_WINDOW < first)
            {
                // Reached the vehicle's first tile without a PREFIX: cannot happen unless a word
                // was overwritten; report it rather than walk on.
                *reinterpret_cast<volatile unsigned*>(err) = 1u;
                done = true;
            }
        }
        lb_publish(&st[d], lb_word(epoch, LB_PREFIX, excl + count));
    }
    s_gbase[d] = excl;

    // Tile-local exclusive digit offsets.
    {
        unsigned incl = count;
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
        s_digoff[d] = wc + incl - count;
    }
    __syncthreads();

#pragma unroll
    for (int j = 0; j < SORT_ITEMS; ++j)
    {
        const int p = tbase + (wid * SORT_ITEMS + j) * WAVE + lane;
        if (p < n)
        {
            const unsigned dj = ((unsigned)key[j] >> shift) & (RADIX - 1);
            const unsigned s = s_digoff[dj] + s_wcnt[wid][dj] + rank[j];
            s_keys[s] = key[j];
            s_vals[s] = val[j];
        }
    }
    __syncthreads();

    const int cnt = min(SORT_TILE, n - tbase);
    // Every destination of the thread's elements is computed before the first store (the output
    // arrays may alias LDS-derived values for the compiler, which would keep each load behind the
    // previous element's stores).
    constexpr int GB = SORT_TILE / BLOCK;
    int kk[GB], vv[GB];
    unsigned gg[GB];
#pragma unroll
    for (int b = 0; b < GB; ++b)
    {
        const int s = min((int)threadIdx.x + b * BLOCK, cnt - 1);
        kk[b] = s_keys[s];
        vv[b] = s_vals[s];
        const unsigned db = ((unsigned)kk[b] >> shift) & (RADIX - 1);
        gg[b] = s_gbase[db] + ((unsigned)s - s_digoff[db]);
    }
    // Uniform weights: every particle's prior weight is weight[0], which predict left in place.
    const float w_prior = GATHER && (gflags & GATHER_AOS) && dv.wuni ? dv.part[pph].weight[0] : 0.0f;
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
            float w;
            if (gflags & GATHER_AOS)
            {
                const float4 st = reinterpret_cast<const float4*>(dv.cdf)[o];
                dst.pos[g] = make_float2(st.x, st.y);
                dst.vel[g] = make_float2(st.z, st.w);
                w = dv.wuni ? predicted_weight(dv, w_prior, st) : src.weight[o];
            }
            else
            {
                w = src.weight[o];
                dst.pos[g] = src.pos[o];
                dst.vel[g] = src.vel[o];
            }
            dst.weight[g] = w;
            dst.assoc[g] = sort_value_assoc(vv[b]);
            dst.idx[g] = kk[b];
            if (gflags & GATHER_WA)
                dv.weight_array[g] = w;
        }
        else
        {
            keys_out[g] = kk[b];
            vals_out[g] = vv[b];
        }
    }
}

static_assert(BLOCK == RADIX, "the sort kernels assign one digit per thread");

void run_sort(Engine* E, int pph, bool hist_done, bool gather, bool write_wa)
{
    const int tiles = E->tiles(SP_P, SORT_Q);
    const int passes = E->passes;
    if (tiles <= 0)
        return;
    const bool onesweep = E->tune.sort == SORT_ONESWEEP;
    // predict of the same step left the state as 16-byte records in the CDF buffer (run_predict)
    const bool aos = hist_done && gather;
    const int gflags = (aos ? GATHER_AOS : 0) | (write_wa ? GATHER_WA : 0);
    if (onesweep && (E->status_plan != E->plan_epoch || passes == 1))
    {
        // New tiles (or a new buffer, or words of the other form): no word may look published.
        run_fill_u32(E, E->sort_status, 0u, (long long)tiles * RADIX);
        E->status_plan = E->plan_epoch;
    }
    if (!onesweep)
        E->status_plan = -1;  // the buffer now holds tile counts
    if (!hist_done)
    {
        if (!E->ghist_clean)
            run_fill_u32(E, E->sort_ghist, 0u, (long long)E->m * 4 * RADIX);
        E->ghist_clean = false;
        launch_begin(E, K_SORT_HIST);
        hipLaunchKernelGGL(k_sort_hist, dim3(tiles), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_P), pph, passes,
                           E->sort_ghist, E->sort_tickets);
        launch_end(E, K_SORT_HIST);
    }

    for (int ps = 0; ps < passes; ++ps)
    {
        // The gathering last pass is timed as the sort's gather (the standalone gather's work).
        const bool g = gather && ps == passes - 1;
        if (onesweep)
        {
            const unsigned epoch = passes == 1 ? 1u : lb_epoch(ps, passes);
            launch_begin(E, g ? K_SORT_GATHER : K_SORT_DOWNSWEEP);
            if (g)
                hipLaunchKernelGGL(k_sort_onesweep<true>, dim3(tiles), dim3(BLOCK), 0, E->stream, E->d_desc,
                                   E->tab(SP_P), pph, ps, passes, E->sort_ghist, E->sort_status, E->sort_tickets,
                                   epoch, E->lookback_spin, E->d_err, gflags);
            else
                hipLaunchKernelGGL(k_sort_onesweep<false>, dim3(tiles), dim3(BLOCK), 0, E->stream, E->d_desc,
                                   E->tab(SP_P), pph, ps, passes, E->sort_ghist, E->sort_status, E->sort_tickets,
                                   epoch, E->lookback_spin, E->d_err, 0);
            launch_end(E, g ? K_SORT_GATHER : K_SORT_DOWNSWEEP);
            continue;
        }
        // With hist_done, predict has also stored the first pass's tile counts.
        if (ps > 0 || !hist_done)
        {
            launch_begin(E, K_SORT_UPSWEEP);
            hipLaunchKernelGGL(k_sort_upsweep, dim3(tiles), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_P), pph,
                               ps, tiles, E->sort_status);
            launch_end(E, K_SORT_UPSWEEP);
        }
        launch_begin(E, K_SORT_SCAN);
        hipLaunchKernelGGL(k_sort_scan, dim3(RADIX, E->m), dim3(BLOCK), 0, E->stream, E->d_desc, E->sort_status, tiles,
                           E->sort_ghist, ps);
        launch_end(E, K_SORT_SCAN);
        launch_begin(E, g ? K_SORT_GATHER : K_SORT_DOWNSWEEP);
        if (g)
            hipLaunchKernelGGL(k_sort_downsweep<true>, dim3(tiles), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_P),
                               pph, ps, tiles, E->sort_status, passes, E->sort_ghist, gflags);
        else
            hipLaunchKernelGGL(k_sort_downsweep<false>, dim3(tiles), dim3(BLOCK), 0, E->stream, E->d_desc,
                               E->tab(SP_P), pph, ps, tiles, E->sort_status, passes, E->sort_ghist, 0);
        launch_end(E, g ? K_SORT_GATHER : K_SORT_DOWNSWEEP);
    }
    E->ghist_clean = true;  // the last pass cleared the histograms
}

}  // namespace dogm_hip
