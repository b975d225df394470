// Device building blocks: exact (128-bit fixed-point) block scans and the exact tile-segmented
// reduction used for every per-cell sum of the filter.
#pragma once

#include "dogm_internal.h"

namespace dogm_hip
{

using namespace dogm_num;

__device__ __forceinline__ fx_t shfl_fx(fx_t x, int src)
{
    long long lo = (long long)(uint64_t)x;
    long long hi = (long long)(x >> 64);
    lo = __shfl(lo, src, WAVE);
    hi = __shfl(hi, src, WAVE);
    return (fx_t)(((ufx_t)(uint64_t)hi << 64) | (ufx_t)(uint64_t)lo);
}

// DPP move of the four dwords of an fx value (lanes whose source is outside the pattern, or whose
// row is masked off, read 0).  CTRL / ROW_MASK are DPP encodings: row_shr:n = 0x110 + n,
// row_bcast:15 = 0x142, row_bcast:31 = 0x143.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ fx_t dpp_fx(fx_t x)
{
    const ufx_t u = (ufx_t)x;
    const int w0 = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROW_MASK, 0xf, false);
    const int w1 = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROW_MASK, 0xf, false);
    const int w2 = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 64), CTRL, ROW_MASK, 0xf, false);
    const int w3 = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 96), CTRL, ROW_MASK, 0xf, false);
    return (fx_t)(((ufx_t)(uint32_t)w3 << 96) | ((ufx_t)(uint32_t)w2 << 64) | ((ufx_t)(uint32_t)w1 << 32) |
                  (ufx_t)(uint32_t)w0);
}

// Inclusive scan over the 64 lanes of a wave: Hillis-Steele inside rows of 16 lanes (row_shr 1, 2,
// 4, 8), then the row totals through row_bcast:15 (rows 1 and 3) and row_bcast:31 (rows 2 and 3).
// Integer adds: the result does not depend on the association.
__device__ __forceinline__ fx_t wave_inclusive_scan_fx(fx_t x)
{
    x += dpp_fx<0x111, 0xf>(x);
    x += dpp_fx<0x112, 0xf>(x);
    x += dpp_fx<0x114, 0xf>(x);
    x += dpp_fx<0x118, 0xf>(x);
    x += dpp_fx<0x142, 0xa>(x);
    x += dpp_fx<0x143, 0xc>(x);
    return x;
}

// DPP moves and the same wave scan for f32 / f64 values (float CDF modes).  The association is
// fixed by the six DPP steps: value + shifted value, rows of 16 first, then the row totals.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_val(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROW_MASK, 0xf, false));
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_val(double x)
{
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROW_MASK, 0xf, false);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ f32c dpp_val(f32c x)
{
    f32c r;
    r.hi = dpp_val<CTRL, ROW_MASK>(x.hi);
    r.lo = dpp_val<CTRL, ROW_MASK>(x.lo);
    return r;
}

// Value of lane - 1 (f32 / f64 / compensated f32); lane 0 gets its own value.
__device__ __forceinline__ float shfl_up1(float x)
{
    return __shfl_up(x, 1, WAVE);
}
__device__ __forceinline__ double shfl_up1(double x)
{
    return __shfl_up(x, 1, WAVE);
}
__device__ __forceinline__ f32c shfl_up1(f32c x)
{
    f32c r;
    r.hi = __shfl_up(x.hi, 1, WAVE);
    r.lo = __shfl_up(x.lo, 1, WAVE);
    return r;
}

template <class T>
__device__ __forceinline__ T wave_inclusive_scan_val(T x)
{
    x = x + dpp_val<0x111, 0xf>(x);
    x = x + dpp_val<0x112, 0xf>(x);
    x = x + dpp_val<0x114, 0xf>(x);
    x = x + dpp_val<0x118, 0xf>(x);
    x = x + dpp_val<0x142, 0xa>(x);
    x = x + dpp_val<0x143, 0xc>(x);
    return x;
}

// Running maximum over the lanes of a wave for non-negative values (max is exact in any order).
template <class T>
__device__ __forceinline__ T wave_inclusive_max_val(T x)
{
    { const T y = dpp_val<0x111, 0xf>(x); x = y > x ? y : x; }
    { const T y = dpp_val<0x112, 0xf>(x); x = y > x ? y : x; }
    { const T y = dpp_val<0x114, 0xf>(x); x = y > x ? y : x; }
    { const T y = dpp_val<0x118, 0xf>(x); x = y > x ? y : x; }
    { const T y = dpp_val<0x142, 0xa>(x); x = y > x ? y : x; }
    { const T y = dpp_val<0x143, 0xc>(x); x = y > x ? y : x; }
    return x;
}

// Exclusive block scan of one fx value per thread (BLOCK threads); returns the exclusive prefix
// and writes the block total.
__device__ __forceinline__ fx_t block_exclusive_scan_fx(fx_t v, fx_t* total)
{
    __shared__ fx_t s_wave[BLOCK / WAVE];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    const fx_t incl = wave_inclusive_scan_fx(v);
    if (lane == WAVE - 1)
        s_wave[wid] = incl;
    __syncthreads();
    fx_t carry = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / WAVE; ++w)
    {
        const fx_t x = s_wave[w];
        if (w < wid)
            carry += x;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return carry + incl - v;
}

// Stores of streaming outputs, which no later work of the same launch reads and which are far out of
// the caches by the time the next stage reads them: vector stores with the nontemporal hint.
template <class T>
__device__ __forceinline__ void stream_store(T* p, T v)
{
    __builtin_nontemporal_store(v, p);
}

__device__ __forceinline__ void stream_store(float2* p, float2 v)
{
    using V = float __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(V{v.x, v.y}, reinterpret_cast<V*>(p));
}

__device__ __forceinline__ void stream_store(float4* p, float4 v)
{
    using V = float __attribute__((ext_vector_type(4)));
    __builtin_nontemporal_store(V{v.x, v.y, v.z, v.w}, reinterpret_cast<V*>(p));
}

__device__ __forceinline__ void stream_store(dogm_measurement_cell* p, dogm_measurement_cell v)
{
    stream_store(reinterpret_cast<float4*>(p), make_float4(v.free_mass, v.occ_mass, v.likelihood, v.p_A));
}

// Predicted weight of a particle (dogm/src/kernel/predict.cu:34-43): persistence probability times
// the prior weight, 0 once the predicted position o leaves the grid.  predict and the gathering sort
// pass (uniform weights) both form it here, so the two give the same bits.
__device__ __forceinline__ float predicted_weight(const VehDesc& d, float w_in, const float4& o)
{
    const float zero = 0.0f;
    const float top = (float)(d.gs - 1);
    float w = d.ps * w_in;
    if ((o.x > top || o.x < zero) || (o.y > top || o.y < zero))
        w = zero;
    return w;
}

// ---------------------------------------------------------------------------------------------
// Vector access to a thread's ITEMS consecutive elements of 32-bit values: 16-byte (ITEMS a
// multiple of 4) or 8-byte (ITEMS even) accesses when the run is complete; runs start at multiples
// of ITEMS.
// ---------------------------------------------------------------------------------------------
template <int IT>
constexpr int run_width()
{
    return IT % 4 == 0 ? 4 : (IT % 2 == 0 ? 2 : 1);
}

template <int IT, class T>
__device__ __forceinline__ void load_full_run(const T* __restrict__ a, int p0, T (&v)[IT])
{
    static_assert(sizeof(T) == 4, "32-bit elements");
    constexpr int W = run_width<IT>();
#pragma unroll
    for (int i = 0; i < IT; i += W)
    {
        if constexpr (W == 4)
        {
            const int4 q = *reinterpret_cast<const int4*>(a + p0 + i);
            __builtin_memcpy(&v[i], &q, 16);
        }
        else if constexpr (W == 2)
        {
            const int2 q = *reinterpret_cast<const int2*>(a + p0 + i);
            __builtin_memcpy(&v[i], &q, 8);
        }
        else
        {
            v[i] = a[p0 + i];
        }
    }
}

template <int IT>
__device__ __forceinline__ void load_run(const float* __restrict__ a, int p0, int cnt, float (&v)[IT])
{
    if (cnt == IT)
    {
        load_full_run<IT, float>(a, p0, v);
    }
    else
    {
#pragma unroll
        for (int i = 0; i < IT; ++i)
            v[i] = i < cnt ? a[p0 + i] : 0.0f;
    }
}

template <int IT>
__device__ __forceinline__ void load_run(const int* __restrict__ a, int p0, int cnt, int (&v)[IT], int fill)
{
    if (cnt == IT)
    {
        load_full_run<IT, int>(a, p0, v);
    }
    else
    {
#pragma unroll
        for (int i = 0; i < IT; ++i)
            v[i] = i < cnt ? a[p0 + i] : fill;
    }
}

template <int IT, class T>
__device__ __forceinline__ void store_run(T* __restrict__ a, int p0, int cnt, const T (&v)[IT])
{
    static_assert(sizeof(T) == 4, "32-bit elements");
    constexpr int W = run_width<IT>();
    if (cnt == IT)
    {
#pragma unroll
        for (int i = 0; i < IT; i += W)
        {
            if constexpr (W == 4)
            {
                int4 q;
                __builtin_memcpy(&q, &v[i], 16);
                *reinterpret_cast<int4*>(a + p0 + i) = q;
            }
            else if constexpr (W == 2)
            {
                int2 q;
                __builtin_memcpy(&q, &v[i], 8);
                *reinterpret_cast<int2*>(a + p0 + i) = q;
            }
            else
            {
                a[p0 + i] = v[i];
            }
        }
    }
    else
    {
#pragma unroll
        for (int i = 0; i < IT; ++i)
            if (i < cnt)
                a[p0 + i] = v[i];
    }
}

// Exclusive block scan of NCH fx values per thread (BLOCK threads).
template <int NCH>
__device__ __forceinline__ void block_exclusive_scan_fx_n(const fx_t (&v)[NCH], fx_t (&ex)[NCH])
{
    __shared__ fx_t s_wave[BLOCK / WAVE][NCH];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    fx_t incl[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        incl[c] = wave_inclusive_scan_fx(v[c]);
    if (lane == WAVE - 1)
    {
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            s_wave[wid][c] = incl[c];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NCH; ++c)
    {
        fx_t carry = 0;
        for (int w = 0; w < wid; ++w)
            carry += s_wave[w][c];
        ex[c] = carry + incl[c] - v[c];
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------------------------
// Radix digits (8 bits per pass) and wave-level digit counting
// ---------------------------------------------------------------------------------------------
constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;

__device__ __forceinline__ uint64_t lanemask_lt()
{
    const int lane = threadIdx.x & (WAVE - 1);
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Adds one to s_h[d] for every active lane, with one LDS atomic per run of equal digits in
// consecutive lanes (keys arrive in cell order, so runs are long: a cell holds about 20 particles
// of the full workload).  Counts only: the order in which runs are added does not matter.
__device__ __forceinline__ void wave_count(unsigned* s_h, unsigned d, bool valid)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t active = __ballot(valid);
    const unsigned prev = __shfl_up(d, 1, WAVE);
    const bool prev_valid = lane > 0 && ((active >> (lane - 1)) & 1ull);
    const bool head = valid && (!prev_valid || prev != d);
    const uint64_t heads = __ballot(head);
    if (head)
    {
        const uint64_t above = lane == WAVE - 1 ? 0ull : heads & (~0ull << (lane + 1));
        const uint64_t upto = above ? (1ull << __builtin_ctzll(above)) - 1ull : ~0ull;
        atomicAdd(&s_h[d], (unsigned)__popcll(active & upto & (~0ull << lane)));
    }
}

// Sort values: the particle index (< 2^31) with its association flag in bit 31.
__device__ __forceinline__ int sort_value(int p, uint8_t assoc)
{
    return (int)((unsigned)p | (assoc ? 0x80000000u : 0u));
}
__device__ __forceinline__ int sort_value_index(int v)
{
    return (int)((unsigned)v & 0x7fffffffu);
}
__device__ __forceinline__ uint8_t sort_value_assoc(int v)
{
    return (uint8_t)((unsigned)v >> 31);
}

// Radix-sort tiles: SORT_ITEMS keys per thread, SORT_Q units per tile (the engine sizes the
// per-tile digit counts for tiles of 8 units).
constexpr int SORT_ITEMS = 8;
constexpr int SORT_TILE = BLOCK * SORT_ITEMS;
constexpr int SORT_Q = SORT_TILE / UNIT;
static_assert(SORT_Q == 8, "sort tiles are 8 units");

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b on XCD b % 8), each with its
// own L2.  xcd_tile(b, n) gives XCD x a contiguous range of the n logical tiles, taken in order, so
// tiles whose scattered writes or gathered reads share cache lines run behind the same L2.
__device__ __forceinline__ int xcd_tile(int b, int n)
{
    const int q = n >> 3, r = n & 7;
    const int x = b & 7, k = b >> 3;
    return x * q + min(x, r) + k;
}

// ---------------------------------------------------------------------------------------------
// Work-table lookup: the vehicle and local element base of this workgroup's tile of q units.
// ---------------------------------------------------------------------------------------------
struct TileRef
{
    int v;     // vehicle index in the step
    int base;  // first local element of the tile
};

__device__ __forceinline__ TileRef tile_ref(const uint32_t* __restrict__ tab, int q, int tile)
{
    const uint32_t e = tab[(size_t)tile * q];
    return TileRef{(int)(e >> TAB_VEH_SHIFT), (int)(e & TAB_UNIT_MASK) * UNIT};
}

// ---------------------------------------------------------------------------------------------
// Exact tile-segmented reduction over a key-sorted array.
//
// Segments are runs of equal keys (cells) inside one vehicle's particles.  A workgroup takes a
// tile of BLOCK * ITEMS consecutive particles of one vehicle (ITEMS consecutive per thread) and
// computes the tile-relative inclusive prefix P of every element with one plain exact block scan;
// the sum of a segment piece [s, e] is then P[e] - P[s - 1].  Integer arithmetic makes the
// subtraction exact.  The start of the piece that reaches a thread's first element is found with
// one ballot (last earlier lane of the wave with a segment start) and a per-wave table for
// earlier waves.  Segments that start and end inside the tile are finalised here; a segment that
// crosses tile boundaries leaves a head partial (its piece in the tile where it ends) and tail
// partials (its pieces in the tiles it continues through), which the fix-up kernel sums from the
// tile where the segment starts.  A vehicle's tiles are consecutive in the table and its last
// tile ends every segment, so no carry crosses vehicles.
//
// Op interface:
//   static constexpr int NCH, ITEMS;                                    // ITEMS even
//   static constexpr int WAVES;                                         // waves per SIMD to fit
//   __device__ const int* keys(const VehDesc& d) const;                 // cell index per particle
//   __device__ void head(const VehDesc& d, int key, int p) const;       // p starts the segment of key
//   __device__ void end(const VehDesc& d, int key, int p) const;        // p ends the segment of key
//   __device__ void cont(const VehDesc& d, int key) const;              // key's segment started in this
//                                                                       // tile and continues past it
//   __device__ void load(const VehDesc& d, int p0, int cnt, const int (&key)[ITEMS],
//                        float (&v)[ITEMS][NCH]) const;
//   __device__ void finalize(const VehDesc& d, int key, const fx_t (&s)[NCH]) const;
// ---------------------------------------------------------------------------------------------
template <class Op>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(Op::WAVES)))
void k_segreduce(Op op, const VehDesc* __restrict__ descs,
                                                     const uint32_t* __restrict__ tab, SegCarry carry)
{
    constexpr int NCH = Op::NCH;
    constexpr int IT = Op::ITEMS;
    constexpr int TILE = BLOCK * IT;
    static_assert(IT % 2 == 0 && BLOCK * IT % UNIT == 0, "runs are loaded with 8- or 16-byte accesses");
    constexpr int QCAP = BLOCK;  // queued segment ends per tile
    __shared__ int s_wh[BLOCK / WAVE];
    __shared__ fx_t s_wb[BLOCK / WAVE][NCH];
    __shared__ int s_nq;
    __shared__ int s_qkey[QCAP];
    __shared__ fx_t s_qsum[QCAP][NCH];

    const int tile = blockIdx.x;
    const TileRef tr = tile_ref(tab, IT * BLOCK / UNIT, tile);
    const VehDesc& d = descs[tr.v];
    const int n = d.N;
    const int tbase = tr.base;
    if (tbase >= n)
    {
        // Padding tile past the vehicle's particles: no segment, no carry.
        if (threadIdx.x == 0)
            carry.flags[tile] = 0;
        return;
    }
    if (threadIdx.x == 0)
        s_nq = 0;
    const int* keys = op.keys(d);
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    const int p0 = tbase + threadIdx.x * IT;
    const int cnt = max(0, min(IT, n - p0));

    int key[IT];
    load_run<IT>(keys, p0, cnt, key, -1);
    const int kprev = (p0 > 0 && p0 - 1 < n) ? keys[p0 - 1] : -1;
    const int knext = (p0 + IT < n) ? keys[p0 + IT] : -1;

    float vf[IT][NCH];
    op.load(d, p0, cnt, key, vf);
    // Thread-local inclusive prefix; the thread's last segment start and the prefix before it.
    fx_t pre[IT][NCH];
    int lh = -1;
#pragma unroll
    for (int i = 0; i < IT; ++i)
    {
#pragma unroll
        for (int c = 0; c < NCH; ++c)
        {
            const fx_t x = i < cnt ? fx_from_f32(vf[i][c]) : (fx_t)0;
            pre[i][c] = (i == 0 ? (fx_t)0 : pre[i - 1][c]) + x;
        }
        const int kp = i == 0 ? kprev : key[i - 1];
        if (i < cnt && (p0 + i == 0 || key[i] != kp))
            lh = i;
    }
    fx_t tot[NCH], ex[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        tot[c] = pre[IT - 1][c];
    block_exclusive_scan_fx_n<NCH>(tot, ex);

    // Prefix (tile-relative, exclusive) at this thread's last segment start.
    fx_t bt[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
    {
        fx_t before = 0;
#pragma unroll
        for (int i = 1; i < IT; ++i)
            if (i == lh)
                before = pre[i - 1][c];
        bt[c] = ex[c] + before;
    }
    const uint64_t hm = __ballot(lh >= 0);
    const uint64_t below = hm & (lane == 0 ? 0ull : (~0ull >> (WAVE - lane)));
    const int src = below ? 63 - __builtin_clzll(below) : lane;
    fx_t base[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        base[c] = shfl_fx(bt[c], src);
    bool in_tile = below != 0;
    if (lane == 0)
        s_wh[wid] = hm != 0;
    if (hm != 0 && lane == 63 - __builtin_clzll(hm))
    {
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            s_wb[wid][c] = bt[c];
    }
    __syncthreads();
    if (!in_tile)
    {
        for (int w = wid - 1; w >= 0; --w)
        {
            if (s_wh[w])
            {
#pragma unroll
                for (int c = 0; c < NCH; ++c)
                    base[c] = s_wb[w][c];
                in_tile = true;
                break;
            }
        }
        if (!in_tile)
        {
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                base[c] = 0;
        }
    }

    const int tlast = min(tbase + TILE, n) - 1;
#pragma unroll
    for (int i = 0; i < IT; ++i)
    {
        if (i >= cnt)
            break;
        const int p = p0 + i;
        const int kp = i == 0 ? kprev : key[i - 1];
        if (p == 0 || key[i] != kp)
        {
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                base[c] = ex[c] + (i == 0 ? (fx_t)0 : pre[i - 1][c]);
            in_tile = true;
            op.head(d, key[i], p);
        }
        const int nk = i + 1 < IT ? key[i + 1] : knext;
        const bool is_end = (p == n - 1) || nk != key[i];
        if (is_end)
            op.end(d, key[i], p);
        if (is_end || p == tlast)
        {
            fx_t s[NCH];
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                s[c] = ex[c] + pre[i][c] - base[c];
            if (is_end)
            {
                if (in_tile)
                {
                    // Queue the segment so that finalisation runs with every lane busy; a tile with
                    // more than QCAP segments finalises the rest in place.
                    const int q = atomicAdd(&s_nq, 1);
                    if (q < QCAP)
                    {
                        s_qkey[q] = key[i];
#pragma unroll
                        for (int c = 0; c < NCH; ++c)
                            s_qsum[q][c] = s[c];
                    }
                    else
                    {
                        op.finalize(d, key[i], s);
                    }
                }
                else
                {
#pragma unroll
                    for (int c = 0; c < NCH; ++c)
                        carry.head_partial[(size_t)tile * NCH + c] = s[c];
                }
            }
            if (p == tlast)
            {
                uint8_t fl = in_tile ? 0 : SEG_NOHEAD;
                if (!is_end)
                {
                    if (in_tile)
                        op.cont(d, key[i]);
                    fl |= SEG_TAIL_CONT;
#pragma unroll
                    for (int c = 0; c < NCH; ++c)
                        carry.tail_partial[(size_t)tile * NCH + c] = s[c];
                }
                carry.flags[tile] = fl;
            }
        }
    }
    __syncthreads();
    const int nq = min(s_nq, QCAP);
    for (int q = threadIdx.x; q < nq; q += BLOCK)
    {
        fx_t s[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            s[c] = s_qsum[q][c];
        op.finalize(d, s_qkey[q], s);
    }
}

// Finalises segments that cross tile boundaries; one thread per tile, the owner is the tile in
// which the segment starts.
template <class Op>
__global__ __launch_bounds__(BLOCK) void k_segfix(Op op, const VehDesc* __restrict__ descs,
                                                  const uint32_t* __restrict__ tab, int ntiles, SegCarry carry)
{
    constexpr int NCH = Op::NCH;
    constexpr int TILE = BLOCK * Op::ITEMS;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= ntiles)
        return;
    const uint8_t fl = carry.flags[t];
    if (!(fl & SEG_TAIL_CONT) || (fl & SEG_NOHEAD))
        return;
    fx_t s[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        s[c] = carry.tail_partial[(size_t)t * NCH + c];
    int u = t + 1;
    while (u < ntiles)
    {
        const uint8_t fu = carry.flags[u];
        if (!((fu & SEG_NOHEAD) && (fu & SEG_TAIL_CONT)))
            break;
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            s[c] += carry.tail_partial[(size_t)u * NCH + c];
        ++u;
    }
    if (u < ntiles)
    {
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            s[c] += carry.head_partial[(size_t)u * NCH + c];
    }
    const TileRef tr = tile_ref(tab, TILE / UNIT, t);
    const VehDesc& d = descs[tr.v];
    const int last = min(tr.base + TILE, d.N) - 1;
    op.finalize(d, op.keys(d)[last], s);
}

// fix = false: the next kernel finalises the segments that cross tiles itself (from E->seg).
template <class Op>
inline void launch_segreduce(Engine* E, int kid, int kid_fix, const Op& op, bool fix = true)
{
    constexpr int TILE = BLOCK * Op::ITEMS;
    const int tiles = E->tiles(SP_P, TILE / UNIT);
    if (tiles <= 0)
        return;
    launch_begin(E, kid);
    hipLaunchKernelGGL(k_segreduce<Op>, dim3(tiles), dim3(BLOCK), 0, E->stream, op, E->d_desc, E->tab(SP_P), E->seg);
    launch_end(E, kid);
    if (fix && tiles > 1)
    {
        launch_begin(E, kid_fix);
        hipLaunchKernelGGL(k_segfix<Op>, dim3(div_up(tiles, BLOCK)), dim3(BLOCK), 0, E->stream, op, E->d_desc,
                           E->tab(SP_P), tiles, E->seg);
        launch_end(E, kid_fix);
    }
}

// First index j in [0, n) with a[j] > v (n if none).
__device__ __forceinline__ int upper_bound_i32(const int* __restrict__ a, int n, int v)
{
    int lo = 0, count = n;
    while (count > 0)
    {
        const int step = count / 2;
        const int mid = lo + step;
        if (a[mid] <= v)
        {
            lo = mid + 1;
            count -= step + 1;
        }
        else
        {
            count = step;
        }
    }
    return lo;
}

// ---------------------------------------------------------------------------------------------
// Exact scans over one space, segmented by vehicle (reduce per tile, per-vehicle scan of the tile
// totals, rescan with carry).  Prefixes are vehicle-relative.
// ---------------------------------------------------------------------------------------------
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = BLOCK * SCAN_ITEMS;
constexpr int SCAN_Q = SCAN_TILE / UNIT;  // units per scan tile

// Src interface:
//   static constexpr bool kNonNeg;                                    // values have a clear sign bit
//   __device__ bool active(const VehDesc& d) const;                   // vehicle takes part
//   __device__ int len(const VehDesc& d) const;                       // elements of the vehicle
//   __device__ void load4(const VehDesc& d, int p0, int cnt, float (&v)[4]) const;
// A CDF segment is one wave's share of a scan tile.  With seg_end set, the tile-local inclusive end
// of every segment is written at seg_end[CDF_SEGS_PER_TILE * tile + segment] (k_scan_top turns
// these into the vehicle's inclusive segment ends: the resampling CDF index).
static_assert(CDF_SEG == WAVE * SCAN_ITEMS && CDF_SEGS_PER_TILE * CDF_SEG == SCAN_TILE, "CDF segments are waves");

// The N + 1 exponentials as a scan source (tile totals only); systematic draws need none.
struct SrcExp
{
    static constexpr bool kNonNeg = true;  // -log(u) >= 0 (or -0)
    __device__ bool active(const VehDesc& d) const { return d.draw_mode == DRAW_MULTINOMIAL; }
    __device__ int len(const VehDesc& d) const { return d.N + 1; }
    __device__ void load4(const VehDesc& d, int p0, int, float (&v)[4]) const
    {
        resample_exponentials4(d.seed, d.inst, d.ep_rs, (uint32_t)(p0 >> 2), v);
    }
};

// Tile `tile` of a space: the tile total into tile_sum[tile] (and the segment ends, see above).
template <class Src>
__device__ __forceinline__ void scan_tile_body(const Src& src, const VehDesc* __restrict__ descs,
                                               const uint32_t* __restrict__ tab, int tile,
                                               fx_t* __restrict__ tile_sum, fx_t* __restrict__ seg_end)
{
    const TileRef tr = tile_ref(tab, SCAN_Q, tile);
    const VehDesc& d = descs[tr.v];
    if (!src.active(d))
        return;  // uniform over the workgroup
    const int n = src.len(d);
    const int p0 = tr.base + threadIdx.x * SCAN_ITEMS;
    const int cnt = max(0, min(SCAN_ITEMS, n - p0));
    float v[SCAN_ITEMS] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (cnt > 0)
        src.load4(d, p0, cnt, v);
    fx_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
        if (i < cnt)
            s += Src::kNonNeg ? fx_from_f32_nonneg(v[i]) : fx_from_f32(v[i]);
    fx_t total;
    const fx_t ex = block_exclusive_scan_fx(s, &total);
    if (threadIdx.x == 0)
        tile_sum[tile] = total;
    if (seg_end && (threadIdx.x & (WAVE - 1)) == WAVE - 1)
        seg_end[(size_t)CDF_SEGS_PER_TILE * tile + threadIdx.x / WAVE] = ex + s;
}

template <class Src>
__global__ __launch_bounds__(BLOCK) void k_scan_tiles(Src src, const VehDesc* __restrict__ descs,
                                                      const uint32_t* __restrict__ tab, fx_t* __restrict__ tile_sum,
                                                      fx_t* __restrict__ seg_end)
{
    scan_tile_body(src, descs, tab, blockIdx.x, tile_sum, seg_end);
}

// The tiles of two independent spaces in one launch: workgroups [0, tiles_a) take space A, the rest
// space B (one launch instead of two; alternating the spaces measured no faster).
template <class SrcA, class SrcB>
__global__ __launch_bounds__(BLOCK) void k_scan_tiles_pair(SrcA a, const uint32_t* __restrict__ tab_a, int tiles_a,
                                                           fx_t* __restrict__ sum_a, fx_t* __restrict__ seg_end_a,
                                                           SrcB b, const uint32_t* __restrict__ tab_b, int tiles_b,
                                                           fx_t* __restrict__ sum_b,
                                                           const VehDesc* __restrict__ descs)
{
    (void)tiles_b;
    if ((int)blockIdx.x < tiles_a)
        scan_tile_body(a, descs, tab_a, blockIdx.x, sum_a, seg_end_a);
    else
        scan_tile_body(b, descs, tab_b, blockIdx.x - tiles_a, sum_b, (fx_t*)nullptr);
}

enum ScanMode
{
    SCAN_ORDERS_N = 0,  // first-frame placement: scale = N / total
    SCAN_ORDERS_B = 1,  // birth placement: scale = B / total
    SCAN_RESAMPLE = 2,  // resampling: joint_max and new weight
    SCAN_EXP = 3,       // exponential spacings: total into exp_total
};

// Vehicle filters of the exact scans.
enum ScanOnly
{
    ONLY_ALL = 0,
    ONLY_INIT = 1,       // vehicles initialising their particles in this step
    ONLY_CDF_EXACT = 2,  // vehicles resampling with the exact CDF
};

__device__ __forceinline__ bool scan_takes(const VehDesc& d, int only)
{
    return only == ONLY_ALL || (only == ONLY_INIT && d.init) || (only == ONLY_CDF_EXACT && d.cdf_mode == CDF_EXACT);
}

// One workgroup per vehicle: exclusive prefix of the vehicle's tile totals (each thread scans a
// contiguous chunk), optionally the inclusive segment ends (tile_incl holds the tile-local ends
// k_scan_tiles wrote; the tile's prefix is added in place), and the scalars derived from the
// vehicle's total.  Vehicles the filter `only` excludes are skipped.
// tile_sum holds `parts` partial sums per tile (consecutive), which are added first.
__device__ __forceinline__ fx_t scan_tile_total(const fx_t* __restrict__ tile_sum, int t, int parts)
{
    if (parts == 1)
        return tile_sum[t];
    fx_t s = 0;
    for (int q = 0; q < parts; ++q)
        s += tile_sum[(size_t)parts * t + q];
    return s;
}

__device__ __forceinline__ void scan_top_body(const VehDesc* __restrict__ descs, int space,
                                              const fx_t* __restrict__ tile_sum, fx_t* __restrict__ tile_pre,
                                              fx_t* __restrict__ tile_incl, DevScalars* __restrict__ scs, int mode,
                                              int only, int parts = 1)
{
    const VehDesc& d = descs[blockIdx.x];
    if (!scan_takes(d, only))
        return;  // uniform over the workgroup
    DevScalars* sc = scs + blockIdx.x;
    const int t0 = d.u0[space] / SCAN_Q;
    const int ntiles = d.nu[space] / SCAN_Q;
    const int per = (ntiles + BLOCK - 1) / BLOCK;
    const int b0 = t0 + min(ntiles, (int)threadIdx.x * per);
    const int b1 = t0 + min(ntiles, (int)threadIdx.x * per + per);
    // Chunks of four independent loads (a loop-carried load per tile would wait for each).
    fx_t s = 0;
    for (int tc = b0; tc < b1; tc += 4)
    {
        fx_t x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            x[k] = tc + k < b1 ? scan_tile_total(tile_sum, tc + k, parts) : (fx_t)0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            s += x[k];
    }
    fx_t carry;
    fx_t run = block_exclusive_scan_fx(s, &carry);
    for (int tc = b0; tc < b1; tc += 4)
    {
        fx_t x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            x[k] = tc + k < b1 ? scan_tile_total(tile_sum, tc + k, parts) : (fx_t)0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            if (tc + k < b1)
            {
                if (tile_pre)
                    tile_pre[tc + k] = run;
                if (tile_incl)
                {
                    fx_t* se = tile_incl + (size_t)CDF_SEGS_PER_TILE * (tc + k);
#pragma unroll
                    for (int q = 0; q < CDF_SEGS_PER_TILE; ++q)
                        se[q] += run;
                }
            }
            run += x[k];
        }
    }
    if (threadIdx.x == 0)
    {
        if (mode == SCAN_EXP)
        {
            // Systematic draws do not use the exponentials: S_N = N * 2^32 (dogm_numerics.h).
            const fx_t s_n = d.draw_mode == DRAW_SYSTEMATIC ? systematic_s(d.N, 0u) : carry;
            sc->exp_total = s_n;
            sc->spacing = spacing_setup(s_n, sc->total);
            return;
        }
        sc->total = carry;
        if (mode == SCAN_ORDERS_N || mode == SCAN_ORDERS_B)
        {
            const int v = mode == SCAN_ORDERS_N ? d.N : d.B;
            sc->scale = carry != 0 ? order_scale(carry, v) : 0.0;
        }
        else
        {
            const float jm = fx_to_f32(carry);
            sc->joint_max = jm;
            sc->new_weight = jm / (float)d.N;
        }
    }
}

template <int Unused = 0>
__global__ __launch_bounds__(BLOCK) void k_scan_top(const VehDesc* __restrict__ descs, int space,
                                                    const fx_t* __restrict__ tile_sum, fx_t* __restrict__ tile_pre,
                                                    fx_t* __restrict__ tile_incl, DevScalars* __restrict__ scs,
                                                    int mode, int only, int parts = 1)
{
    scan_top_body(descs, space, tile_sum, tile_pre, tile_incl, scs, mode, only, parts);
}

// Sink interface:
//   __device__ void apply4(const VehDesc& d, const DevScalars* sc, int p0, int cnt,
//                          const fx_t (&prefix)[4], const fx_t (&own)[4]) const;
template <class Src, class Sink>
__global__ __launch_bounds__(BLOCK) void k_scan_apply(Src src, Sink sink, const VehDesc* __restrict__ descs,
                                                      const uint32_t* __restrict__ tab,
                                                      const fx_t* __restrict__ tile_pre,
                                                      const DevScalars* __restrict__ scs)
{
    const TileRef tr = tile_ref(tab, SCAN_Q, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    if (!src.active(d))
        return;
    const int n = src.len(d);
    if (tr.base >= n)
        return;
    const int p0 = tr.base + threadIdx.x * SCAN_ITEMS;
    const int cnt = max(0, min(SCAN_ITEMS, n - p0));
    float f[SCAN_ITEMS] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (cnt > 0)
        src.load4(d, p0, cnt, f);
    fx_t v[SCAN_ITEMS];
    fx_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
    {
        v[i] = i < cnt ? (Src::kNonNeg ? fx_from_f32_nonneg(f[i]) : fx_from_f32(f[i])) : (fx_t)0;
        s += v[i];
    }
    fx_t total;
    fx_t run = block_exclusive_scan_fx(s, &total) + tile_pre[blockIdx.x];
    fx_t pre[SCAN_ITEMS];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
    {
        run += v[i];
        pre[i] = run;
    }
    if (cnt > 0)
        sink.apply4(d, scs + tr.v, p0, cnt, pre, v);
}

// unit_sums: the sources' exact sums per 256-element unit were left there by an earlier kernel of the
// step (SCAN_Q per tile), so the tile pass is skipped.
template <class Src, class Sink>
inline void run_exact_scan(Engine* E, int space, const Src& src, const Sink& sink, int mode, int kid_apply,
                           int only, fx_t* tile_incl = nullptr, const fx_t* unit_sums = nullptr)
{
    const int tiles = E->tiles(space, SCAN_Q);
    if (tiles <= 0)
        return;
    if (!unit_sums)
    {
        launch_begin(E, K_SCAN_TILES);
        hipLaunchKernelGGL(k_scan_tiles<Src>, dim3(tiles), dim3(BLOCK), 0, E->stream, src, E->d_desc, E->tab(space),
                           E->scan_tile_sum, tile_incl);
        launch_end(E, K_SCAN_TILES);
    }
    launch_begin(E, K_SCAN_TOP);
    hipLaunchKernelGGL(k_scan_top<0>, dim3(E->m), dim3(BLOCK), 0, E->stream, E->d_desc, space,
                       unit_sums ? unit_sums : E->scan_tile_sum, E->scan_tile_pre, tile_incl, E->scalars, mode, only,
                       unit_sums ? SCAN_Q : 1);
    launch_end(E, K_SCAN_TOP);
    launch_begin(E, kid_apply);
    hipLaunchKernelGGL((k_scan_apply<Src, Sink>), dim3(tiles), dim3(BLOCK), 0, E->stream, src, sink, E->d_desc,
                       E->tab(space), E->scan_tile_pre, E->scalars);
    launch_end(E, kid_apply);
}

}  // namespace dogm_hip
