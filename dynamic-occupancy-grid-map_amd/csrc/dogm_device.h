// Device building blocks: exact (128-bit fixed-point) block scans and the exact tile-segmented
// reduction used for every per-cell sum of the filter.
#pragma once

#include "dogm_internal.h"

namespace dogm_hip
{

using namespace dogm_num;

__device__ __forceinline__ fx_t shfl_up_fx(fx_t x, int d)
{
    long long lo = (long long)(uint64_t)x;
    long long hi = (long long)(x >> 64);
    lo = __shfl_up(lo, d, WAVE);
    hi = __shfl_up(hi, d, WAVE);
    return (fx_t)(((ufx_t)(uint64_t)hi << 64) | (ufx_t)(uint64_t)lo);
}

__device__ __forceinline__ fx_t shfl_fx(fx_t x, int src)
{
    long long lo = (long long)(uint64_t)x;
    long long hi = (long long)(x >> 64);
    lo = __shfl(lo, src, WAVE);
    hi = __shfl(hi, src, WAVE);
    return (fx_t)(((ufx_t)(uint64_t)hi << 64) | (ufx_t)(uint64_t)lo);
}

// Exclusive block scan of one fx value per thread (BLOCK threads); returns the exclusive prefix
// and writes the block total.  Integer adds: the result does not depend on the association.
__device__ __forceinline__ fx_t block_exclusive_scan_fx(fx_t v, fx_t* total)
{
    __shared__ fx_t s_wave[BLOCK / WAVE];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    fx_t incl = v;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1)
    {
        const fx_t u = shfl_up_fx(incl, d);
        if (lane >= d)
            incl += u;
    }
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

// ---------------------------------------------------------------------------------------------
// Exact tile-segmented reduction over a key-sorted array.
//
// Segments are runs of equal keys (cells).  A workgroup reduces a tile of BLOCK * ITEMS
// consecutive elements: every segment that starts and ends inside the tile is finalised there;
// a segment that crosses tile boundaries leaves a head partial (its piece in the tile where it
// ends) and tail partials (its pieces in the tiles it continues through), and the fix-up kernel
// sums those pieces from the tile where the segment starts.  All sums are 128-bit integers, so
// the split changes nothing in the result.
//
// Op interface:
//   static constexpr int NCH, ITEMS;
//   __device__ void load(int pos, int key, float (&v)[NCH]) const;   // values of element pos
//   __device__ void finalize(int key, const fx_t (&s)[NCH]) const;   // exact sums of a segment
// ---------------------------------------------------------------------------------------------
template <class Op>
__global__ __launch_bounds__(BLOCK) void k_segreduce(Op op, const int* __restrict__ keys, int n, SegCarry carry)
{
    constexpr int NCH = Op::NCH;
    constexpr int IT = Op::ITEMS;
    constexpr int TILE = BLOCK * IT;
    const int tile = blockIdx.x;
    const int tbase = tile * TILE;
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    const int p0 = tbase + threadIdx.x * IT;

    int key[IT];
    bool head[IT];
    fx_t val[IT][NCH];
    const int kprev = (p0 > 0 && p0 - 1 < n) ? keys[p0 - 1] : -1;
    const int knext = (p0 + IT < n) ? keys[p0 + IT] : -1;
#pragma unroll
    for (int i = 0; i < IT; ++i)
    {
        const int p = p0 + i;
        key[i] = p < n ? keys[p] : -1;
    }
#pragma unroll
    for (int i = 0; i < IT; ++i)
    {
        const int p = p0 + i;
        const int kp = i == 0 ? kprev : key[i - 1];
        head[i] = p < n && (p == 0 || key[i] != kp);
        if (p < n)
        {
            float v[NCH];
            op.load(p, key[i], v);
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                val[i][c] = fx_from_f32(v[c]);
        }
        else
        {
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                val[i][c] = 0;
        }
    }

    // Thread aggregate (flag = a segment starts in this thread; value = sum since the last start).
    bool f = false;
    fx_t v[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        v[c] = 0;
#pragma unroll
    for (int i = 0; i < IT; ++i)
    {
        if (head[i])
        {
            f = true;
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                v[c] = val[i][c];
        }
        else
        {
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                v[c] += val[i][c];
        }
    }

    // Wave inclusive segmented scan of (f, v).
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1)
    {
        const int fu = __shfl_up((int)f, d, WAVE);
        fx_t vu[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            vu[c] = shfl_up_fx(v[c], d);
        if (lane >= d)
        {
            if (!f)
            {
#pragma unroll
                for (int c = 0; c < NCH; ++c)
                    v[c] += vu[c];
            }
            f = f || (fu != 0);
        }
    }
    // Exclusive within the wave.
    bool ef = __shfl_up((int)f, 1, WAVE) != 0;
    fx_t ev[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        ev[c] = shfl_up_fx(v[c], 1);
    if (lane == 0)
    {
        ef = false;
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            ev[c] = 0;
    }
    __shared__ fx_t s_wv[BLOCK / WAVE][NCH];
    __shared__ int s_wf[BLOCK / WAVE];
    if (lane == WAVE - 1)
    {
        s_wf[wid] = f ? 1 : 0;
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            s_wv[wid][c] = v[c];
    }
    __syncthreads();
    bool cf = false;
    fx_t cv[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        cv[c] = 0;
    for (int w = 0; w < wid; ++w)
    {
        if (s_wf[w])
        {
            cf = true;
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                cv[c] = s_wv[w][c];
        }
        else
        {
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                cv[c] += s_wv[w][c];
        }
    }
    if (!ef)
    {
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            ev[c] += cv[c];
    }
    ef = ef || cf;

    // Second pass: running inclusive value per element; emit at segment ends and at the tile end.
    const int tlast = min(tbase + TILE, n) - 1;
    bool rf = ef;
    fx_t r[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        r[c] = ev[c];
#pragma unroll
    for (int i = 0; i < IT; ++i)
    {
        const int p = p0 + i;
        if (p < n)
        {
            if (head[i])
            {
                rf = true;
#pragma unroll
                for (int c = 0; c < NCH; ++c)
                    r[c] = val[i][c];
            }
            else
            {
#pragma unroll
                for (int c = 0; c < NCH; ++c)
                    r[c] += val[i][c];
            }
            const int nk = i + 1 < IT ? key[i + 1] : knext;
            const bool is_end = (p == n - 1) || nk != key[i];
            if (is_end)
            {
                if (rf)
                {
                    op.finalize(key[i], r);
                }
                else
                {
#pragma unroll
                    for (int c = 0; c < NCH; ++c)
                        carry.head_partial[(size_t)tile * NCH + c] = r[c];
                }
            }
            if (p == tlast)
            {
                uint8_t fl = rf ? 0 : SEG_NOHEAD;
                if (!is_end)
                {
                    fl |= SEG_TAIL_CONT;
#pragma unroll
                    for (int c = 0; c < NCH; ++c)
                        carry.tail_partial[(size_t)tile * NCH + c] = r[c];
                }
                carry.flags[tile] = fl;
            }
        }
    }
}

// Finalises segments that cross tile boundaries; one thread per tile, the owner is the tile in
// which the segment starts.
template <class Op>
__global__ __launch_bounds__(BLOCK) void k_segfix(Op op, const int* __restrict__ keys, int n, int ntiles, SegCarry carry)
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
    const int last = min((t + 1) * TILE, n) - 1;
    op.finalize(keys[last], s);
}

template <class Op>
inline void launch_segreduce(dogm_handle* h, int kid, int kid_fix, const Op& op, const int* keys, int n)
{
    if (n <= 0)
        return;
    constexpr int TILE = BLOCK * Op::ITEMS;
    const int tiles = div_up(n, TILE);
    launch_begin(h, kid);
    hipLaunchKernelGGL(k_segreduce<Op>, dim3(tiles), dim3(BLOCK), 0, h->stream, op, keys, n, h->seg);
    launch_end(h, kid);
    if (tiles > 1)
    {
        launch_begin(h, kid_fix);
        hipLaunchKernelGGL(k_segfix<Op>, dim3(div_up(tiles, BLOCK)), dim3(BLOCK), 0, h->stream, op, keys, n, tiles,
                           h->seg);
        launch_end(h, kid_fix);
    }
}

// Binary search shared with the checker: first index in [0, n) with a[index] >= v
// (n if none).  Written out so that it is the same search on any input, sorted or not.
__device__ __forceinline__ int lower_bound_fx(const fx_t* __restrict__ a, int n, fx_t v)
{
    int lo = 0, count = n;
    while (count > 0)
    {
        const int step = count / 2;
        const int mid = lo + step;
        if (a[mid] < v)
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

}  // namespace dogm_hip
