// Multinomial resampling on gfx950 (dogm/src/dogm.cu:386-423, resampling.cu:17-68).
//
// 1. Exact CDF over [posterior weights | birth weights]: tile totals and their scan give the
//    inclusive CDF value at the end of every 256-entry segment (the coarse index, four per
//    1024-entry scan tile); entries are not stored but recomputed per staged segment in the merge.
// 2. Sorted uniforms by exponential spacings: E_k for k = 0..N, exact prefix S_k, draw k is the
//    fraction S_k / S_N of the CDF total (spacing_draw in dogm_numerics.h).  Only the tile totals of E are
//    stored; the merge kernel regenerates E from the counter-based RNG (four E per Philox call).
// 3. Merge: a workgroup owns 1024 consecutive (hence sorted) draws.  Its ancestors lie between the
//    CDF segments holding its first and last draw; those are found by a two-level k-ary search over
//    the coarse index (two dependent loads instead of a 22-level binary search), the CDF entries of
//    those segments are rebuilt in LDS from the weights (one wave per segment: exact wave scan +
//    the index entry before the segment), and every draw finds its ancestor there (the reference's
//    lower_bound).  The reference's resamplingKernel gather is deferred: only the ancestors (and the
//    new weight) are written, the next full step's prediction reads each state through its
//    ancestor, and anything else that reads the particles first has k_materialize copy them.
// No atomics on global memory and no global sort: the result is a pure function of the inputs.
// Every vehicle of a step resamples from its own CDF with its own draws.
#include "dogm_device.h"

namespace dogm_hip
{

constexpr int RS_ITEMS = 4;
constexpr int RS_TILE = BLOCK * RS_ITEMS;  // draws per workgroup
// CDF entries staged in LDS per window: 8 segments (2048 entries) for launches of up to
// RS_WIDE_MAX_TILES workgroups (fewer windows per workgroup shortens the critical path), 4 segments
// for larger launches (half the LDS, more workgroups per CU).  The float modes stage whole tiles
// (RS_FLOAT_SLICE_TILES per window).
constexpr int RS_WINDOW_NARROW = 4;
constexpr int RS_WINDOW_WIDE = 8;
constexpr int RS_WIDE_MAX_TILES = 4096;
constexpr int RS_FLOAT_SLICE_TILES = 2;
// Doubles per scan tile in the float modes' index (the envelope at the tile end and before it),
// spaced like the exact mode's segment ends so that vehicles of both kinds in one step keep
// disjoint index ranges.
constexpr int ENV_STRIDE = 2 * CDF_SEGS_PER_TILE;

static_assert(RS_TILE == SCAN_TILE, "draw tiles and CDF tiles share the exact-scan tile size");

struct SrcJointPos
{
    static constexpr bool kNonNeg = true;  // cdf_weight: +0 or positive
    __device__ bool active(const VehDesc& d) const { return d.cdf_mode == CDF_EXACT; }
    __device__ int len(const VehDesc& d) const { return d.N + d.B; }
    __device__ void load4(const VehDesc& d, int p0, int cnt, float (&v)[4]) const
    {
        const int N = d.N;
        if (p0 + cnt <= N)
            load_run<4>(d.weight_array, p0, cnt, v);
        else if (p0 >= N && ((p0 - N) & 3) == 0)
            load_run<4>(d.birth.weight, p0 - N, cnt, v);
        else
        {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                v[i] = i < cnt ? (p0 + i < N ? d.weight_array[p0 + i] : d.birth.weight[p0 + i - N]) : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
            v[i] = cdf_weight(v[i]);
    }
};

// E for the four consecutive draws k = p0 .. p0 + 3 of one thread (p0 is a multiple of 4).
__device__ __forceinline__ void exp4(const VehDesc& d, int p0, int n, fx_t* e)
{
    float f[4];
    resample_exponentials4(d.seed, d.inst, d.ep_rs, (uint32_t)(p0 >> 2), f);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        e[i] = p0 + i < n ? fx_from_f32_nonneg(f[i]) : (fx_t)0;  // -log(u) >= 0 (or -0)
}

// ---------------------------------------------------------------------------------------------
// Float CDF modes (precision study).  Per tile of 1024 joint weights: 4-element sequential folds
// per thread, the 64-lane DPP wave scan, sequential carries across the 4 waves, and a sequential
// prefix over the vehicle's tiles; element = ((tile prefix + wave carry) + thread's exclusive wave
// prefix) + thread-local prefix.  The running maximum (monotone envelope) is what the ancestor
// search sees.  The CDF is stored as doubles in the vehicle's CDF buffer; per tile, the coarse
// index holds (envelope at the tile end, envelope before the tile) as two doubles.
// ---------------------------------------------------------------------------------------------
template <class T>
struct CdfFloatMode;
template <>
struct CdfFloatMode<double>
{
    static constexpr int mode = CDF_FP64;
};
template <>
struct CdfFloatMode<float>
{
    static constexpr int mode = CDF_FP32;
};
template <>
struct CdfFloatMode<f32c>
{
    static constexpr int mode = CDF_FP32C;
};

// Thread fold, wave scan and wave carries of one tile: the thread's local prefixes, its wave
// carry (sum of the earlier waves' totals), its exclusive prefix inside the wave, and the tile total
// (on every thread).
template <class T>
__device__ __forceinline__ void tile_float_scan(const VehDesc& d, int p0, int cnt, T (&loc)[4], T* carry_out,
                                                T* excl_out, T* tile_total)
{
    __shared__ T s_w[BLOCK / WAVE];
    float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (cnt > 0)
        SrcJointPos{}.load4(d, p0, cnt, f);
    T run = (T)0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
        run = run + (T)(i < cnt ? f[i] : 0.0f);
        loc[i] = run;
    }
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    const T incl = wave_inclusive_scan_val(run);
    T excl = shfl_up1(incl);
    if (lane == 0)
        excl = (T)0;
    if (lane == WAVE - 1)
        s_w[wid] = incl;
    __syncthreads();
    T carry = (T)0;
    for (int w = 0; w < wid; ++w)
        carry = carry + s_w[w];
    T tot = (T)0;
#pragma unroll
    for (int w = 0; w < BLOCK / WAVE; ++w)
        tot = tot + s_w[w];
    __syncthreads();
    *carry_out = carry;
    *excl_out = excl;
    *tile_total = tot;
}

template <class T>
__global__ __launch_bounds__(BLOCK) void k_cdff_tiles(const VehDesc* __restrict__ descs,
                                                      const uint32_t* __restrict__ tab, double* __restrict__ tile_sum)
{
    const TileRef tr = tile_ref(tab, SCAN_Q, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    if (d.cdf_mode != CdfFloatMode<T>::mode)
        return;
    const int n = d.N + d.B;
    const int p0 = tr.base + threadIdx.x * SCAN_ITEMS;
    const int cnt = max(0, min(SCAN_ITEMS, n - p0));
    T loc[4];
    T carry, ex, tot;
    tile_float_scan<T>(d, p0, cnt, loc, &carry, &ex, &tot);
    if (threadIdx.x == 0)
        tile_sum[blockIdx.x] = (double)tot;
}

// One thread per vehicle: sequential prefix over the vehicle's tile totals.
template <class T>
__global__ void k_cdff_top(const VehDesc* __restrict__ descs, int m, const double* __restrict__ tile_sum,
                           double* __restrict__ tile_pre)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m)
        return;
    const VehDesc& d = descs[v];
    if (d.cdf_mode != CdfFloatMode<T>::mode)
        return;
    const int t0 = d.u0[SP_J] / SCAN_Q;
    const int nt = d.nu[SP_J] / SCAN_Q;
    T run = (T)0;
    for (int t = 0; t < nt; ++t)
    {
        tile_pre[t0 + t] = (double)run;
        run = run + (T)tile_sum[t0 + t];
    }
}

// Values, the envelope inside the tile, and the tile maximum (into env[ENV_STRIDE t]).
template <class T>
__global__ __launch_bounds__(BLOCK) void k_cdff_apply(const VehDesc* __restrict__ descs,
                                                      const uint32_t* __restrict__ tab,
                                                      const double* __restrict__ tile_pre, double* __restrict__ env)
{
    __shared__ T s_m[BLOCK / WAVE];
    const TileRef tr = tile_ref(tab, SCAN_Q, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    if (d.cdf_mode != CdfFloatMode<T>::mode)
        return;
    const int n = d.N + d.B;
    const int p0 = tr.base + threadIdx.x * SCAN_ITEMS;
    const int cnt = max(0, min(SCAN_ITEMS, n - p0));
    T loc[4];
    T carry, ex, tot;
    tile_float_scan<T>(d, p0, cnt, loc, &carry, &ex, &tot);
    const T tp = (T)tile_pre[blockIdx.x];
    const int wid = threadIdx.x / WAVE;
    const T base = (tp + carry) + ex;
    T val[4];
    T mx = (T)0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
        val[i] = base + loc[i];
        if (i < cnt && val[i] > mx)
            mx = val[i];
    }
    // Exclusive running maximum over the earlier threads of the tile.
    const T wmax = wave_inclusive_max_val(mx);
    T before = shfl_up1(wmax);
    if ((threadIdx.x & (WAVE - 1)) == 0)
        before = (T)0;
    if ((threadIdx.x & (WAVE - 1)) == WAVE - 1)
        s_m[wid] = wmax;
    __syncthreads();
    for (int w = 0; w < wid; ++w)
        before = s_m[w] > before ? s_m[w] : before;
    T tmax = (T)0;
#pragma unroll
    for (int w = 0; w < BLOCK / WAVE; ++w)
        tmax = s_m[w] > tmax ? s_m[w] : tmax;
    double* cdf = reinterpret_cast<double*>(d.cdf);
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
        if (i < cnt)
        {
            const T e = val[i] > before ? val[i] : before;
            cdf[p0 + i] = (double)e;
            before = e;
        }
    }
    if (threadIdx.x == 0)
        env[ENV_STRIDE * (size_t)blockIdx.x] = (double)tmax;
}

// One thread per vehicle: the envelope across tiles, the total and the new particle weight.
__global__ void k_cdff_env(const VehDesc* __restrict__ descs, int m, double* __restrict__ env,
                           DevScalars* __restrict__ scs)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m)
        return;
    const VehDesc& d = descs[v];
    if (d.cdf_mode == CDF_EXACT)
        return;
    const int t0 = d.u0[SP_J] / SCAN_Q;
    const int nt = d.nu[SP_J] / SCAN_Q;
    double run = 0.0;
    for (int t = 0; t < nt; ++t)
    {
        const double tm = env[ENV_STRIDE * (size_t)(t0 + t)];
        env[ENV_STRIDE * (size_t)(t0 + t) + 1] = run;
        run = tm > run ? tm : run;
        env[ENV_STRIDE * (size_t)(t0 + t)] = run;
    }
    DevScalars* sc = scs + v;
    sc->ftotal = run;
    const float jm = (float)run;
    sc->joint_max = jm;
    sc->new_weight = jm / (float)d.N;
}

// Ancestor search over the float modes' envelope (doubles): the draw, the coarse index entry of a
// tile and a staged CDF value.
struct CdfFloatTr
{
    using V = double;
    __device__ static bool takes(const VehDesc& d) { return d.cdf_mode != CDF_EXACT; }
    __device__ static V draw(fx_t s_k, const Spacing& sp, const DevScalars* sc, int mode)
    {
        return cdf_float_draw(sc->ftotal, spacing_fraction(s_k, sp), mode);
    }
    __device__ static V incl(const void* ci, int t0, int t)
    {
        return static_cast<const double*>(ci)[ENV_STRIDE * (size_t)(t0 + t)];
    }
    __device__ static V value(const VehDesc& d, const void* ci, int t0, int j)
    {
        const double c = reinterpret_cast<const double*>(d.cdf)[j];
        const double b = static_cast<const double*>(ci)[ENV_STRIDE * (size_t)(t0 + j / SCAN_TILE) + 1];
        return c > b ? c : b;
    }
};

// A chunk of CDF entries staged in LDS.  Exact entries (128-bit) are kept as separate high and low
// 64-bit words: the lower-bound search compares the high words (8-byte LDS reads, half the bytes
// and bank traffic of 16-byte ones) and reads a low word only where the high words tie.  Entry i
// sits at slot i + i / 32: the rebuild writes four consecutive entries per thread, and without the
// skew the lanes of a wave would hit the same banks eight at a time (entries 32 words apart).
__device__ __forceinline__ int cdf_slot(int i)
{
    return i + (i >> 5);
}

template <class V, int LEN>
struct CdfStage
{
    V v[LEN + LEN / 32];
    __device__ void put(int i, const V& x) { v[cdf_slot(i)] = x; }
    __device__ V get(int i) const { return v[cdf_slot(i)]; }
    __device__ bool less(int i, const V& r) const { return v[cdf_slot(i)] < r; }
};

template <int LEN>
struct CdfStage<fx_t, LEN>
{
    int64_t hi[LEN + LEN / 32];
    uint64_t lo[LEN + LEN / 32];
    __device__ void put(int i, const fx_t& x)
    {
        const int s = cdf_slot(i);
        hi[s] = (int64_t)(x >> 64);
        lo[s] = (uint64_t)x;
    }
    __device__ fx_t get(int i) const
    {
        const int s = cdf_slot(i);
        return (fx_t)(((ufx_t)(uint64_t)hi[s] << 64) | (ufx_t)lo[s]);
    }
    // Both words are read at once (one LDS round trip per probe) and compared without a branch.
    __device__ bool less(int i, const fx_t& r) const
    {
        const int s = cdf_slot(i);
        const int64_t h = hi[s], rh = (int64_t)(r >> 64);
        const uint64_t l = lo[s];
        return (h < rh) | ((h == rh) & (l < (uint64_t)r));
    }
};

// Index of the first of n non-decreasing values value(0 .. n-1) that is >= r (n - 1 if none): a
// two-level k-ary search by the workgroup (the last value of each of up to BLOCK groups, then the
// values of one group).
template <class V, class Get>
__device__ int find_first_not_below(const Get& value, int n, V r, int* s_cnt)
{
    const int G = (n + BLOCK - 1) / BLOCK;  // values per group; at most BLOCK groups
    const int ng = (n + G - 1) / G;
    if (threadIdx.x == 0)
        s_cnt[0] = 0;
    __syncthreads();
    const int i = threadIdx.x;
    const bool below = i < ng && value(min(n - 1, (i + 1) * G - 1)) < r;
    const int groups_below = __syncthreads_count(below);
    const int g = min(groups_below, ng - 1);
    for (int j = threadIdx.x; j < G; j += BLOCK)
    {
        const int t = g * G + j;
        if (t < n && value(t) < r)
            atomicAdd(&s_cnt[0], 1);
    }
    __syncthreads();
    const int t = min(g * G + s_cnt[0], n - 1);
    __syncthreads();
    return t;
}

// The workgroup's draws r_k, k = base + threadIdx.x + i * BLOCK (item i of a thread is strided, so
// at every item a wave searches 64 consecutive draws, whose lower bounds are neighbouring CDF
// entries: the probes of a wave read neighbouring LDS words, and the ancestors come out
// wave-contiguous for the gather), and s_rb = {first draw, last draw}.  Multinomial draws: the
// spacings' prefix runs over four consecutive draws per thread (one Philox call each); the draws
// are then moved to their strided owners through the staging buffer.
template <class V, class Draw, class Stage>
__device__ __forceinline__ void merge_draws(const VehDesc& d, const DevScalars* sc, int base, int n,
                                            const fx_t* __restrict__ exp_tile_pre, const Draw& draw, Stage& stage,
                                            V (&r)[RS_ITEMS], V* s_rb)
{
    const int last = min(base + RS_TILE, n) - 1;
    const Spacing sp = sc->spacing;
    if (d.draw_mode == DRAW_SYSTEMATIC)
    {
        const uint32_t u = systematic_u32(d.seed, d.inst, d.ep_rs);
#pragma unroll
        for (int i = 0; i < RS_ITEMS; ++i)
            r[i] = draw(systematic_s(base + (int)threadIdx.x + i * BLOCK, u), sp);
    }
    else
    {
        const int q0 = base + threadIdx.x * RS_ITEMS;
        fx_t e[4];
        exp4(d, q0, n, e);
        fx_t tot;
        fx_t run = block_exclusive_scan_fx(e[0] + e[1] + e[2] + e[3], &tot) +
                   exp_tile_pre[d.u0[SP_E] / SCAN_Q + base / RS_TILE];
#pragma unroll
        for (int i = 0; i < RS_ITEMS; ++i)
        {
            run += e[i];
            stage.put(threadIdx.x * RS_ITEMS + i, draw(run, sp));
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < RS_ITEMS; ++i)
            r[i] = stage.get(threadIdx.x + i * BLOCK);
    }
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i)
    {
        const int k = base + (int)threadIdx.x + i * BLOCK;
        if (k == base)
            s_rb[0] = r[i];
        if (k == last)
            s_rb[1] = r[i];
    }
    __syncthreads();
}

// The smallest draw of the workgroup that has no ancestor yet (n if none), and its value in s_rb[0].
template <class V>
__device__ __forceinline__ int merge_next_draw(int base, int n, const int (&anc)[RS_ITEMS], const V (&r)[RS_ITEMS],
                                               int* s_kmin, V* s_rb)
{
    if (threadIdx.x == 0)
        *s_kmin = n;
    __syncthreads();
    int mine = n;
#pragma unroll
    for (int i = RS_ITEMS - 1; i >= 0; --i)
        if (anc[i] < 0 && base + (int)threadIdx.x + i * BLOCK < n)
            mine = base + (int)threadIdx.x + i * BLOCK;
    if (mine < n)
        atomicMin(s_kmin, mine);
    __syncthreads();
    const int kmin = *s_kmin;
    if (kmin < n)
    {
#pragma unroll
        for (int i = 0; i < RS_ITEMS; ++i)
            if (base + (int)threadIdx.x + i * BLOCK == kmin)
                s_rb[0] = r[i];
    }
    __syncthreads();
    return kmin;
}

// The resampled particles k = base + threadIdx.x + i * BLOCK: ancestor index, state copied from the
// ancestor (posterior particle or new-born one) and the uniform new weight.  A draw without an
// ancestor (above the CDF total by rounding) takes the last joint entry.  A deferred set
// (d.rs_defer) gets the ancestors and weight[0] only: the next prediction reads the states through
// the ancestors, and every other reader first writes the set out (k_materialize).
__device__ __forceinline__ void merge_write(const VehDesc& d, const DevScalars* sc, int pph, int base, int n, int J,
                                            const int (&anc)[RS_ITEMS])
{
    const float nw = sc->new_weight;
    const DevParticles P = d.part[pph];
    const DevParticles Bp = d.birth;
    const DevParticles out = d.part[pph ^ 1];
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i)
    {
        const int k = base + i * BLOCK + (int)threadIdx.x;
        if (k >= n)
            break;
        const int a = anc[i] < 0 ? J - 1 : min(anc[i], J - 1);
        d.ancestors[k] = a;
        if (d.rs_defer)
        {
            if (k == 0)
                out.weight[0] = nw;
            continue;
        }
        const bool from_p = a < n;
        const int o = from_p ? a : a - n;
        const float2* ps = from_p ? P.pos : Bp.pos;
        const float2* vs = from_p ? P.vel : Bp.vel;
        const int* ix = from_p ? P.idx : Bp.idx;
        const uint8_t* as = from_p ? P.assoc : Bp.assoc;
        out.pos[k] = ps[o];
        out.vel[k] = vs[o];
        out.idx[k] = ix[o];
        out.assoc[k] = as[o];
        out.weight[k] = nw;
    }
}

// Index of the first of n non-decreasing values value(0 .. n-1) that is >= r (n - 1 if none), for
// two values at once, found by one wave without a barrier: the lanes probe every s-th value (s =
// ceil(n / 64)); the probes below r form a prefix of k lanes, so the answer lies in [k s, (k + 1) s),
// whose values below r are then counted (ceil(s / 64) probes per lane).
template <class V, class Get>
__device__ __forceinline__ void wave_find2(const Get& value, int n, V r0, V r1, int* c0, int* c1)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int s = (n + WAVE - 1) / WAVE;
    const V p = value(min(n - 1, (lane + 1) * s - 1));
    const int b0 = __popcll(__ballot(p < r0)) * s;
    const int b1 = __popcll(__ballot(p < r1)) * s;
    int n0 = 0, n1 = 0;
    for (int q0 = 0; q0 < s; q0 += WAVE)
    {
        const int q = q0 + lane;
        const bool in0 = q < s && b0 + q < n, in1 = q < s && b1 + q < n;
        const V v0 = in0 ? value(b0 + q) : V(0), v1 = in1 ? value(b1 + q) : V(0);
        n0 += __popcll(__ballot(in0 && v0 < r0));
        n1 += __popcll(__ballot(in1 && v1 < r1));
    }
    *c0 = min(b0 + n0, n - 1);
    *c1 = min(b1 + n1, n - 1);
}

// Exact-CDF merge.  The workgroup's ancestors lie in the CDF segments [s_lo, s_hi] holding its first
// and last draw; they are staged WSEG segments at a time.  Each wave rebuilds whole segments: the 64
// lanes' four joint weights in fixed point, a 64-lane exact scan, and the index entry before the
// segment as its start (integer adds, so the entries equal those of any other association).  A
// draw is resolved in the first window whose last entry is >= its value (the CDF is non-decreasing,
// so its lower bound is there), by a branch-free binary search over the window padded to a power
// of two; windows no remaining draw lands in are skipped with one coarse-index search.  Every wave
// runs the coarse-index searches itself (the same answer in each), so the only barriers are those of
// the draws' prefix and two per window plus one between windows.
// Seven waves per SIMD for the 4-segment window; the 8-segment form is bounded by its LDS.
template <int WSEG>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WSEG <= RS_WINDOW_NARROW ? 7 : 4)))
void k_rs_merge_exact(const VehDesc* __restrict__ descs, const uint32_t* __restrict__ tab, int pph,
                      const fx_t* __restrict__ exp_tile_pre, const fx_t* __restrict__ seg_incl,
                      const DevScalars* __restrict__ scs)
{
    constexpr int WLEN = WSEG * CDF_SEG;
    constexpr int NWAVE = BLOCK / WAVE;
    static_assert(WLEN >= RS_TILE, "the window also stages the workgroup's draws");
    __shared__ CdfStage<fx_t, WLEN> s_cdf;
    __shared__ fx_t s_rb[2];
    __shared__ int s_wmin[NWAVE];
    __shared__ fx_t s_wr[NWAVE];

    const TileRef tr = tile_ref(tab, RS_TILE / UNIT, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    const int n = d.N;
    const int base = tr.base;
    if (base >= n || d.cdf_mode != CDF_EXACT)
        return;
    const DevScalars* sc = scs + tr.v;
    const int J = n + d.B;
    fx_t r[RS_ITEMS];
    merge_draws<fx_t>(d, sc, base, n, exp_tile_pre,
                      [](fx_t s_k, const Spacing& sp) { return spacing_draw_nonneg(s_k, sp); }, s_cdf, r, s_rb);
    const int S = (J + CDF_SEG - 1) / CDF_SEG;  // segments holding entries
    const fx_t* si = seg_incl + (size_t)(d.u0[SP_J] / SCAN_Q) * CDF_SEGS_PER_TILE;
    const auto seg_end = [si](int s) { return si[s]; };
    int s_lo, s_hi;
    wave_find2<fx_t>(seg_end, S, s_rb[0], s_rb[1], &s_lo, &s_hi);
    const int wid = threadIdx.x / WAVE;
    const int lane = threadIdx.x & (WAVE - 1);
    int anc[RS_ITEMS];
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i)
        anc[i] = -1;
    for (int s0 = s_lo; s0 <= s_hi;)
    {
        const int s1 = min(s0 + WSEG, s_hi + 1);  // window: segments [s0, s1)
        const int c0 = s0 * CDF_SEG;
        const int clen = min(J, s1 * CDF_SEG) - c0;
        __syncthreads();  // the previous window's searches (and the draws' staging) are done
        for (int s = s0 + wid; s < s1; s += NWAVE)
        {
            const int q0 = s * CDF_SEG + lane * SCAN_ITEMS;
            const int cnt = max(0, min(SCAN_ITEMS, J - q0));
            float f[SCAN_ITEMS] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (cnt > 0)
                SrcJointPos{}.load4(d, q0, cnt, f);
            fx_t v[SCAN_ITEMS];
            fx_t sum = 0;
#pragma unroll
            for (int i = 0; i < SCAN_ITEMS; ++i)
            {
                v[i] = i < cnt ? fx_from_f32_nonneg(f[i]) : (fx_t)0;  // cdf_weight: sign bit clear
                sum += v[i];
            }
            fx_t run = wave_inclusive_scan_fx(sum) - sum + (s > 0 ? si[s - 1] : (fx_t)0);
            const int at = (s - s0) * CDF_SEG + lane * SCAN_ITEMS;
#pragma unroll
            for (int i = 0; i < SCAN_ITEMS; ++i)
            {
                run += v[i];
                if (i < cnt)
                    s_cdf.put(at + i, run);
            }
        }
        __syncthreads();
        const fx_t cmax = s_cdf.get(clen - 1);
        const int span = clen > 1 ? 1 << (32 - __clz(clen - 1)) : 1;  // power of two >= clen, <= WLEN
#pragma unroll
        for (int i = 0; i < RS_ITEMS; ++i)
        {
            if (anc[i] < 0 && base + (int)threadIdx.x + i * BLOCK < n && !(cmax < r[i]))
            {
                // lo counts the entries < r[i]: after the probe at lo + step - 1 the count lies in
                // [lo, lo + step).  Entries at clen and beyond count as +infinity; probing the last
                // entry (>= r[i]) in their place gives the same answer without a branch.
                int lo = 0;
                for (int step = span >> 1; step > 0; step >>= 1)
                {
                    const int j = min(lo + step, clen) - 1;
                    if (s_cdf.less(j, r[i]))
                        lo += step;
                }
                anc[i] = c0 + lo;
            }
        }
        if (s1 > s_hi)
            break;
        // The smallest unresolved draw: each wave's minimum (draw indices are distinct, so one lane
        // holds it) with its value, then the minimum over the waves.
        int mine = n;
        fx_t rmine = 0;
#pragma unroll
        for (int i = RS_ITEMS - 1; i >= 0; --i)
        {
            const int k = base + (int)threadIdx.x + i * BLOCK;
            if (anc[i] < 0 && k < n)
            {
                mine = k;
                rmine = r[i];
            }
        }
        int wmin = mine;
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1)
            wmin = min(wmin, __shfl_xor(wmin, o, WAVE));
        if (wmin < n && mine == wmin)
        {
            s_wmin[wid] = wmin;
            s_wr[wid] = rmine;
        }
        if (wmin >= n && lane == 0)
            s_wmin[wid] = n;
        __syncthreads();
        int kmin = n, kw = 0;
#pragma unroll
        for (int w = 0; w < NWAVE; ++w)
        {
            const int x = s_wmin[w];
            if (x < kmin)
            {
                kmin = x;
                kw = w;
            }
        }
        if (kmin >= n)
            break;
        const fx_t rmin = s_wr[kw];
        const int next_last = min(s1 - 1 + WSEG, s_hi);
        if (!(si[next_last] < rmin))
        {
            s0 = s1;
        }
        else
        {
            int c, unused;
            wave_find2<fx_t>(seg_end, S, rmin, rmin, &c, &unused);
            s0 = max(s1, c);
        }
    }
    merge_write(d, sc, pph, base, n, J, anc);
}

// Float-mode merge (precision study): the envelope is stored, so a window stages
// RS_FLOAT_SLICE_TILES tiles of it; the search and the window walk are those of the exact merge
// at tile granularity.
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
void k_rs_merge_float(const VehDesc* __restrict__ descs, const uint32_t* __restrict__ tab, int pph,
                      const fx_t* __restrict__ exp_tile_pre, const void* __restrict__ cdf_incl,
                      const DevScalars* __restrict__ scs)
{
    using Tr = CdfFloatTr;
    constexpr int SLICE = RS_FLOAT_SLICE_TILES * SCAN_TILE;
    __shared__ CdfStage<double, SLICE> s_cdf;
    __shared__ double s_rb[2];
    __shared__ int s_cnt[1];
    __shared__ int s_kmin;

    const TileRef tr = tile_ref(tab, RS_TILE / UNIT, blockIdx.x);
    const VehDesc& d = descs[tr.v];
    const int n = d.N;
    const int base = tr.base;
    if (base >= n || !Tr::takes(d))
        return;
    const DevScalars* sc = scs + tr.v;
    const int J = n + d.B;
    double r[RS_ITEMS];
    merge_draws<double>(
        d, sc, base, n, exp_tile_pre,
        [sc, &d](fx_t s_k, const Spacing& sp) { return Tr::draw(s_k, sp, sc, d.cdf_mode); }, s_cdf, r, s_rb);
    const int T = (J + SCAN_TILE - 1) / SCAN_TILE;
    const int t0 = d.u0[SP_J] / SCAN_Q;
    const auto tile_end = [cdf_incl, t0](int t) { return Tr::incl(cdf_incl, t0, t); };
    const int t_lo = find_first_not_below<double>(tile_end, T, s_rb[0], s_cnt);
    const int t_hi = find_first_not_below<double>(tile_end, T, s_rb[1], s_cnt);
    int anc[RS_ITEMS];
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i)
        anc[i] = -1;
    for (int tc = t_lo; tc <= t_hi;)
    {
        const int tc_end = min(tc + RS_FLOAT_SLICE_TILES, t_hi + 1) - 1;
        const int c0 = tc * SCAN_TILE;
        const int clen = min(J, (tc_end + 1) * SCAN_TILE) - c0;
        __syncthreads();
        for (int i = threadIdx.x; i < clen; i += BLOCK)
            s_cdf.put(i, Tr::value(d, cdf_incl, t0, c0 + i));
        __syncthreads();
        const double cmax = s_cdf.get(clen - 1);
        const int span = clen > 1 ? 1 << (32 - __clz(clen - 1)) : 1;
#pragma unroll
        for (int i = 0; i < RS_ITEMS; ++i)
        {
            if (anc[i] < 0 && base + (int)threadIdx.x + i * BLOCK < n && !(cmax < r[i]))
            {
                int lo = 0;
                for (int step = span >> 1; step > 0; step >>= 1)
                {
                    const int j = lo + step - 1;
                    if (j < clen && s_cdf.less(j, r[i]))
                        lo += step;
                }
                anc[i] = c0 + lo;
            }
        }
        if (tc_end >= t_hi)
            break;
        const int kmin = merge_next_draw<double>(base, n, anc, r, &s_kmin, s_rb);
        if (kmin >= n)
            break;
        const double rmin = s_rb[0];
        const int next_end = min(tc_end + RS_FLOAT_SLICE_TILES, t_hi);
        if (!(tile_end(next_end) < rmin))
            tc = tc_end + 1;
        else
            tc = max(tc_end + 1, find_first_not_below<double>(tile_end, T, rmin, s_cnt));
    }
    merge_write(d, sc, pph, base, n, J, anc);
}

// The resampling's two top-level scans in one launch, in their order (the spacings read the CDF
// total): the joint CDF of exact-CDF vehicles (segment ends in place, no tile prefixes), then the
// exponentials of every vehicle (tile prefixes into exp_pre).
// Writes out a deferred resampled set (see merge_write): particle k of dst copied from its ancestor
// (particle anc[k] of src below n, birth particle anc[k] - n above) with the new weight, which the
// resampling left in dst.weight[0].
__global__ __launch_bounds__(BLOCK) void k_materialize(DevParticles dst, DevParticles src, DevParticles birth,
                                                       const int* __restrict__ anc, int n)
{
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n)
        return;
    const int a = anc[k];
    const bool from_p = a < n;
    const int o = from_p ? a : a - n;
    const DevParticles& s = from_p ? src : birth;
    dst.pos[k] = s.pos[o];
    dst.vel[k] = s.vel[o];
    dst.idx[k] = s.idx[o];
    dst.assoc[k] = s.assoc[o];
    if (k > 0)
        dst.weight[k] = dst.weight[0];
}

void run_materialize(Engine* E, const Vehicle* v)
{
    if (v->N <= 0)
        return;
    launch_begin(E, K_MATERIALIZE);
    hipLaunchKernelGGL(k_materialize, dim3(div_up(v->N, BLOCK)), dim3(BLOCK), 0, E->stream, v->part[0], v->part[1],
                       v->birth, v->ancestors, v->N);
    launch_end(E, K_MATERIALIZE);
}

__global__ __launch_bounds__(BLOCK) void k_scan_top_resample(const VehDesc* __restrict__ descs,
                                                             const fx_t* __restrict__ cdf_sum,
                                                             fx_t* __restrict__ cdf_incl,
                                                             const fx_t* __restrict__ exp_sum,
                                                             fx_t* __restrict__ exp_pre, DevScalars* __restrict__ scs)
{
    scan_top_body(descs, SP_J, cdf_sum, (fx_t*)nullptr, cdf_incl, scs, SCAN_RESAMPLE, ONLY_CDF_EXACT);
    __syncthreads();
    scan_top_body(descs, SP_E, exp_sum, exp_pre, (fx_t*)nullptr, scs, SCAN_EXP, ONLY_ALL);
}

void run_resampling(Engine* E, int pph, int modes, bool exp_done)
{
    const bool exact = modes & (1 << CDF_EXACT);
    const int jtiles = E->tiles(SP_J, SCAN_Q);
    const int etiles = exp_done ? 0 : E->tiles(SP_E, SCAN_Q);
    // Tile totals of the joint CDF (exact-CDF vehicles, with the tile-local segment ends of the coarse
    // index; the merge recomputes entries) and of the exponentials (multinomial draws) in one launch,
    // then both top-level scans in one launch.
    launch_begin(E, K_SCAN_TILES);
    hipLaunchKernelGGL((k_scan_tiles_pair<SrcJointPos, SrcExp>), dim3(jtiles + etiles), dim3(BLOCK), 0, E->stream,
                       SrcJointPos{}, E->tab(SP_J), jtiles, E->scan_tile_sum, E->cdf_tile_incl, SrcExp{}, E->tab(SP_E),
                       etiles, E->exp_tile_sum, E->d_desc);
    launch_end(E, K_SCAN_TILES);
    launch_begin(E, K_SCAN_TOP);
    hipLaunchKernelGGL(k_scan_top_resample, dim3(E->m), dim3(BLOCK), 0, E->stream, E->d_desc, E->scan_tile_sum,
                       E->cdf_tile_incl, E->exp_tile_sum, E->scan_tile_pre, E->scalars);
    launch_end(E, K_SCAN_TOP);
    // Float modes: tile totals and prefixes as doubles in the two consumed total buffers (the
    // exponentials' prefixes stay in scan_tile_pre for the merge).
    double* tsum = reinterpret_cast<double*>(E->scan_tile_sum);
    double* tpre = reinterpret_cast<double*>(E->exp_tile_sum);
    double* env = reinterpret_cast<double*>(E->cdf_tile_incl);
    const int vb = div_up(E->m, 64);
    auto float_mode = [&](auto tag) {
        using T = decltype(tag);
        launch_begin(E, K_CDF_WRITE);
        hipLaunchKernelGGL(k_cdff_tiles<T>, dim3(jtiles), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_J), tsum);
        hipLaunchKernelGGL(k_cdff_top<T>, dim3(vb), dim3(64), 0, E->stream, E->d_desc, E->m, tsum, tpre);
        hipLaunchKernelGGL(k_cdff_apply<T>, dim3(jtiles), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_J), tpre,
                           env);
        launch_end(E, K_CDF_WRITE);
    };
    if (modes & (1 << CDF_FP64))
        float_mode(0.0);
    if (modes & (1 << CDF_FP32))
        float_mode(0.0f);
    if (modes & (1 << CDF_FP32C))
        float_mode(f32c(0));
    if (modes & ~(1 << CDF_EXACT))
        hipLaunchKernelGGL(k_cdff_env, dim3(vb), dim3(64), 0, E->stream, E->d_desc, E->m, env, E->scalars);

    const int mt = E->tiles(SP_P, RS_TILE / UNIT);
    const bool wide = E->tune.merge_slice == 0 ? mt <= RS_WIDE_MAX_TILES : E->tune.merge_slice == 2;
    launch_begin(E, K_RS_MERGE);
    if (exact)
    {
        if (wide)
            hipLaunchKernelGGL(k_rs_merge_exact<RS_WINDOW_WIDE>, dim3(mt), dim3(BLOCK), 0, E->stream, E->d_desc,
                               E->tab(SP_P), pph, E->scan_tile_pre, E->cdf_tile_incl, E->scalars);
        else
            hipLaunchKernelGGL(k_rs_merge_exact<RS_WINDOW_NARROW>, dim3(mt), dim3(BLOCK), 0, E->stream, E->d_desc,
                               E->tab(SP_P), pph, E->scan_tile_pre, E->cdf_tile_incl, E->scalars);
    }
    if (modes & ~(1 << CDF_EXACT))
        hipLaunchKernelGGL(k_rs_merge_float, dim3(mt), dim3(BLOCK), 0, E->stream, E->d_desc, E->tab(SP_P), pph,
                           E->scan_tile_pre, (const void*)E->cdf_tile_incl, E->scalars);
    launch_end(E, K_RS_MERGE);
}

}  // namespace dogm_hip
