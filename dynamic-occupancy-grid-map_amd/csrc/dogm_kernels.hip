// Stage kernels of the DOGM particle-filter update on gfx950.
//
// Layout in HBM: particles structure-of-arrays (float4 state, int cell index, float weight,
// uint8 association), grid cells structure-of-arrays (one array per field), measurement cells as
// given by the caller (16-byte AoS).  Every per-cell sum is an exact 128-bit fixed-point
// segmented reduction over the cell-sorted particles (dogm_device.h).
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

void run_fill_u32(dogm_handle* h, unsigned* p, unsigned v, long long n)
{
    if (n <= 0)
        return;
    launch_begin(h, K_FILL);
    hipLaunchKernelGGL(k_fill_u32, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, h->stream, p, v, n);
    launch_end(h, K_FILL);
}

void run_fill_i32(dogm_handle* h, int* p, int v, long long n)
{
    run_fill_u32(h, reinterpret_cast<unsigned*>(p), (unsigned)v, n);
}

// ---------------------------------------------------------------------------------------------
// Ego-motion compensation (dogm/src/kernel/ego_motion_compensation.cu:16-43)
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_move_particles(DevParticles P, int n, float mx, float my)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n)
        return;
    float4 s = P.state[i];
    s.x = s.x - mx;
    s.y = s.y - my;
    P.state[i] = s;
}

void run_move_particles(dogm_handle* h, float mx, float my)
{
    launch_begin(h, K_MOVE_PARTICLES);
    hipLaunchKernelGGL(k_move_particles, dim3(div_up(h->N, BLOCK)), dim3(BLOCK), 0, h->stream, h->part[h->cur],
                       h->N, mx, my);
    launch_end(h, K_MOVE_PARTICLES);
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

void run_shift_field(dogm_handle* h, const void* src, void* dst, int ix, int iy)
{
    launch_begin(h, K_SHIFT_FIELD);
    hipLaunchKernelGGL(k_shift_field, dim3(div_up(h->C, BLOCK)), dim3(BLOCK), 0, h->stream,
                       static_cast<const uint32_t*>(src), static_cast<uint32_t*>(dst), h->gs, ix, iy);
    launch_end(h, K_SHIFT_FIELD);
}

// ---------------------------------------------------------------------------------------------
// Prediction (dogm/src/kernel/predict.cu:16-52), with the pending ego shift of the frame fused in
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

__global__ __launch_bounds__(BLOCK) void k_predict(DevParticles P, int n, float dt, float sp, float sv, float ps,
                                                   int gs, uint64_t seed, uint32_t inst, uint32_t epoch, int move,
                                                   float mx, float my)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n)
        return;
    float4 s = P.state[i];
    const float w_in = P.weight[i];
    if (move)
    {
        s.x = s.x - mx;
        s.y = s.y - my;
    }
    const u32x4 r = draw(seed, inst, TAG_PREDICT, epoch, (uint32_t)i);
    float n0, n1, n2, n3;
    box_muller(r.x, r.y, &n0, &n1);
    box_muller(r.z, r.w, &n2, &n3);
    const float zero = 0.0f, one = 1.0f;
    const float npx = normal_scale(n0, zero, sp), npy = normal_scale(n1, zero, sp);
    const float nvx = normal_scale(n2, zero, sv), nvy = normal_scale(n3, zero, sv);
    // glm mat4 * vec4 association ((c0 x + c1 y) + (c2 vx + c3 vy)) of the transposed constant
    // velocity matrix (dogm.cu:246-253).
    const float sx = (one * s.x + zero * s.y) + (dt * s.z + zero * s.w);
    const float sy = (zero * s.x + one * s.y) + (zero * s.z + dt * s.w);
    const float svx = (zero * s.x + zero * s.y) + (one * s.z + zero * s.w);
    const float svy = (zero * s.x + zero * s.y) + (zero * s.z + one * s.w);
    float4 o;
    o.x = sx + npx;
    o.y = sy + npy;
    o.z = svx + nvx;
    o.w = svy + nvy;
    float w = ps * w_in;
    if ((o.x > (float)(gs - 1) || o.x < zero) || (o.y > (float)(gs - 1) || o.y < zero))
        w = zero;
    P.state[i] = o;
    P.weight[i] = w;
    P.idx[i] = clamp_cell(o.x, gs) + gs * clamp_cell(o.y, gs);
}

void run_predict(dogm_handle* h, float dt, bool move, float mx, float my)
{
    const dogm_params& p = h->params;
    launch_begin(h, K_PREDICT);
    hipLaunchKernelGGL(k_predict, dim3(div_up(h->N, BLOCK)), dim3(BLOCK), 0, h->stream, h->part[h->cur], h->N, dt,
                       p.stddev_process_noise_position, p.stddev_process_noise_velocity, p.persistence_prob, h->gs,
                       h->seed, h->instance, h->epoch, move ? 1 : 0, mx, my);
    launch_end(h, K_PREDICT);
    ++h->epoch;
}

// ---------------------------------------------------------------------------------------------
// Assignment (dogm/src/dogm.cu:262-281, init.cu:86-93, particle_to_grid.cu:16-44)
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_reset_bounds(int* __restrict__ start, int* __restrict__ end, int C)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c < C)
    {
        start[c] = -1;
        end[c] = -1;
    }
}

void run_reset_bounds(dogm_handle* h)
{
    launch_begin(h, K_RESET_BOUNDS);
    hipLaunchKernelGGL(k_reset_bounds, dim3(div_up(h->C, BLOCK)), dim3(BLOCK), 0, h->stream, h->grid.start,
                       h->grid.end, h->C);
    launch_end(h, K_RESET_BOUNDS);
}

// Gathers the payload into sorted order and records first/last occurrence of every cell.
__global__ __launch_bounds__(BLOCK) void k_sort_gather(DevParticles src, DevParticles dst,
                                                       const int* __restrict__ skeys, const int* __restrict__ perm,
                                                       float* __restrict__ weight_array, int* __restrict__ start,
                                                       int* __restrict__ end, int n)
{
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n)
        return;
    const int o = perm[k];
    const int key = skeys[k];
    const float w = src.weight[o];
    dst.state[k] = src.state[o];
    dst.weight[k] = w;
    dst.assoc[k] = src.assoc[o];
    dst.idx[k] = key;
    weight_array[k] = w;
    if (k == 0 || skeys[k - 1] != key)
        start[key] = k;
    if (k == n - 1 || skeys[k + 1] != key)
        end[key] = k;
}

void run_sort_gather(dogm_handle* h, const int* skeys, const int* perm)
{
    const int other = h->cur ^ 1;
    launch_begin(h, K_SORT_GATHER);
    hipLaunchKernelGGL(k_sort_gather, dim3(div_up(h->N, BLOCK)), dim3(BLOCK), 0, h->stream, h->part[h->cur],
                       h->part[other], skeys, perm, h->weight_array, h->grid.start, h->grid.end, h->N);
    launch_end(h, K_SORT_GATHER);
    h->cur = other;
}

// ---------------------------------------------------------------------------------------------
// Occupancy / mass update (dogm/src/dogm.cu:283-298, mass_update.cu:16-93)
// ---------------------------------------------------------------------------------------------
struct OccOp
{
    static constexpr int NCH = 1;
    static constexpr int ITEMS = 8;
    const float* weight_array;
    float* m_raw;
    __device__ void load(int p0, int cnt, const int (&)[ITEMS], float (&v)[ITEMS][NCH]) const
    {
        float w[ITEMS];
        load_run<ITEMS>(weight_array, p0, cnt, w);
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
            v[i][0] = w[i];
    }
    __device__ void finalize(int key, const fx_t (&s)[1]) const { m_raw[key] = fx_to_f32(s[0]); }
};

__global__ __launch_bounds__(BLOCK) void k_mass_update(DevGrid G, const float* __restrict__ free_src,
                                                       float* __restrict__ free_dst,
                                                       const dogm_measurement_cell* __restrict__ meas,
                                                       const float* __restrict__ m_raw, float* __restrict__ born,
                                                       int C, int gs, float alpha, float pB, int shifted, int ix,
                                                       int iy)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= C)
        return;
    const float one = 1.0f;
    float m_occ_pred = 0.0f;
    if (G.start[c] != -1)
        m_occ_pred = m_raw[c];
    if (m_occ_pred > one)
        m_occ_pred = one;
    float fprev;
    if (shifted)
    {
        const int x = c % gs, y = c / gs;
        const int sx = x + ix, sy = y + iy;
        fprev = (sx > 0 && sx < gs && sy > 0 && sy < gs) ? free_src[sx + gs * sy] : 0.0f;
    }
    else
    {
        fprev = free_src[c];
    }
    const dogm_measurement_cell mc = meas[c];
    const float m_free_pred = fminf(alpha * fprev, one - m_occ_pred);
    const float unknown_pred = one - m_occ_pred - m_free_pred;
    const float meas_unknown = one - mc.free_mass - mc.occ_mass;
    const float K = m_free_pred * mc.occ_mass + m_occ_pred * mc.free_mass;
    const float occ_up =
        (m_occ_pred * meas_unknown + unknown_pred * mc.occ_mass + m_occ_pred * mc.occ_mass) / (one - K);
    const float free_up =
        (m_free_pred * meas_unknown + unknown_pred * mc.free_mass + m_free_pred * mc.free_mass) / (one - K);
    const float rho_b = (occ_up * pB * (one - m_occ_pred)) / (m_occ_pred + pB * (one - m_occ_pred));
    const float rho_p = occ_up - rho_b;
    born[c] = rho_b;
    G.f[GF_PERS][c] = rho_p;
    G.f[GF_NEW_BORN][c] = rho_b;
    free_dst[c] = free_up;
    G.f[GF_OCC][c] = occ_up;
    G.f[GF_PRED][c] = m_occ_pred;
}

// normalize_weights (mass_update.cu:51-59) for cells whose predicted occupancy exceeds 1.
__global__ __launch_bounds__(BLOCK) void k_normalize(DevParticles P, float* __restrict__ weight_array,
                                                     const float* __restrict__ m_raw, int n)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n)
        return;
    const float m = m_raw[P.idx[i]];
    if (m > 1.0f)
    {
        const float w = weight_array[i] / m;
        weight_array[i] = w;
        P.weight[i] = w;
    }
}

void run_occupancy(dogm_handle* h, float alpha, bool shifted, int ix, int iy)
{
    const DevParticles P = h->part[h->cur];
    OccOp op{h->weight_array, h->m_raw};
    launch_segreduce(h, K_SEG_OCC, K_SEG_OCC_FIX, op, P.idx, h->N);
    float* free_src = h->grid.f[GF_FREE];
    float* free_dst = h->free_alt;
    launch_begin(h, K_MASS_UPDATE);
    hipLaunchKernelGGL(k_mass_update, dim3(div_up(h->C, BLOCK)), dim3(BLOCK), 0, h->stream, h->grid, free_src,
                       free_dst, h->meas, h->m_raw, h->born, h->C, h->gs, alpha, h->params.birth_prob,
                       shifted ? 1 : 0, ix, iy);
    launch_end(h, K_MASS_UPDATE);
    h->grid.f[GF_FREE] = free_dst;
    h->free_alt = free_src;
    launch_begin(h, K_NORMALIZE);
    hipLaunchKernelGGL(k_normalize, dim3(div_up(h->N, BLOCK)), dim3(BLOCK), 0, h->stream, P, h->weight_array,
                       h->m_raw, h->N);
    launch_end(h, K_NORMALIZE);
}

// ---------------------------------------------------------------------------------------------
// Persistent particle update (dogm/src/dogm.cu:300-321, update_persistent_particles.cu:16-87)
// ---------------------------------------------------------------------------------------------
// Kernel1 (lw = likelihood * w, written to weight_array) is fused into the reduction's load.
struct PersOp
{
    static constexpr int NCH = 1;
    static constexpr int ITEMS = 8;
    const float* weight;
    float* weight_array;
    const dogm_measurement_cell* meas;
    DevGrid G;
    __device__ void load(int p0, int cnt, const int (&key)[ITEMS], float (&v)[ITEMS][NCH]) const
    {
        float w[ITEMS], lw[ITEMS];
        load_run<ITEMS>(weight, p0, cnt, w);
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
        {
            lw[i] = i < cnt ? meas[key[i]].likelihood * w[i] : 0.0f;
            v[i][0] = lw[i];
        }
        store_run<ITEMS>(weight_array, p0, cnt, lw);
    }
    __device__ void finalize(int key, const fx_t (&s)[1]) const
    {
        const float acc = fx_to_f32(s[0]);
        const float rho_p = G.f[GF_PERS][key];
        G.f[GF_MU_A][key] = acc > 0.0f ? rho_p / acc : 0.0f;
        const float pred = G.f[GF_PRED][key];
        G.f[GF_MU_UA][key] = pred > 0.0f ? rho_p / pred : 0.0f;
    }
};

// Writes 0 into up to five fields of every cell without particles.
__global__ __launch_bounds__(BLOCK) void k_empty_cells(const int* __restrict__ start, float* f0, float* f1,
                                                       float* f2, float* f3, float* f4, int nf, int C)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= C || start[c] != -1)
        return;
    f0[c] = 0.0f;
    if (nf > 1)
        f1[c] = 0.0f;
    if (nf > 2)
        f2[c] = 0.0f;
    if (nf > 3)
        f3[c] = 0.0f;
    if (nf > 4)
        f4[c] = 0.0f;
}

__global__ __launch_bounds__(BLOCK) void k_post_weights(DevParticles P, float* __restrict__ weight_array,
                                                        const dogm_measurement_cell* __restrict__ meas, DevGrid G,
                                                        int n)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n)
        return;
    const int c = P.idx[i];
    const float pA = meas[c].p_A;
    weight_array[i] = pA * G.f[GF_MU_A][c] * weight_array[i] + (1.0f - pA) * G.f[GF_MU_UA][c] * P.weight[i];
}

void run_persistent(dogm_handle* h)
{
    const DevParticles P = h->part[h->cur];
    PersOp op{P.weight, h->weight_array, h->meas, h->grid};
    launch_segreduce(h, K_SEG_PERS, K_SEG_PERS_FIX, op, P.idx, h->N);
    launch_begin(h, K_EMPTY_CELLS);
    hipLaunchKernelGGL(k_empty_cells, dim3(div_up(h->C, BLOCK)), dim3(BLOCK), 0, h->stream, h->grid.start,
                       h->grid.f[GF_MU_A], h->grid.f[GF_MU_UA], nullptr, nullptr, nullptr, 2, h->C);
    launch_end(h, K_EMPTY_CELLS);
    launch_begin(h, K_POST_WEIGHTS);
    hipLaunchKernelGGL(k_post_weights, dim3(div_up(h->N, BLOCK)), dim3(BLOCK), 0, h->stream, P, h->weight_array,
                       h->meas, h->grid, h->N);
    launch_end(h, K_POST_WEIGHTS);
}

struct SrcArray
{
    const float* a;
    __device__ void load4(int p0, int cnt, float (&v)[4]) const { load_run<4>(a, p0, cnt, v); }
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
    int* pend;
    const DevScalars* sc;
    int C, v;
    __device__ void apply4(int p0, int cnt, const fx_t (&pre)[4], const fx_t (&)[4]) const
    {
        int e[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            e[i] = i < cnt ? order_end_at(p0 + i, pre[i], sc, C, v) : 0;
        store_run<4>(pend, p0, cnt, e);
    }
};

// Birth placement and weights per cell (init_new_particles.cu:126-155).
struct SinkBirth
{
    int* pend;
    const DevScalars* sc;
    const dogm_measurement_cell* meas;
    const float* born;
    float* wA;
    float* wUA;
    int C, v;
    __device__ void apply4(int p0, int cnt, const fx_t (&pre)[4], const fx_t (&own)[4]) const
    {
        float bm[4];
        load_run<4>(born, p0, cnt, bm);
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
                const int s = j == 0 ? 0 : order_end_at(j - 1, pre[i] - own[i], sc, C, v);
                const int n = e[i] - s;
                const float pA = meas[j].p_A;
                const int nu_A = (int)roundf((float)n * pA);
                const int nu_UA = n - nu_A;
                a[i] = nu_A > 0 ? (pA * bm[i]) / (float)nu_A : 0.0f;
                u[i] = nu_UA > 0 ? ((1.0f - pA) * bm[i]) / (float)nu_UA : 0.0f;
            }
        }
        store_run<4>(pend, p0, cnt, e);
        store_run<4>(wA, p0, cnt, a);
        store_run<4>(wUA, p0, cnt, u);
    }
};

// ---------------------------------------------------------------------------------------------
// Birth particles (dogm/src/dogm.cu:323-348, init.cu:46-67, init_new_particles.cu:157-195)
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_birth_particles(DevParticles Bp, int nb, const int* __restrict__ pend,
                                                           int C, int gs, const dogm_measurement_cell* __restrict__ meas,
                                                           const float* __restrict__ wA,
                                                           const float* __restrict__ wUA, float sd, uint64_t seed,
                                                           uint32_t inst, uint32_t epoch)
{
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= nb)
        return;
    const int j = upper_bound_i32(pend, C, k);
    const int s = j > 0 ? pend[j - 1] : 0;
    const int n = pend[j] - s;
    const int nu_A = (int)roundf((float)n * meas[j].p_A);
    const bool assoc = (k - s) <= nu_A;
    const u32x4 r = draw(seed, inst, TAG_BIRTH, epoch, (uint32_t)k);
    float n0, n1;
    box_muller(r.x, r.y, &n0, &n1);
    float4 st;
    st.x = (float)(j % gs) + 0.5f;
    st.y = (float)j / (float)gs + 0.5f;
    st.z = normal_scale(n0, 0.0f, sd);
    st.w = normal_scale(n1, 0.0f, sd);
    Bp.state[k] = st;
    Bp.idx[k] = j;
    Bp.weight[k] = assoc ? wA[j] : wUA[j];
    Bp.assoc[k] = assoc ? 1 : 0;
}

void run_births(dogm_handle* h)
{
    // With B = 0 the cell pass still runs (every cell gets n_j = 0 and w_A = w_UA = 0) and the
    // stage still consumes its RNG epoch, so the stream of later stages does not depend on B.
    SrcArray src{h->born};
    SinkBirth sink{h->pend, h->scalars, h->meas, h->born, h->grid.f[GF_W_A], h->grid.f[GF_W_UA], h->C, h->B};
    run_exact_scan(h, src, sink, h->C, SCAN_ORDERS, h->B, K_BIRTH_CELLS);
    if (h->B <= 0)
    {
        ++h->epoch;
        return;
    }
    launch_begin(h, K_BIRTH_PARTICLES);
    hipLaunchKernelGGL(k_birth_particles, dim3(div_up(h->B, BLOCK)), dim3(BLOCK), 0, h->stream, h->birth, h->B,
                       h->pend, h->C, h->gs, h->meas, h->grid.f[GF_W_A], h->grid.f[GF_W_UA],
                       h->params.stddev_velocity, h->seed, h->instance, h->epoch);
    launch_end(h, K_BIRTH_PARTICLES);
    ++h->epoch;
}

// ---------------------------------------------------------------------------------------------
// First-measurement initialisation (dogm/src/dogm.cu:217-241, init_new_particles.cu:76-124)
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_copy_masses(const dogm_measurement_cell* __restrict__ meas,
                                                       float* __restrict__ born, int C)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c < C)
        born[c] = meas[c].occ_mass;
}

__global__ __launch_bounds__(BLOCK) void k_init_particles(DevParticles P, int n, const int* __restrict__ pend, int C,
                                                          int gs, float v, float new_weight, uint64_t seed,
                                                          uint32_t inst, uint32_t epoch)
{
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n)
        return;
    const int j = upper_bound_i32(pend, C, k);
    const u32x4 r = draw(seed, inst, TAG_INIT_PARTICLES, epoch, (uint32_t)k);
    float4 st;
    st.x = (float)(j % gs) + 0.5f;
    st.y = (float)(j / gs) + 0.5f;
    st.z = -v + (v - -v) * uniform01(r.x);
    st.w = -v + (v - -v) * uniform01(r.y);
    P.state[k] = st;
    P.idx[k] = j;
    P.weight[k] = new_weight;
    P.assoc[k] = 0;
}

void run_init_particles(dogm_handle* h)
{
    launch_begin(h, K_COPY_MASSES);
    hipLaunchKernelGGL(k_copy_masses, dim3(div_up(h->C, BLOCK)), dim3(BLOCK), 0, h->stream, h->meas, h->born, h->C);
    launch_end(h, K_COPY_MASSES);
    SrcArray src{h->born};
    SinkOrders sink{h->pend, h->scalars, h->C, h->N};
    run_exact_scan(h, src, sink, h->C, SCAN_ORDERS, h->N, K_INIT_CELLS);
    const float new_weight = 1.0f / (float)h->N;
    launch_begin(h, K_INIT_PARTICLES);
    hipLaunchKernelGGL(k_init_particles, dim3(div_up(h->N, BLOCK)), dim3(BLOCK), 0, h->stream, h->part[h->cur],
                       h->N, h->pend, h->C, h->gs, h->params.init_max_velocity, new_weight, h->seed, h->instance,
                       h->epoch);
    launch_end(h, K_INIT_PARTICLES);
    ++h->epoch;
}

// ---------------------------------------------------------------------------------------------
// Statistical moments (dogm/src/dogm.cu:350-384, statistical_moments.cu:16-106)
// ---------------------------------------------------------------------------------------------
struct MomOp
{
    static constexpr int NCH = 5;
    static constexpr int ITEMS = 4;
    const float* weight_array;
    const float4* state;
    DevGrid G;
    __device__ void load(int p0, int cnt, const int (&)[ITEMS], float (&v)[ITEMS][NCH]) const
    {
        float w[ITEMS];
        load_run<ITEMS>(weight_array, p0, cnt, w);
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
        {
            const float4 s = i < cnt ? state[p0 + i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            v[i][0] = w[i] * s.z;
            v[i][1] = w[i] * s.w;
            v[i][2] = w[i] * s.z * s.z;
            v[i][3] = w[i] * s.w * s.w;
            v[i][4] = w[i] * s.z * s.w;
        }
    }
    __device__ void finalize(int key, const fx_t (&s)[5]) const
    {
        const float rho_p = G.f[GF_PERS][key];
        float mx = 0.0f, my = 0.0f, vx = 0.0f, vy = 0.0f, cv = 0.0f;
        if (rho_p > 0.0f)
        {
            const float inv = 1.0f / rho_p;
            mx = inv * fx_to_f32(s[0]);
            my = inv * fx_to_f32(s[1]);
            vx = inv * fx_to_f32(s[2]) - mx * mx;
            vy = inv * fx_to_f32(s[3]) - my * my;
            cv = inv * fx_to_f32(s[4]) - mx * my;
        }
        G.f[GF_MEAN_X][key] = mx;
        G.f[GF_MEAN_Y][key] = my;
        G.f[GF_VAR_X][key] = vx;
        G.f[GF_VAR_Y][key] = vy;
        G.f[GF_COVAR][key] = cv;
    }
};

void run_moments(dogm_handle* h)
{
    const DevParticles P = h->part[h->cur];
    MomOp op{h->weight_array, P.state, h->grid};
    launch_segreduce(h, K_SEG_MOM, K_SEG_MOM_FIX, op, P.idx, h->N);
    launch_begin(h, K_EMPTY_CELLS);
    hipLaunchKernelGGL(k_empty_cells, dim3(div_up(h->C, BLOCK)), dim3(BLOCK), 0, h->stream, h->grid.start,
                       h->grid.f[GF_MEAN_X], h->grid.f[GF_MEAN_Y], h->grid.f[GF_VAR_X], h->grid.f[GF_VAR_Y],
                       h->grid.f[GF_COVAR], 5, h->C);
    launch_end(h, K_EMPTY_CELLS);
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

void run_pack_grid(dogm_handle* h)
{
    launch_begin(h, K_PACK_GRID);
    hipLaunchKernelGGL(k_pack_grid, dim3(div_up(h->C, BLOCK)), dim3(BLOCK), 0, h->stream, h->grid, h->aos_tmp, h->C);
    launch_end(h, K_PACK_GRID);
}

void run_unpack_grid(dogm_handle* h)
{
    launch_begin(h, K_UNPACK_GRID);
    hipLaunchKernelGGL(k_unpack_grid, dim3(div_up(h->C, BLOCK)), dim3(BLOCK), 0, h->stream, h->grid, h->aos_tmp,
                       h->C);
    launch_end(h, K_UNPACK_GRID);
}

}  // namespace dogm_hip
