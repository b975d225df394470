// CPU restatement of the reference DOGM particle-filter update (dogm::DOGM::updateGrid).
//
// TEST INFRASTRUCTURE ONLY.  This is the checker for the gfx950 HIP path: it is loaded by
// tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and never by the
// product library.  It follows the reference's stage order (dogm/src/dogm.cu:115-131) and
// per-element formulas, with the deterministic resolutions of the reference's races and
// float-order dependences listed at each function.  It shares only dogm_numerics.h (Philox,
// Box-Muller, exact fixed-point sums) with the product; those primitives are pinned by their
// own known-answer tests.
//
// Beside the update it restates the steps before and after it: the two-level lidar measurement grid
// (oracle_lidar_grid) and the demo's result extraction (oracle_occupancy, oracle_cells_with_velocity).
// The velocity test restates OpenCV's f32 evaluation of v^T S^-1 v; closed forms pin it, but the last-bit
// agreement with OpenCV itself on arbitrary inputs is parity unpinned (OpenCV is not installed and the
// reference's tests hold no vectors for it).
//
// Two modes:
//   T = float   bit-exact twin of the GPU path (same f32 operations, same order, no contraction);
//   T = double  the same algorithm with every per-element formula in double, used to report the
//               error of the f32 results against an FP64 evaluation.
//
// Loops over independent particles or cells run on OpenMP threads.  Results do not depend on the
// thread count: every parallel loop writes disjoint elements, prefixes and sums are exact integer
// (fixed-point) arithmetic, and the sort is a stable LSD radix sort (the permutation of a stable
// sort is unique).
#include "../include/dogm/dogm_c.h"
#include "../dynamic-occupancy-grid-map_amd/csrc/dogm_numerics.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>

using namespace dogm_num;

namespace
{

// Contiguous chunk [b, e) of n elements for thread t of nt.
inline void chunk(long long n, int t, int nt, long long* b, long long* e)
{
    *b = n * t / nt;
    *e = n * (t + 1) / nt;
}

// Inclusive exact prefix out[i] = value(0) + ... + value(i).  Integer addition, so the chunked
// parallel evaluation equals the sequential one.
template <typename F>
void fx_prefix(int n, F value, fx_t* out)
{
    std::vector<fx_t> carry((size_t)omp_get_max_threads() + 1, 0);
#pragma omp parallel
    {
        const int t = omp_get_thread_num(), nt = omp_get_num_threads();
        long long b, e;
        chunk(n, t, nt, &b, &e);
        fx_t s = 0;
        for (long long i = b; i < e; ++i)
        {
            s += value((int)i);
            out[i] = s;
        }
        carry[t + 1] = s;
#pragma omp barrier
#pragma omp single
        for (int k = 1; k <= nt; ++k)
            carry[k] += carry[k - 1];
        const fx_t off = carry[t];
        for (long long i = b; i < e; ++i)
            out[i] += off;
    }
}

// Stable sort of 0..n-1 by key[] in [0, max_key]: LSD radix with 11-bit digits; per pass each thread
// counts its contiguous chunk, offsets run digit-major then thread order, and each thread scatters
// its chunk in order, so equal keys keep their input order.
void stable_sort_by_key(const int* key, int n, int max_key, std::vector<int>& perm)
{
    constexpr int BITS = 11, R = 1 << BITS;
    perm.resize(n);
    std::iota(perm.begin(), perm.end(), 0);
    int bits = 0;
    while (bits < 31 && (max_key >> bits) != 0)
        ++bits;
    std::vector<int> tmp(n);
    const int nt_max = omp_get_max_threads();
    std::vector<long long> cnt((size_t)nt_max * R);
    for (int shift = 0; shift < bits; shift += BITS)
    {
        std::fill(cnt.begin(), cnt.end(), 0);
        int nt_used = 1;
#pragma omp parallel
        {
            const int t = omp_get_thread_num(), nt = omp_get_num_threads();
#pragma omp single
            nt_used = nt;
            long long b, e;
            chunk(n, t, nt, &b, &e);
            long long* c = &cnt[(size_t)t * R];
            for (long long i = b; i < e; ++i)
                ++c[(key[perm[i]] >> shift) & (R - 1)];
#pragma omp barrier
#pragma omp single
            {
                long long run = 0;
                for (int d = 0; d < R; ++d)
                    for (int u = 0; u < nt_used; ++u)
                    {
                        const long long v = cnt[(size_t)u * R + d];
                        cnt[(size_t)u * R + d] = run;
                        run += v;
                    }
            }
            for (long long i = b; i < e; ++i)
                tmp[c[(key[perm[i]] >> shift) & (R - 1)]++] = perm[i];
        }
        perm.swap(tmp);
    }
}

template <typename T>
struct ParticleSet
{
    std::vector<T> x, y, vx, vy, w;
    std::vector<int> idx;
    std::vector<uint8_t> assoc;

    void resize(int n)
    {
        x.assign(n, T(0));
        y.assign(n, T(0));
        vx.assign(n, T(0));
        vy.assign(n, T(0));
        w.assign(n, T(0));
        idx.assign(n, 0);
        assoc.assign(n, 0);
    }
};

enum GridField
{
    GF_NEW_BORN,
    GF_PERS,
    GF_FREE,
    GF_OCC,
    GF_PRED,
    GF_MU_A,
    GF_MU_UA,
    GF_W_A,
    GF_W_UA,
    GF_MEAN_X,
    GF_MEAN_Y,
    GF_VAR_X,
    GF_VAR_Y,
    GF_COVAR,
    GF_COUNT
};

struct OracleBase
{
    virtual ~OracleBase() = default;
    virtual int update_measurement_grid(const dogm_measurement_cell* meas) = 0;
    virtual void update_pose(float x, float y, float yaw) = 0;
    virtual void initialize_particles() = 0;
    virtual void particle_prediction(float dt) = 0;
    virtual void particle_assignment() = 0;
    virtual void grid_cell_occupancy_update(float dt) = 0;
    virtual void update_persistent_particles() = 0;
    virtual void initialize_new_particles() = 0;
    virtual void statistical_moments() = 0;
    virtual void resampling() = 0;
    virtual int get(int field, double* out) = 0;
    virtual int set(int field, const double* in) = 0;
    virtual int get_i32(int field, int* out) = 0;
    virtual int set_i32(int field, const int* in) = 0;
    float pos_x = 0.0f, pos_y = 0.0f, yaw = 0.0f;
    bool first_pose = false, first_meas = false;
    int ego_mode = 0;                                // 0 reference, 1 metric (opt-in extension)
    double ref_x = 0.0, ref_y = 0.0, ref_yaw = 0.0;  // metric mode: pose of the grid frame
    uint32_t epoch = 0;
    int cdf_mode = 0;  // CdfMode of the resampling scan (f32 checker only; the f64 twin is exact)
    int draw_mode = 0;  // DrawMode of the resampling draws

    // Draw prefixes S_0 .. S_N of a resampling: exponential spacings (multinomial, the reference's
    // draw) or k * 2^32 + u32 (systematic; dogm_numerics.h, systematic_s).
    void draw_prefixes(int N, uint64_t seed, uint32_t instance, std::vector<fx_t>& spac) const
    {
        spac.assign(N + 1, 0);
        if (draw_mode == DRAW_SYSTEMATIC)
        {
            const uint32_t u = systematic_u32(seed, instance, epoch);
            for (int k = 0; k < N; ++k)
                spac[k] = systematic_s(k, u);
            spac[N] = systematic_s(N, 0u);
            return;
        }
        fx_prefix(
            N + 1,
            [&](int k) {
                float e4[4];
                resample_exponentials4(seed, instance, epoch, (uint32_t)(k >> 2), e4);
                return fx_from_f32(e4[k & 3]);
            },
            spac.data());
    }
};

// Velocity-moment sums of the f32 path: the device's fixed f64 association (dogm_kernels.hip,
// k_moments).  A grid's sorted particles are cut into chunks of 256 (chunk k = [256 k, 256 k + 256)),
// one wave64 per chunk, 4 consecutive particles per lane; every value is the f32 channel product
// widened to f64.
//   lane:  sequential fold that restarts where a cell starts: r = x there, r + x otherwise; the
//          lane's first particle always starts the fold (r = x);
//   wave:  segmented inclusive scan of the lanes' last folds in six simultaneous steps -- row shifts
//          1, 2, 4, 8 inside 16-lane rows, lane 15 of rows 0 / 2 into rows 1 / 3, lane 31 into rows
//          2 / 3 -- each adding the moved value only while no cell starts in the lanes between;
//   cell:  ending in lane l, its lane-local fold if it starts in lane l, else (scan of lane l - 1) +
//          fold (lane 0: the fold);
//   chunks: a cell crossing chunks is the left fold ((piece in its first chunk + whole middle chunk)
//          + ...) + piece in its last chunk.
static void wave_segscan64(double* x, const bool* head)
{
    bool f[64];
    std::memcpy(f, head, sizeof(f));
    auto step = [&](auto src_of) {
        double nx[64];
        bool nf[64];
        for (int l = 0; l < 64; ++l)
        {
            nx[l] = x[l];
            nf[l] = f[l];
            const int s = src_of(l);
            if (s >= 0)
            {
                if (!f[l])
                    nx[l] = x[s] + x[l];
                nf[l] = f[l] || f[s];
            }
        }
        std::memcpy(x, nx, sizeof(nx));
        std::memcpy(f, nf, sizeof(nf));
    };
    for (int sh : {1, 2, 4, 8})
        step([sh](int l) { return (l & 15) >= sh ? l - sh : -1; });
    step([](int l) {
        const int r = l >> 4;
        return (r == 1 || r == 3) ? 16 * r - 1 : -1;
    });
    step([](int l) { return (l >> 4) >= 2 ? 31 : -1; });
}

constexpr int MOM_CH = 5;
constexpr int MOM_CHUNK = 256;

// out[k * MOM_CH + c]: channel c's sum of the cell whose particles end at particle k (written at the
// last particle of every cell).
static void moment_sums_wave(const int* key, int N, const float* const* ch, double* out)
{
    const int chunks = (N + MOM_CHUNK - 1) / MOM_CHUNK;
    // per chunk: the piece of its last cell (scan of lane 63) and of its first cell when that cell
    // started in an earlier chunk and ends in this one
    std::vector<double> tail((size_t)chunks * MOM_CH), headp((size_t)chunks * MOM_CH);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < chunks; ++k)
    {
        const int base = k * MOM_CHUNK;
        const bool from_before = base > 0 && key[base] == key[base - 1];
        for (int c = 0; c < MOM_CH; ++c)
        {
            double r[MOM_CHUNK];
            double t[64];
            bool lh[64];
            for (int l = 0; l < 64; ++l)
            {
                t[l] = 0.0;
                lh[l] = false;
                for (int i = 0; i < 4; ++i)
                {
                    const int p = base + 4 * l + i;
                    if (p >= N)
                        break;
                    const double x = (double)ch[c][p];
                    const bool h = p == 0 || key[p] != key[p - 1];
                    lh[l] = lh[l] || h;
                    r[4 * l + i] = (i == 0 || h) ? x : r[4 * l + i - 1] + x;
                    t[l] = r[4 * l + i];
                }
            }
            double S[64];
            std::memcpy(S, t, sizeof(S));
            wave_segscan64(S, lh);
            bool seen_end = false;
            for (int l = 0; l < 64; ++l)
            {
                bool started = false;  // a cell starts in lane l at or before particle i
                for (int i = 0; i < 4; ++i)
                {
                    const int p = base + 4 * l + i;
                    if (p >= N)
                        break;
                    started = started || p == 0 || key[p] != key[p - 1];
                    if (p != N - 1 && key[p + 1] == key[p])
                        continue;
                    const double v = (started || l == 0) ? r[4 * l + i] : S[l - 1] + r[4 * l + i];
                    if (!seen_end && from_before)
                        headp[(size_t)k * MOM_CH + c] = v;
                    else
                        out[(size_t)p * MOM_CH + c] = v;
                    seen_end = true;
                }
            }
            tail[(size_t)k * MOM_CH + c] = S[63];
        }
    }
    // Cells crossing chunks, folded from the chunk where they start.
#pragma omp parallel for schedule(static)
    for (int k = 0; k < chunks - 1; ++k)
    {
        const int base = k * MOM_CHUNK, last = base + MOM_CHUNK - 1;
        const int cell = key[last];
        if (key[last + 1] != cell || (base > 0 && key[base - 1] == cell))
            continue;  // the chunk's last cell ends here, or started in an earlier chunk
        double acc[MOM_CH];
        for (int c = 0; c < MOM_CH; ++c)
            acc[c] = tail[(size_t)k * MOM_CH + c];
        int j = k + 1;
        // middle chunks: every particle of chunk j belongs to the cell and the cell goes on
        while (j < chunks - 1 && key[j * MOM_CHUNK + MOM_CHUNK - 1] == cell && key[j * MOM_CHUNK + MOM_CHUNK] == cell)
        {
            for (int c = 0; c < MOM_CH; ++c)
                acc[c] = acc[c] + tail[(size_t)j * MOM_CH + c];
            ++j;
        }
        int e = j * MOM_CHUNK;
        while (e + 1 < N && key[e + 1] == cell)
            ++e;
        for (int c = 0; c < MOM_CH; ++c)
            out[(size_t)e * MOM_CH + c] = acc[c] + headp[(size_t)j * MOM_CH + c];
    }
}

// Float CDF modes: restatement of the device association (dogm_resample.hip, k_cdff_*): per tile of
// 1024 entries, 4-element sequential folds per thread, the 64-lane wave scan of the thread totals as
// six simultaneous steps (row shifts 1, 2, 4, 8 inside 16-lane rows, then lane 15 of rows 0 / 2 into
// rows 1 / 3 and lane 31 into rows 2 / 3; every step adds the moved value), sequential carries
// across the 4 waves and across tiles; element = ((tile prefix + wave carry) + exclusive wave
// prefix) + thread-local prefix; E = running maximum.
template <typename F>
static void float_cdf(const std::vector<float>& w, std::vector<double>& E)
{
    const int J = (int)w.size();
    const int ntiles = (J + 1023) / 1024;
    F tp = (F)0;
    double emax = 0.0;
    E.assign(J, 0.0);
    for (int k = 0; k < ntiles; ++k)
    {
        F loc[256][4];
        F incl[256];
        for (int t = 0; t < 256; ++t)
        {
            F run = (F)0;
            for (int i = 0; i < 4; ++i)
            {
                const int j = k * 1024 + 4 * t + i;
                run = run + (F)(j < J ? w[j] : 0.0f);
                loc[t][i] = run;
            }
            incl[t] = run;
        }
        F W[4];
        for (int wv = 0; wv < 4; ++wv)
        {
            F* x = incl + 64 * wv;
            for (int d = 1; d <= 8; d <<= 1)
            {
                F y[64];
                for (int l = 0; l < 64; ++l)
                    y[l] = (l % 16 >= d) ? x[l] + x[l - d] : x[l];
                for (int l = 0; l < 64; ++l)
                    x[l] = y[l];
            }
            {
                F y[64];
                for (int l = 0; l < 64; ++l)
                {
                    const int row = l / 16;
                    y[l] = (row == 1 || row == 3) ? x[l] + x[16 * row - 1] : x[l];
                }
                for (int l = 0; l < 64; ++l)
                    x[l] = y[l];
            }
            {
                F y[64];
                for (int l = 0; l < 64; ++l)
                    y[l] = (l / 16 >= 2) ? x[l] + x[31] : x[l];
                for (int l = 0; l < 64; ++l)
                    x[l] = y[l];
            }
            W[wv] = x[63];
        }
        F tot = (F)0;
        for (int wv = 0; wv < 4; ++wv)
            tot = tot + W[wv];
        for (int t = 0; t < 256; ++t)
        {
            const int wv = t / 64, l = t % 64;
            F carry = (F)0;
            for (int u = 0; u < wv; ++u)
                carry = carry + W[u];
            const F excl = l == 0 ? (F)0 : incl[t - 1];
            const F base = (tp + carry) + excl;
            for (int i = 0; i < 4; ++i)
            {
                const int j = k * 1024 + 4 * t + i;
                if (j >= J)
                    continue;
                const double v = (double)(base + loc[t][i]);
                emax = v > emax ? v : emax;
                E[j] = emax;
            }
        }
        tp = tp + tot;
    }
}

// Field ids for oracle_get / oracle_set (double transport) and oracle_get_i32 / oracle_set_i32.
enum Field
{
    F_P_X = 0,
    F_P_Y,
    F_P_VX,
    F_P_VY,
    F_P_W,
    F_B_X,
    F_B_Y,
    F_B_VX,
    F_B_VY,
    F_B_W,
    F_WEIGHT_ARRAY,
    F_BORN,
    F_GRID0 = 16,  // F_GRID0 + GridField
    I_P_IDX = 100,
    I_P_ASSOC,
    I_B_IDX,
    I_B_ASSOC,
    I_G_START,
    I_G_END,
    I_ANCESTORS,
};

template <typename T>
T to_T_from_fx(fx_t v);
template <>
float to_T_from_fx<float>(fx_t v)
{
    return fx_to_f32(v);
}
template <>
double to_T_from_fx<double>(fx_t v)
{
    return fx_to_f64(v);
}
static inline fx_t fx_of(float v)
{
    return fx_from_f32(v);
}
static inline fx_t fx_of(double v)
{
    return fx_from_f64(v);
}

// Clamped cell coordinate: clamp(static_cast<int>(x), 0, grid_size - 1) of predict.cu:46-47 with
// the out-of-range and NaN conversions (undefined in C++) pinned to the clamp bounds.
template <typename T>
static inline int clamp_cell(T x, int gs)
{
    if (x >= (T)(gs - 1))
        return gs - 1;
    if (x > (T)0)
        return (int)x;
    return 0;
}

// Branch-free-equivalent of CUDA's fminf as used by predict_free_mass (mass_update.cu:18).
template <typename T>
static inline T tmin(T a, T b)
{
    return std::fmin(a, b);
}

template <typename T>
struct Oracle final : OracleBase
{
    dogm_params p;
    int gs, C, N, B;
    uint64_t seed;
    uint32_t instance;

    ParticleSet<T> part, birth;
    std::vector<int> start, end;
    std::vector<T> g[GF_COUNT];
    std::vector<dogm_measurement_cell> meas;
    std::vector<T> weight_array, born;
    std::vector<int> ancestors;

    Oracle(const dogm_params& params, uint64_t seed_, uint32_t instance_)
        : p(params), seed(seed_), instance(instance_)
    {
        gs = (int)(params.size / params.resolution);  // dogm.cu:34 (truncation)
        C = gs * gs;
        N = params.particle_count;
        B = params.new_born_particle_count;
        part.resize(N);
        birth.resize(B);
        // initGridCellsKernel (init.cu:69-84): start = end = -1, masses 0; measurement (0, 0, 1, 1).
        // The remaining fields are zero-initialised here (uninitialised device memory there).
        start.assign(C, -1);
        end.assign(C, -1);
        for (auto& f : g)
            f.assign(C, T(0));
        meas.assign(C, dogm_measurement_cell{0.0f, 0.0f, 1.0f, 1.0f});
        weight_array.assign(N, T(0));
        born.assign(C, T(0));
        ancestors.assign(N, 0);
    }

    // ---- dogm.cu:205-215 -----------------------------------------------------------------------
    int update_measurement_grid(const dogm_measurement_cell* m) override
    {
        std::memcpy(meas.data(), m, sizeof(dogm_measurement_cell) * (size_t)C);
        if (!first_meas)
        {
            initialize_particles();
            first_meas = true;
        }
        return 0;
    }

    // ---- dogm.cu:161-203, ego_motion_compensation.cu:16-43 --------------------------------------
    // First pose is stored only.  A move of more than one resolution in x or y shifts particles by
    // the integer cell offset (x -= x_move) and the grid by new[x, y] = old[x + x_move, y + y_move]
    // when 0 < source < grid_size (strict lower bound, ego_motion_compensation.cu:38), else zero.
    // Yaw is stored but never applied; the sub-cell remainder is dropped.
    // Metric ego compensation (the library's opt-in DOGM_EGO_METRIC; no counterpart in the
    // reference).  Grid frame = sensor frame at `ref`: a world point w at cell
    // c + R(-yaw)(w - ref) / resolution.  A pose change applies the whole cells of the translation,
    // expressed in the new frame, and rotates by -dpsi about c = grid_size / 2: particles exactly,
    // every grid field by the nearest source cell of each destination cell centre.
    // Copy of the cell bounds and grid fields (the source of a shift).
    void snapshot_grid(std::vector<int>& os, std::vector<int>& oe, std::vector<T> (&og)[GF_COUNT]) const
    {
        os.resize(C);
        oe.resize(C);
        for (int f = 0; f < GF_COUNT; ++f)
            og[f].resize(C);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < C; ++i)
        {
            os[i] = start[i];
            oe[i] = end[i];
            for (int f = 0; f < GF_COUNT; ++f)
                og[f][i] = g[f][i];
        }
    }

    void update_pose_metric(float new_x, float new_y, float new_yaw)
    {
        if (!first_pose)
        {
            ref_x = new_x;
            ref_y = new_y;
            ref_yaw = new_yaw;
            pos_x = new_x;
            pos_y = new_y;
            yaw = new_yaw;
            first_pose = true;
            return;
        }
        const double res = p.resolution;
        const double two_pi = 2.0 * 3.14159265358979323846;
        const double dpsi = std::remainder((double)new_yaw - ref_yaw, two_pi);
        const double cy = std::cos((double)new_yaw), sy = std::sin((double)new_yaw);
        const double wx = (double)new_x - ref_x, wy = (double)new_y - ref_y;
        // displacement in cells along the new frame's axes, whole cells only
        const double fx = (cy * wx + sy * wy) / res, fy = (-sy * wx + cy * wy) / res;
        const double nx = std::trunc(fx), ny = std::trunc(fy);
        ref_yaw = new_yaw;
        ref_x += (cy * nx - sy * ny) * res;
        ref_y += (sy * nx + cy * ny) * res;
        pos_x = (float)ref_x;
        pos_y = (float)ref_y;
        yaw = new_yaw;
        if (dpsi == 0.0 && nx == 0.0 && ny == 0.0)
            return;
        const float c = 0.5f * (float)gs;
        const float cs = (float)std::cos(dpsi), sn = (float)std::sin(dpsi);
        const float tx = (float)(-nx), ty = (float)(-ny);
        const bool translate_only = sn == 0.0f && cs == 1.0f;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i)
        {
            if (translate_only)
            {
                part.x[i] = part.x[i] + (T)tx;
                part.y[i] = part.y[i] + (T)ty;
                continue;
            }
            const T dx = part.x[i] - (T)c, dy = part.y[i] - (T)c;
            const T rx = (T)cs * dx + (T)sn * dy;
            const T ry = (T)cs * dy - (T)sn * dx;
            const T vx = (T)cs * part.vx[i] + (T)sn * part.vy[i];
            const T vy = (T)cs * part.vy[i] - (T)sn * part.vx[i];
            part.x[i] = ((T)c + rx) + (T)tx;
            part.y[i] = ((T)c + ry) + (T)ty;
            part.vx[i] = vx;
            part.vy[i] = vy;
        }
        std::vector<int> os, oe;
        std::vector<T> og[GF_COUNT];
        snapshot_grid(os, oe, og);
#pragma omp parallel for schedule(static)
        for (int yy = 0; yy < gs; ++yy)
            for (int xx = 0; xx < gs; ++xx)
            {
                const float qx = ((float)xx + 0.5f - c) - tx;
                const float qy = ((float)yy + 0.5f - c) - ty;
                const float sxf = std::floor(c + (cs * qx - sn * qy));
                const float syf = std::floor(c + (sn * qx + cs * qy));
                const bool in = sxf >= 0.0f && sxf < (float)gs && syf >= 0.0f && syf < (float)gs;
                const int i = xx + gs * yy;
                const int j = in ? (int)sxf + gs * (int)syf : 0;
                start[i] = in ? os[j] : 0;
                end[i] = in ? oe[j] : 0;
                for (int f = 0; f < GF_COUNT; ++f)
                    g[f][i] = in ? og[f][j] : T(0);
            }
    }

    void update_pose(float new_x, float new_y, float new_yaw) override
    {
        if (ego_mode == 1)
        {
            update_pose_metric(new_x, new_y, new_yaw);
            return;
        }
        if (!first_pose)
        {
            pos_x = new_x;
            pos_y = new_y;
            yaw = new_yaw;
            first_pose = true;
            return;
        }
        const float x_diff = new_x - pos_x;
        const float y_diff = new_y - pos_y;
        if (std::fabs(x_diff) > p.resolution || std::fabs(y_diff) > p.resolution)
        {
            const int x_move = -static_cast<int>(x_diff / p.resolution);
            const int y_move = -static_cast<int>(y_diff / p.resolution);
#pragma omp parallel for schedule(static)
            for (int i = 0; i < N; ++i)
            {
                part.x[i] -= (T)(float)x_move;
                part.y[i] -= (T)(float)y_move;
            }
            std::vector<int> os, oe;
            std::vector<T> og[GF_COUNT];
            snapshot_grid(os, oe, og);
#pragma omp parallel for schedule(static)
            for (int y = 0; y < gs; ++y)
                for (int x = 0; x < gs; ++x)
                {
                    const int i = x + gs * y;
                    const int nx = x + x_move, ny = y + y_move;
                    const bool in = nx > 0 && nx < gs && ny > 0 && ny < gs;
                    const int j = nx + gs * ny;
                    start[i] = in ? os[j] : 0;
                    end[i] = in ? oe[j] : 0;
                    for (int f = 0; f < GF_COUNT; ++f)
                        g[f][i] = in ? og[f][j] : T(0);
                }
            pos_x = new_x;
            pos_y = new_y;
            yaw = new_yaw;
        }
    }

    // Exact inclusive prefix over per-cell masses, then cell j owns slots [P_{j-1}, P_j).
    // Restates normalize_particle_orders + calc_start_idx/calc_end_idx
    // (init_new_particles.cu:30-43, 66-74) with an exact prefix (see dogm_numerics.h).
    void particle_orders(const std::vector<T>& masses, int v, std::vector<int>& pend)
    {
        pend.resize(C);
        std::vector<fx_t> pref(C);
        fx_prefix(C, [&](int j) { return fx_of(masses[j]); }, pref.data());
        const fx_t s = C > 0 ? pref[C - 1] : (fx_t)0;
        if (s == 0)
        {
            for (int j = 0; j < C; ++j)
                pend[j] = order_end_uniform(j, C, v);
            return;
        }
        const double scale = order_scale(s, v);
#pragma omp parallel for schedule(static)
        for (int j = 0; j < C; ++j)
            pend[j] = order_end(pref[j], scale, v);
        pend[C - 1] = v;
    }

    // ---- dogm.cu:217-241, init_new_particles.cu:76-124 -----------------------------------------
    void initialize_particles() override
    {
        std::vector<T> masses(C);
        for (int j = 0; j < C; ++j)
            masses[j] = (T)meas[j].occ_mass;  // copyMassesKernel
        std::vector<int> pend;
        particle_orders(masses, N, pend);
        born = masses;  // born_masses_array holds the measurement occupancy after this step
        for (int j = 0, k = 0; j < C; ++j)
            for (; k < pend[j]; ++k)
                part.idx[k] = j;  // initParticlesKernel1
        const T new_weight = (T)1 / (T)N;  // 1.0f / particle_count
        const float v = p.init_max_velocity;
#pragma omp parallel for schedule(static)
        for (int k = 0; k < N; ++k)
        {  // initParticlesKernel2
            const int cell = part.idx[k];
            const u32x4 r = draw(seed, instance, TAG_INIT_PARTICLES, epoch, (uint32_t)k);
            part.x[k] = (T)(float)(cell % gs) + (T)0.5f;
            part.y[k] = (T)(float)(cell / gs) + (T)0.5f;
            part.vx[k] = (T)(-v) + ((T)v - (T)(-v)) * (T)uniform01(r.x);
            part.vy[k] = (T)(-v) + ((T)v - (T)(-v)) * (T)uniform01(r.y);
            part.w[k] = new_weight;
            part.assoc[k] = 0;
        }
        ++epoch;
    }

    // ---- dogm.cu:243-260, predict.cu:16-52 ------------------------------------------------------
    void particle_prediction(float dt_f) override
    {
        const T dt = (T)dt_f;
        const T one = 1, zero = 0;
        const T sp = (T)p.stddev_process_noise_position, sv = (T)p.stddev_process_noise_velocity;
        const T ps = (T)p.persistence_prob;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i)
        {
            const u32x4 r = draw(seed, instance, TAG_PREDICT, epoch, (uint32_t)i);
            float n0, n1, n2, n3;
            box_muller(r.x, r.y, &n0, &n1);
            box_muller(r.z, r.w, &n2, &n3);
            const T npx = (T)n0 * sp + zero, npy = (T)n1 * sp + zero;
            const T nvx = (T)n2 * sv + zero, nvy = (T)n3 * sv + zero;
            const T x = part.x[i], y = part.y[i], vx = part.vx[i], vy = part.vy[i];
            // transition_matrix * state with glm's mat4 * vec4 association
            // ((c0 x + c1 y) + (c2 vx + c3 vy)), transposed matrix of dogm.cu:246-253.
            const T sx = (one * x + zero * y) + (dt * vx + zero * vy);
            const T sy = (zero * x + one * y) + (zero * vx + dt * vy);
            const T svx = (zero * x + zero * y) + (one * vx + zero * vy);
            const T svy = (zero * x + zero * y) + (zero * vx + one * vy);
            const T nx = sx + npx, ny = sy + npy;
            part.x[i] = nx;
            part.y[i] = ny;
            part.vx[i] = svx + nvx;
            part.vy[i] = svy + nvy;
            T w = ps * part.w[i];
            if ((nx > (T)(gs - 1) || nx < zero) || (ny > (T)(gs - 1) || ny < zero))
                w = zero;
            part.w[i] = w;
            part.idx[i] = clamp_cell(nx, gs) + gs * clamp_cell(ny, gs);
        }
        ++epoch;
    }

    // ---- dogm.cu:262-281, init.cu:86-93, particle_to_grid.cu:16-44 ------------------------------
    // Stable sort by cell index (thrust::sort_by_key stability is not guaranteed by its API; the
    // restatement defines it as stable), first/last occurrence bounds, weight_array = weight.
    void particle_assignment() override
    {
        std::fill(start.begin(), start.end(), -1);
        std::fill(end.begin(), end.end(), -1);
        std::vector<int> perm;
        stable_sort_by_key(part.idx.data(), N, C - 1, perm);
        ParticleSet<T> s;
        s.resize(N);
#pragma omp parallel for schedule(static)
        for (int k = 0; k < N; ++k)
        {
            const int o = perm[k];
            s.x[k] = part.x[o];
            s.y[k] = part.y[o];
            s.vx[k] = part.vx[o];
            s.vy[k] = part.vy[o];
            s.w[k] = part.w[o];
            s.idx[k] = part.idx[o];
            s.assoc[k] = part.assoc[o];
        }
        part = std::move(s);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i)
        {
            const int c = part.idx[i];
            if (i == 0 || part.idx[i - 1] != c)
                start[c] = i;
            if (i == N - 1 || part.idx[i + 1] != c)
                end[c] = i;
            weight_array[i] = part.w[i];
        }
    }

    T cell_sum(const std::vector<T>& v, int s, int e) const
    {
        fx_t acc = 0;
        for (int k = s; k <= e; ++k)
            acc += fx_of(v[k]);
        return to_T_from_fx<T>(acc);
    }

    // ---- dogm.cu:283-298, mass_update.cu:16-93 --------------------------------------------------
    // m_occ_pred is the exact sum of the cell's weights (the reference takes an f32 prefix
    // difference, common.h:24-32).
    void grid_cell_occupancy_update(float dt) override
    {
        const T alpha = freespace_factor(dt);
        const T one = 1, zero = 0;
        const T pB = (T)p.birth_prob;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < C; ++i)
        {
            const int s = start[i], e = end[i];
            T m_occ_pred = zero;
            if (s != -1)
                m_occ_pred = cell_sum(weight_array, s, e);
            if (m_occ_pred > one)
            {  // normalize_weights (mass_update.cu:51-59)
                for (int k = s; k < e + 1; ++k)
                {
                    weight_array[k] = weight_array[k] / m_occ_pred;
                    part.w[k] = weight_array[k];
                }
                m_occ_pred = one;
            }
            const T m_free_pred = tmin(alpha * g[GF_FREE][i], one - m_occ_pred);
            const T mfree = (T)meas[i].free_mass, mocc = (T)meas[i].occ_mass;
            const T unknown_pred = one - m_occ_pred - m_free_pred;
            const T meas_unknown = one - mfree - mocc;
            const T K = m_free_pred * mocc + m_occ_pred * mfree;
            const T occ_up = (m_occ_pred * meas_unknown + unknown_pred * mocc + m_occ_pred * mocc) / (one - K);
            const T free_up = (m_free_pred * meas_unknown + unknown_pred * mfree + m_free_pred * mfree) / (one - K);
            const T rho_b = (occ_up * pB * (one - m_occ_pred)) / (m_occ_pred + pB * (one - m_occ_pred));
            const T rho_p = occ_up - rho_b;
            born[i] = rho_b;
            g[GF_PERS][i] = rho_p;
            g[GF_NEW_BORN][i] = rho_b;
            g[GF_FREE][i] = free_up;
            g[GF_OCC][i] = occ_up;
            g[GF_PRED][i] = m_occ_pred;
        }
    }

    T freespace_factor(float dt) const;

    // ---- dogm.cu:300-321, update_persistent_particles.cu:16-87 ----------------------------------
    // mu_A / mu_UA are written for every cell (0 for cells without particles; the reference leaves
    // those cells' previous values in place).
    void update_persistent_particles() override
    {
        const T zero = 0, one = 1;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i)
            weight_array[i] = (T)meas[part.idx[i]].likelihood * part.w[i];
#pragma omp parallel for schedule(static)
        for (int c = 0; c < C; ++c)
        {
            T mu_A = zero, mu_UA = zero;
            if (start[c] != -1)
            {
                const T acc = cell_sum(weight_array, start[c], end[c]);
                const T rho_p = g[GF_PERS][c];
                mu_A = acc > zero ? rho_p / acc : zero;
                const T pred = g[GF_PRED][c];
                mu_UA = pred > zero ? g[GF_PERS][c] / pred : zero;
            }
            g[GF_MU_A][c] = mu_A;
            g[GF_MU_UA][c] = mu_UA;
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i)
        {
            const int c = part.idx[i];
            const T pA = (T)meas[c].p_A;
            weight_array[i] = pA * g[GF_MU_A][c] * weight_array[i] + (one - pA) * g[GF_MU_UA][c] * part.w[i];
        }
    }

    // ---- dogm.cu:323-348, init.cu:46-67, init_new_particles.cu:18-195 ---------------------------
    // Birth counts from the exact prefix (Sum n_j = B exactly).  Slot association: the first
    // min(nu_A + 1, n_j) slots of a cell are associated; the reference's loop
    // [start, start + nu_A] (init_new_particles.cu:145-148) writes one slot more than nu_A and
    // spills into the next cell's first slot; the owner of a slot wins here.
    void initialize_new_particles() override
    {
        const T zero = 0, one = 1;
        std::vector<int> pend;
        particle_orders(born, B, pend);
#pragma omp parallel for schedule(static)
        for (int j = 0; j < C; ++j)
        {
            const int s = j == 0 ? 0 : pend[j - 1];
            const int n = pend[j] - s;
            const float pA_f = meas[j].p_A;
            const T pA = (T)pA_f;
            const int nu_A = num_assoc(n, pA_f);
            const int nu_UA = n - nu_A;
            g[GF_W_A][j] = nu_A > 0 ? (pA * born[j]) / (T)nu_A : zero;
            g[GF_W_UA][j] = nu_UA > 0 ? ((one - pA) * born[j]) / (T)nu_UA : zero;
            for (int k = s; k < pend[j]; ++k)
            {
                birth.idx[k] = j;
                birth.assoc[k] = (k - s) <= nu_A ? 1 : 0;
            }
        }
        const T sd = (T)p.stddev_velocity;
#pragma omp parallel for schedule(static)
        for (int k = 0; k < B; ++k)
        {
            const int cell = birth.idx[k];
            const u32x4 r = draw(seed, instance, TAG_BIRTH, epoch, (uint32_t)k);
            float n0, n1;
            box_muller(r.x, r.y, &n0, &n1);
            birth.x[k] = (T)(float)(cell % gs) + (T)0.5f;
            birth.y[k] = (T)(float)cell / (T)(float)gs + (T)0.5f;  // float division (init_new_particles.cu:173)
            birth.vx[k] = (T)n0 * sd + zero;
            birth.vy[k] = (T)n1 * sd + zero;
            birth.w[k] = birth.assoc[k] ? g[GF_W_A][cell] : g[GF_W_UA][cell];
        }
        ++epoch;
    }

    int num_assoc(int n, float pA) const;

    // ---- dogm.cu:350-384, statistical_moments.cu:16-106 ----------------------------------------
    void statistical_moments() override
    {
        const T zero = 0, one = 1;
        std::vector<T> m1(N), m2(N), m3(N), m4(N), m5(N);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i)
        {
            const T w = weight_array[i], vx = part.vx[i], vy = part.vy[i];
            m1[i] = w * vx;
            m2[i] = w * vy;
            m3[i] = w * vx * vx;
            m4[i] = w * vy * vy;
            m5[i] = w * vx * vy;
        }
        // f32 path: the device's f64 association (moment_sums_wave); FP64 twin: exact sums.
        std::vector<double> wsum;
        if constexpr (std::is_same<T, float>::value)
        {
            wsum.resize((size_t)N * MOM_CH);
            const float* ch[MOM_CH] = {m1.data(), m2.data(), m3.data(), m4.data(), m5.data()};
            moment_sums_wave(part.idx.data(), N, ch, wsum.data());
        }
        auto sum = [&](int ch, const std::vector<T>& m, int s, int e) -> T {
            if constexpr (std::is_same<T, float>::value)
                return (float)wsum[(size_t)e * MOM_CH + ch];
            else
                return cell_sum(m, s, e);
        };
#pragma omp parallel for schedule(static)
        for (int c = 0; c < C; ++c)
        {
            T mx = zero, my = zero, vx = zero, vy = zero, cv = zero;
            const T rho_p = g[GF_PERS][c];
            if (start[c] != -1 && rho_p > zero)
            {
                const int s = start[c], e = end[c];
                const T inv = one / rho_p;
                mx = inv * sum(0, m1, s, e);
                my = inv * sum(1, m2, s, e);
                vx = inv * sum(2, m3, s, e) - mx * mx;
                vy = inv * sum(3, m4, s, e) - my * my;
                cv = inv * sum(4, m5, s, e) - mx * my;
            }
            g[GF_MEAN_X][c] = mx;
            g[GF_MEAN_Y][c] = my;
            g[GF_VAR_X][c] = vx;
            g[GF_VAR_Y][c] = vy;
            g[GF_COVAR][c] = cv;
        }
    }

    // ---- dogm.cu:386-423, resampling.cu:17-68 ---------------------------------------------------
    // Multinomial resampling: exact CDF over [posterior weights | birth weights] (negative weights
    // count as 0); N sorted uniforms on [0, total) from exponential spacings (dogm_numerics.h);
    // ancestor = first CDF entry >= draw (lower bound, resampling.cu:44-46); the next particle
    // copies the ancestor with weight total / N.
    static T pos_weight(T w) { return w > T(0) ? w : T(0); }

    void resampling() override
    {
        if (cdf_mode != CDF_EXACT && sizeof(T) == sizeof(float))
        {
            resampling_float();
            return;
        }
        const int J = N + B;
        std::vector<fx_t> cdf(J);
        fx_prefix(
            J, [&](int k) { return fx_of(pos_weight(k < N ? weight_array[k] : birth.w[k - N])); }, cdf.data());
        const fx_t total = J > 0 ? cdf[J - 1] : (fx_t)0;
        std::vector<fx_t> spac;
        draw_prefixes(N, seed, instance, spac);
        const Spacing sp = spacing_setup(spac[N], total);
        // Lower bound of each draw in the CDF.  Draws are non-decreasing in k, so each thread takes
        // the lower bound of its first draw and walks forward (the same indices as a binary search
        // per draw).  Draws are < total (or 0 with total 0), so the index is always < J.
#pragma omp parallel
        {
            const int t = omp_get_thread_num(), nt = omp_get_num_threads();
            long long b, e;
            chunk(N, t, nt, &b, &e);
            int lo = 0;
            for (long long k = b; k < e; ++k)
            {
                const fx_t rv = spacing_draw(spac[k], sp);
                if (k == b)
                    lo = (int)(std::lower_bound(cdf.begin(), cdf.end(), rv) - cdf.begin());
                else
                    while (lo < J && cdf[lo] < rv)
                        ++lo;
                ancestors[k] = lo < J ? lo : J - 1;
            }
        }
        const T joint_max = to_T_from_fx<T>(total);
        const T new_weight = joint_max / (T)N;
        ParticleSet<T> next;
        next.resize(N);
#pragma omp parallel for schedule(static)
        for (int k = 0; k < N; ++k)
        {
            const int a = ancestors[k];
            const ParticleSet<T>& src = a < N ? part : birth;
            const int o = a < N ? a : a - N;
            next.x[k] = src.x[o];
            next.y[k] = src.y[o];
            next.vx[k] = src.vx[o];
            next.vy[k] = src.vy[o];
            next.idx[k] = src.idx[o];
            next.assoc[k] = src.assoc[o];
            next.w[k] = new_weight;
        }
        part = std::move(next);
        ++epoch;
    }

    // Float CDF modes (precision study): same draws u_k, CDF from float_cdf, ancestor = first j
    // with E[j] >= r_k, r_k = total * u_k in the mode's precision (dogm_numerics.h, cdf_float_draw).
    void resampling_float()
    {
        const int J = N + B;
        std::vector<float> wj(J);
        for (int k = 0; k < N; ++k)
            wj[k] = cdf_weight((float)weight_array[k]);
        for (int k = 0; k < B; ++k)
            wj[N + k] = cdf_weight((float)birth.w[k]);
        std::vector<double> E;
        if (cdf_mode == CDF_FP64)
            float_cdf<double>(wj, E);
        else if (cdf_mode == CDF_FP32C)
            float_cdf<f32c>(wj, E);
        else
            float_cdf<float>(wj, E);
        const double total = J > 0 ? E[J - 1] : 0.0;
        std::vector<fx_t> spac;
        draw_prefixes(N, seed, instance, spac);
        const Spacing sp = spacing_setup(spac[N], 0);
        for (int k = 0; k < N; ++k)
        {
            const double rv = cdf_float_draw(total, spacing_fraction(spac[k], sp), cdf_mode);
            const int lo = (int)(std::lower_bound(E.begin(), E.end(), rv) - E.begin());
            ancestors[k] = lo < J ? lo : J - 1;
        }
        const float joint_max = (float)total;
        const float new_weight = joint_max / (float)N;
        ParticleSet<T> next;
        next.resize(N);
#pragma omp parallel for schedule(static)
        for (int k = 0; k < N; ++k)
        {
            const int a = ancestors[k];
            const ParticleSet<T>& src = a < N ? part : birth;
            const int o = a < N ? a : a - N;
            next.x[k] = src.x[o];
            next.y[k] = src.y[o];
            next.vx[k] = src.vx[o];
            next.vy[k] = src.vy[o];
            next.idx[k] = src.idx[o];
            next.assoc[k] = src.assoc[o];
            next.w[k] = (T)new_weight;
        }
        part = std::move(next);
        ++epoch;
    }

    // ---- field transport -----------------------------------------------------------------------
    std::vector<T>* fvec(int field)
    {
        switch (field)
        {
        case F_P_X: return &part.x;
        case F_P_Y: return &part.y;
        case F_P_VX: return &part.vx;
        case F_P_VY: return &part.vy;
        case F_P_W: return &part.w;
        case F_B_X: return &birth.x;
        case F_B_Y: return &birth.y;
        case F_B_VX: return &birth.vx;
        case F_B_VY: return &birth.vy;
        case F_B_W: return &birth.w;
        case F_WEIGHT_ARRAY: return &weight_array;
        case F_BORN: return &born;
        default:
            if (field >= F_GRID0 && field < F_GRID0 + GF_COUNT)
                return &g[field - F_GRID0];
            return nullptr;
        }
    }
    int get(int field, double* out) override
    {
        auto* v = fvec(field);
        if (!v)
            return -1;
        for (size_t i = 0; i < v->size(); ++i)
            out[i] = (double)(*v)[i];
        return (int)v->size();
    }
    int set(int field, const double* in) override
    {
        auto* v = fvec(field);
        if (!v)
            return -1;
        for (size_t i = 0; i < v->size(); ++i)
            (*v)[i] = (T)in[i];
        return (int)v->size();
    }
    int get_i32(int field, int* out) override
    {
        switch (field)
        {
        case I_P_IDX: std::copy(part.idx.begin(), part.idx.end(), out); return N;
        case I_P_ASSOC: std::copy(part.assoc.begin(), part.assoc.end(), out); return N;
        case I_B_IDX: std::copy(birth.idx.begin(), birth.idx.end(), out); return B;
        case I_B_ASSOC: std::copy(birth.assoc.begin(), birth.assoc.end(), out); return B;
        case I_G_START: std::copy(start.begin(), start.end(), out); return C;
        case I_G_END: std::copy(end.begin(), end.end(), out); return C;
        case I_ANCESTORS: std::copy(ancestors.begin(), ancestors.end(), out); return N;
        default: return -1;
        }
    }
    int set_i32(int field, const int* in) override
    {
        switch (field)
        {
        case I_P_IDX: std::copy(in, in + N, part.idx.begin()); return N;
        case I_P_ASSOC:
            for (int i = 0; i < N; ++i)
                part.assoc[i] = in[i] ? 1 : 0;
            return N;
        case I_B_IDX: std::copy(in, in + B, birth.idx.begin()); return B;
        case I_B_ASSOC:
            for (int i = 0; i < B; ++i)
                birth.assoc[i] = in[i] ? 1 : 0;
            return B;
        case I_G_START: std::copy(in, in + C, start.begin()); return C;
        case I_G_END: std::copy(in, in + C, end.begin()); return C;
        default: return -1;
        }
    }
};

// alpha = freespace_discount^dt: std::pow(float, float) -> powf on the host (dogm.cu:291).
template <>
float Oracle<float>::freespace_factor(float dt) const
{
    return std::pow(p.freespace_discount, dt);
}
template <>
double Oracle<double>::freespace_factor(float dt) const
{
    return std::pow((double)p.freespace_discount, (double)dt);
}

// nu_A = int(roundf(n * p_A)) (init_new_particles.cu:45-48).
template <>
int Oracle<float>::num_assoc(int n, float pA) const
{
    return (int)roundf((float)n * pA);
}
template <>
int Oracle<double>::num_assoc(int n, float pA) const
{
    return (int)std::round((double)n * (double)pA);
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// C ABI used by the tests (ctypes) and the host-core timing leg of bench.py.
// ------------------------------------------------------------------------------------------------
// ---------------------------------------------------------------------------------------------
// Result extraction (dogm/demo/utils/image_creation.cpp), over structure-of-arrays cell fields.
// ---------------------------------------------------------------------------------------------

// image_creation.cpp:19-22.
static float pignistic_f32(float free_mass, float occ_mass)
{
    const float unknown = 1.0f - occ_mass - free_mass;
    return occ_mass + 0.5f * unknown;
}

// OpenCV cv::Mat::inv() (cv::invert, DECOMP_LU) of a 2x2 CV_32F matrix: the closed form of
// modules/core/src/lapack.cpp for n == 2 -- determinant in double, d = 1 / det, each element
// stored as float(element * d) at its adjugate position; a zero matrix when det == 0.
static void cv_invert_2x2_f32(const float m[2][2], float r[2][2])
{
    double det = (double)m[0][0] * m[1][1] - (double)m[0][1] * m[1][0];
    r[0][0] = r[0][1] = r[1][0] = r[1][1] = 0.0f;
    if (det != 0.0)
    {
        const double d = 1.0 / det;
        r[1][1] = (float)(m[0][0] * d);
        r[0][0] = (float)(m[1][1] * d);
        r[0][1] = (float)(-m[0][1] * d);
        r[1][0] = (float)(-m[1][0] * d);
    }
}

// cv::gemm of CV_32F operands: every output element is a dot product accumulated in double from
// zero and stored as float.  A is rows x inner, B is inner x cols, all row-major.
static void cv_gemm_f32(const float* a, const float* b, float* c, int rows, int inner, int cols)
{
    for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j)
        {
            double acc = 0.0;
            for (int k = 0; k < inner; ++k)
                acc += (double)a[i * inner + k] * (double)b[k * cols + j];
            c[i * cols + j] = (float)acc;
        }
}

// velocity_mean.t() * velocity_covar.inv() * velocity_mean (image_creation.cpp:35-50), evaluated
// left to right with a float temporary between the two products.
static float velocity_normalized_by_variance(float mx, float my, float vxx, float vyy, float cxy)
{
    const float mean[2] = {mx, my};
    const float cov[2][2] = {{vxx, cxy}, {cxy, vyy}};
    float inv[2][2];
    cv_invert_2x2_f32(cov, inv);
    float row[2], out[1];
    cv_gemm_f32(mean, &inv[0][0], row, 1, 2, 2);
    cv_gemm_f32(row, mean, out, 1, 2, 1);
    return out[0];
}

extern "C" {

void* oracle_create(const dogm_params* params, uint64_t seed, uint32_t instance, int fp64)
{
    if (!params)
        return nullptr;
    if (fp64)
        return static_cast<OracleBase*>(new Oracle<double>(*params, seed, instance));
    return static_cast<OracleBase*>(new Oracle<float>(*params, seed, instance));
}

void oracle_destroy(void* h)
{
    delete static_cast<OracleBase*>(h);
}

#define O static_cast<OracleBase*>(h)
int oracle_update_measurement_grid(void* h, const dogm_measurement_cell* m) { return O->update_measurement_grid(m); }
void oracle_update_pose(void* h, float x, float y, float yaw) { O->update_pose(x, y, yaw); }
void oracle_initialize_particles(void* h) { O->initialize_particles(); }
void oracle_particle_prediction(void* h, float dt) { O->particle_prediction(dt); }
void oracle_particle_assignment(void* h) { O->particle_assignment(); }
void oracle_grid_cell_occupancy_update(void* h, float dt) { O->grid_cell_occupancy_update(dt); }
void oracle_update_persistent_particles(void* h) { O->update_persistent_particles(); }
void oracle_initialize_new_particles(void* h) { O->initialize_new_particles(); }
void oracle_statistical_moments(void* h) { O->statistical_moments(); }
void oracle_resampling(void* h) { O->resampling(); }

// dogm.cu:115-131 stage order.
int oracle_update_grid(void* h, const dogm_measurement_cell* m, float x, float y, float yaw, float dt)
{
    O->update_measurement_grid(m);
    O->update_pose(x, y, yaw);
    O->particle_prediction(dt);
    O->particle_assignment();
    O->grid_cell_occupancy_update(dt);
    O->update_persistent_particles();
    O->initialize_new_particles();
    O->statistical_moments();
    O->resampling();
    return 0;
}

int oracle_get(void* h, int field, double* out) { return O->get(field, out); }
int oracle_set(void* h, int field, const double* in) { return O->set(field, in); }
int oracle_get_i32(void* h, int field, int* out) { return O->get_i32(field, out); }
int oracle_set_i32(void* h, int field, const int* in) { return O->set_i32(field, in); }
uint32_t oracle_get_epoch(void* h) { return O->epoch; }
int oracle_threads(void) { return omp_get_max_threads(); }
void oracle_set_epoch(void* h, uint32_t e) { O->epoch = e; }
float oracle_position_x(void* h) { return O->pos_x; }
float oracle_position_y(void* h) { return O->pos_y; }
float oracle_yaw(void* h) { return O->yaw; }
void oracle_set_cdf_mode(void* h, int mode) { O->cdf_mode = mode; }
void oracle_set_draw_mode(void* h, int mode) { O->draw_mode = mode; }
void oracle_set_ego_mode(void* h, int mode)
{
    if (mode == 1 && O->ego_mode != 1 && O->first_pose)
    {
        O->ref_x = O->pos_x;
        O->ref_y = O->pos_y;
        O->ref_yaw = O->yaw;
    }
    O->ego_mode = mode;
}
void oracle_set_flags(void* h, int first_pose, int first_meas)
{
    O->first_pose = first_pose != 0;
    O->first_meas = first_meas != 0;
}
#undef O

// The N draw fractions u_k of one resampling (seed, instance, epoch): exponential spacings with
// the exact prefix, as every CDF mode uses them.
void oracle_spacing_fractions(uint64_t seed, uint32_t instance, uint32_t epoch, int n, double* out)
{
    std::vector<fx_t> spac((size_t)n + 1);
    fx_t e = 0;
    float e4[4];
    for (int k = 0; k <= n; ++k)
    {
        if ((k & 3) == 0)
            resample_exponentials4(seed, instance, epoch, (uint32_t)(k >> 2), e4);
        e += fx_from_f32(e4[k & 3]);
        spac[k] = e;
    }
    const Spacing sp = spacing_setup(spac[n], 0);
    for (int k = 0; k < n; ++k)
        out[k] = spacing_fraction(spac[k], sp);
}

// Numerics entry points for the known-answer tests of dogm_numerics.h.
void oracle_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t* out)
{
    const u32x4 r = philox4x32_10(u32x4{c0, c1, c2, c3}, k0, k1);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}
float oracle_log_f32(float x) { return log_f32(x); }
void oracle_sincos_2pi(float t, float* s, float* c) { sincos_2pi_f32(t, s, c); }
void oracle_box_muller(uint32_t a, uint32_t b, float* z0, float* z1) { box_muller(a, b, z0, z1); }
float oracle_uniform01(uint32_t x) { return uniform01(x); }
float oracle_uniform01_oc(uint32_t x) { return uniform01_oc(x); }
// Fixed point: raw values travel as (hi, lo) 64-bit halves.
void oracle_fx_from_f32(float v, int64_t* hi, uint64_t* lo)
{
    const fx_t r = fx_from_f32(v);
    *hi = (int64_t)(r >> 64);
    *lo = (uint64_t)r;
}
// All n floats of bit patterns bits[0..n) through fx_from_f32_nonneg (the kernels' form for
// weights with a clear sign bit); returns how many differ from fx_from_f32.
int oracle_fx_from_f32_nonneg_mismatches(const uint32_t* bits, int n)
{
    int bad = 0;
    for (int i = 0; i < n; ++i)
    {
        float v;
        std::memcpy(&v, &bits[i], sizeof v);
        bad += fx_from_f32_nonneg(v) != fx_from_f32(v);
    }
    return bad;
}
// The same for doubles and fx_from_f64_nonneg.
int oracle_fx_from_f64_nonneg_mismatches(const uint64_t* bits, int n)
{
    int bad = 0;
    for (int i = 0; i < n; ++i)
    {
        double v;
        std::memcpy(&v, &bits[i], sizeof v);
        bad += fx_from_f64_nonneg(v) != fx_from_f64(v);
    }
    return bad;
}
// Every exponential -log(u) of the resampling draws, u = uniform01_oc(x) over all 2^24 values it
// takes: returns how many are below zero or convert differently through fx_from_f32_nonneg.
int oracle_exponential_sign_violations(void)
{
    int bad = 0;
    for (uint32_t q = 0; q < (1u << 24); ++q)
    {
        const float e = -log_f32(uniform01_oc(q << 8));
        bad += e < 0.0f || fx_from_f32_nonneg(e) != fx_from_f32(e);
    }
    return bad;
}
void oracle_fx_from_f64(double v, int64_t* hi, uint64_t* lo)
{
    const fx_t r = fx_from_f64(v);
    *hi = (int64_t)(r >> 64);
    *lo = (uint64_t)r;
}
float oracle_fx_to_f32(int64_t hi, uint64_t lo) { return fx_to_f32((fx_t)(((ufx_t)(uint64_t)hi << 64) | lo)); }
double oracle_fx_to_f64(int64_t hi, uint64_t lo) { return fx_to_f64((fx_t)(((ufx_t)(uint64_t)hi << 64) | lo)); }
void oracle_fx_scale_u64(int64_t hi, uint64_t lo, uint64_t u, int64_t* ohi, uint64_t* olo)
{
    const fx_t r = fx_scale_u64((fx_t)(((ufx_t)(uint64_t)hi << 64) | lo), u);
    *ohi = (int64_t)(r >> 64);
    *olo = (uint64_t)r;
}
int oracle_order_end(int64_t hi, uint64_t lo, int64_t thi, uint64_t tlo, int v)
{
    const fx_t pref = (fx_t)(((ufx_t)(uint64_t)hi << 64) | lo);
    const fx_t tot = (fx_t)(((ufx_t)(uint64_t)thi << 64) | tlo);
    return order_end(pref, order_scale(tot, v), v);
}

// Pignistic occupancy and the 8-bit image (image_creation.cpp:19-22, 123-125; floor clamped to
// [0, 255]) of n cells.  Either output may be null.
void oracle_occupancy(const float* occ, const float* free_mass, int n, float* p, uint8_t* img)
{
    for (int c = 0; c < n; ++c)
    {
        const float q = pignistic_f32(free_mass[c], occ[c]);
        if (p)
            p[c] = q;
        if (img)
        {
            float g = std::floor(q * 255.0f);
            if (!(g >= 0.0f))
                g = 0.0f;
            if (g > 255.0f)
                g = 255.0f;
            img[c] = (uint8_t)(255 - (int)g);
        }
    }
}

// computeCellsWithVelocity (image_creation.cpp:24-66): indices of the passing cells in the demo's
// y-then-x loop order; returns their number.  out_idx holds up to n entries.
int oracle_cells_with_velocity(int gs, const float* occ, const float* free_mass, const float* mean_x,
                               const float* mean_y, const float* var_x, const float* var_y, const float* covar,
                               float min_occ, float min_vel, int* out_idx)
{
    int k = 0;
    for (int y = 0; y < gs; y++)
        for (int x = 0; x < gs; x++)
        {
            const int index = y * gs + x;
            const float o = pignistic_f32(free_mass[index], occ[index]);
            const float m = velocity_normalized_by_variance(mean_x[index], mean_y[index], var_x[index], var_y[index],
                                                           covar[index]);
            if (o >= min_occ && m >= min_vel)
                out_idx[k++] = index;
        }
    return k;
}

float oracle_velocity_normalized(float mx, float my, float vxx, float vyy, float cxy)
{
    return velocity_normalized_by_variance(mx, my, vxx, vyy, cxy);
}

// Measurement grid of one lidar scan: the two-level ray cast that dogm_lidar_generate runs on the
// device (the step before updateGrid; the reference's LaserMeasurementGrid::generateGrid,
// dogm/demo/simulator/mapping/laser_to_meas_grid.cu:25-70).  Sensor at the grid centre; samples every
// resolution / 2 metres while the distance is below min(range, max_range) mark free_mass; the cell of
// a return within max_range gets occ_mass and free 0; likelihood = p_A = 1 (measurement_grid.cu:129-130).
// Velocity-moment sums of the f32 path in the device's association (moment_sums_wave): for keys[n]
// sorted, channel values x[c * n + k], out[c * n + k] = the sum of channel c over the cell that ends at
// particle k (other entries untouched).
void oracle_moment_sums(const int* keys, int n, const float* x, double* out)
{
    const float* ch[MOM_CH];
    for (int c = 0; c < MOM_CH; ++c)
        ch[c] = x + (size_t)c * n;
    std::vector<double> tmp((size_t)n * MOM_CH, 0.0);
    moment_sums_wave(keys, n, ch, tmp.data());
    for (int k = 0; k < n; ++k)
        for (int c = 0; c < MOM_CH; ++c)
            if (k == n - 1 || keys[k + 1] != keys[k])
                out[(size_t)c * n + k] = tmp[(size_t)k * MOM_CH + c];
}

void oracle_lidar_grid(int gs, int beams, double res, double max_range, float occ_mass, float free_mass,
                       const double* cosv, const double* sinv, const double* ranges, dogm_measurement_cell* out)
{
    const size_t cells = (size_t)gs * (size_t)gs;
    std::vector<unsigned char> is_free(cells, 0), is_occ(cells, 0);
    const double centre = 0.5 * gs * res;
    const double step = 0.5 * res;
    const int steps = (int)std::ceil(max_range / step) + 1;
    auto cell_of = [&](double x, double y, size_t* c) {
        const double fx = std::floor(x / res), fy = std::floor(y / res);
        if (!(fx >= 0.0 && fx < (double)gs && fy >= 0.0 && fy < (double)gs))
            return false;
        *c = (size_t)fx + (size_t)gs * (size_t)fy;
        return true;
    };
    for (int b = 0; b < beams; ++b)
    {
        const double r = ranges[b];
        const double lim = std::isnan(r) ? r : std::min(r, max_range);
        for (int k = 0; k < steps; ++k)
        {
            const double dist = k * step;
            if (!(dist < lim))
                break;
            size_t c;
            if (cell_of(centre + cosv[b] * dist, centre + sinv[b] * dist, &c))
                is_free[c] = 1;
        }
        size_t c;
        if (r <= max_range && cell_of(centre + cosv[b] * r, centre + sinv[b] * r, &c))
            is_occ[c] = 1;
    }
    for (size_t c = 0; c < cells; ++c)
    {
        out[c].free_mass = (is_free[c] && !is_occ[c]) ? free_mass : 0.0f;
        out[c].occ_mass = is_occ[c] ? occ_mass : 0.0f;
        out[c].likelihood = 1.0f;
        out[c].p_A = 1.0f;
    }
}

// Compensated f32 primitives (dogm_numerics.h) for their known-answer tests.
void oracle_two_sum_f32(float a, float b, float* s, float* e) { two_sum_f32(a, b, s, e); }
void oracle_f32c_add(float ahi, float alo, float bhi, float blo, float* hi, float* lo)
{
    f32c a, b;
    a.hi = ahi;
    a.lo = alo;
    b.hi = bhi;
    b.lo = blo;
    const f32c r = a + b;
    *hi = r.hi;
    *lo = r.lo;
}

}  // extern "C"
