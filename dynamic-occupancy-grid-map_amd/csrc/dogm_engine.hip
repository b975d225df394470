// Step engine: per-vehicle allocation, step descriptors and work tables, shared scratch, the stage
// sequence of dogm::DOGM::updateGrid (dogm/src/dogm.cu:115-131) over any list of vehicles, and
// the optional per-kernel event timing.
#include "dogm_internal.h"

#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace dogm_hip
{

const char* const kKernelNames[K_COUNT] = {
    "predict",        "move_particles",  "shift_field",     "sort_hist",       "sort_upsweep",
    "sort_scan",      "sort_downsweep",  "sort_gather",     "reset_bounds",    "seg_occ",
    "seg_occ_fix",    "mass_update",     "normalize",       "seg_pers",        "seg_pers_fix",
    "post_weights",   "empty_cells",     "scan_tiles",      "scan_top",        "birth_cells",
    "birth_particles", "init_cells",     "init_particles",  "copy_masses",     "seg_moments",
    "seg_moments_fix", "cdf_write",      "exp_tiles",       "rs_merge",        "fill",            "pack_grid",
    "unpack_grid",    "build_tables",    "extract_mark",    "extract_scan",    "extract_compact",
    "extract_gather", "ego_particles",   "ego_field",       "pack_state",      "unpack_state",
    "materialize",
};

// ---- timing ------------------------------------------------------------------------------------

static hipEvent_t take_event(Engine* E)
{
    if (!E->event_pool.empty())
    {
        hipEvent_t e = E->event_pool.back();
        E->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

void launch_begin(Engine* E, int kid)
{
    if (!E->timing_on[kid])
        return;
    TimedLaunch t{kid, take_event(E), take_event(E)};
    if (t.a)
        (void)hipEventRecord(t.a, E->stream);
    E->timed_pending.push_back(t);
}

void launch_end(Engine* E, int kid)
{
    if (!E->timing_on[kid] || E->timed_pending.empty())
        return;
    TimedLaunch& t = E->timed_pending.back();
    if (t.b)
        (void)hipEventRecord(t.b, E->stream);
}

void engine_collect_timing(Engine* E)
{
    for (auto& t : E->timed_pending)
    {
        float ms = 0.0f;
        if (t.a && t.b && hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess)
        {
            E->timing_ms[t.kid] += ms;
            E->timing_n[t.kid] += 1;
        }
        if (t.a)
            E->event_pool.push_back(t.a);
        if (t.b)
            E->event_pool.push_back(t.b);
    }
    E->timed_pending.clear();
}

// ---- allocation helpers ------------------------------------------------------------------------

long long units_for(long long len)
{
    const long long u = (len + UNIT - 1) / UNIT;
    return (u + UNIT_ALIGN - 1) / UNIT_ALIGN * UNIT_ALIGN;
}

namespace
{

int set_err(std::string* why, int code, const std::string& msg)
{
    if (why)
        *why = msg;
    return code;
}

template <typename T>
int dalloc(T** p, size_t count, const char* what, std::string* why, size_t* bytes = nullptr)
{
    *p = nullptr;
    if (count == 0)
        return DOGM_OK;
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
    if (e != hipSuccess)
        return set_err(why, e == hipErrorOutOfMemory ? DOGM_ERR_OUT_OF_MEMORY : DOGM_ERR_HIP,
                       std::string("allocating ") + what + ": " + hipGetErrorString(e));
    if (bytes)
        *bytes += count * sizeof(T);
    return DOGM_OK;
}

// Grows a device buffer to hold `need` elements (contents are not kept).
template <typename T>
int grow(T** p, long long* cap, long long need, const char* what, std::string* why)
{
    if (need <= *cap)
        return DOGM_OK;
    (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    int st = dalloc(p, (size_t)need, what, why);
    if (st == DOGM_OK)
        *cap = need;
    return st;
}

int bits_for(long long x)  // number of bits needed to represent x >= 0
{
    int b = 0;
    while (b < 62 && (1ll << b) <= x)
        ++b;
    return b;
}

long long space_len(const Vehicle* v, int s)
{
    switch (s)
    {
    case SP_P: return v->N;
    case SP_B: return v->B;
    case SP_C: return v->C;
    case SP_J: return (long long)v->N + v->B;
    default: return (long long)v->N + 1;
    }
}

void free_particles(DevParticles& p)
{
    (void)hipFree(p.pos);  // pos and vel share this allocation
    (void)hipFree(p.idx);
    (void)hipFree(p.weight);
    (void)hipFree(p.assoc);
    p = DevParticles{};
}

int alloc_particles(DevParticles& p, int n, std::string* why, size_t* bytes)
{
    int st;
    // vel starts on a 16-byte boundary (an odd count pads pos by one entry) so the moments can load
    // two velocities per 16-byte access.
    const size_t vel_off = (size_t)n + (size_t)(n & 1);
    if ((st = dalloc(&p.pos, vel_off + (size_t)n, "particle state", why, bytes)) ||
        (st = dalloc(&p.idx, n, "particle index", why, bytes)) ||
        (st = dalloc(&p.weight, n, "particle weight", why, bytes)) ||
        (st = dalloc(&p.assoc, n, "particle flags", why, bytes)))
        return st;
    p.vel = p.pos ? p.pos + vel_off : nullptr;
    if (n > 0 && (hipMemset(p.pos, 0, sizeof(float2) * (vel_off + n)) != hipSuccess ||
                  hipMemset(p.idx, 0, sizeof(int) * n) != hipSuccess ||
                  hipMemset(p.weight, 0, sizeof(float) * n) != hipSuccess || hipMemset(p.assoc, 0, n) != hipSuccess))
        return set_err(why, DOGM_ERR_HIP, "clearing particle buffers");
    return DOGM_OK;
}

uint64_t next_uid()
{
    static uint64_t uid = 0;
    return ++uid;
}

}  // namespace

// ---- engine ------------------------------------------------------------------------------------

namespace
{

// A roctx range around the host-side enqueue of one stage (rocprofv3 --marker-trace shows them
// on the host timeline beside the stage's kernels); a no-op when no tool listens.
struct StageRange
{
    explicit StageRange(const char* name) { roctxRangePush(name); }
    ~StageRange() { roctxRangePop(); }
    StageRange(const StageRange&) = delete;
    StageRange& operator=(const StageRange&) = delete;
};

// Element counts of the plan-sized step buffers of a step over `veh`.
struct PlanSizes
{
    long long tot[SP_COUNT];  // table units per space
    long long tab;            // work-table entries (all spaces)
    long long scan_tiles;     // 1024-element tiles of the longest scanned space
    long long seg_tiles;      // segmented-reduction carries: one per 256-particle unit (the moments'
                              // chunks; the other reductions take 512 or more particles per tile)
    long long sort_tiles;     // 2048-particle sort tiles
    int passes;               // radix passes for the largest grid
};

int plan_sizes(Vehicle* const* veh, int m, PlanSizes* ps, std::string* why)
{
    *ps = PlanSizes{};
    ps->passes = 1;
    for (int s = 0; s < SP_COUNT; ++s)
        for (int i = 0; i < m; ++i)
        {
            ps->tot[s] += units_for(space_len(veh[i], s));
            if (ps->tot[s] > (long long)INT32_MAX)
                return set_err(why, DOGM_ERR_INVALID_ARGUMENT, "step too large for the work tables");
        }
    for (int i = 0; i < m; ++i)
        ps->passes = std::max(ps->passes, std::max(1, (bits_for(veh[i]->C - 1) + 7) / 8));
    for (int s = 0; s < SP_COUNT; ++s)
        ps->tab += ps->tot[s];
    ps->scan_tiles = std::max({ps->tot[SP_C], ps->tot[SP_J], ps->tot[SP_E]}) / 4;
    ps->seg_tiles = ps->tot[SP_P];
    ps->sort_tiles = ps->tot[SP_P] / 8;
    return DOGM_OK;
}

// Byte offsets of the step buffers in one region (256-byte aligned pieces).
struct ScratchLayout
{
    size_t tab, scan_sum, scan_pre, exp_sum, cdf_incl, seg_flags, seg_head, seg_tail, status, total;
};

ScratchLayout scratch_layout(const PlanSizes& ps)
{
    ScratchLayout L{};
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t o = at;
        at += (bytes + 255) / 256 * 256;
        return o;
    };
    L.tab = take(sizeof(uint32_t) * (size_t)ps.tab);
    L.scan_sum = take(sizeof(fx_t) * (size_t)ps.scan_tiles);
    L.scan_pre = take(sizeof(fx_t) * (size_t)ps.scan_tiles);
    L.exp_sum = take(sizeof(fx_t) * (size_t)ps.scan_tiles);
    L.cdf_incl = take(sizeof(fx_t) * CDF_SEGS_PER_TILE * (size_t)ps.scan_tiles);
    L.seg_flags = take((size_t)ps.seg_tiles);
    L.seg_head = take(sizeof(fx_t) * 5 * (size_t)ps.seg_tiles);
    L.seg_tail = take(sizeof(fx_t) * 5 * (size_t)ps.seg_tiles);
    L.status = take(sizeof(unsigned) * 256 * (size_t)ps.sort_tiles);
    L.total = at;
    return L;
}

}  // namespace

int engine_step_scratch_bytes(Vehicle* const* veh, int m, size_t* bytes, std::string* why)
{
    PlanSizes ps;
    const int st = plan_sizes(veh, m, &ps, why);
    if (st != DOGM_OK)
        return st;
    *bytes = scratch_layout(ps).total;
    return DOGM_OK;
}

int engine_bind_scratch(Engine* E, void* base, size_t bytes, Vehicle* const* veh, int m, std::string* why)
{
    if (E->bound)
        return set_err(why, DOGM_ERR_INVALID_ARGUMENT, "the engine's step buffers are already bound");
    PlanSizes ps;
    int st = plan_sizes(veh, m, &ps, why);
    if (st != DOGM_OK)
        return st;
    const ScratchLayout L = scratch_layout(ps);
    if (L.total > bytes)
        return set_err(why, DOGM_ERR_INVALID_ARGUMENT,
                       "the step needs " + std::to_string(L.total) + " bytes of scratch, " + std::to_string(bytes) +
                           " are bound");
    E->own = Engine::OwnStepBuffers{E->d_tab,         E->tab_cap,        E->scan_tile_sum, E->scan_tile_pre,
                                    E->exp_tile_sum,  E->cdf_tile_incl,  E->scan_tiles_cap, E->seg,
                                    E->seg_tiles_cap,
                                    E->sort_status,    E->sort_tiles_cap};
    char* b = static_cast<char*>(base);
    E->d_tab = reinterpret_cast<uint32_t*>(b + L.tab);
    E->tab_cap = ps.tab;
    E->scan_tile_sum = reinterpret_cast<fx_t*>(b + L.scan_sum);
    E->scan_tile_pre = reinterpret_cast<fx_t*>(b + L.scan_pre);
    E->exp_tile_sum = reinterpret_cast<fx_t*>(b + L.exp_sum);
    E->cdf_tile_incl = reinterpret_cast<fx_t*>(b + L.cdf_incl);
    E->scan_tiles_cap = ps.scan_tiles;
    E->seg.flags = reinterpret_cast<uint8_t*>(b + L.seg_flags);
    E->seg.head_partial = reinterpret_cast<fx_t*>(b + L.seg_head);
    E->seg.tail_partial = reinterpret_cast<fx_t*>(b + L.seg_tail);
    E->seg_tiles_cap = ps.seg_tiles;
    E->sort_status = reinterpret_cast<unsigned*>(b + L.status);
    E->sort_tiles_cap = ps.sort_tiles * 256;
    E->bound = true;
    // The tables of the last plan are in the engine's own buffers: rebuild them in the bound memory
    // (and recapture a step graph afterwards).
    E->plan_veh.clear();
    ++E->plan_epoch;
    return DOGM_OK;
}

void engine_unbind_scratch(Engine* E)
{
    if (!E->bound)
        return;
    const Engine::OwnStepBuffers& o = E->own;
    E->d_tab = o.d_tab;
    E->tab_cap = o.tab_cap;
    E->scan_tile_sum = o.scan_tile_sum;
    E->scan_tile_pre = o.scan_tile_pre;
    E->exp_tile_sum = o.exp_tile_sum;
    E->cdf_tile_incl = o.cdf_tile_incl;
    E->scan_tiles_cap = o.scan_tiles_cap;
    E->seg = o.seg;
    E->seg_tiles_cap = o.seg_tiles_cap;
    E->sort_status = o.sort_status;
    E->sort_tiles_cap = o.sort_tiles_cap;
    E->own = Engine::OwnStepBuffers{};
    E->bound = false;
    E->plan_veh.clear();
    ++E->plan_epoch;
}

int engine_init(Engine* E, std::string* why)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_err(why, DOGM_ERR_NO_DEVICE, "no HIP device");
    if (hipGetDevice(&E->device) != hipSuccess)
        return set_err(why, DOGM_ERR_NO_DEVICE, "hipGetDevice failed");
    if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess)
        return set_err(why, DOGM_ERR_HIP, "hipStreamCreateWithFlags failed");
    if (hipHostMalloc(reinterpret_cast<void**>(&E->h_err), sizeof(unsigned),
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
    {
        E->h_err = nullptr;
        return set_err(why, DOGM_ERR_OUT_OF_MEMORY, "allocating the device-check flag");
    }
    *E->h_err = 0u;
    if (hipHostGetDevicePointer(reinterpret_cast<void**>(&E->d_err), E->h_err, 0) != hipSuccess)
        return set_err(why, DOGM_ERR_HIP, "mapping the device-check flag");
    // Polls before a lookback wait gives up; a negative value gives up at once (a host-forced
    // failure for tests of that path).
    if (const char* spin = std::getenv("DOGM_LOOKBACK_SPIN"))
        E->lookback_spin = std::atoi(spin);
    return DOGM_OK;
}

int engine_device_check(Engine* E)
{
    if (!E->h_err || *reinterpret_cast<volatile unsigned*>(E->h_err) == 0u)
        return DOGM_OK;
    *reinterpret_cast<volatile unsigned*>(E->h_err) = 0u;
    return set_err(&E->last_error, DOGM_ERR_DEVICE_CHECK,
                   "a radix-sort lookback wait ran out of polls (an earlier tile never published)");
}

void engine_release(Engine* E)
{
    if (E->stream)
        (void)hipStreamSynchronize(E->stream);
    engine_unbind_scratch(E);
    if (E->graph_exec)
        (void)hipGraphExecDestroy(E->graph_exec);
    E->graph_exec = nullptr;
    if (E->h_desc)
        (void)hipHostFree(E->h_desc);
    if (E->h_upre)
        (void)hipHostFree(E->h_upre);
    E->h_desc = nullptr;
    E->h_upre = nullptr;
    if (E->h_err)
        (void)hipHostFree(E->h_err);
    E->h_err = nullptr;
    E->d_err = nullptr;
    {
        void* bufs[] = {E->d_desc,        E->d_upre,         E->d_tab,            E->scan_tile_sum,
                        E->scan_tile_pre, E->exp_tile_sum,   E->cdf_tile_incl,    E->seg.flags,
                        E->seg.head_partial,
                        E->seg.tail_partial, E->sort_status,  E->sort_ghist,       E->scalars};
        for (void* b : bufs)
            (void)hipFree(b);
    }
    for (auto& t : E->timed_pending)
    {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (auto e : E->event_pool)
        (void)hipEventDestroy(e);
    if (E->stream)
        (void)hipStreamDestroy(E->stream);
    E->stream = nullptr;
}

int engine_reserve(Engine* E, int m, std::string* why)
{
    if (m > E->m_cap)
    {
        const int cap = std::max(m, 2 * E->m_cap);
        if (E->h_desc)
            (void)hipHostFree(E->h_desc);
        if (E->h_upre)
            (void)hipHostFree(E->h_upre);
        (void)hipFree(E->d_desc);
        (void)hipFree(E->d_upre);
        (void)hipFree(E->sort_ghist);
        (void)hipFree(E->scalars);
        E->h_desc = nullptr;
        E->h_upre = nullptr;
        E->d_desc = nullptr;
        E->d_upre = nullptr;
        E->sort_ghist = nullptr;
        E->sort_tickets = nullptr;
        E->scalars = nullptr;
        E->m_cap = 0;
        if (hipHostMalloc(reinterpret_cast<void**>(&E->h_desc), sizeof(VehDesc) * cap) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void**>(&E->h_upre), sizeof(int) * SP_COUNT * (cap + 1)) != hipSuccess)
            return set_err(why, DOGM_ERR_OUT_OF_MEMORY, "allocating pinned step descriptors");
        int st;
        if ((st = dalloc(&E->d_desc, cap, "step descriptors", why)) ||
            (st = dalloc(&E->d_upre, (size_t)SP_COUNT * (cap + 1), "step tables", why)) ||
            (st = dalloc(&E->sort_ghist, (size_t)cap * 4 * 256 + 64, "sort histograms", why)) ||
            (st = dalloc(&E->scalars, cap, "step scalars", why)))
            return st;
        // Cleared on the engine's own stream and waited for: hipMemset runs on the null stream, which
        // a non-blocking stream does not wait for, so the next step's predict could otherwise count
        // digits into histograms still holding a freed buffer's contents.
        if (hipMemsetAsync(E->scalars, 0, sizeof(DevScalars) * cap, E->stream) != hipSuccess ||
            hipMemsetAsync(E->sort_ghist, 0, sizeof(unsigned) * ((size_t)cap * 4 * 256 + 64), E->stream) != hipSuccess ||
            hipStreamSynchronize(E->stream) != hipSuccess)
            return set_err(why, DOGM_ERR_HIP, "clearing step scalars");
        E->sort_tickets = E->sort_ghist + (size_t)cap * 4 * 256;
        E->ghist_clean = true;
        E->m_cap = cap;
    }
    return DOGM_OK;
}

int engine_plan(Engine* E, Vehicle* const* veh, int m, int stages)
{
    std::string why;
    if (m < 1 || m > MAX_STEP_VEHICLES)
        return set_err(&E->last_error, DOGM_ERR_INVALID_ARGUMENT,
                       "a step takes 1 to " + std::to_string(MAX_STEP_VEHICLES) + " vehicles");
    int st = engine_reserve(E, m, &why);
    if (st != DOGM_OK)
        return set_err(&E->last_error, st, why);

    bool same = (int)E->plan_veh.size() == m && E->m == m;
    for (int i = 0; same && i < m; ++i)
        same = E->plan_veh[i] == veh[i]->uid;
    if (!same)
    {
        PlanSizes ps;
        if ((st = plan_sizes(veh, m, &ps, &why)))
            return set_err(&E->last_error, st, why);
        const long long* tot = ps.tot;
        for (int s = 0; s < SP_COUNT; ++s)
        {
            int* pre = E->h_upre + s * (m + 1);
            pre[0] = 0;
            for (int i = 0; i < m; ++i)
                pre[i + 1] = pre[i] + (int)units_for(space_len(veh[i], s));
        }
        const int passes = ps.passes;
        const long long total = ps.tab;
        for (int s = 0; s < SP_COUNT; ++s)
        {
            E->units[s] = tot[s];
            E->tab_off[s] = s == 0 ? 0 : E->tab_off[s - 1] + tot[s - 1];
        }
        const long long scan_tiles = ps.scan_tiles;
        const long long seg_tiles = ps.seg_tiles;
        const long long sort_tiles = ps.sort_tiles;
        {
            if ((st = grow(&E->d_tab, &E->tab_cap, total, "work tables", &why)))
                return set_err(&E->last_error, st, why);
            long long c1 = E->scan_tiles_cap, c2 = E->scan_tiles_cap, c4 = E->scan_tiles_cap,
                      c3 = CDF_SEGS_PER_TILE * E->scan_tiles_cap;
            if ((st = grow(&E->scan_tile_sum, &c1, scan_tiles, "scan", &why)) ||
                (st = grow(&E->scan_tile_pre, &c2, scan_tiles, "scan", &why)) ||
                (st = grow(&E->exp_tile_sum, &c4, scan_tiles, "scan", &why)) ||
                (st = grow(&E->cdf_tile_incl, &c3, CDF_SEGS_PER_TILE * scan_tiles, "cdf index", &why)))
                return set_err(&E->last_error, st, why);
            E->scan_tiles_cap = std::min({c1, c2, c4, c3 / CDF_SEGS_PER_TILE});
            if (seg_tiles > E->seg_tiles_cap)
            {
                long long a = 0, b = 0, c = 0;
                (void)hipFree(E->seg.flags);
                (void)hipFree(E->seg.head_partial);
                (void)hipFree(E->seg.tail_partial);
                E->seg = SegCarry{};
                E->seg_tiles_cap = 0;
                if ((st = grow(&E->seg.flags, &a, seg_tiles, "segments", &why)) ||
                    (st = grow(&E->seg.head_partial, &b, seg_tiles * 5, "segments", &why)) ||
                    (st = grow(&E->seg.tail_partial, &c, seg_tiles * 5, "segments", &why)))
                    return set_err(&E->last_error, st, why);
                E->seg_tiles_cap = seg_tiles;
            }
            if ((st = grow(&E->sort_status, &E->sort_tiles_cap, sort_tiles * 256, "sort", &why)))
                return set_err(&E->last_error, st, why);
        }
        E->passes = passes;
        E->m = m;
        ++E->plan_epoch;
        E->plan_veh.assign(m, 0);
        for (int i = 0; i < m; ++i)
            E->plan_veh[i] = veh[i]->uid;
        if (hipMemcpyAsync(E->d_upre, E->h_upre, sizeof(int) * SP_COUNT * (m + 1), hipMemcpyHostToDevice,
                           E->stream) != hipSuccess)
            return set_err(&E->last_error, DOGM_ERR_HIP, "uploading step tables");
        run_build_tables(E);
    }

    // Per-step descriptors.  A deferred resampled set is read through its ancestors by a prediction
    // that the sort and the occupancy update follow in the same step (engine_run writes it out before
    // any other step that reads particles).
    const bool full_predict = (stages & ST_PREDICT) && (stages & ST_ASSIGN) && (stages & ST_OCC);
    for (int i = 0; i < m; ++i)
    {
        Vehicle* v = veh[i];
        VehDesc& d = E->h_desc[i];
        std::memset(&d, 0, sizeof(d));
        d.gs = v->gs;
        d.C = v->C;
        d.N = v->N;
        d.B = v->B;
        d.seed = v->seed;
        d.inst = v->instance;
        uint32_t e = v->epoch;
        d.init = (stages & ST_INIT) && v->init_pending ? 1 : 0;
        d.wuni = v->w_uniform || d.init ? 1 : 0;  // initialisation sets every weight to 1 / N
        d.rs_gather = v->rs_pending && !d.init && full_predict ? 1 : 0;
        d.rs_defer = (stages & ST_RESAMPLE) ? 1 : 0;
        d.ep_init = e;
        e += d.init ? 1u : 0u;
        d.ep_pred = e;
        e += (stages & ST_PREDICT) ? 1u : 0u;
        d.ep_birth = e;
        e += (stages & ST_BIRTH) ? 1u : 0u;
        d.ep_rs = e;
        const dogm_params& p = v->params;
        d.ps = p.persistence_prob;
        d.sp = p.stddev_process_noise_position;
        d.sv = p.stddev_process_noise_velocity;
        d.pB = p.birth_prob;
        d.sd_birth = p.stddev_velocity;
        d.vmax = p.init_max_velocity;
        d.dt = v->dt;
        // freespace_discount^dt with std::pow(float, float) as in dogm.cu:291.
        d.alpha = std::pow(p.freespace_discount, v->dt);
        d.move = (stages & ST_PREDICT) && v->pend_part_shift ? 1 : 0;
        d.mx = v->pend_px;
        d.my = v->pend_py;
        d.shifted = (stages & ST_OCC) && v->pend_grid_shift ? 1 : 0;
        d.gx = v->pend_gx;
        d.gy = v->pend_gy;
        d.xform_p = (stages & ST_PREDICT) && v->pend_xp ? 1 : 0;
        d.xform_g = (stages & ST_OCC) && v->pend_xg ? 1 : 0;
        d.xt = v->pend_xt;
        d.cdf_mode = v->cdf_mode;
        d.draw_mode = v->draw_mode;
        d.part[0] = v->part[0];
        d.part[1] = v->part[1];
        d.birth = v->birth;
        d.start = v->grid.start;
        d.end = v->grid.end;
        for (int f = 0; f < GF_COUNT; ++f)
            d.f[f] = v->grid.f[f];
        d.free[0] = v->grid.f[GF_FREE];
        d.free[1] = v->free_alt;
        d.meas = v->meas;
        d.meas_in = (stages & ST_OCC) ? v->meas_src : nullptr;
        d.weight_array = v->weight_array;
        d.cellrec = v->cellrec;
        d.cellin = v->cellin;
        d.born = v->born;
        d.m_raw = v->m_raw;
        d.pend = v->pend;
        d.ancestors = v->ancestors;
        d.sort_keys[0] = v->sort_keys[0];
        d.sort_keys[1] = v->sort_keys[1];
        d.sort_vals[0] = v->sort_vals[0];
        d.sort_vals[1] = v->sort_vals[1];
        d.cdf = v->cdf;
        d.x_occ = v->x_occ;
        d.x_img = v->x_img;
        d.x_mask = v->x_mask;
        d.x_cnt = v->x_cnt;
        d.x_off = v->x_off;
        d.x_idx = v->x_idx;
        d.nz = v->nz;
        d.nz_valid = v->nz_valid ? 1 : 0;
    }
    // Table offsets per vehicle (units); the prefixes are those the tables were built from.
    {
        long long off[SP_COUNT] = {};
        for (int i = 0; i < m; ++i)
            for (int s = 0; s < SP_COUNT; ++s)
            {
                const long long u = units_for(space_len(veh[i], s));
                E->h_desc[i].u0[s] = (int)off[s];
                E->h_desc[i].nu[s] = (int)u;
                off[s] += u;
            }
    }
    if (hipMemcpyAsync(E->d_desc, E->h_desc, sizeof(VehDesc) * m, hipMemcpyHostToDevice, E->stream) != hipSuccess)
        return set_err(&E->last_error, DOGM_ERR_HIP, "uploading step descriptors");
    return DOGM_OK;
}

int engine_run(Engine* E, Vehicle* const* veh, int m, int stages)
{
    if (stages & ST_EXTRACT)
    {
        const ExtractCfg& x = E->ext;
        if ((x.what & ~(DOGM_EXTRACT_OCCUPANCY | DOGM_EXTRACT_IMAGE | DOGM_EXTRACT_VELOCITY)) || x.what == 0)
            return set_err(&E->last_error, DOGM_ERR_INVALID_ARGUMENT, "unknown extraction flags");
        if (x.source != DOGM_SOURCE_GRID && x.source != DOGM_SOURCE_MEASUREMENT)
            return set_err(&E->last_error, DOGM_ERR_INVALID_ARGUMENT, "unknown extraction source");
        if ((x.what & DOGM_EXTRACT_VELOCITY) && x.source != DOGM_SOURCE_GRID)
            return set_err(&E->last_error, DOGM_ERR_INVALID_ARGUMENT, "the velocity test reads grid cells only");
        for (int i = 0; i < m; ++i)
        {
            std::string why;
            const int st = vehicle_ext_alloc(veh[i], x.what, &why);
            if (st != DOGM_OK)
                return set_err(&E->last_error, st, why);
        }
    }
    for (int i = 0; i < m; ++i)
    {
        // A pending device grid that this step's stages do not copy is copied here.
        Vehicle* v = veh[i];
        if (v->meas_src && !(stages & ST_OCC))
        {
            const dogm_measurement_cell* src = v->meas_src;
            v->meas_src = nullptr;
            const int cs = vehicle_set_measurement(E, v, src, 1);
            if (cs != DOGM_OK)
                return cs;
        }
    }
    // Deferred resampled sets (Vehicle::rs_pending): dropped where this step initialises the particles,
    // read through their ancestors by a full step's prediction, written out before any other stage
    // that reads or writes particles.
    {
        const bool full_predict = (stages & ST_PREDICT) && (stages & ST_ASSIGN) && (stages & ST_OCC);
        const int particle_stages = ST_PREDICT | ST_ASSIGN | ST_OCC | ST_PERS | ST_BIRTH | ST_MOMENTS | ST_RESAMPLE;
        for (int i = 0; i < m; ++i)
        {
            Vehicle* v = veh[i];
            if (!v->rs_pending)
                continue;
            if ((stages & ST_INIT) && v->init_pending)
                v->rs_pending = false;  // initialisation writes every particle
            else if (!full_predict && (stages & particle_stages))
                vehicle_materialize(E, v);
        }
    }
    if (!E->tune_locked)
    {
        long long particles = 0;
        for (int i = 0; i < m; ++i)
            particles += veh[i]->N;
        E->tune = tuning_for(m, particles);
    }
    int st = engine_plan(E, veh, m, stages);
    for (int i = 0; i < m; ++i)
        veh[i]->meas_src = nullptr;  // consumed by the descriptors (or dropped with the failed step)
    if (st != DOGM_OK)
        return st;
    bool any_init = false;
    for (int i = 0; i < m; ++i)
        any_init = any_init || ((stages & ST_INIT) && veh[i]->init_pending);
    int modes = 0;
    for (int i = 0; i < m; ++i)
        modes |= 1 << veh[i]->cdf_mode;
    bool timing = false;
    for (int k = 0; k < K_COUNT; ++k)
        timing = timing || E->timing_on[k];
    // Launch signature of the step: everything besides the plan that selects kernels or their
    // by-value arguments (descriptors live in device memory and are uploaded before the launches).
    uint32_t occ_bits = 0, vel_bits = 0;
    std::memcpy(&occ_bits, &E->ext.min_occ, 4);
    std::memcpy(&vel_bits, &E->ext.min_vel, 4);
    const uint64_t sig[3] = {(uint64_t)stages | ((uint64_t)any_init << 16) | ((uint64_t)modes << 17) |
                                 ((uint64_t)(stages & ST_EXTRACT ? E->ext.what : 0) << 24) |
                                 ((uint64_t)(stages & ST_EXTRACT ? E->ext.source : 0) << 32),
                             (stages & ST_EXTRACT) ? ((uint64_t)occ_bits << 32 | vel_bits) : 0,
                             tuning_signature(E->tune)};
    const bool graph = E->use_graph && !timing;
    bool replay = graph && E->graph_exec && E->graph_plan == E->plan_epoch && sig[0] == E->graph_sig[0] &&
                  sig[1] == E->graph_sig[1] && sig[2] == E->graph_sig[2];
    if (graph && !replay)
    {
        if (E->graph_exec)
            (void)hipGraphExecDestroy(E->graph_exec);
        E->graph_exec = nullptr;
        if (hipStreamBeginCapture(E->stream, hipStreamCaptureModeRelaxed) != hipSuccess)
            return set_err(&E->last_error, DOGM_ERR_HIP, "hipStreamBeginCapture failed");
    }
    int pph = 0;
    const bool hist = (stages & ST_PREDICT) && (stages & ST_ASSIGN);
    const bool occ_pers = (stages & ST_OCC) && (stages & ST_PERS);
    const bool assign_occ = (stages & ST_ASSIGN) && (stages & ST_OCC);
    const bool pers_mom = (stages & ST_PERS) && (stages & ST_MOMENTS);
    const bool occ_birth = (stages & ST_OCC) && (stages & ST_BIRTH);
    const bool occ_resample = (stages & ST_OCC) && (stages & ST_RESAMPLE);
    const int zero_empty = (stages & ST_OCC) ? (((stages & ST_PERS) ? 1 : 0) | ((stages & ST_MOMENTS) ? 2 : 0)) : 0;
    StageRange step_range(replay ? "dogm step (graph replay)" : "dogm step");
    if (!replay)
    {
        // pph below follows the same sequence in capture and replay
        if (any_init)
        {
            StageRange r("initializeParticles");
            run_init_particles(E);
        }
        if (stages & ST_PREDICT)
        {
            StageRange r("particlePrediction");
            run_predict(E, pph, hist, hist && assign_occ);
        }
    }
    if (stages & ST_ASSIGN)
    {
        if (!replay)
        {
            StageRange r("particleAssignment");
            run_assignment(E, pph, hist, assign_occ, !occ_pers);
        }
        pph ^= 1;
    }
    if (!replay)
    {
        if (stages & ST_OCC)
        {
            StageRange r("gridCellOccupancyUpdate");
            run_occupancy(E, pph, zero_empty, occ_pers, assign_occ, occ_birth, occ_resample);
        }
        if (stages & ST_PERS)
        {
            StageRange r("updatePersistentParticles");
            run_persistent(E, pph, occ_pers, (zero_empty & 1) != 0, pers_mom);
        }
        if (stages & ST_BIRTH)
        {
            StageRange r("initializeNewParticles");
            run_births(E, occ_birth);
        }
        if (stages & ST_MOMENTS)
        {
            StageRange r("statisticalMoments");
            run_moments(E, pph, pers_mom, (zero_empty & 2) != 0);
        }
    }
    if (stages & ST_RESAMPLE)
    {
        if (!replay)
        {
            StageRange r("resampling");
            run_resampling(E, pph, modes, occ_resample);
        }
        pph ^= 1;
    }
    if ((stages & ST_EXTRACT) && !replay)
    {
        StageRange r("extraction");
        run_extract(E, (stages & ST_OCC) ? 1 : 0);
    }
    if (graph)
    {
        if (!replay)
        {
            hipGraph_t g = nullptr;
            if (hipStreamEndCapture(E->stream, &g) != hipSuccess || !g)
                return set_err(&E->last_error, DOGM_ERR_HIP, "capturing the step graph failed");
            const hipError_t ie = hipGraphInstantiateWithFlags(&E->graph_exec, g, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess)
            {
                E->graph_exec = nullptr;
                return set_err(&E->last_error, DOGM_ERR_HIP, std::string("instantiating the step graph: ") +
                                                                 hipGetErrorString(ie));
            }
            E->graph_plan = E->plan_epoch;
            E->graph_sig[0] = sig[0];
            E->graph_sig[1] = sig[1];
            E->graph_sig[2] = sig[2];
        }
        if (hipGraphLaunch(E->graph_exec, E->stream) != hipSuccess)
            return set_err(&E->last_error, DOGM_ERR_HIP, "launching the step graph failed");
    }
    // Host state: the descriptors already hold this step's buffers, so swaps take effect for the
    // next step.
    for (int i = 0; i < m; ++i)
    {
        Vehicle* v = veh[i];
        const VehDesc& d = E->h_desc[i];
        uint32_t e = d.ep_rs;
        if (stages & ST_RESAMPLE)
            ++e;
        v->epoch = e;
        if (d.init)
            v->init_pending = false;
        // Resampling leaves every weight at the new weight; prediction, the sort and the persistent
        // update leave them differing.
        if (stages & ST_RESAMPLE)
            v->w_uniform = true;
        else if (stages & (ST_PREDICT | ST_ASSIGN | ST_PERS))
            v->w_uniform = false;
        else
            v->w_uniform = d.wuni != 0;
        // A resampling leaves a deferred set (merge_write); a prediction that read one consumed it.
        if (d.rs_defer)
            v->rs_pending = true;
        else if (d.rs_gather)
            v->rs_pending = false;
        if (pph)
            std::swap(v->part[0], v->part[1]);
        if (stages & ST_OCC)
            std::swap(v->grid.f[GF_FREE], v->free_alt);
        if (stages & ST_PREDICT)
            v->pend_part_shift = v->pend_xp = false;
        if (stages & ST_OCC)
            v->pend_grid_shift = v->pend_xg = false;
        if (stages & ST_ALL)
            v->nz_valid = zero_empty == 3;  // full updates rewrite the bits (k_mass_update)
        if (stages & ST_EXTRACT)
        {
            v->ext_what = E->ext.what;
            v->ext_source = E->ext.source;
            v->ext_min_occ = E->ext.min_occ;
            v->ext_min_vel = E->ext.min_vel;
        }
        else if (stages & ST_ALL)
        {
            v->ext_what = 0;
        }
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_err(&E->last_error, DOGM_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return DOGM_OK;
}

namespace
{

// Measurement copies, lazy pose updates and the stage sequence of one step, without waiting.
int frame_enqueue(Engine* E, Vehicle* const* veh, const FrameIn* in, int m, bool extract)
{
    for (int i = 0; i < m; ++i)
    {
        Vehicle* v = veh[i];
        if (in[i].meas)
        {
            int st = vehicle_take_measurement(E, v, in[i].meas, in[i].is_device);
            if (st != DOGM_OK)
                return st;
        }
        else if (!v->first_meas)
        {
            v->init_pending = true;
            v->first_meas = true;
        }
        vehicle_update_pose(E, v, in[i].x, in[i].y, in[i].yaw, true);
        v->dt = in[i].dt;
    }
    return engine_run(E, veh, m, ST_ALL | (extract ? ST_EXTRACT : 0));
}

int frame_wait(Engine* E, int st)
{
    const hipError_t e = hipStreamSynchronize(E->stream);
    const int dc = engine_device_check(E);
    if (st != DOGM_OK)
        return st;
    if (e != hipSuccess)
        return set_err(&E->last_error, DOGM_ERR_HIP, std::string("step: ") + hipGetErrorString(e));
    if (dc != DOGM_OK)
        return dc;
    engine_collect_timing(E);
    return DOGM_OK;
}

}  // namespace

int engine_frame(Engine* E, Vehicle* const* veh, const FrameIn* in, int m, bool extract)
{
    return frame_wait(E, frame_enqueue(E, veh, in, m, extract));
}

static int lanes_run(Engine* const* lanes, int nl, Vehicle* const* veh, const FrameIn* in, int m, bool extract);

int engine_frame_lanes(Engine* const* lanes, int nl, Vehicle* const* veh, const FrameIn* in, int m, bool extract)
{
    long long particles = 0;
    for (int i = 0; i < m; ++i)
        particles += veh[i]->N;
    const Tuning t = tuning_for(m, particles);
    if (nl == 0)
        nl = t.lanes;
    const int used = nl < 2 || m < 2 ? 1 : 2;
    for (int l = 0; l < used; ++l)
    {
        lanes[l]->tune = t;
        lanes[l]->tune_locked = true;
    }
    const int st = lanes_run(lanes, used, veh, in, m, extract);
    for (int l = 0; l < used; ++l)
        lanes[l]->tune_locked = false;
    return st;
}

static int lanes_run(Engine* const* lanes, int nl, Vehicle* const* veh, const FrameIn* in, int m, bool extract)
{
    if (nl < 2)
        return engine_frame(lanes[0], veh, in, m, extract);
    // Each vehicle goes to the lane with less stage-boundary traffic so far (bytes per frame
    // 172 N + 29 B + 128 C); a given vehicle list always splits the same way, so each lane keeps
    // its work tables from step to step.
    std::vector<Vehicle*> lv[2];
    std::vector<FrameIn> li[2];
    double load[2] = {0.0, 0.0};
    for (int i = 0; i < m; ++i)
    {
        const int l = load[1] < load[0] ? 1 : 0;
        load[l] += 172.0 * veh[i]->N + 29.0 * veh[i]->B + 128.0 * veh[i]->C;
        lv[l].push_back(veh[i]);
        li[l].push_back(in[i]);
    }
    int st[2];
    for (int l = 0; l < 2; ++l)
    {
        lanes[l]->ext = lanes[0]->ext;
        st[l] = frame_enqueue(lanes[l], lv[l].data(), li[l].data(), (int)lv[l].size(), extract);
    }
    for (int l = 0; l < 2; ++l)
        st[l] = frame_wait(lanes[l], st[l]);
    if (st[0] != DOGM_OK)
        return st[0];
    if (st[1] != DOGM_OK)
    {
        lanes[0]->last_error = lanes[1]->last_error;
        return st[1];
    }
    return DOGM_OK;
}

// ---- vehicles ----------------------------------------------------------------------------------

int vehicle_set_measurement(Engine* E, Vehicle* v, const dogm_measurement_cell* m, int is_device)
{
    const hipError_t e = hipMemcpyAsync(v->meas, m, sizeof(dogm_measurement_cell) * (size_t)v->C,
                                        is_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, E->stream);
    if (e != hipSuccess)
        return set_err(&E->last_error, DOGM_ERR_HIP, std::string("copying the measurement grid: ") + hipGetErrorString(e));
    // dogm.cu:205-215: the first measurement initialises the particles.
    if (!v->first_meas)
    {
        v->init_pending = true;
        v->first_meas = true;
    }
    return DOGM_OK;
}

int vehicle_take_measurement(Engine* E, Vehicle* v, const dogm_measurement_cell* m, int is_device)
{
    if (is_device && v->first_meas && !v->init_pending)
    {
        // The step's occupancy update copies the grid as it reads it (no separate copy).
        v->meas_src = m;
        return DOGM_OK;
    }
    return vehicle_set_measurement(E, v, m, is_device);
}

static void shift_grid_eager(Engine* E, Vehicle* v, int ix, int iy)
{
    v->nz_valid = false;
    auto shift_ptr = [&](void** field) {
        run_shift_field(E, v, *field, v->field_tmp, ix, iy);
        void* old = *field;
        *field = v->field_tmp;
        v->field_tmp = static_cast<float*>(old);
    };
    shift_ptr(reinterpret_cast<void**>(&v->grid.start));
    shift_ptr(reinterpret_cast<void**>(&v->grid.end));
    for (int f = 0; f < GF_COUNT; ++f)
        shift_ptr(reinterpret_cast<void**>(&v->grid.f[f]));
}

// Metric ego compensation (no counterpart in the reference, which ignores yaw and drops the sub-cell
// remainder of every shift, dogm.cu:172-200).  The grid frame is the sensor frame: x, y along the
// vehicle axes at yaw, y pointing left of x, the sensor at the grid centre.  A world point w lies at
// cell c + R(-yaw) (w - ref) / resolution, so a pose change moves it to
//   c + R(-dpsi) (g - c) - R(-yaw') (pos' - ref) / resolution.
// Only whole cells of the translation are applied; `ref` advances by exactly the applied distance,
// so the remainder carries into later frames.  The rotation is applied at every yaw change:
// particle positions and velocities continuously, grid fields by nearest-cell resampling.
// A metric transform still pending from a pose update whose step did not consume it is applied
// here, eagerly, before another pose update.
static void flush_pending_xform(Engine* E, Vehicle* v)
{
    if (v->pend_xp)
    {
        vehicle_materialize(E, v);
        run_ego_particles(E, v, v->pend_xt);
    }
    if (v->pend_xg)
    {
        run_ego_field(E, v, v->grid.f[GF_FREE], v->field_tmp, v->pend_xt);
        std::swap(v->grid.f[GF_FREE], v->field_tmp);
    }
    v->pend_xp = v->pend_xg = false;
}

static void ego_metric_update(Engine* E, Vehicle* v, float new_x, float new_y, float new_yaw, bool lazy)
{
    flush_pending_xform(E, v);
    if (!v->first_pose)
    {
        v->ref_x = new_x;
        v->ref_y = new_y;
        v->ref_yaw = new_yaw;
        v->pos_x = new_x;
        v->pos_y = new_y;
        v->yaw = new_yaw;
        v->first_pose = true;
        return;
    }
    const double res = v->params.resolution;
    const double dpsi = std::remainder((double)new_yaw - v->ref_yaw, 2.0 * M_PI);
    const double c = std::cos((double)new_yaw), s = std::sin((double)new_yaw);
    const double wx = (double)new_x - v->ref_x, wy = (double)new_y - v->ref_y;
    const double nx = std::trunc((c * wx + s * wy) / res), ny = std::trunc((-s * wx + c * wy) / res);
    v->ref_yaw = new_yaw;
    v->ref_x += (c * nx - s * ny) * res;
    v->ref_y += (s * nx + c * ny) * res;
    v->pos_x = (float)v->ref_x;
    v->pos_y = (float)v->ref_y;
    v->yaw = new_yaw;
    if (dpsi == 0.0 && nx == 0.0 && ny == 0.0)
        return;
    const EgoXform t{0.5f * (float)v->gs, (float)std::cos(dpsi), (float)std::sin(dpsi), (float)(-nx), (float)(-ny)};
    if (lazy)
    {
        // Inside a frame: particles in the step's predict, the free mass (the only grid field read
        // before being rewritten in a frame) in its occupancy update.
        v->pend_xp = v->pend_xg = true;
        v->pend_xt = t;
        return;
    }
    vehicle_materialize(E, v);
    run_ego_particles(E, v, t);
    auto xform = [&](void** field) {
        run_ego_field(E, v, *field, v->field_tmp, t);
        void* old = *field;
        *field = v->field_tmp;
        v->field_tmp = static_cast<float*>(old);
    };
    v->nz_valid = false;
    xform(reinterpret_cast<void**>(&v->grid.start));
    xform(reinterpret_cast<void**>(&v->grid.end));
    for (int f = 0; f < GF_COUNT; ++f)
        xform(reinterpret_cast<void**>(&v->grid.f[f]));
}

// dogm.cu:161-203.  lazy: the shift is applied inside the next step's predict (particles) and
// occupancy update (free mass, the only grid field read before being rewritten in a frame).
void vehicle_update_pose(Engine* E, Vehicle* v, float new_x, float new_y, float new_yaw, bool lazy)
{
    flush_pending_xform(E, v);
    if (v->ego_mode == DOGM_EGO_METRIC)
    {
        ego_metric_update(E, v, new_x, new_y, new_yaw, lazy);
        return;
    }
    if (!v->first_pose)
    {
        v->pos_x = new_x;
        v->pos_y = new_y;
        v->yaw = new_yaw;
        v->first_pose = true;
        return;
    }
    const float x_diff = new_x - v->pos_x;
    const float y_diff = new_y - v->pos_y;
    const float res = v->params.resolution;
    if (std::fabs(x_diff) > res || std::fabs(y_diff) > res)
    {
        const int x_move = -static_cast<int>(x_diff / res);
        const int y_move = -static_cast<int>(y_diff / res);
        if (lazy)
        {
            v->pend_part_shift = true;
            v->pend_px = (float)x_move;
            v->pend_py = (float)y_move;
            v->pend_grid_shift = true;
            v->pend_gx = x_move;
            v->pend_gy = y_move;
        }
        else
        {
            vehicle_materialize(E, v);
            run_move_particles(E, v, (float)x_move, (float)y_move);
            shift_grid_eager(E, v, x_move, y_move);
        }
        v->pos_x = new_x;
        v->pos_y = new_y;
        v->yaw = new_yaw;
    }
}


void vehicle_materialize(Engine* E, Vehicle* v)
{
    if (!v->rs_pending)
        return;
    run_materialize(E, v);
    v->rs_pending = false;
}

int vehicle_validate(const dogm_params* p, std::string* why)
{
    if (!p)
        return set_err(why, DOGM_ERR_INVALID_ARGUMENT, "params is null");
    if (!(p->size > 0.0f) || !(p->resolution > 0.0f) || !std::isfinite(p->size) || !std::isfinite(p->resolution))
        return set_err(why, DOGM_ERR_INVALID_ARGUMENT, "size and resolution must be positive and finite");
    const double gsd = (double)(p->size / p->resolution);
    if (!(gsd >= 1.0) || gsd > 16383.0)
        return set_err(why, DOGM_ERR_INVALID_ARGUMENT, "grid_size = int(size / resolution) must be in [1, 16383]");
    if (p->particle_count < 1 || p->new_born_particle_count < 0)
        return set_err(why, DOGM_ERR_INVALID_ARGUMENT,
                       "particle_count must be >= 1 and new_born_particle_count >= 0");
    if ((long long)p->particle_count + p->new_born_particle_count + 1 > MAX_SPACE_LEN)
        return set_err(why, DOGM_ERR_INVALID_ARGUMENT,
                       "particle_count + new_born_particle_count must be below 2^28");
    return DOGM_OK;
}

size_t vehicle_footprint(const dogm_params* p)
{
    const long long gs = (long long)(p->size / p->resolution);
    const long long C = gs * gs, N = p->particle_count, B = p->new_born_particle_count;
    const long long part = 16 + 4 + 4 + 1;
    // particles x2 + births, 16 grid fields + free_alt + field_tmp + meas(16) + born + m_raw + pend,
    // weights + ancestors + sort (16 B), CDF (16 B per joint entry), cell factors (16 B) and inputs
    // (8 B), occupancy bits.
    const long long pad = 8 * (2 * (N & 1) + (B & 1));  // the velocity arrays start 16-byte aligned
    return (size_t)(2 * N * part + B * part + pad + C * 4 * (GF_COUNT + 2 + 2) + C * 16 + C * 12 + N * 8 +
                    N * 16 + (N + B) * 16 + C * 24 + (C + 63) / 64 * 8);
}

int vehicle_alloc(Vehicle* v, const dogm_params* params, uint64_t seed, uint32_t instance, std::string* why)
{
    int st = vehicle_validate(params, why);
    if (st != DOGM_OK)
        return st;
    v->params = *params;
    v->gs = static_cast<int>(params->size / params->resolution);  // dogm.cu:34
    v->C = v->gs * v->gs;
    v->N = params->particle_count;
    v->B = params->new_born_particle_count;
    v->seed = seed;
    v->instance = instance;
    v->uid = next_uid();
    const int C = v->C, N = v->N, B = v->B;
    size_t* by = &v->bytes;

    if ((st = alloc_particles(v->part[0], N, why, by)) || (st = alloc_particles(v->part[1], N, why, by)) ||
        (st = alloc_particles(v->birth, B, why, by)))
        return st;
    if ((st = dalloc(&v->grid.start, C, "grid", why, by)) || (st = dalloc(&v->grid.end, C, "grid", why, by)))
        return st;
    for (int f = 0; f < GF_COUNT; ++f)
        if ((st = dalloc(&v->grid.f[f], C, "grid", why, by)))
            return st;
    if ((st = dalloc(&v->free_alt, C, "grid", why, by)) || (st = dalloc(&v->field_tmp, C, "grid", why, by)) ||
        (st = dalloc(&v->meas, C, "measurement grid", why, by)) ||
        (st = dalloc(&v->weight_array, N, "weights", why, by)) || (st = dalloc(&v->born, C, "born masses", why, by)) ||
        (st = dalloc(&v->m_raw, C, "cell masses", why, by)) || (st = dalloc(&v->pend, C, "particle orders", why, by)) ||
        (st = dalloc(&v->ancestors, N, "ancestors", why, by)) ||
        (st = dalloc(&v->sort_keys[0], N, "sort", why, by)) || (st = dalloc(&v->sort_keys[1], N, "sort", why, by)) ||
        (st = dalloc(&v->sort_vals[0], N, "sort", why, by)) || (st = dalloc(&v->sort_vals[1], N, "sort", why, by)) ||
        (st = dalloc(&v->cdf, (size_t)N + B, "cdf", why, by)) || (st = dalloc(&v->cellrec, C, "cell factors", why, by)) ||
        (st = dalloc(&v->cellin, C, "cell inputs", why, by)) ||
        (st = dalloc(&v->nz, ((size_t)C + 63) / 64, "occupancy bits", why, by)))
        return st;

    // initGridCellsKernel (init.cu:69-84): start = end = -1, masses 0, measurement (0, 0, 1, 1);
    // the remaining grid fields and buffers start at zero.
    if (hipMemset(v->grid.start, 0xff, sizeof(int) * C) != hipSuccess ||
        hipMemset(v->grid.end, 0xff, sizeof(int) * C) != hipSuccess)
        return set_err(why, DOGM_ERR_HIP, "clearing the grid");
    for (int f = 0; f < GF_COUNT; ++f)
        if (hipMemset(v->grid.f[f], 0, sizeof(float) * C) != hipSuccess)
            return set_err(why, DOGM_ERR_HIP, "clearing the grid");
    float* zbufs[] = {v->free_alt, v->weight_array, v->born, v->m_raw};
    const size_t zcnt[] = {(size_t)C, (size_t)N, (size_t)C, (size_t)C};
    for (int i = 0; i < 4; ++i)
        if (zcnt[i] && hipMemset(zbufs[i], 0, sizeof(float) * zcnt[i]) != hipSuccess)
            return set_err(why, DOGM_ERR_HIP, "clearing buffers");
    // Sort buffers start as valid cell / particle indices (0), so a key or value slot that a sort
    // pass failed to write (a lookback that gave up) still indexes inside the vehicle's arrays.
    if (hipMemset(v->ancestors, 0, sizeof(int) * N) != hipSuccess ||
        hipMemset(v->sort_keys[0], 0, sizeof(int) * N) != hipSuccess ||
        hipMemset(v->sort_keys[1], 0, sizeof(int) * N) != hipSuccess ||
        hipMemset(v->sort_vals[0], 0, sizeof(int) * N) != hipSuccess ||
        hipMemset(v->sort_vals[1], 0, sizeof(int) * N) != hipSuccess)
        return set_err(why, DOGM_ERR_HIP, "clearing buffers");
    {
        std::vector<dogm_measurement_cell> m0((size_t)C, dogm_measurement_cell{0.0f, 0.0f, 1.0f, 1.0f});
        if (hipMemcpy(v->meas, m0.data(), sizeof(dogm_measurement_cell) * C, hipMemcpyHostToDevice) != hipSuccess)
            return set_err(why, DOGM_ERR_HIP, "initialising the measurement grid");
    }
    if (hipDeviceSynchronize() != hipSuccess)
        return set_err(why, DOGM_ERR_HIP, "hipDeviceSynchronize");
    return DOGM_OK;
}

void vehicle_free(Vehicle* v)
{
    free_particles(v->part[0]);
    free_particles(v->part[1]);
    free_particles(v->birth);
    (void)hipFree(v->grid.start);
    (void)hipFree(v->grid.end);
    for (int f = 0; f < GF_COUNT; ++f)
        (void)hipFree(v->grid.f[f]);
    void* bufs[] = {v->free_alt, v->field_tmp,    v->meas,         v->weight_array, v->born,
                    v->m_raw,    v->pend,         v->ancestors,    v->sort_keys[0], v->sort_keys[1],
                    v->sort_vals[0], v->sort_vals[1], v->cdf,      v->aos_tmp,      v->state_tmp,
                    v->x_occ,    v->x_img,        v->x_mask,       v->x_cnt,        v->x_off,
                    v->x_idx,    v->x_rec,        v->nz,           v->cellrec,      v->cellin};
    for (void* b : bufs)
        (void)hipFree(b);
    *v = Vehicle{};
}

size_t extract_footprint(int C, int what)
{
    const size_t units = (size_t)units_for(C);
    size_t b = 0;
    if (what & DOGM_EXTRACT_OCCUPANCY)
        b += sizeof(float) * (size_t)C;
    if (what & DOGM_EXTRACT_IMAGE)
        b += (size_t)C;
    if (what & DOGM_EXTRACT_VELOCITY)
        b += units * (UNIT / WAVE * 8 + 8) + 4 + sizeof(int) * (size_t)C +
             sizeof(dogm_velocity_cell) * (size_t)std::min(C, EXTRACT_REC_CHUNK);
    return b;
}

int vehicle_ext_alloc(Vehicle* v, int what, std::string* why)
{
    const long long units = units_for(v->C);
    int st;
    if ((what & DOGM_EXTRACT_OCCUPANCY) && !v->x_occ)
    {
        if ((st = dalloc(&v->x_occ, (size_t)v->C, "occupancy output", why)))
            return st;
        v->x_bytes += extract_footprint(v->C, DOGM_EXTRACT_OCCUPANCY);
    }
    if ((what & DOGM_EXTRACT_IMAGE) && !v->x_img)
    {
        if ((st = dalloc(&v->x_img, (size_t)v->C, "occupancy image", why)))
            return st;
        v->x_bytes += extract_footprint(v->C, DOGM_EXTRACT_IMAGE);
    }
    if ((what & DOGM_EXTRACT_VELOCITY) && !v->x_idx)
    {
        if ((st = dalloc(&v->x_mask, (size_t)units * (UNIT / WAVE), "velocity ballots", why)) ||
            (st = dalloc(&v->x_cnt, (size_t)units, "velocity counts", why)) ||
            (st = dalloc(&v->x_off, (size_t)units + 1, "velocity offsets", why)) ||
            (st = dalloc(&v->x_idx, (size_t)v->C, "velocity cells", why)) ||
            (st = dalloc(&v->x_rec, (size_t)std::min(v->C, EXTRACT_REC_CHUNK), "velocity records", why)))
            return st;
        v->x_rec_cap = std::min(v->C, EXTRACT_REC_CHUNK);
        v->x_bytes += extract_footprint(v->C, DOGM_EXTRACT_VELOCITY);
    }
    return DOGM_OK;
}

}  // namespace dogm_hip
