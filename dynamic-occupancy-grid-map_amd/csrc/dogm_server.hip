// Serving: vehicles admitted against a device-memory budget keep their filter state resident;
// frames from any of them are queued with their arrival times, and each step runs the oldest queued
// frame of every vehicle that has one (up to max_batch) in one batched update (continuous
// batching: frames that arrive while a step runs join the next one).  No counterpart in the
// reference, which runs one dogm::DOGM per grid, one blocking updateGrid per frame
// (dogm/src/dogm.cu:115-131).
//
// Two ways to drive it:
//  - dogm_server_step(now): one step on the caller's thread, stamped on the caller's clock (start =
//    now, done = now + the step's measured duration), for virtual-time simulation and tests;
//  - dogm_server_start(): a worker thread that steps whenever frames are queued, stamped with
//    dogm_server_clock().
#include "dogm_internal.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <set>
#include <thread>

using namespace dogm_hip;

struct dogm_server
{
    Engine eng;
    Engine aux;  // second lane of steps over two or more vehicles (engine_frame_lanes)
    std::mutex eng_mu;  // serialises steps and read-backs on the engine's stream
    std::mutex mu;      // guards everything below
    std::condition_variable cv_work, cv_done;
    std::map<int, std::unique_ptr<Vehicle>> veh;
    std::map<int, int> inflight;  // queued + running frames per vehicle
    std::map<int, uint64_t> reserved;  // budget bytes held per vehicle
    ExtractCfg ext;                    // extraction inside every step (what = 0: none)
    uint64_t budget = 0, resident = 0;
    int max_batch = 0;
    int next_id = 1;
    uint64_t next_ticket = 1;
    struct Frame
    {
        uint64_t ticket;
        int vid;
        FrameIn in;
        double t_arrival;
        dogm_lidar* lidar;           // frames submitted as a scan: the generator of their grid
        std::vector<double> ranges;  // the scan (beam_count distances)
    };
    std::deque<Frame> queue;
    std::vector<dogm_frame_result> done;
    std::thread worker;
    bool running = false, stopping = false;
    std::string last_error;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    long long steps = 0, frames = 0;
};

namespace
{

int sfail(dogm_server* s, int code, const std::string& msg)
{
    std::lock_guard<std::mutex> g(s->mu);
    s->last_error = msg;
    return code;
}

double clock_now(const dogm_server* s)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

// Takes the next step's frames from the queue (caller holds s->mu).
std::vector<dogm_server::Frame> take_frames(dogm_server* s)
{
    std::vector<dogm_server::Frame> out;
    std::set<int> seen;
    for (auto it = s->queue.begin(); it != s->queue.end();)
    {
        if (s->max_batch > 0 && (int)out.size() >= s->max_batch)
            break;
        if (seen.insert(it->vid).second)
        {
            out.push_back(*it);
            it = s->queue.erase(it);
        }
        else
        {
            ++it;
        }
    }
    return out;
}

// Generates the measurement grids of the step's scan frames into their vehicles' measurement
// buffers, one generate_many call per generator.
int generate_scans(dogm_server* s, const std::vector<dogm_server::Frame>& fr, const std::vector<Vehicle*>& list)
{
    std::vector<dogm_lidar*> gens;
    for (const auto& f : fr)
        if (f.lidar && std::find(gens.begin(), gens.end(), f.lidar) == gens.end())
            gens.push_back(f.lidar);
    for (dogm_lidar* l : gens)
    {
        const int beams = lidar_beam_count(l);
        std::vector<double> ranges;
        std::vector<dogm_measurement_cell*> outs;
        for (size_t i = 0; i < fr.size(); ++i)
        {
            if (fr[i].lidar != l)
                continue;
            ranges.insert(ranges.end(), fr[i].ranges.begin(), fr[i].ranges.begin() + beams);
            outs.push_back(list[i]->meas);
        }
        for (size_t o = 0; o < outs.size(); o += 65535)
        {
            const int n = (int)std::min<size_t>(65535, outs.size() - o);
            const int st = dogm_lidar_generate_many(l, n, ranges.data() + o * (size_t)beams, outs.data() + o);
            if (st != DOGM_OK)
            {
                s->eng.last_error = std::string("generating measurement grids: ") + dogm_lidar_last_error(l);
                return st;
            }
        }
    }
    return DOGM_OK;
}

// Runs one step over `fr`; wall duration in *secs.
int run_step(dogm_server* s, const std::vector<dogm_server::Frame>& fr, double* secs)
{
    std::vector<Vehicle*> list(fr.size());
    std::vector<FrameIn> in(fr.size());
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (size_t i = 0; i < fr.size(); ++i)
        {
            list[i] = s->veh.at(fr[i].vid).get();
            in[i] = fr[i].in;
        }
    }
    std::lock_guard<std::mutex> ge(s->eng_mu);
    const auto a = std::chrono::steady_clock::now();
    const int gst = generate_scans(s, fr, list);
    if (gst != DOGM_OK)
    {
        *secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
        std::lock_guard<std::mutex> g(s->mu);
        s->last_error = s->eng.last_error;
        return gst;
    }
    s->eng.ext = s->ext;
    Engine* lanes[2] = {&s->eng, &s->aux};
    const int st = engine_frame_lanes(lanes, 0, list.data(), in.data(), (int)fr.size(), s->ext.what != 0);
    *secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
    if (st != DOGM_OK)
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->last_error = s->eng.last_error;
    }
    return st;
}

void record(dogm_server* s, const std::vector<dogm_server::Frame>& fr, int st, double t_start, double t_done)
{
    std::lock_guard<std::mutex> g(s->mu);
    for (const auto& f : fr)
    {
        dogm_frame_result r{};
        r.ticket = f.ticket;
        r.vehicle_id = f.vid;
        r.status = st;
        r.batch_size = (int)fr.size();
        r.t_arrival = f.t_arrival;
        r.t_start = t_start;
        r.t_done = t_done;
        s->done.push_back(r);
        if (--s->inflight[f.vid] == 0)
            s->inflight.erase(f.vid);
    }
    s->steps += 1;
    s->frames += (long long)fr.size();
    s->cv_done.notify_all();
}

void worker_loop(dogm_server* s)
{
    for (;;)
    {
        std::vector<dogm_server::Frame> fr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stopping || !s->queue.empty(); });
            if (s->queue.empty())
                return;  // stopping and drained
            fr = take_frames(s);
        }
        const double t_start = clock_now(s);
        double secs = 0.0;
        const int st = run_step(s, fr, &secs);
        record(s, fr, st, t_start, clock_now(s));
    }
}

}  // namespace

extern "C" {

int dogm_server_create(const dogm_server_config* cfg, dogm_server** out)
{
    if (!cfg || !out)
        return DOGM_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (cfg->max_batch < 0)
        return DOGM_ERR_INVALID_ARGUMENT;
    dogm_server* s = new (std::nothrow) dogm_server();
    if (!s)
        return DOGM_ERR_OUT_OF_MEMORY;
    std::string why;
    int st = engine_init(&s->eng, &why);
    if (st == DOGM_OK)
        st = engine_init(&s->aux, &why);
    if (st != DOGM_OK)
    {
        engine_release(&s->aux);
        engine_release(&s->eng);
        delete s;
        return st;
    }
    s->budget = cfg->hbm_budget_bytes;
    s->max_batch = cfg->max_batch;
    *out = s;
    return DOGM_OK;
}

void dogm_server_destroy(dogm_server* s)
{
    if (!s)
        return;
    (void)dogm_server_stop(s);
    {
        std::lock_guard<std::mutex> ge(s->eng_mu);
        for (auto& kv : s->veh)
            vehicle_free(kv.second.get());
        s->veh.clear();
        engine_release(&s->aux);
        engine_release(&s->eng);
    }
    delete s;
}

const char* dogm_server_last_error(const dogm_server* s)
{
    return s ? s->last_error.c_str() : "null server";
}

double dogm_server_clock(const dogm_server* s)
{
    return s ? clock_now(s) : 0.0;
}

int dogm_server_admit(dogm_server* s, const dogm_params* params, uint64_t seed, uint32_t instance, int* vehicle_id)
{
    if (!s || !vehicle_id)
        return DOGM_ERR_INVALID_ARGUMENT;
    *vehicle_id = 0;
    std::string why;
    int st = vehicle_validate(params, &why);
    if (st != DOGM_OK)
        return sfail(s, st, why);
    uint64_t fp = (uint64_t)vehicle_footprint(params);
    int ext_what = 0;
    {
        std::lock_guard<std::mutex> g(s->mu);
        ext_what = s->ext.what;
        if (ext_what)
        {
            const int gs = (int)(params->size / params->resolution);
            fp += (uint64_t)extract_footprint(gs * gs, ext_what);
        }
        if (s->resident + fp > s->budget)
        {
            s->last_error = "admission would exceed the device-memory budget (" + std::to_string(s->resident) +
                            " + " + std::to_string(fp) + " > " + std::to_string(s->budget) + " bytes)";
            return DOGM_ERR_OUT_OF_MEMORY;
        }
        s->resident += fp;  // reserved while allocating
    }
    std::unique_ptr<Vehicle> v(new (std::nothrow) Vehicle());
    if (!v)
        st = DOGM_ERR_OUT_OF_MEMORY;
    else
        st = vehicle_alloc(v.get(), params, seed, instance, &why);
    if (st == DOGM_OK && ext_what)
        st = vehicle_ext_alloc(v.get(), ext_what, &why);
    std::lock_guard<std::mutex> g(s->mu);
    if (st != DOGM_OK)
    {
        if (v)
            vehicle_free(v.get());
        s->resident -= fp;
        s->last_error = why;
        return st;
    }
    const int id = s->next_id++;
    s->veh[id] = std::move(v);
    s->reserved[id] = fp;
    *vehicle_id = id;
    return DOGM_OK;
}

int dogm_server_release(dogm_server* s, int vehicle_id)
{
    if (!s)
        return DOGM_ERR_INVALID_ARGUMENT;
    std::unique_ptr<Vehicle> v;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->veh.find(vehicle_id);
        if (it == s->veh.end())
        {
            s->last_error = "unknown vehicle id " + std::to_string(vehicle_id);
            return DOGM_ERR_INVALID_ARGUMENT;
        }
        if (s->inflight.count(vehicle_id))
        {
            if (!s->running)
            {
                s->last_error = "vehicle " + std::to_string(vehicle_id) + " has queued frames";
                return DOGM_ERR_INVALID_ARGUMENT;
            }
            s->cv_done.wait(lk, [&] { return !s->inflight.count(vehicle_id); });
        }
        v = std::move(it->second);
        s->veh.erase(it);
        s->resident -= s->reserved[vehicle_id];
        s->reserved.erase(vehicle_id);
    }
    std::lock_guard<std::mutex> ge(s->eng_mu);
    vehicle_free(v.get());
    return DOGM_OK;
}

uint64_t dogm_server_resident_bytes(const dogm_server* s)
{
    if (!s)
        return 0;
    std::lock_guard<std::mutex> g(const_cast<dogm_server*>(s)->mu);
    return s->resident;
}

int dogm_server_submit(dogm_server* s, int vehicle_id, const dogm_measurement_cell* meas, int is_device, float x,
                       float y, float yaw, float dt, double t_arrival, uint64_t* ticket)
{
    if (!s)
        return DOGM_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(s->mu);
    if (!s->veh.count(vehicle_id))
    {
        s->last_error = "unknown vehicle id " + std::to_string(vehicle_id);
        return DOGM_ERR_INVALID_ARGUMENT;
    }
    const uint64_t t = s->next_ticket++;
    s->queue.push_back(
        dogm_server::Frame{t, vehicle_id, FrameIn{meas, is_device, x, y, yaw, dt}, t_arrival, nullptr, {}});
    s->inflight[vehicle_id] += 1;
    if (ticket)
        *ticket = t;
    s->cv_work.notify_one();
    return DOGM_OK;
}

int dogm_server_submit_scan(dogm_server* s, int vehicle_id, dogm_lidar* lidar, const double* ranges, float x,
                            float y, float yaw, float dt, double t_arrival, uint64_t* ticket)
{
    if (!s)
        return DOGM_ERR_INVALID_ARGUMENT;
    if (!lidar || !ranges)
        return sfail(s, DOGM_ERR_INVALID_ARGUMENT, "a scan needs a generator and its ranges");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->veh.find(vehicle_id);
    if (it == s->veh.end())
    {
        s->last_error = "unknown vehicle id " + std::to_string(vehicle_id);
        return DOGM_ERR_INVALID_ARGUMENT;
    }
    if (lidar_grid_size(lidar) != it->second->gs)
    {
        s->last_error = "the generator's grid size " + std::to_string(lidar_grid_size(lidar)) +
                        " differs from the vehicle's " + std::to_string(it->second->gs);
        return DOGM_ERR_INVALID_ARGUMENT;
    }
    const uint64_t t = s->next_ticket++;
    const int beams = lidar_beam_count(lidar);
    // The grid is generated into the vehicle's measurement buffer at the start of the step.
    s->queue.push_back(dogm_server::Frame{t, vehicle_id, FrameIn{nullptr, 1, x, y, yaw, dt}, t_arrival, lidar,
                                          std::vector<double>(ranges, ranges + beams)});
    s->inflight[vehicle_id] += 1;
    if (ticket)
        *ticket = t;
    s->cv_work.notify_one();
    return DOGM_OK;
}

int dogm_server_queued(const dogm_server* s)
{
    if (!s)
        return 0;
    std::lock_guard<std::mutex> g(const_cast<dogm_server*>(s)->mu);
    return (int)s->queue.size();
}

int dogm_server_step(dogm_server* s, double now, int* frames_run, double* step_seconds)
{
    if (!s)
        return DOGM_ERR_INVALID_ARGUMENT;
    std::vector<dogm_server::Frame> fr;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->running)
        {
            s->last_error = "the worker thread is running";
            return DOGM_ERR_INVALID_ARGUMENT;
        }
        fr = take_frames(s);
    }
    if (frames_run)
        *frames_run = (int)fr.size();
    if (step_seconds)
        *step_seconds = 0.0;
    if (fr.empty())
        return DOGM_OK;
    double secs = 0.0;
    const int st = run_step(s, fr, &secs);
    record(s, fr, st, now, now + secs);
    if (step_seconds)
        *step_seconds = secs;
    return st;
}

int dogm_server_start(dogm_server* s)
{
    if (!s)
        return DOGM_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(s->mu);
    if (s->running)
        return DOGM_OK;
    s->stopping = false;
    s->running = true;
    s->worker = std::thread(worker_loop, s);
    return DOGM_OK;
}

int dogm_server_stop(dogm_server* s)
{
    if (!s)
        return DOGM_ERR_INVALID_ARGUMENT;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (!s->running)
            return DOGM_OK;
        s->stopping = true;
        s->cv_work.notify_all();
    }
    s->worker.join();
    std::lock_guard<std::mutex> g(s->mu);
    s->running = false;
    s->stopping = false;
    return DOGM_OK;
}

int dogm_server_poll(dogm_server* s, dogm_frame_result* out, int max_results, int* count)
{
    if (!s || !count || (max_results > 0 && !out))
        return DOGM_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(s->mu);
    const int n = std::min<int>(max_results, (int)s->done.size());
    std::copy(s->done.begin(), s->done.begin() + n, out);
    s->done.erase(s->done.begin(), s->done.begin() + n);
    *count = n;
    return DOGM_OK;
}

int dogm_server_wait_idle(dogm_server* s)
{
    if (!s)
        return DOGM_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    if (!s->running && !s->queue.empty())
    {
        s->last_error = "frames are queued and no worker is running";
        return DOGM_ERR_INVALID_ARGUMENT;
    }
    s->cv_done.wait(lk, [&] { return s->inflight.empty(); });
    return DOGM_OK;
}

int dogm_server_get_grid_cells(dogm_server* s, int vehicle_id, dogm_grid_cell* out)
{
    if (!s || !out)
        return DOGM_ERR_INVALID_ARGUMENT;
    Vehicle* v = nullptr;
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->veh.find(vehicle_id);
        if (it == s->veh.end())
        {
            s->last_error = "unknown vehicle id " + std::to_string(vehicle_id);
            return DOGM_ERR_INVALID_ARGUMENT;
        }
        v = it->second.get();
    }
    std::lock_guard<std::mutex> ge(s->eng_mu);
    Engine* E = &s->eng;
    if (!v->aos_tmp && hipMalloc(reinterpret_cast<void**>(&v->aos_tmp), sizeof(dogm_grid_cell) * (size_t)v->C) !=
                           hipSuccess)
    {
        v->aos_tmp = nullptr;
        return sfail(s, DOGM_ERR_OUT_OF_MEMORY, "allocating the grid packing buffer");
    }
    run_pack_grid(E, v);
    if (hipMemcpyAsync(out, v->aos_tmp, sizeof(dogm_grid_cell) * (size_t)v->C, hipMemcpyDeviceToHost, E->stream) !=
            hipSuccess ||
        hipStreamSynchronize(E->stream) != hipSuccess)
        return sfail(s, DOGM_ERR_HIP, "reading back the grid");
    return DOGM_OK;
}

int dogm_server_get_particles(dogm_server* s, int vehicle_id, dogm_vec4* state, int* grid_cell_idx, float* weight,
                              uint8_t* associated)
{
    if (!s)
        return DOGM_ERR_INVALID_ARGUMENT;
    Vehicle* v = nullptr;
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->veh.find(vehicle_id);
        if (it == s->veh.end())
        {
            s->last_error = "unknown vehicle id " + std::to_string(vehicle_id);
            return DOGM_ERR_INVALID_ARGUMENT;
        }
        v = it->second.get();
    }
    std::lock_guard<std::mutex> ge(s->eng_mu);
    Engine* E = &s->eng;
    const int st = read_particles(E, v, v->part[0], v->N, state, grid_cell_idx, weight, associated);
    if (st != DOGM_OK)
        return sfail(s, st, "reading back particles: " + E->last_error);
    return DOGM_OK;
}

int dogm_server_set_extraction(dogm_server* s, int what, float min_occupancy, float min_velocity)
{
    if (!s)
        return DOGM_ERR_INVALID_ARGUMENT;
    if (what & ~(DOGM_EXTRACT_OCCUPANCY | DOGM_EXTRACT_IMAGE | DOGM_EXTRACT_VELOCITY))
        return sfail(s, DOGM_ERR_INVALID_ARGUMENT, "unknown extraction flags");
    std::lock_guard<std::mutex> g(s->mu);
    if (!s->veh.empty())
    {
        s->last_error = "set the extraction before admitting vehicles (its buffers count against the budget)";
        return DOGM_ERR_INVALID_ARGUMENT;
    }
    s->ext = ExtractCfg{what, DOGM_SOURCE_GRID, min_occupancy, min_velocity};
    return DOGM_OK;
}

}  // extern "C"

namespace
{

// Runs `read` on the vehicle's extraction results for `what` under the engine lock.
template <typename Read>
int server_extract(dogm_server* s, int vehicle_id, int what, Read read)
{
    if (!s)
        return DOGM_ERR_INVALID_ARGUMENT;
    Vehicle* v = nullptr;
    ExtractCfg want;
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->veh.find(vehicle_id);
        if (it == s->veh.end())
        {
            s->last_error = "unknown vehicle id " + std::to_string(vehicle_id);
            return DOGM_ERR_INVALID_ARGUMENT;
        }
        v = it->second.get();
        want = ExtractCfg{what, DOGM_SOURCE_GRID, s->ext.min_occ, s->ext.min_vel};
    }
    std::lock_guard<std::mutex> ge(s->eng_mu);
    Engine* E = &s->eng;
    int st = extract_ensure(E, v, want, true);
    if (st == DOGM_OK)
        st = read(E, v);
    if (st != DOGM_OK)
        return sfail(s, st, E->last_error);
    return DOGM_OK;
}

}  // namespace

extern "C" {

int dogm_server_get_occupancy(dogm_server* s, int vehicle_id, float* out)
{
    return server_extract(s, vehicle_id, DOGM_EXTRACT_OCCUPANCY,
                          [&](Engine* E, Vehicle* v) { return extract_read_occupancy(E, v, out); });
}

int dogm_server_get_occupancy_image(dogm_server* s, int vehicle_id, uint8_t* out)
{
    return server_extract(s, vehicle_id, DOGM_EXTRACT_IMAGE,
                          [&](Engine* E, Vehicle* v) { return extract_read_image(E, v, out); });
}

int dogm_server_get_cells_with_velocity(dogm_server* s, int vehicle_id, dogm_velocity_cell* out, int capacity,
                                        int* count)
{
    return server_extract(s, vehicle_id, DOGM_EXTRACT_VELOCITY, [&](Engine* E, Vehicle* v) {
        return extract_read_velocity(E, v, out, capacity, count);
    });
}

int dogm_server_stats(const dogm_server* s, long long* steps, long long* frames)
{
    if (!s)
        return DOGM_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(const_cast<dogm_server*>(s)->mu);
    if (steps)
        *steps = s->steps;
    if (frames)
        *frames = s->frames;
    return DOGM_OK;
}

}  // extern "C"
