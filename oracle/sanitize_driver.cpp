// Host sanitizer run (TEST INFRASTRUCTURE): the checker (oracle_dogm.cpp) and the library's host-side
// JSON reader (csrc/dogm_json.h) built with AddressSanitizer and UndefinedBehaviorSanitizer by
// `make -C oracle sanitize` and run by tests/test_sanitizers.py.  Every stage of the update runs in
// both precisions, every CDF and draw mode and both ego modes over several frames with a grid shift,
// on a measurement grid from the checker's lidar generator, with every field read back; the result
// extraction and moment-sum helpers run too; the JSON reader takes valid, truncated, deeply nested and
// out-of-range inputs.  Any report aborts the run (-fno-sanitize-recover=all).
#include "../dynamic-occupancy-grid-map_amd/csrc/dogm_json.h"
#include "../include/dogm/dogm_c.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

extern "C" {
void* oracle_create(const dogm_params* params, uint64_t seed, uint32_t instance, int fp64);
void oracle_destroy(void* h);
int oracle_update_grid(void* h, const dogm_measurement_cell* m, float x, float y, float yaw, float dt);
int oracle_get(void* h, int field, double* out);
int oracle_get_i32(void* h, int field, int* out);
void oracle_set_cdf_mode(void* h, int mode);
void oracle_set_draw_mode(void* h, int mode);
void oracle_set_ego_mode(void* h, int mode);
void oracle_lidar_grid(int gs, int beams, double res, double max_range, float occ_mass, float free_mass,
                       const double* cosv, const double* sinv, const double* ranges, dogm_measurement_cell* out);
void oracle_occupancy(const float* occ, const float* free_mass, int n, float* p, uint8_t* img);
int oracle_cells_with_velocity(int gs, const float* occ, const float* free_mass, const float* mean_x,
                               const float* mean_y, const float* var_x, const float* var_y, const float* covar,
                               float min_occ, float min_vel, int* out_idx);
void oracle_moment_sums(const int* keys, int n, const float* x, double* out);
}

namespace
{

// Field ids of oracle_get / oracle_get_i32 (oracle_ctypes.py).
constexpr int F_P_X = 0, F_B_X = 5, F_WEIGHT_ARRAY = 10, F_BORN = 11, F_GRID0 = 16;
constexpr int I_P_IDX = 100, I_P_ASSOC = 101, I_B_IDX = 102, I_B_ASSOC = 103, I_G_START = 104, I_G_END = 105,
              I_ANCESTORS = 106;
constexpr int GRID_FLOAT_FIELDS = 14;

int failures = 0;

void check(bool ok, const char* what)
{
    if (!ok)
    {
        std::fprintf(stderr, "check failed: %s\n", what);
        ++failures;
    }
}

std::vector<dogm_measurement_cell> lidar_grid(int gs, double res, int beams, int frame)
{
    std::vector<double> c(beams), s(beams), r(beams);
    for (int b = 0; b < beams; ++b)
    {
        const double a = 2.0 * M_PI * b / beams;
        c[b] = std::cos(a);
        s[b] = std::sin(a);
        // a few boxes at frame-dependent ranges, misses (inf) and out-of-range returns
        r[b] = (b % 17 == 0) ? INFINITY : (b % 29 == 0) ? 1e9 : 2.0 + std::fmod(0.37 * b + 0.5 * frame, 0.45 * gs * res);
    }
    std::vector<dogm_measurement_cell> m((size_t)gs * gs);
    oracle_lidar_grid(gs, beams, res, 0.5 * gs * res, 0.7f, 0.4f, c.data(), s.data(), r.data(), m.data());
    return m;
}

void run_filter(int fp64, int cdf_mode, int draw_mode, int ego_mode)
{
    const int gs = 48;
    const double res = 0.5;
    dogm_params p{};
    p.size = (float)(gs * res);
    p.resolution = (float)res;
    p.particle_count = 9000;
    p.new_born_particle_count = 900;
    p.persistence_prob = 0.99f;
    p.stddev_process_noise_position = 0.1f;
    p.stddev_process_noise_velocity = 2.0f;
    p.birth_prob = 0.02f;
    p.stddev_velocity = 30.0f;
    p.init_max_velocity = 15.0f;
    p.freespace_discount = 0.01f;
    void* h = oracle_create(&p, 99, 3, fp64);
    check(h != nullptr, "oracle_create");
    oracle_set_cdf_mode(h, cdf_mode);
    oracle_set_draw_mode(h, draw_mode);
    oracle_set_ego_mode(h, ego_mode);
    const int C = gs * gs, N = p.particle_count, B = p.new_born_particle_count;
    std::vector<double> f((size_t)N);
    std::vector<int> iv((size_t)N);
    for (int fr = 0; fr < 4; ++fr)
    {
        const std::vector<dogm_measurement_cell> m = lidar_grid(gs, res, 180, fr);
        // 3 m/s with a yaw rate: the grid shifts on every frame after the first
        oracle_update_grid(h, m.data(), 0.3f * fr, 0.05f * fr, 0.1f * fr, 0.1f);
        // oracle_get and oracle_get_i32 return the element count
        for (int k = 0; k < 5; ++k)
            check(oracle_get(h, F_P_X + k, f.data()) == N, "particle field");
        for (int k = 0; k < 5; ++k)
        {
            std::vector<double> fb((size_t)B);
            check(oracle_get(h, F_B_X + k, fb.data()) == B, "birth field");
        }
        check(oracle_get(h, F_WEIGHT_ARRAY, f.data()) == N, "weight array");
        std::vector<double> g((size_t)C);
        check(oracle_get(h, F_BORN, g.data()) == C, "born masses");
        for (int k = 0; k < GRID_FLOAT_FIELDS; ++k)
            check(oracle_get(h, F_GRID0 + k, g.data()) == C, "grid field");
        std::vector<int> gi((size_t)C), bi((size_t)B);
        check(oracle_get_i32(h, I_G_START, gi.data()) == C && oracle_get_i32(h, I_G_END, gi.data()) == C, "bounds");
        check(oracle_get_i32(h, I_P_IDX, iv.data()) == N && oracle_get_i32(h, I_P_ASSOC, iv.data()) == N, "idx");
        check(oracle_get_i32(h, I_B_IDX, bi.data()) == B && oracle_get_i32(h, I_B_ASSOC, bi.data()) == B, "births");
        check(oracle_get_i32(h, I_ANCESTORS, iv.data()) == N, "ancestors");
        for (int k = 0; k < N; ++k)
            if (iv[k] < 0 || iv[k] >= N + B)
            {
                check(false, "ancestor in range");
                break;
            }
    }
    oracle_destroy(h);
}

void run_helpers()
{
    const int gs = 40, C = gs * gs;
    std::vector<float> occ(C), fr(C), mx(C), my(C), vx(C), vy(C), cv(C), pg(C);
    std::vector<uint8_t> img(C);
    for (int c = 0; c < C; ++c)
    {
        occ[c] = (float)((c * 37) % 101) / 100.0f;
        fr[c] = 1.0f - occ[c];
        mx[c] = (float)((c % 13) - 6);
        my[c] = (float)((c % 7) - 3);
        vx[c] = 1.0f + (c % 5);
        vy[c] = 1.0f + (c % 3);
        cv[c] = (c % 11 == 0) ? 0.0f : 0.25f;
    }
    oracle_occupancy(occ.data(), fr.data(), C, pg.data(), img.data());
    std::vector<int> cells(C);
    const int n = oracle_cells_with_velocity(gs, occ.data(), fr.data(), mx.data(), my.data(), vx.data(), vy.data(),
                                             cv.data(), 0.5f, 2.0f, cells.data());
    check(n >= 0 && n <= C, "cells with velocity");
    // five channels of 1000 values (x[c * n + k]) over sorted keys
    std::vector<int> keys(1000);
    std::vector<float> x((size_t)5 * 1000);
    for (int k = 0; k < 1000; ++k)
        keys[k] = k / 7;
    for (int k = 0; k < 5 * 1000; ++k)
        x[k] = 0.001f * (k % 997) - 0.3f;
    std::vector<double> out((size_t)5 * 1000);
    oracle_moment_sums(keys.data(), 1000, x.data(), out.data());
}

void run_json()
{
    const char* inputs[] = {
        R"({"entries": [{"vehicles": [1, 1], "particles": [0, 1e12], "sort": "onesweep", "tried": [{"a": null}]}]})",
        R"({"entries": [{"vehicles": [1e300, -1e300], "particles": [NaN, 1]}]})",
        R"({"entries": [{"vehicles": [1, 1], "particles": [0, 1], "merge_slice": 1e99}]})",
        R"({"entries": [)",
        R"({"entries": "abc)",
        R"({"a": "\)",
        R"([[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]])",
        R"({"k": tru})",
        "",
        "   ",
        R"({"entries": [1, "x", true, false, null, -0.0e-3]})",
    };
    for (const char* in : inputs)
    {
        const std::string text(in);
        dogm_json::Parser p{text};
        const dogm_json::JVal v = p.value();
        if (v.kind == dogm_json::JVal::OBJ)
            if (const dogm_json::JVal* e = v.get("entries"))
                for (const auto& x : e->arr)
                    if (const dogm_json::JVal* r = x.get("vehicles"))
                        for (const auto& y : r->arr)
                            (void)dogm_json::clamp_ll(y.num);
    }
}

}  // namespace

int main()
{
    for (int fp64 = 0; fp64 < 2; ++fp64)
        for (int mode = 0; mode < 4; ++mode)
            run_filter(fp64, mode, mode & 1, (mode >> 1) & 1);
    run_helpers();
    run_json();
    if (failures)
    {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("sanitizer run: checker stages (f32 and f64, 4 CDF modes, both draw and ego modes), helpers "
                "and the JSON reader clean\n");
    return 0;
}
