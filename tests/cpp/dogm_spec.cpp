// C++ re-expression of the reference's filter specs (dogm/test/dogm_spec.cpp) against the drop-in
// header include/dogm/dogm.h.  The reference's versions read particle memory that was never
// initialised and pass a null measurement grid (dogm_spec.cpp:24-37); here particles are set
// explicitly so every expectation is deterministic.  GoogleTest is not available offline, so a
// minimal EXPECT/TEST harness is used.  Needs a GPU; run by tests/test_gpu_cpp_spec.py.
#include "dogm/dogm.h"

#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

static int g_failures = 0;
static std::vector<std::pair<std::string, std::function<void()>>>& registry()
{
    static std::vector<std::pair<std::string, std::function<void()>>> r;
    return r;
}
struct Registrar
{
    Registrar(const char* n, std::function<void()> f) { registry().emplace_back(n, std::move(f)); }
};
#define TEST(suite, name)                                                        \
    static void suite##_##name();                                                \
    static Registrar reg_##suite##_##name(#suite "." #name, suite##_##name);     \
    static void suite##_##name()
#define EXPECT_TRUE(c)                                                                         \
    do                                                                                         \
    {                                                                                          \
        if (!(c))                                                                              \
        {                                                                                      \
            ++g_failures;                                                                      \
            std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);                       \
        }                                                                                      \
    } while (0)
#define EXPECT_EQ(a, b) EXPECT_TRUE((a) == (b))

static dogm::DOGM::Params spec_params()
{
    // dogm_spec.cpp:12-22; freespace_discount is left uninitialised there.
    dogm::DOGM::Params p;
    p.size = 10.0f;
    p.resolution = 1.0f;
    p.particle_count = 2;
    p.new_born_particle_count = 1;
    p.persistence_prob = 0.5f;
    p.stddev_process_noise_position = 0.0f;
    p.stddev_process_noise_velocity = 0.0f;
    p.birth_prob = 0.02f;
    p.stddev_velocity = 10.0f;
    p.init_max_velocity = 30.0f;
    p.freespace_discount = 0.01f;
    return p;
}

static void set_two_particles(dogm::DOGM& d)
{
    dogm::ParticlesSoA p(2, false);
    p.state[0] = dogm::vec4(2.5f, 3.5f, 1.25f, -0.75f);
    p.state[1] = dogm::vec4(7.25f, 1.5f, -2.0f, 0.5f);
    p.grid_cell_idx[0] = 2 + 10 * 3;
    p.grid_cell_idx[1] = 7 + 10 * 1;
    p.weight[0] = 0.25f;
    p.weight[1] = 0.5f;
    p.associated[0] = true;
    p.associated[1] = false;
    d.setParticles(p);
    p.free();
}

// dogm_spec.cpp:10-66 (ego-motion compensation, dogm.cu:161-203).
TEST(DOGM, EgoMotionCompensation)
{
    dogm::DOGM dogm(spec_params());
    set_two_particles(dogm);
    dogm::ParticlesSoA particles = dogm.getParticles();
    dogm::vec4 old_state = particles.state[0];
    float pose_x = 10.0f, pose_y = 10.0f;

    // Set initial pose (no position update).
    EXPECT_EQ(dogm_update_pose(dogm.handle(), pose_x, pose_y, 0.0f), DOGM_OK);
    EXPECT_EQ(pose_x, dogm.getPositionX());
    EXPECT_EQ(pose_y, dogm.getPositionY());
    dogm::ParticlesSoA new_particles = dogm.getParticles();
    EXPECT_EQ(old_state, new_particles.state[0]);

    // A change below the resolution does not update the pose after the initial one is set.
    EXPECT_EQ(dogm_update_pose(dogm.handle(), pose_x + 0.5f, pose_y + 0.5f, 0.0f), DOGM_OK);
    EXPECT_EQ(pose_x, dogm.getPositionX());
    EXPECT_EQ(pose_y, dogm.getPositionY());
    dogm::ParticlesSoA new_particles2 = dogm.getParticles();
    EXPECT_EQ(old_state, new_particles2.state[0]);

    // A 3-cell change moves the particles by +3 in x: the move is negative and is subtracted.
    const float x_change = 3.0f;
    pose_x += x_change;
    EXPECT_EQ(dogm_update_pose(dogm.handle(), pose_x, pose_y, 0.0f), DOGM_OK);
    EXPECT_EQ(pose_x, dogm.getPositionX());
    EXPECT_EQ(pose_y, dogm.getPositionY());
    old_state.x += x_change;
    dogm::ParticlesSoA new_particles3 = dogm.getParticles();
    EXPECT_EQ(old_state, new_particles3.state[0]);

    particles.free();
    new_particles.free();
    new_particles2.free();
    new_particles3.free();
}

// dogm_spec.cpp:68-103 (constant-velocity prediction without noise, predict.cu:16-52).
TEST(DOGM, Predict)
{
    dogm::DOGM dogm(spec_params());
    set_two_particles(dogm);
    const float delta_time = 0.1f;
    dogm::ParticlesSoA particles = dogm.getParticles();
    dogm::vec4 old_state = particles.state[0];
    dogm::vec4 pred_state = old_state + delta_time * dogm::vec4(old_state[2], old_state[3], 0, 0);
    float old_weight = particles.weight[0];

    dogm.particlePrediction(delta_time);
    dogm::ParticlesSoA new_particles = dogm.getParticles();
    EXPECT_EQ(pred_state, new_particles.state[0]);
    EXPECT_EQ(old_weight * dogm.params.persistence_prob, new_particles.weight[0]);
    EXPECT_EQ(new_particles.grid_cell_idx[0], 2 + 10 * 3);
    particles.free();
    new_particles.free();
}

// A few full frames through the drop-in surface, from host and from device measurement grids.
TEST(DOGM, UpdateGridHostAndDeviceMeasurements)
{
    dogm::DOGM::Params p;
    p.size = 20.0f;
    p.resolution = 0.5f;
    p.particle_count = 4000;
    p.new_born_particle_count = 400;
    dogm::DOGM dogm(p);
    const int gs = dogm.getGridSize();
    EXPECT_EQ(gs, 40);
    std::vector<dogm::MeasurementCell> meas(gs * gs, dogm::MeasurementCell{0.0f, 0.0f, 1.0f, 1.0f});
    for (int x = 10; x < 14; ++x)
        meas[x + gs * 20] = dogm::MeasurementCell{0.0f, 0.7f, 1.0f, 1.0f};
    for (int y = 0; y < 20; ++y)
        meas[12 + gs * y].free_mass = 0.4f;
    dogm.updateGrid(meas.data(), 0.0f, 0.0f, 0.0f, 0.1f, false);

    void* dmeas = nullptr;
    EXPECT_EQ(dogm_device_malloc(&dmeas, meas.size() * sizeof(dogm::MeasurementCell)), DOGM_OK);
    EXPECT_EQ(dogm_memcpy(dmeas, meas.data(), meas.size() * sizeof(dogm::MeasurementCell), 0), DOGM_OK);
    for (int f = 1; f < 4; ++f)
        dogm.updateGrid(static_cast<dogm::MeasurementCell*>(dmeas), 0.6f * f, 0.0f, 0.0f, 0.1f);
    dogm_device_free(dmeas);

    dogm::ParticlesSoA parts = dogm.getParticles();
    bool equal_weights = true;
    for (int i = 1; i < parts.size; ++i)
        equal_weights = equal_weights && parts.weight[i] == parts.weight[0];
    EXPECT_TRUE(equal_weights);  // resampling sets every weight to total / N (resampling.cu:66)
    parts.free();

    const std::vector<dogm::GridCell> cells = dogm.getGridCells();
    float max_occ = 0.0f;
    bool bounded = true;
    for (const auto& c : cells)
    {
        max_occ = std::fmax(max_occ, c.occ_mass);
        bounded = bounded && c.occ_mass >= 0.0f && c.free_mass >= 0.0f && c.occ_mass + c.free_mass <= 1.0f + 1e-6f;
    }
    EXPECT_TRUE(bounded);
    EXPECT_TRUE(max_occ > 0.5f);
    EXPECT_EQ(dogm.getMeasurementCells()[12 + gs * 20].occ_mass, 0.7f);
}

TEST(DOGM, InvalidArgumentsThrow)
{
    dogm::DOGM::Params p = spec_params();
    p.particle_count = 0;
    bool threw = false;
    try
    {
        dogm::DOGM d(p);
    }
    catch (const std::runtime_error&)
    {
        threw = true;
    }
    EXPECT_TRUE(threw);
    dogm::DOGM ok(spec_params());
    threw = false;
    try
    {
        ok.updateGrid(nullptr, 0.0f, 0.0f, 0.0f, 0.1f);
    }
    catch (const std::runtime_error&)
    {
        threw = true;
    }
    EXPECT_TRUE(threw);
}

int main()
{
    for (auto& t : registry())
    {
        const int before = g_failures;
        std::printf("[ RUN      ] %s\n", t.first.c_str());
        std::fflush(stdout);
        try
        {
            t.second();
        }
        catch (const std::exception& e)
        {
            ++g_failures;
            std::printf("  exception: %s\n", e.what());
        }
        std::printf("[ %s ] %s\n", g_failures == before ? "      OK" : " FAILED ", t.first.c_str());
        std::fflush(stdout);
    }
    std::printf("%d failure(s)\n", g_failures);
    return g_failures == 0 ? 0 : 1;
}
