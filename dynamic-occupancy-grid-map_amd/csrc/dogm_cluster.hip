// Accuracy harness (host code): DBSCAN over the cells that pass the velocity test, cluster means,
// assignment to ground-truth objects and MAE / RMSE of position and velocity.  Restates the
// reference demo's evaluation (dogm/demo/utils/dbscan.cpp:24-114, precision_evaluator.cpp:37-122,
// metrics.cpp:7-53).
//
// DBSCAN here is the textbook algorithm (Ester et al., KDD'96) with the reference's conventions --
// a point's neighbourhood includes the point, core points have at least min_cells neighbours within
// eps (Euclidean distance of the cell coordinates, sqrtf as dbscan.cpp:19-22), clusters are numbered
// in the order of their first core point in the input, expansion is breadth-first in input order.
// Neighbourhoods come from a hash of the integer cell coordinates (the cells lie on the grid), so a
// query touches the (2 floor(eps) + 1)^2 window instead of every point.  Border points keep the
// first cluster that reaches them.  The reference's expandCluster marks the first |seeds| entries
// of the point array instead of the seeds (dbscan.cpp:62-65); that is not reproduced.
#include "../../include/dogm/dogm_c.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <deque>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

namespace
{

constexpr int kUnclassified = -2;  // dbscan.h:20-21
constexpr int kNoise = -1;

uint64_t cell_key(long long x, long long y)
{
    return ((uint64_t)(uint32_t)(int32_t)x << 32) | (uint64_t)(uint32_t)(int32_t)y;
}

struct Index
{
    std::unordered_map<uint64_t, std::vector<int>> at;  // cell -> points on it (input order)
    int reach = 0;                                       // floor(eps): window half-width in cells
};

// Points within eps of point i, in input order (dbscan.cpp:100-114).
void region(const dogm_velocity_cell* c, const Index& ix, float eps, int i, std::vector<int>* out)
{
    out->clear();
    const long long x0 = (long long)std::floor(c[i].x), y0 = (long long)std::floor(c[i].y);
    for (long long dy = -ix.reach - 1; dy <= ix.reach + 1; ++dy)
        for (long long dx = -ix.reach - 1; dx <= ix.reach + 1; ++dx)
        {
            auto it = ix.at.find(cell_key(x0 + dx, y0 + dy));
            if (it == ix.at.end())
                continue;
            for (int j : it->second)
            {
                const float ddx = c[i].x - c[j].x, ddy = c[i].y - c[j].y;
                if (sqrtf(ddx * ddx + ddy * ddy) <= eps)
                    out->push_back(j);
            }
        }
    std::sort(out->begin(), out->end());
}

}  // namespace

struct dogm_evaluator
{
    float resolution = 0.0f;
    float max_distance = 0.0f;
    float eps = 0.0f;
    int min_cells = 0;
    double abs_sum[4] = {0, 0, 0, 0};  // |error| of x, y, vx, vy (metrics.cpp:7-20)
    double sq_sum[4] = {0, 0, 0, 0};   // error^2 (metrics.cpp:33-46)
    long long detections = 0;
    long long unassigned = 0;
    long long clusters = 0;
    std::string last_error;
};

extern "C" {

int dogm_dbscan(const dogm_velocity_cell* cells, int n, float eps, int min_cells, int* cluster_id, int* clusters)
{
    if (n < 0 || (n > 0 && (!cells || !cluster_id)) || !clusters || !(eps >= 0.0f) || !std::isfinite(eps) ||
        eps > 1e6f)
        return DOGM_ERR_INVALID_ARGUMENT;
    Index ix;
    ix.reach = (int)std::floor(eps);
    for (int i = 0; i < n; ++i)
    {
        if (!std::isfinite(cells[i].x) || !std::isfinite(cells[i].y))
            return DOGM_ERR_INVALID_ARGUMENT;
        ix.at[cell_key((long long)std::floor(cells[i].x), (long long)std::floor(cells[i].y))].push_back(i);
    }
    for (int i = 0; i < n; ++i)
        cluster_id[i] = kUnclassified;
    int cid = 0;
    std::vector<int> nb, nb2;
    std::deque<int> seeds;
    for (int p = 0; p < n; ++p)
    {
        if (cluster_id[p] != kUnclassified)
            continue;
        region(cells, ix, eps, p, &nb);
        if ((int)nb.size() < min_cells)
        {
            cluster_id[p] = kNoise;
            continue;
        }
        for (int q : nb)
        {
            if (cluster_id[q] == kUnclassified)
                seeds.push_back(q);
            if (cluster_id[q] == kUnclassified || cluster_id[q] == kNoise)
                cluster_id[q] = cid;
        }
        while (!seeds.empty())
        {
            const int q = seeds.front();
            seeds.pop_front();
            region(cells, ix, eps, q, &nb2);
            if ((int)nb2.size() < min_cells)
                continue;
            for (int r : nb2)
            {
                if (cluster_id[r] == kUnclassified)
                    seeds.push_back(r);
                if (cluster_id[r] == kUnclassified || cluster_id[r] == kNoise)
                    cluster_id[r] = cid;
            }
        }
        ++cid;
    }
    *clusters = cid;
    return DOGM_OK;
}

int dogm_cluster_means(const dogm_velocity_cell* cells, int n, const int* cluster_id, int clusters,
                       dogm_cluster* out)
{
    if (n < 0 || clusters < 0 || (n > 0 && (!cells || !cluster_id)) || (clusters > 0 && !out))
        return DOGM_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < clusters; ++k)
        out[k] = dogm_cluster{0.0f, 0.0f, 0.0f, 0.0f, 0};
    // f32 sums in input order, then divided by the size (precision_evaluator.cpp:98-110).
    for (int i = 0; i < n; ++i)
    {
        const int k = cluster_id[i];
        if (k < 0)
            continue;
        if (k >= clusters)
            return DOGM_ERR_INVALID_ARGUMENT;
        out[k].x += cells[i].x;
        out[k].y += cells[i].y;
        out[k].vx += cells[i].data.mean_x_vel;
        out[k].vy += cells[i].data.mean_y_vel;
        out[k].cells += 1;
    }
    for (int k = 0; k < clusters; ++k)
    {
        const float m = (float)out[k].cells;
        if (out[k].cells > 0)
        {
            out[k].x /= m;
            out[k].y /= m;
            out[k].vx /= m;
            out[k].vy /= m;
        }
    }
    return DOGM_OK;
}

int dogm_evaluator_create(float resolution, float max_assignment_distance, float eps, int min_cells,
                          dogm_evaluator** out)
{
    if (!out || !(resolution > 0.0f) || !(max_assignment_distance >= 0.0f) || !(eps >= 0.0f) || min_cells < 1)
        return DOGM_ERR_INVALID_ARGUMENT;
    dogm_evaluator* e = new (std::nothrow) dogm_evaluator();
    if (!e)
        return DOGM_ERR_OUT_OF_MEMORY;
    e->resolution = resolution;
    e->max_distance = max_assignment_distance;
    e->eps = eps;
    e->min_cells = min_cells;
    *out = e;
    return DOGM_OK;
}

void dogm_evaluator_destroy(dogm_evaluator* e)
{
    delete e;
}

int dogm_evaluator_add_step(dogm_evaluator* e, const dogm_velocity_cell* cells, int n, const dogm_object* truth,
                            int objects, dogm_cluster* clusters_out, int clusters_cap, int* clusters)
{
    if (!e || n < 0 || objects < 0 || (objects > 0 && !truth) || (n > 0 && !cells) || clusters_cap < 0 ||
        (clusters_cap > 0 && !clusters_out))
        return DOGM_ERR_INVALID_ARGUMENT;
    if (clusters)
        *clusters = 0;
    // precision_evaluator.cpp:41: a step with no cells or no objects adds nothing.
    if (n == 0 || objects == 0)
        return DOGM_OK;
    std::vector<int> id((size_t)n);
    int k = 0;
    int st = dogm_dbscan(cells, n, e->eps, e->min_cells, id.data(), &k);
    if (st != DOGM_OK)
        return st;
    std::vector<dogm_cluster> means((size_t)k);
    st = dogm_cluster_means(cells, n, id.data(), k, means.data());
    if (st != DOGM_OK)
        return st;
    for (int c = 0; c < k; ++c)
    {
        // Cell units to metres (precision_evaluator.cpp:111-114); the grid's y axis points up here,
        // so no flip.
        dogm_cluster m = means[c];
        m.x *= e->resolution;
        m.y *= e->resolution;
        m.vx *= e->resolution;
        m.vy *= e->resolution;
        if (c < clusters_cap)
            clusters_out[c] = m;
        // Closest object within the assignment distance (precision_evaluator.cpp:49-75).
        int best = -1;
        float best_d = 0.0f;
        for (int o = 0; o < objects; ++o)
        {
            const float dx = m.x - truth[o].x, dy = m.y - truth[o].y;
            const float d = sqrtf(dx * dx + dy * dy);
            if (d < e->max_distance && (best < 0 || d < best_d))
            {
                best = o;
                best_d = d;
            }
        }
        if (best < 0)
        {
            ++e->unassigned;
            continue;
        }
        const double err[4] = {(double)(m.x - truth[best].x), (double)(m.y - truth[best].y),
                               (double)(m.vx - truth[best].vx), (double)(m.vy - truth[best].vy)};
        for (int a = 0; a < 4; ++a)
        {
            e->abs_sum[a] += std::fabs(err[a]);
            e->sq_sum[a] += err[a] * err[a];
        }
        ++e->detections;
    }
    e->clusters += k;
    if (clusters)
        *clusters = k;
    return DOGM_OK;
}

int dogm_evaluator_summary(const dogm_evaluator* e, dogm_object* mae, dogm_object* rmse, long long* detections,
                           long long* unassigned)
{
    if (!e)
        return DOGM_ERR_INVALID_ARGUMENT;
    const double n = e->detections > 0 ? (double)e->detections : NAN;
    double v[4], r[4];
    for (int a = 0; a < 4; ++a)
    {
        v[a] = e->abs_sum[a] / n;
        r[a] = std::sqrt(e->sq_sum[a] / n);
    }
    if (mae)
        *mae = dogm_object{(float)v[0], (float)v[1], (float)v[2], (float)v[3]};
    if (rmse)
        *rmse = dogm_object{(float)r[0], (float)r[1], (float)r[2], (float)r[3]};
    if (detections)
        *detections = e->detections;
    if (unassigned)
        *unassigned = e->unassigned;
    return DOGM_OK;
}

}  // extern "C"
