// validate_traversal on the GPU, through the C++ API (dmap_amd.hpp): the
// reference's validation invariants (validate_traversal.cpp:18-25 checks
// (3) and (4), test_traversal_cost.cpp:268-297) on a flat quad with a random
// 64^2 map and on a quad with interpolated normals, over two ray
// distributions (from above towards the quad, and near the surface in
// random directions, as traversal_test_util.cpp:83-111 draws them):
//  (3) box, slab, both and slab levels 0, 1, 2 give the same hits, t within
//      1e-5 relative (hit/miss may differ only where the incidence is
//      grazing, |cos| < 1e-2);
//  (4) never-miss: rays aimed at sampled surface points of a flat quad under
//      a smooth map (slopes below 0.3) hit at or before the point (1e-4
//      relative: the rays are fp32), except where the hit is grazing
//      (|cos| < 1e-2). On the random map of (3) the entry and exit roots of
//      a nearly tangent ray can fall in one depth-5 sub-square, where the
//      reference's Newton leaf (displaced_intersector.cpp:239-323) converges
//      to one of them; the double oracle then reports the same exit root.
//  (5) the both test costs no more node tests than the box test, summed.
//  (6) path tracing through the C++ API: a white furnace (albedo 1,
//      environment 1, no light) returns at most 1 per pixel and about 1 on
//      average, and the ray counts add up.
// Prints the counts and "validate_traversal_gpu: PASS" or "FAIL"; exit 0 on
// PASS.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "dmap_amd.hpp"

using namespace dmap_amd;

namespace
{

std::vector<double> flat_quad()
{
    // two triangles over the unit square, normal +z, uv = xy
    return {0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 1, 1,
            0, 0, 0, 1, 1, 0, 0, 1, 0, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 1, 0, 1};
}

std::vector<double> curved_quad()
{
    std::vector<double> t = flat_quad();
    for (int k = 0; k < 2; ++k)
        for (int v = 0; v < 3; ++v) {
            double x = t[24 * k + 3 * v], y = t[24 * k + 3 * v + 1];
            double n[3] = {0.35 * (x - 0.5), 0.25 * (y - 0.5), 1.0};
            double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            for (int q = 0; q < 3; ++q)
                t[24 * k + 9 + 3 * v + q] = n[q] / l;
        }
    return t;
}

std::vector<RayQuery> ray_set(int n, std::mt19937_64 &rng)
{
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<RayQuery> rays(n);
    for (int r = 0; r < n; ++r) {
        RayQuery &q = rays[r];
        if (r < n / 2) {
            double o[3] = {-0.2 + 1.4 * U(rng), -0.2 + 1.4 * U(rng), 0.5 + U(rng)};
            double tgt[3] = {U(rng), U(rng), 0.0};
            for (int k = 0; k < 3; ++k) {
                q.o[k] = o[k];
                q.d[k] = tgt[k] - o[k];
            }
        } else {
            double o[3] = {U(rng), U(rng), -0.1 + 0.2 * U(rng)};
            double d[3] = {2 * U(rng) - 1, 2 * U(rng) - 1, 2 * U(rng) - 1};
            for (int k = 0; k < 3; ++k) {
                q.o[k] = o[k];
                q.d[k] = d[k];
            }
        }
        // the GPU takes fp32 rays: keep the queries exactly representable
        for (int k = 0; k < 3; ++k) {
            q.o[k] = (double)(float)q.o[k];
            q.d[k] = (double)(float)q.d[k];
        }
    }
    return rays;
}

double incidence(const RayQuery &q, const DisplacedHit &h)
{
    double l = std::sqrt(q.d[0] * q.d[0] + q.d[1] * q.d[1] + q.d[2] * q.d[2]);
    return std::fabs((q.d[0] * h.ng[0] + q.d[1] * h.ng[1] + q.d[2] * h.ng[2]) / l);
}

// HeightGrid::h on the closed grid (displaced_surface.cpp:12-39), repeat.
double height(const std::vector<float> &nodes, int W, double strength, double offset, double x, double y)
{
    x -= std::floor(x);
    y -= std::floor(y);
    double fx = x * (W - 1), fy = y * (W - 1);
    int i = std::min((int)fx, W - 2), j = std::min((int)fy, W - 2);
    double s = fx - i, t = fy - j;
    auto v = [&](int jj, int ii) { return (double)nodes[(size_t)jj * W + ii]; };
    return strength * ((1 - s) * (1 - t) * v(j, i) + s * (1 - t) * v(j, i + 1) + (1 - s) * t * v(j + 1, i) +
                       s * t * v(j + 1, i + 1)) +
           offset;
}

} // namespace

int main()
{
    bool pass = true;
    try {
        Context ctx(0);
        std::mt19937_64 rng(2027);
        const int n_tex = 64;
        std::vector<float> tex((size_t)n_tex * n_tex);
        std::uniform_real_distribution<float> Uf(0.f, 1.f);
        for (float &v : tex)
            v = Uf(rng);
        const double strength = 0.05;
        DisplacementTile tile(ctx, tex, n_tex, strength);
        for (int kind = 0; kind < 2; ++kind) {
            ObjectIntersector obj(ctx, tile, kind == 0 ? flat_quad() : curved_quad());
            std::vector<RayQuery> rays = ray_set(20000, rng);
            TraversalOptions base;
            std::vector<TraceResult> ref = obj.intersect(rays, base);
            // (3) every bound mode and slab level
            int64_t flips = 0, grazing = 0, t_bad = 0, hits = 0;
            for (int mode = 0; mode < 3; ++mode)
                for (int level = 0; level <= 2; ++level) {
                    TraversalOptions o;
                    o.bound = (BoundMode)mode;
                    o.slab_max_level = level;
                    std::vector<TraceResult> got = obj.intersect(rays, o);
                    for (size_t r = 0; r < rays.size(); ++r) {
                        const TraceResult &a = got[r], &b = ref[r];
                        if (a.hit != b.hit) {
                            const DisplacedHit &h = a.hit ? a.h : b.h;
                            if (incidence(rays[r], h) < 1e-2)
                                ++grazing;
                            else
                                ++flips;
                        } else if (a.hit) {
                            ++hits;
                            if (std::fabs(a.h.t - b.h.t) > 1e-5 * std::max(1.0, b.h.t))
                                ++t_bad;
                        }
                    }
                }
            std::printf("(3) %s: 9 mode/level policies x %zu rays: %lld hits agree, %lld flips, %lld grazing, "
                        "%lld t out of 1e-5\n",
                        kind == 0 ? "flat" : "curved", rays.size(), (long long)hits, (long long)flips,
                        (long long)grazing, (long long)t_bad);
            pass = pass && flips == 0 && t_bad == 0;
            // (5) both costs no more node tests than box
            TraversalOptions box, both;
            box.bound = BoundMode::Box;
            std::vector<TraceResult> sb = obj.intersect(rays, box, true), sbo = obj.intersect(rays, both, true);
            int64_t nb = 0, nbo = 0;
            for (size_t r = 0; r < rays.size(); ++r) {
                nb += sb[r].stats.node_tests;
                nbo += sbo[r].stats.node_tests;
            }
            std::printf("(5) %s: node tests box %lld, both %lld\n", kind == 0 ? "flat" : "curved", (long long)nb,
                        (long long)nbo);
            pass = pass && nbo <= nb;
            // (6) the node-test records of the box walk from the roots (TraversalStats::records): one per node
            // test, each ray's numbered from 0, and every cell whose box test passed above the leaves has its
            // children tested later in the same walk (in box mode the box's decision is the walk's).
            {
                std::vector<dmap_node_record> rec;
                TraversalOptions roots = box;
                roots.from_roots = true;
                std::vector<TraceResult> sr = obj.intersect(rays, roots, true, -1, &rec, (int64_t)(4 * nb + 1024));
                int64_t nr = 0;
                for (const TraceResult &x : sr)
                    nr += x.stats.node_tests;
                std::vector<std::vector<dmap_node_record>> per(rays.size());
                for (const dmap_node_record &x : rec)
                    per[(size_t)x.ray].push_back(x);
                int64_t bad = 0, descents = 0;
                for (size_t r = 0; r < rays.size(); ++r) {
                    std::vector<dmap_node_record> &v = per[r];
                    std::sort(v.begin(), v.end(),
                              [](const dmap_node_record &a, const dmap_node_record &b) { return a.seq < b.seq; });
                    if ((int64_t)v.size() != sr[r].stats.node_tests)
                        ++bad;
                    for (size_t q = 0; q < v.size(); ++q) {
                        if (v[q].seq != (int32_t)q)
                            ++bad;
                        if (!v[q].box_pass || v[q].level == 0)
                            continue;
                        ++descents;
                        bool child = false;
                        for (size_t u = q + 1; u < v.size() && !child; ++u)
                            child = v[u].prim == v[q].prim && v[u].level == v[q].level - 1 &&
                                    (v[u].i >> 1) == v[q].i && (v[u].j >> 1) == v[q].j;
                        if (!child)
                            ++bad;
                    }
                }
                std::printf("(6) %s: %zu node-test records of %lld node tests, %lld passing cells above the "
                            "leaves, %lld without their children or out of order\n",
                            kind == 0 ? "flat" : "curved", rec.size(), (long long)nr, (long long)descents,
                            (long long)bad);
                pass = pass && (int64_t)rec.size() == nr && bad == 0 && descents > 0;
            }
        }
        // (4) never-miss on the flat quad under a smooth map
        {
            std::vector<float> smooth((size_t)n_tex * n_tex);
            for (int j = 0; j < n_tex; ++j)
                for (int i = 0; i < n_tex; ++i)
                    smooth[(size_t)j * n_tex + i] =
                        (float)(0.5 + 0.25 * std::sin(2 * M_PI * 3 * i / n_tex) * std::cos(2 * M_PI * 2 * j / n_tex));
            DisplacementTile smooth_tile(ctx, smooth, n_tex, strength);
            ObjectIntersector obj(ctx, smooth_tile, flat_quad());
            std::vector<float> nodes = smooth_tile.nodes();
            std::uniform_real_distribution<double> U(0.0, 1.0);
            std::vector<RayQuery> aimed(20000);
            std::vector<double> dist(aimed.size());
            for (size_t r = 0; r < aimed.size(); ++r) {
                double s = 0.02 + 0.96 * U(rng), t = 0.02 + 0.96 * U(rng);
                double S[3] = {s, t, height(nodes, n_tex + 1, strength, 0.0, s, t)};
                double zc = 0.2 + 0.8 * U(rng), phi = 2 * M_PI * U(rng), sz = std::sqrt(1 - zc * zc);
                double dir[3] = {sz * std::cos(phi), sz * std::sin(phi), zc};
                double len = 0.3 + U(rng);
                RayQuery &q = aimed[r];
                for (int k = 0; k < 3; ++k) {
                    q.o[k] = (double)(float)(S[k] + len * dir[k]);
                    q.d[k] = (double)(float)(S[k] - q.o[k]);
                }
                // the aimed point in units of d: 1, up to the fp32 rounding of o and d
                dist[r] = 1.0;
            }
            std::vector<TraceResult> got = obj.intersect(aimed, TraversalOptions{});
            int64_t miss = 0, beyond = 0, grazing = 0;
            for (size_t r = 0; r < aimed.size(); ++r) {
                if (!got[r].hit)
                    ++miss;
                else if (got[r].h.t > dist[r] * (1.0 + 1e-4)) {
                    if (incidence(aimed[r], got[r].h) < 1e-2) {
                        ++grazing;
                    } else {
                        ++beyond;
                        if (beyond <= 5)
                            std::printf("    beyond: o (%.9g %.9g %.9g) d (%.9g %.9g %.9g) t %.9g texel (%lld %lld) "
                                        "st (%.9g %.9g) cos %.4f\n",
                                        aimed[r].o[0], aimed[r].o[1], aimed[r].o[2], aimed[r].d[0], aimed[r].d[1],
                                        aimed[r].d[2], got[r].h.t, (long long)got[r].h.i, (long long)got[r].h.j,
                                        got[r].h.st[0], got[r].h.st[1], incidence(aimed[r], got[r].h));
                    }
                }
            }
            std::printf("(4) never-miss: %zu aimed rays, %lld misses, %lld hits beyond the aimed point, "
                        "%lld grazing past it\n",
                        aimed.size(), (long long)miss, (long long)beyond, (long long)grazing);
            pass = pass && miss == 0 && beyond == 0;
        }
        // (6) white furnace through the wavefront path tracer
        {
            ObjectIntersector obj(ctx, tile, flat_quad());
            std::vector<RayQuery> cam(4096);
            std::uniform_real_distribution<double> U(0.0, 1.0);
            for (RayQuery &q : cam) {
                double o[3] = {0.5, -0.3, 0.8}, tgt[3] = {0.1 + 0.8 * U(rng), 0.1 + 0.8 * U(rng), 0.0};
                for (int k = 0; k < 3; ++k) {
                    q.o[k] = o[k];
                    q.d[k] = tgt[k] - o[k];
                }
            }
            ObjectIntersector::PathScene white;
            white.radiance = 0.0;
            white.env = 1.0;
            white.albedo = 1.0;
            for (int k = 0; k < 3; ++k)
                white.light_corner[k] = 1e3;
            int64_t n_closest = 0, n_shadow = 0;
            std::vector<double> img = obj.pathtrace(cam, 4, 16, white, 5, false, TraversalOptions{}, &n_closest,
                                                    &n_shadow);
            double mx = 0, mean = 0;
            for (double v : img) {
                mx = std::max(mx, v);
                mean += v / img.size();
            }
            std::printf("(6) white furnace: %zu pixels x 4 spp, %lld closest + %lld shadow rays, max %.6f mean %.6f\n",
                        img.size(), (long long)n_closest, (long long)n_shadow, mx, mean);
            pass = pass && mx <= 1.0 + 1e-12 && mean > 0.99 && n_closest >= (int64_t)(4 * cam.size()) &&
                   n_shadow == 0;
        }
    } catch (const Error &e) {
        std::printf("error: %s\n", e.what());
        pass = false;
    }
    std::printf("validate_traversal_gpu: %s\n", pass ? "PASS" : "FAIL");
    return pass ? 0 : 1;
}
