// C++ drop-in check of include/antipodal_hip/gwn_mesh_hip.hh: the reference's own test cases (test/test_gwn_mesh.cc,
// test/test_intersector.cc, test/test_gwn_mesh_robust.cc, test/test_gwn_mesh_opengl.cc), written against the
// reference's API and types, with HipDispatcher / HipMeshIntersector / RobustMeshGwn / HipGridGwn swapped in.
// Built against the reference headers where they are present (tests/cpp/Makefile); run by tests/test_gpu_cpp.py.
#include <antipodal/dispatcher/dispatcher.hh>
#include <antipodal/gwn_mesh.hh>
#include <antipodal/intersector/intersector.hh>
#include <antipodal/math/common.hh>
#include <antipodal_hip/gwn_mesh_hip.hh>

#include "gwn_baseline.hh"

#include <cmath>
#include <cstdio>
#include <vector>

static int g_fail = 0, g_pass = 0;
#define CHECK(cond)                                                                                                    \
    do                                                                                                                 \
    {                                                                                                                  \
        if (cond)                                                                                                      \
            ++g_pass;                                                                                                  \
        else                                                                                                           \
        {                                                                                                              \
            ++g_fail;                                                                                                  \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);                                                \
        }                                                                                                              \
    } while (0)

using namespace antipodal;

template <class T>
std::vector<vec3<T>> cube_vertices()
{
    using V = vec3<T>;
    return {V{T(-.5), T(-.5), T(-.5)}, V{T(.5), T(-.5), T(-.5)}, V{T(-.5), T(.5), T(-.5)}, V{T(.5), T(.5), T(-.5)},
            V{T(-.5), T(-.5), T(.5)},  V{T(.5), T(-.5), T(.5)},  V{T(-.5), T(.5), T(.5)},  V{T(.5), T(.5), T(.5)}};
}
static std::vector<int> const closed = {0, 2, 1, 1, 2, 3, 4, 5, 6, 6, 5, 7, 0, 4, 2, 2, 4, 6,
                                        1, 3, 5, 5, 3, 7, 0, 1, 4, 4, 1, 5, 2, 6, 3, 3, 6, 7};
static std::vector<int> const open_y(closed.begin(), closed.begin() + 30);

template <class T>
std::vector<weighted_segment3<T>> open_y_boundary()
{
    using V = vec3<T>;
    V const v2{T(-.5), T(.5), T(-.5)}, v3{T(.5), T(.5), T(-.5)}, v6{T(-.5), T(.5), T(.5)}, v7{T(.5), T(.5), T(.5)};
    return {{{v6, v2}, T(1)}, {{v7, v6}, T(1)}, {{v3, v7}, T(1)}, {{v2, v3}, T(1)}};
}

template <class T>
std::vector<vec3<T>> sample_points()
{
    using V = vec3<T>;
    return {V{0, 0, 0},       V{T(.1), T(-.2), T(.3)}, V{T(-.4), T(.4), T(-.4)}, V{T(.49), 0, 0},
            V{T(2), 0, 0},    V{T(-2), T(-2), T(-2)},  V{0, T(1.5), 0},          V{T(1), T(1), T(1)}};
}

template <class T>
void plain_cases(T tol)
{
    auto const verts = cube_vertices<T>();
    auto const x0 = normalize(vec3<T>{T(0.31), T(0.57), T(0.76)});
    auto const pts = sample_points<T>();
    std::span<vec3<T> const> const P{pts};
    for (int open = 0; open < 2; ++open)
    {
        auto const& idx = open ? open_y : closed;
        auto const bnd = open ? open_y_boundary<T>() : std::vector<weighted_segment3<T>>{};
        std::span<weighted_segment3<T> const> const B{bnd};
        NaiveIntersector<T> naive{verts, idx};
        HipMeshIntersector<T> hip{verts, idx};
        HipDispatcher hd;
        SinglethreadDispatcher sd;
        std::vector<T> got(pts.size()), ref(pts.size()), got2(pts.size());
        std::vector<int> gi(pts.size()), ri(pts.size());
        // reference CPU vs drop-in GPU: same call, different dispatcher / intersector
        eval_gwnr_mesh_batch(sd, naive, P, std::span<T>{ref}, x0, B);
        eval_gwnr_mesh_batch(hd, hip, P, std::span<T>{got}, x0, B);
        eval_gwnr_mesh_batch(hd, naive, P, std::span<T>{got2}, x0, B); // HipDispatcher + NaiveIntersector route
        eval_gwnr_mesh_batch_integer(sd, naive, P, std::span<int>{ri}, x0);
        eval_gwnr_mesh_batch_integer(hd, hip, P, std::span<int>{gi}, x0);
        for (size_t i = 0; i < pts.size(); ++i)
        {
            T const base = antipodal_test::eval_gwn_baseline<T>(pts[i], verts, idx);
            CHECK(std::abs(got[i] - ref[i]) < tol);
            CHECK(got2[i] == got[i]);
            CHECK(std::abs(got[i] - base) < tol);
            CHECK(gi[i] == ri[i]);
            // single-query Intersector concept, with x0 and with other directions (unnormalised ones included)
            CHECK(hip.signed_intersection_count(pts[i], x0) == naive.signed_intersection_count(pts[i], x0));
            CHECK(eval_gwnr_mesh_single_integer(pts[i], x0, hip) == ri[i]);
            for (vec3<T> const d : {vec3<T>{T(1), T(0), T(0)}, vec3<T>{T(0), T(-2), T(0)}, vec3<T>{T(0), T(0), T(0.5)},
                                    vec3<T>{T(-0.3), T(0.9), T(-0.2)}, vec3<T>{T(0.7), T(-0.1), T(0.45)}})
                CHECK(hip.signed_intersection_count(pts[i], d) == naive.signed_intersection_count(pts[i], d));
        }
        std::vector<T> fr(pts.size()), frr(pts.size());
        eval_gwnr_mesh_batch_fractional(hd, P, std::span<T>{fr}, x0, B);
        eval_gwnr_mesh_batch_fractional(sd, P, std::span<T>{frr}, x0, B);
        for (size_t i = 0; i < pts.size(); ++i)
            CHECK(std::abs(fr[i] - frr[i]) < tol);
    }
}

template <class T>
void robust_cases(T tol)
{
    auto const verts = cube_vertices<T>();
    auto const pts = sample_points<T>();
    auto const bnd = open_y_boundary<T>();
    // the reference's RobustMeshGwn name (compiled out without Embree) resolves to the GPU evaluator
    RobustMeshGwn<T> const robust{std::span<vec3<T> const>{verts}, std::span<int const>{open_y},
                                  std::span<weighted_segment3<T> const>{bnd}};
    SinglethreadDispatcher sd;
    std::vector<T> out(pts.size());
    eval_gwnr_mesh_batch_robust(sd, robust, std::span<vec3<T> const>{pts}, std::span<T>{out});
    for (size_t i = 0; i < pts.size(); ++i)
    {
        T const base = antipodal_test::eval_gwn_baseline<T>(pts[i], verts, open_y);
        CHECK(std::abs(out[i] - base) < tol);
        CHECK(robust.eval(pts[i]) == out[i]);
        CHECK(robust.fractional(pts[i]) + T(robust.signed_intersection_count(pts[i])) == out[i]);
    }
    CHECK(RobustMeshGwn<T>::axis().x == T(-1));
}

void grid_case()
{
    auto const verts = cube_vertices<float>();
    auto const bnd = build_boundary_segments<float>(std::span<fvec3 const>{verts}, std::span<int const>{open_y});
    HipGridGwn gpu;
    gpu.set_scene(verts, open_y, bnd);
    auto const ex = normalize(fvec3{0.60f, 0.70f, 0.39f});
    auto const ez = normalize(cross(ex, fvec3{-0.40f, 0.20f, 0.90f}));
    auto const ey = cross(ez, ex);
    float const e = 0.15f;
    int const nx = 20, ny = 20, nz = 6;
    fvec3 const dx = ex * e, dy = ey * e, dz = ez * e;
    fvec3 const start = dx * (-0.5f * float(nx - 1)) + dy * (-0.5f * float(ny - 1)) + dz * (-0.5f * float(nz - 1));
    std::vector<float> got(size_t(nx) * ny * nz);
    gpu.eval_grid(start, dx, dy, dz, nx, ny, nz, std::span<float>{got});
    // every cell against the per-triangle solid angle away from the surface
    int checked = 0;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix)
            {
                fvec3 const p = start + dx * float(ix) + dy * float(iy) + dz * float(iz);
                bool near = false;
                for (float c : {p.x, p.y, p.z})
                    near = near || std::abs(std::abs(c) - 0.5f) < 2 * e;
                if (near)
                    continue;
                double const base = antipodal_test::eval_gwn_baseline<double>(
                    dvec3{p.x, p.y, p.z},
                    std::vector<dvec3>{{-.5, -.5, -.5}, {.5, -.5, -.5}, {-.5, .5, -.5}, {.5, .5, -.5}, {-.5, -.5, .5},
                                       {.5, -.5, .5}, {-.5, .5, .5}, {.5, .5, .5}},
                    open_y);
                CHECK(std::abs(got[(size_t(iz) * ny + iy) * nx + ix] - base) < 3e-3);
                ++checked;
            }
    CHECK(checked > 100);
}

int main()
{
    try
    {
        plain_cases<float>(1e-4f);
        plain_cases<double>(1e-9);
        robust_cases<float>(1e-4f);
        robust_cases<double>(1e-5);
        grid_case();
    }
    catch (std::exception const& e)
    {
        std::printf("FAIL exception: %s\n", e.what());
        return 2;
    }
    std::printf("%d checks passed, %d failed\n", g_pass, g_fail);
    return g_fail ? 1 : 0;
}
