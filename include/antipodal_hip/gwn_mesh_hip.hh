#pragma once

// antipodal_hip -- a HIP backend for antipodal's batch mesh-GWN API (MI355X / gfx950), as a plugin to the reference.
//
// Include it after (or instead of) the reference's own headers; it includes <antipodal/gwn_mesh.hh> and
// <antipodal/intersector/intersector.hh> for the reference's types (vec3, weighted_segment3, NaiveIntersector, the
// batch templates) and adds, in namespace antipodal:
//
//   HipDispatcher            a Dispatcher that selects the overloads below. It runs no host lambda: calling
//                            parallel_for on it with any other template is a compile-time error, never a silent
//                            CPU loop.
//   HipMeshIntersector<T>    the Intersector concept (intersector.hh:8-43) on a GPU LBVH; a single query is a
//                            batch of one.
//   eval_gwnr_mesh_batch_fractional / _integer / eval_gwnr_mesh_batch (gwn_mesh.hh:67-83, 127-143, 201-219)
//                            overloads taking HipDispatcher& -- more specialised than the reference's templates,
//                            so existing call sites pick them by changing only the dispatcher. With a
//                            NaiveIntersector<T> the triangle spans it holds are uploaded once per call.
//   HipRobustMeshGwn<T, P>   RobustMeshGwn's shape (gwn_mesh_robust.hh:170-424): ctor(vertices, indices, boundary),
//                            axis(), eval(p), fractional(p), signed_intersection_count(p), non-copyable/movable.
//                            Where the reference's own RobustMeshGwn is compiled out (no Embree), RobustMeshGwn
//                            names this class, so code written against the reference compiles unchanged.
//   eval_gwnr_mesh_batch_robust(Dispatcher&, HipRobustMeshGwn const&, positions, out) (gwn_mesh_robust.hh:441-456)
//   HipGridGwn               OpenGLMeshGwn's shape (gwn_mesh_opengl.hh:426-643): set_scene, eval_grid.
//
// Error behaviour: the reference asserts on span sizes (gwn_mesh.hh:74,134,209) -- so do these; a failing device call
// throws std::runtime_error with the library's message, as the reference's GL evaluator does on GL failures
// (gwn_mesh_opengl.hh:163,668,681). Link with libagwn.so (antipodal_amd/libagwn.so).

#include <antipodal/gwn_mesh.hh>
#include <antipodal/intersector/intersector.hh>
#include <antipodal/math/common.hh>

#include "../agwn.h"

#include <cassert>
#include <span>
#include <stdexcept>
#include <string>
#include <type_traits>

namespace antipodal
{
namespace hip_detail
{
inline void check(int status)
{
    if (status != AGWN_OK)
        throw std::runtime_error(std::string("antipodal_hip: ") + agwn_last_error());
}

template <class T>
inline constexpr bool is_scalar_v = std::is_same_v<T, float> || std::is_same_v<T, double>;

static_assert(sizeof(vec3<float>) == 3 * sizeof(float) && sizeof(vec3<double>) == 3 * sizeof(double),
              "vec3<T> must be three packed T (math/common.hh:32-45)");
static_assert(sizeof(weighted_segment3<float>) == 7 * sizeof(float) &&
                  sizeof(weighted_segment3<double>) == 7 * sizeof(double),
              "weighted_segment3<T> must be seven packed T with the weight last (math/common.hh:50-65)");

template <class T>
T const* data(std::span<vec3<T> const> s)
{
    return reinterpret_cast<T const*>(s.data());
}
template <class T>
T const* data(std::span<weighted_segment3<T> const> s)
{
    return s.empty() ? nullptr : reinterpret_cast<T const*>(s.data());
}

struct mesh_handle
{
    agwn_mesh* h = nullptr;
    mesh_handle() = default;
    template <class T>
    mesh_handle(std::span<vec3<T> const> vertices, std::span<int const> indices)
    {
        assert(indices.size() % 3 == 0);
        static_assert(sizeof(int) == sizeof(int32_t));
        auto const* idx = reinterpret_cast<int32_t const*>(indices.data());
        if constexpr (std::is_same_v<T, float>)
            check(agwn_mesh_create_f32(data(vertices), int64_t(vertices.size()), idx, int64_t(indices.size()), &h));
        else
            check(agwn_mesh_create_f64(data(vertices), int64_t(vertices.size()), idx, int64_t(indices.size()), &h));
    }
    mesh_handle(mesh_handle const&) = delete;
    mesh_handle& operator=(mesh_handle const&) = delete;
    mesh_handle(mesh_handle&& o) noexcept : h(o.h) { o.h = nullptr; }
    ~mesh_handle()
    {
        if (h)
            agwn_mesh_destroy(h);
    }
};
} // namespace hip_detail

// ---------------------------------------------------------------------------------------------------------------------
struct HipDispatcher
{
    agwn_stream stream = nullptr; // used by the *_dev entry points; the span overloads below are synchronous

    template <class F>
    void parallel_for(int, int, F&&)
    {
        static_assert(sizeof(F) == 0, "HipDispatcher runs only the batch GWN entry points of antipodal_hip; it "
                                      "executes no host lambda (use a CPU dispatcher for other templates)");
    }
};

// ---------------------------------------------------------------------------------------------------------------------
template <class T>
class HipMeshIntersector
{
    static_assert(hip_detail::is_scalar_v<T>, "float or double");

public:
    HipMeshIntersector(std::span<vec3<T> const> vertices, std::span<int const> indices) : m_mesh(vertices, indices) {}

    // Intersector concept: sum of sign(N . dir) over the triangles the ray (p, dir) hits at t >= 0, with
    // NaiveIntersector<T>'s arithmetic (a batch of one).
    [[nodiscard]] int signed_intersection_count(vec3<T> p, vec3<T> dir) const
    {
        int32_t out = 0;
        T const d[3] = {dir.x, dir.y, dir.z};
        T const q[3] = {p.x, p.y, p.z};
        if constexpr (std::is_same_v<T, float>)
            hip_detail::check(agwn_int_batch_f32(m_mesh.h, q, 1, d, &out));
        else
            hip_detail::check(agwn_int_batch_f64(m_mesh.h, q, 1, d, &out));
        return out;
    }

    [[nodiscard]] agwn_mesh const* handle() const { return m_mesh.h; }

private:
    hip_detail::mesh_handle m_mesh;
};

// --- gwn_mesh.hh overloads --------------------------------------------------------------------------------------------
template <class T>
void eval_gwnr_mesh_batch_fractional(HipDispatcher&,
                                     std::span<vec3<T> const> positions,
                                     std::span<T> out_wnrs,
                                     vec3<T> x0,
                                     std::span<weighted_segment3<T> const> boundary)
{
    assert(positions.size() == out_wnrs.size());
    T const x[3] = {x0.x, x0.y, x0.z};
    auto const n = int64_t(positions.size());
    if constexpr (std::is_same_v<T, float>)
        hip_detail::check(agwn_frac_batch_f32(hip_detail::data(positions), n, hip_detail::data(boundary),
                                              int64_t(boundary.size()), x, out_wnrs.data()));
    else
        hip_detail::check(agwn_frac_batch_f64(hip_detail::data(positions), n, hip_detail::data(boundary),
                                              int64_t(boundary.size()), x, out_wnrs.data()));
}

template <class T>
void eval_gwnr_mesh_batch_integer(HipDispatcher&,
                                  HipMeshIntersector<T> const& intersector,
                                  std::span<vec3<T> const> positions,
                                  std::span<int> out_counts,
                                  vec3<T> x0)
{
    assert(positions.size() == out_counts.size());
    T const x[3] = {x0.x, x0.y, x0.z};
    auto* out = reinterpret_cast<int32_t*>(out_counts.data());
    auto const n = int64_t(positions.size());
    if constexpr (std::is_same_v<T, float>)
        hip_detail::check(agwn_int_batch_f32(intersector.handle(), hip_detail::data(positions), n, x, out));
    else
        hip_detail::check(agwn_int_batch_f64(intersector.handle(), hip_detail::data(positions), n, x, out));
}

template <class T>
void eval_gwnr_mesh_batch(HipDispatcher&,
                          HipMeshIntersector<T> const& intersector,
                          std::span<vec3<T> const> positions,
                          std::span<T> out_wnrs,
                          vec3<T> x0,
                          std::span<weighted_segment3<T> const> boundary)
{
    assert(positions.size() == out_wnrs.size());
    T const x[3] = {x0.x, x0.y, x0.z};
    auto const n = int64_t(positions.size());
    if constexpr (std::is_same_v<T, float>)
        hip_detail::check(agwn_full_batch_f32(intersector.handle(), hip_detail::data(positions), n, x,
                                              hip_detail::data(boundary), int64_t(boundary.size()), out_wnrs.data()));
    else
        hip_detail::check(agwn_full_batch_f64(intersector.handle(), hip_detail::data(positions), n, x,
                                              hip_detail::data(boundary), int64_t(boundary.size()), out_wnrs.data()));
}

// HipDispatcher + the reference's NaiveIntersector<T>: its public spans (intersector.hh:58-59) are uploaded into a
// transient LBVH; the counts are NaiveIntersector's, bit for bit.
template <class T>
void eval_gwnr_mesh_batch_integer(HipDispatcher& d,
                                  NaiveIntersector<T> const& naive,
                                  std::span<vec3<T> const> positions,
                                  std::span<int> out_counts,
                                  vec3<T> x0)
{
    HipMeshIntersector<T> const isect{naive.vertices, naive.indices};
    eval_gwnr_mesh_batch_integer(d, isect, positions, out_counts, x0);
}

template <class T>
void eval_gwnr_mesh_batch(HipDispatcher& d,
                          NaiveIntersector<T> const& naive,
                          std::span<vec3<T> const> positions,
                          std::span<T> out_wnrs,
                          vec3<T> x0,
                          std::span<weighted_segment3<T> const> boundary)
{
    HipMeshIntersector<T> const isect{naive.vertices, naive.indices};
    eval_gwnr_mesh_batch(d, isect, positions, out_wnrs, x0, boundary);
}

// --- gwn_mesh_robust.hh -----------------------------------------------------------------------------------------------
namespace hip_robust
{
// The reference's policy tags live in robust_predicates.hh; these carry the same names and map onto the C ABI.
struct int64_grid
{
    static constexpr int code = AGWN_INT64_GRID;
};
struct int128_grid
{
    static constexpr int code = AGWN_INT128_GRID;
};
struct exact_float
{
    static constexpr int code = AGWN_EXACT_FLOAT;
};
template <class T>
using default_precision = std::conditional_t<std::is_same_v<T, float>, int64_grid, exact_float>;

template <class P>
struct code_of
{
    static constexpr int value = P::code;
};
} // namespace hip_robust

template <class T, class Precision = hip_robust::default_precision<T>>
class HipRobustMeshGwn
{
    static_assert(hip_detail::is_scalar_v<T>, "float or double");

public:
    using precision = Precision;

    HipRobustMeshGwn(std::span<vec3<T> const> vertices,
                     std::span<int const> indices,
                     std::span<weighted_segment3<T> const> boundary)
    {
        assert(indices.size() % 3 == 0);
        hip_detail::check(agwn_robust_create(int(sizeof(T)), hip_robust::code_of<Precision>::value,
                                             vertices.data(), int64_t(vertices.size()),
                                             reinterpret_cast<int32_t const*>(indices.data()),
                                             int64_t(indices.size()), hip_detail::data(boundary),
                                             int64_t(boundary.size()), &m_h));
    }
    ~HipRobustMeshGwn()
    {
        if (m_h)
            agwn_robust_destroy(m_h);
    }
    HipRobustMeshGwn(HipRobustMeshGwn const&) = delete;
    HipRobustMeshGwn& operator=(HipRobustMeshGwn const&) = delete;
    HipRobustMeshGwn(HipRobustMeshGwn&&) = delete;
    HipRobustMeshGwn& operator=(HipRobustMeshGwn&&) = delete;

    [[nodiscard]] static constexpr vec3<T> axis() { return {T(-1), T(0), T(0)}; }

    [[nodiscard]] int signed_intersection_count(vec3<T> p) const
    {
        int32_t out = 0;
        hip_detail::check(agwn_robust_count_batch(m_h, &p, 1, &out));
        return out;
    }
    [[nodiscard]] T fractional(vec3<T> p) const
    {
        T out{};
        hip_detail::check(agwn_robust_fractional_batch(m_h, &p, 1, &out));
        return out;
    }
    [[nodiscard]] T eval(vec3<T> p) const
    {
        T out{};
        hip_detail::check(agwn_robust_eval_batch(m_h, &p, 1, &out));
        return out;
    }

    [[nodiscard]] agwn_robust const* handle() const { return m_h; }

private:
    agwn_robust* m_h = nullptr;
};

// The batch runs on the GPU whatever dispatcher is passed (the evaluator is a device object).
template <class Dispatcher, class T, class Precision>
void eval_gwnr_mesh_batch_robust(Dispatcher&,
                                 HipRobustMeshGwn<T, Precision> const& robust,
                                 std::span<vec3<T> const> positions,
                                 std::span<T> out_wnrs)
{
    assert(positions.size() == out_wnrs.size());
    hip_detail::check(agwn_robust_eval_batch(robust.handle(), positions.data(), int64_t(positions.size()),
                                             out_wnrs.data()));
}

#if !(defined(ANTIPODAL_HAS_EMBREE) && ANTIPODAL_HAS_EMBREE)
// Without Embree the reference compiles RobustMeshGwn out (gwn_mesh_robust.hh:29); the GPU evaluator takes its name.
template <class T, class Precision = hip_robust::default_precision<T>>
using RobustMeshGwn = HipRobustMeshGwn<T, Precision>;
#endif

// --- gwn_mesh_opengl.hh -----------------------------------------------------------------------------------------------
class HipGridGwn
{
public:
    HipGridGwn() { hip_detail::check(agwn_grid_create(&m_h)); }
    ~HipGridGwn()
    {
        if (m_h)
            agwn_grid_destroy(m_h);
    }
    HipGridGwn(HipGridGwn const&) = delete;
    HipGridGwn& operator=(HipGridGwn const&) = delete;
    HipGridGwn(HipGridGwn&&) = delete;
    HipGridGwn& operator=(HipGridGwn&&) = delete;

    void set_scene(std::span<fvec3 const> vertices,
                   std::span<int const> indices,
                   std::span<weighted_fsegment3 const> boundary)
    {
        assert(indices.size() % 3 == 0);
        hip_detail::check(agwn_grid_set_scene(m_h, hip_detail::data(vertices), int64_t(vertices.size()),
                                              reinterpret_cast<int32_t const*>(indices.data()),
                                              int64_t(indices.size()), hip_detail::data(boundary),
                                              int64_t(boundary.size())));
    }

    void eval_grid(fvec3 start, fvec3 dx, fvec3 dy, fvec3 dz, int nx, int ny, int nz, std::span<float> out)
    {
        assert(nx > 0 && ny > 0 && nz > 0);
        assert(out.size() == std::size_t(nx) * std::size_t(ny) * std::size_t(nz));
        hip_detail::check(agwn_grid_eval(m_h, &start.x, &dx.x, &dy.x, &dz.x, nx, ny, nz, out.data()));
    }

private:
    agwn_grid* m_h = nullptr;
};
} // namespace antipodal
