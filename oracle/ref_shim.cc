// =====================================================================================================================
// TEST INFRASTRUCTURE ONLY -- a C-ABI shim over the REFERENCE's own headers, compiled from where they lie
// (REF_SRC = <reference>/src and <reference>/test, see oracle/Makefile) into oracle/_ref/libref.so.
//
// It gives the tests and bench.py's cpu_baseline leg the reference's actual code for the parts of the batch-GWN path
// that build in this image without any missing dependency:
//   gwn_mesh.hh (eval_gwnr_mesh_batch_fractional / _integer / eval_gwnr_mesh_batch), intersector.hh
//   (NaiveIntersector), dispatcher.hh / dispatcher_openmp.hh / dispatcher_threadpool.hh, math/common.hh
//   (build_boundary_segments, signed_spherical_tri_area_half_unorm), math/robust_predicates.hh, test/gwn_baseline.hh.
// RobustMeshGwn / EmbreeIntersector need Embree 4, which this image lacks: they are not built (no stand-in headers).
// Nothing here is product code; the product never loads this library.
// =====================================================================================================================

#include <antipodal/dispatcher/dispatcher.hh>
#include <antipodal/dispatcher/dispatcher_openmp.hh>
#include <antipodal/dispatcher/dispatcher_threadpool.hh>
#include <antipodal/gwn_mesh.hh>
#include <antipodal/intersector/intersector.hh>
#include <antipodal/math/common.hh>
#include <antipodal/math/robust_predicates.hh>

#include "gwn_baseline.hh"

#include <cstdint>
#include <span>
#include <thread>
#include <vector>

using namespace antipodal;

namespace
{
// mode 0: SinglethreadDispatcher, 1: OpenMPDispatcher, 2: PoolDispatcher(hardware_concurrency)
template <class F>
void with_dispatcher(int mode, F&& f)
{
    if (mode == 1)
    {
        OpenMPDispatcher d;
        f(d);
    }
    else if (mode == 2)
    {
        PoolDispatcher d{int(std::max(1u, std::thread::hardware_concurrency()))};
        f(d);
    }
    else
    {
        SinglethreadDispatcher d;
        f(d);
    }
}

template <class T>
void frac(T const* pos, int64_t n, T const* seg7, int64_t nseg, T const* x0, T* out, int mode)
{
    std::span<vec3<T> const> const P{reinterpret_cast<vec3<T> const*>(pos), size_t(n)};
    std::span<weighted_segment3<T> const> const S{reinterpret_cast<weighted_segment3<T> const*>(seg7), size_t(nseg)};
    vec3<T> const d{x0[0], x0[1], x0[2]};
    with_dispatcher(mode, [&](auto& disp) { eval_gwnr_mesh_batch_fractional(disp, P, std::span<T>{out, size_t(n)}, d, S); });
}

template <class T>
void count(T const* verts, int64_t nv, int32_t const* idx, int64_t nidx, T const* pos, int64_t n, T const* dir,
           int32_t* out, int mode)
{
    NaiveIntersector<T> const isect{std::span<vec3<T> const>{reinterpret_cast<vec3<T> const*>(verts), size_t(nv)},
                                    std::span<int const>{idx, size_t(nidx)}};
    std::span<vec3<T> const> const P{reinterpret_cast<vec3<T> const*>(pos), size_t(n)};
    vec3<T> const d{dir[0], dir[1], dir[2]};
    with_dispatcher(mode,
                    [&](auto& disp) { eval_gwnr_mesh_batch_integer(disp, isect, P, std::span<int>{out, size_t(n)}, d); });
}

template <class T>
void full(T const* verts, int64_t nv, int32_t const* idx, int64_t nidx, T const* pos, int64_t n, T const* x0,
          T const* seg7, int64_t nseg, T* out, int mode)
{
    NaiveIntersector<T> const isect{std::span<vec3<T> const>{reinterpret_cast<vec3<T> const*>(verts), size_t(nv)},
                                    std::span<int const>{idx, size_t(nidx)}};
    std::span<vec3<T> const> const P{reinterpret_cast<vec3<T> const*>(pos), size_t(n)};
    std::span<weighted_segment3<T> const> const S{reinterpret_cast<weighted_segment3<T> const*>(seg7), size_t(nseg)};
    vec3<T> const d{x0[0], x0[1], x0[2]};
    with_dispatcher(mode, [&](auto& disp) { eval_gwnr_mesh_batch(disp, isect, P, std::span<T>{out, size_t(n)}, d, S); });
}

template <class T>
void baseline(T const* verts, int64_t nv, int32_t const* idx, int64_t nidx, T const* pos, int64_t n, T* out)
{
    std::span<vec3<T> const> const V{reinterpret_cast<vec3<T> const*>(verts), size_t(nv)};
    std::span<int const> const I{idx, size_t(nidx)};
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < n; ++i)
        out[i] = antipodal_test::eval_gwn_baseline<T>(vec3<T>{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}, V, I);
}

template <class T>
int64_t boundary(T const* verts, int64_t nv, int32_t const* idx, int64_t nidx, T* out7, int64_t cap)
{
    auto const segs = build_boundary_segments<T>(
        std::span<vec3<T> const>{reinterpret_cast<vec3<T> const*>(verts), size_t(nv)},
        std::span<int const>{idx, size_t(nidx)});
    for (size_t i = 0; i < segs.size() && int64_t(i) < cap; ++i)
    {
        auto const& s = segs[i];
        T* o = out7 + 7 * i;
        o[0] = s.segment.pos0.x;
        o[1] = s.segment.pos0.y;
        o[2] = s.segment.pos0.z;
        o[3] = s.segment.pos1.x;
        o[4] = s.segment.pos1.y;
        o[5] = s.segment.pos1.z;
        o[6] = s.weight;
    }
    return int64_t(segs.size());
}

template <class C>
void preds(C const* cases, int64_t n, int32_t* es, int32_t* ns, int32_t* ins, double* w)
{
    for (int64_t i = 0; i < n; ++i)
    {
        C const* c = cases + 12 * i;
        vec3<C> const q{c[0], c[1], c[2]}, a{c[3], c[4], c[5]}, b{c[6], c[7], c[8]}, t{c[9], c[10], c[11]};
        es[i] = robust::edge_sign(q, a, b);
        auto const nm = robust::atan2_num(q, a, b);
        ns[i] = nm > 0 ? 1 : (nm < 0 ? -1 : 0);
        auto const h = robust::project_into_tri(q, a, b, t);
        ins[i] = h.sign;
        w[3 * i + 0] = h.w0;
        w[3 * i + 1] = h.w1;
        w[3 * i + 2] = h.w2;
    }
}
} // namespace

extern "C" {
float ref_pair_half_area_f32(float const* x1, float const* v0, float const* v1)
{
    return signed_spherical_tri_area_half_unorm<float>({x1[0], x1[1], x1[2]}, {v0[0], v0[1], v0[2]}, {v1[0], v1[1], v1[2]});
}
double ref_pair_half_area_f64(double const* x1, double const* v0, double const* v1)
{
    return signed_spherical_tri_area_half_unorm<double>({x1[0], x1[1], x1[2]}, {v0[0], v0[1], v0[2]},
                                                        {v1[0], v1[1], v1[2]});
}
void ref_frac_batch_f32(float const* p, int64_t n, float const* s, int64_t ns, float const* x0, float* o, int mode)
{
    frac(p, n, s, ns, x0, o, mode);
}
void ref_frac_batch_f64(double const* p, int64_t n, double const* s, int64_t ns, double const* x0, double* o, int mode)
{
    frac(p, n, s, ns, x0, o, mode);
}
void ref_naive_count_f32(float const* v, int64_t nv, int32_t const* i, int64_t ni, float const* p, int64_t n,
                         float const* d, int32_t* o, int mode)
{
    count(v, nv, i, ni, p, n, d, o, mode);
}
void ref_naive_count_f64(double const* v, int64_t nv, int32_t const* i, int64_t ni, double const* p, int64_t n,
                         double const* d, int32_t* o, int mode)
{
    count(v, nv, i, ni, p, n, d, o, mode);
}
void ref_full_batch_f32(float const* v, int64_t nv, int32_t const* i, int64_t ni, float const* p, int64_t n,
                        float const* x0, float const* s, int64_t ns, float* o, int mode)
{
    full(v, nv, i, ni, p, n, x0, s, ns, o, mode);
}
void ref_full_batch_f64(double const* v, int64_t nv, int32_t const* i, int64_t ni, double const* p, int64_t n,
                        double const* x0, double const* s, int64_t ns, double* o, int mode)
{
    full(v, nv, i, ni, p, n, x0, s, ns, o, mode);
}
void ref_gwn_baseline_f32(float const* v, int64_t nv, int32_t const* i, int64_t ni, float const* p, int64_t n, float* o)
{
    baseline(v, nv, i, ni, p, n, o);
}
void ref_gwn_baseline_f64(double const* v, int64_t nv, int32_t const* i, int64_t ni, double const* p, int64_t n,
                          double* o)
{
    baseline(v, nv, i, ni, p, n, o);
}
int64_t ref_boundary_f32(float const* v, int64_t nv, int32_t const* i, int64_t ni, float* o, int64_t cap)
{
    return boundary(v, nv, i, ni, o, cap);
}
int64_t ref_boundary_f64(double const* v, int64_t nv, int32_t const* i, int64_t ni, double* o, int64_t cap)
{
    return boundary(v, nv, i, ni, o, cap);
}
void ref_pred_batch_i32(int32_t const* c, int64_t n, int32_t* es, int32_t* ns, int32_t* ins, double* w)
{
    preds(c, n, es, ns, ins, w);
}
void ref_pred_batch_i64(int64_t const* c, int64_t n, int32_t* es, int32_t* ns, int32_t* ins, double* w)
{
    preds(c, n, es, ns, ins, w);
}
void ref_pred_batch_f64(double const* c, int64_t n, int32_t* es, int32_t* ns, int32_t* ins, double* w)
{
    preds(c, n, es, ns, ins, w);
}
int ref_omp_max_threads(void)
{
#if defined(_OPENMP)
    return omp_get_max_threads();
#else
    return 1;
#endif
}
}
