// ref_shim.cpp -- a C entry layer over the compiled reference (fockl/FMMA),
// built by oracle/Makefile into oracle/_ref/ only.  TEST INFRASTRUCTURE ONLY:
// it is linked with the reference's own src/*.cpp (compiled where they lie)
// and lets tests/ and bench.py's cpu_baseline leg call the real reference
// `fmm`/`exact` solvers (src/fmm.cpp:15-144, src/fmma.cpp:79-94) on plain
// arrays.  Nothing here is shipped or measured as the product.
#include "fmma/fmma.hpp"

#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Kernels with the numbering of include/fmma_hip.h.  Each is written like the
// reference's default kernel (include/fmma/fmma.hpp:55-63).
template <typename T, std::size_t D>
std::function<T(const std::array<T, D>&, const std::array<T, D>&)> make_kernel(int id, const double* prm)
{
    using A = std::array<T, D>;
    double p0 = prm ? prm[0] : 0.0;
    switch (id) {
    case 0:
        return [](const A& x, const A& y) {
            double r2 = 0.0;
            for (std::size_t d = 0; d < D; ++d) { double t = x[d] - y[d]; r2 += t * t; }
            return (T)(1.0 / std::sqrt(r2));
        };
    case 1:
        return [](const A& x, const A& y) {
            double r2 = 0.0;
            for (std::size_t d = 0; d < D; ++d) { double t = x[d] - y[d]; r2 += t * t; }
            return (T)(r2 == 0.0 ? 0.0 : 1.0 / std::sqrt(r2));
        };
    case 2:
        return [](const A& x, const A& y) { return (T)(1.0 / (x[0] - y[0])); };
    case 3:
        return [](const A& x, const A& y) {
            double r2 = 0.0;
            for (std::size_t d = 0; d < D; ++d) { double t = x[d] - y[d]; r2 += t * t; }
            return (T)(r2 == 0.0 ? 0.0 : 0.5 * std::log(r2));
        };
    case 4:
        return [p0](const A& x, const A& y) {
            double r2 = 0.0;
            for (std::size_t d = 0; d < D; ++d) { double t = x[d] - y[d]; r2 += t * t; }
            return (T)std::exp(-r2 / (2.0 * p0 * p0));
        };
    case 5:
        return [p0](const A& x, const A& y) {
            double r2 = 0.0;
            for (std::size_t d = 0; d < D; ++d) { double t = x[d] - y[d]; r2 += t * t; }
            if (r2 == 0.0) return (T)0.0;
            double r = std::sqrt(r2);
            return (T)(std::exp(-p0 * r) / r);
        };
    case 6:
        return [](const A& x, const A& y) {
            double r2 = 0.0, s = 0.0;
            for (std::size_t d = 0; d < D; ++d) { double t = x[d] - y[d]; r2 += t * t; s += t; }
            return (T)(s / r2);
        };
    case 8:  // the Python tutorials' kernel (tutorials/tutorial0{1..5}/*.py)
        return [](const A& x, const A& y) {
            double r2 = 0.0;
            for (std::size_t d = 0; d < D; ++d) { double t = x[d] - y[d]; r2 += t * t; }
            return (T)(1.0 / r2);
        };
    default:
        return [](const A& x, const A& y) {
            double r2 = 0.0, s2 = 0.0;
            for (std::size_t d = 0; d < D; ++d) {
                double t = x[d] - y[d], u = x[d] + y[d];
                r2 += t * t; s2 += u * u;
            }
            return (T)(s2 / r2);
        };
    }
}

template <typename T, std::size_t D>
std::vector<std::array<T, D>> to_aos(long n, const T* p)
{
    std::vector<std::array<T, D>> v(n);
    for (long i = 0; i < n; ++i)
        for (std::size_t d = 0; d < D; ++d) v[i][d] = p[i * D + d];
    return v;
}

int g_nrn_N = -1;  // set_nrn_N of the next solves (<= 0: the reference's auto rule)

template <typename T, std::size_t D>
int run(const char* solver, int poly_ord, int depth, int trans_sym, int kid, const double* prm,
        long nt, const T* t, long ns, const T* w, const T* s, T* ans)
{
    fmma::FMMA<T, D> f;
    f.set_fn(make_kernel<T, D>(kid, prm));
    f.set_solver_type(solver);
    f.set_poly_ord(poly_ord);
    f.set_Depth(depth);
    f.set_trans_sym_flag(trans_sym);
    f.set_nrn_N(g_nrn_N);
    auto tv = to_aos<T, D>(nt, t), sv = to_aos<T, D>(ns, s);
    std::vector<T> wv(w, w + ns), av(ans, ans + nt);
    f.solve(tv, wv, sv, av);
    std::memcpy(ans, av.data(), sizeof(T) * nt);
    return 0;
}

// Reaches the protected stage operators the way the reference's own
// operator test does (test/src/fmma/m2m.cpp:13-16).
template <typename T, std::size_t D>
struct Stages : public fmma::FMMA<T, D> {
    int levels(int poly_ord, int depth, int trans_sym, int kid, const double* prm,
               long nt, const T* t, long ns, const T* w, const T* s, T* Wm_out, T* Wl_out)
    {
        this->set_fn(make_kernel<T, D>(kid, prm));
        this->set_poly_ord(poly_ord);
        this->set_trans_sym_flag(trans_sym);
        auto tv = to_aos<T, D>(nt, t), sv = to_aos<T, D>(ns, s);
        std::vector<T> wv(w, w + ns);
        std::array<T, D> origin;
        T Len;
        this->get_origin_and_length(tv, sv, origin, Len);
        std::vector<std::vector<std::vector<T>>> Wm(depth), Wl(depth);
        std::vector<std::vector<std::size_t>> lists;
        std::vector<std::array<T, D>> nodes;
        std::size_t N = std::size_t(1) << (depth - 1);
        this->P2M(wv, sv, N, origin, Len / N, lists, Wm[depth - 1], nodes);
        for (int k = 0; k + 3 < depth; ++k) {
            this->M2M(N, nodes, Wm[depth - k - 1], Wm[depth - k - 2]);
            N /= 2;
        }
        for (int l = 2; l < depth; ++l) {
            Wl[l].assign(Wm[l].size(), std::vector<T>(nodes.size(), 0.0));
            this->M2L(std::size_t(1) << l, origin, Len, nodes, Wm[l], Wl[l]);
        }
        for (int l = 2; l + 1 < depth; ++l) this->L2L(std::size_t(2) << l, nodes, Wl[l], Wl[l + 1]);
        std::size_t off = 0, P = nodes.size();
        for (int l = 0; l < depth; ++l) {
            std::size_t nb = 1;
            for (std::size_t d = 0; d < D; ++d) nb *= (std::size_t(1) << l);
            for (std::size_t b = 0; b < nb; ++b)
                for (std::size_t k = 0; k < P; ++k) {
                    Wm_out[off + b * P + k] = b < Wm[l].size() ? Wm[l][b][k] : 0;
                    Wl_out[off + b * P + k] = b < Wl[l].size() ? Wl[l][b][k] : 0;
                }
            off += nb * P;
        }
        return 0;
    }

    // The reference's own leaf box of each of `pts` (row-major index): P2M's
    // binning (src/fmm_core.cpp:255-267) applied to pts as sources, with the
    // cube of targets and sources (get_origin_and_length, src/fmm_core.cpp:160-173).
    int leaf_boxes(int depth, long nt, const T* t, long ns, const T* s, long np, const T* pts, long* out)
    {
        this->set_poly_ord(1);
        auto tv = to_aos<T, D>(nt, t), sv = to_aos<T, D>(ns, s), pv = to_aos<T, D>(np, pts);
        std::array<T, D> origin;
        T Len;
        this->get_origin_and_length(tv, sv, origin, Len);
        std::vector<T> wv(np, T(0));
        std::vector<std::vector<T>> Wm;
        std::vector<std::vector<std::size_t>> lists;
        std::vector<std::array<T, D>> nodes;
        const std::size_t N = std::size_t(1) << (depth - 1);
        this->P2M(wv, pv, N, origin, Len / N, lists, Wm, nodes);
        for (long i = 0; i < np; ++i) out[i] = -1;
        for (std::size_t b = 0; b < lists.size(); ++b)
            for (std::size_t i : lists[b]) out[i] = (long)b;
        return 0;
    }
};

template <typename T>
int dispatch(int dim, const char* solver, int poly_ord, int depth, int trans_sym, int kid, const double* prm,
             long nt, const T* t, long ns, const T* w, const T* s, T* ans)
{
    switch (dim) {
    case 1: return run<T, 1>(solver, poly_ord, depth, trans_sym, kid, prm, nt, t, ns, w, s, ans);
    case 2: return run<T, 2>(solver, poly_ord, depth, trans_sym, kid, prm, nt, t, ns, w, s, ans);
    case 3: return run<T, 3>(solver, poly_ord, depth, trans_sym, kid, prm, nt, t, ns, w, s, ans);
    }
    return -1;
}

}  // namespace

extern "C" {

// solver: "exact" or "fmm" (any name `solve` accepts).  `ans` is in/out.
int ref_solve_f64(int dim, const char* solver, int poly_ord, int depth, int trans_sym, int kid,
                  const double* prm, long nt, const double* t, long ns, const double* w,
                  const double* s, double* ans)
{
    return dispatch<double>(dim, solver, poly_ord, depth, trans_sym, kid, prm, nt, t, ns, w, s, ans);
}

int ref_solve_f32(int dim, const char* solver, int poly_ord, int depth, int trans_sym, int kid,
                  const double* prm, long nt, const float* t, long ns, const float* w,
                  const float* s, float* ans)
{
    return dispatch<float>(dim, solver, poly_ord, depth, trans_sym, kid, prm, nt, t, ns, w, s, ans);
}

int ref_levels_f64(int dim, int poly_ord, int depth, int trans_sym, int kid, const double* prm,
                   long nt, const double* t, long ns, const double* w, const double* s,
                   double* Wm, double* Wl)
{
    switch (dim) {
    case 1: return Stages<double, 1>().levels(poly_ord, depth, trans_sym, kid, prm, nt, t, ns, w, s, Wm, Wl);
    case 2: return Stages<double, 2>().levels(poly_ord, depth, trans_sym, kid, prm, nt, t, ns, w, s, Wm, Wl);
    case 3: return Stages<double, 3>().levels(poly_ord, depth, trans_sym, kid, prm, nt, t, ns, w, s, Wm, Wl);
    }
    return -1;
}

int ref_leaf_boxes_f64(int dim, int depth, long nt, const double* t, long ns, const double* s, long np,
                       const double* pts, long* out)
{
    switch (dim) {
    case 1: return Stages<double, 1>().leaf_boxes(depth, nt, t, ns, s, np, pts, out);
    case 2: return Stages<double, 2>().leaf_boxes(depth, nt, t, ns, s, np, pts, out);
    case 3: return Stages<double, 3>().leaf_boxes(depth, nt, t, ns, s, np, pts, out);
    }
    return -1;
}

int ref_leaf_boxes_f32(int dim, int depth, long nt, const float* t, long ns, const float* s, long np,
                       const float* pts, long* out)
{
    switch (dim) {
    case 1: return Stages<float, 1>().leaf_boxes(depth, nt, t, ns, s, np, pts, out);
    case 2: return Stages<float, 2>().leaf_boxes(depth, nt, t, ns, s, np, pts, out);
    case 3: return Stages<float, 3>().leaf_boxes(depth, nt, t, ns, s, np, pts, out);
    }
    return -1;
}

int ref_set_nrn_N(int n)
{
    g_nrn_N = n;
    return 0;
}

int ref_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

}  // extern "C"
