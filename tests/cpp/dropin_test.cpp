// Drop-in check of fmma::FMMA<TYPE,DIM> (include/fmma/fmma.hpp) used the way
// the reference's own tests and tutorials use it (test/src/fmma/fmm.cpp,
// tutorials/tutorial0*/): std::vector<std::array> inputs, set_fn with a
// functor, set_solver_type, solve.  Built with hipcc; run on the GPU by
// tests/test_gpu_dropin.py.  Prints one "name value" line per check and exits
// non-zero on a failed check.
#include <fmma/fmma.hpp>

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <vector>

static uint64_t sm_state = 0;
static double uni()  // SplitMix64 uniform in [0, 1)
{
    uint64_t z = (sm_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (z >> 11) * (1.0 / 9007199254740992.0);
}

template <typename T, std::size_t D>
static std::vector<std::array<T, D>> pts(size_t n)
{
    std::vector<std::array<T, D>> v(n);
    for (auto &p : v)
        for (auto &c : p) c = (T)uni();
    return v;
}

template <typename T>
static double rel(const std::vector<T> &a, const std::vector<T> &b)
{
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += ((double)a[i] - b[i]) * ((double)a[i] - b[i]);
        den += (double)b[i] * b[i];
    }
    return std::sqrt(num / den);
}

// f(x - y) = d_0 exp(-50 |d|^2): odd in d = x - y and finite at d = 0.
struct OddKernel {
    static constexpr int symmetry = -1;
    __host__ __device__ double operator()(const std::array<double, 3> &d) const
    {
        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        return d[0] * std::exp(-50.0 * r2);
    }
};

static int fails = 0;
static void check(const char *name, double v, double tol)
{
    printf("%s %.3e\n", name, v);
    if (!(v <= tol)) {
        printf("FAIL %s > %.1e\n", name, tol);
        ++fails;
    }
}

int main()
{
    sm_state = 7;
    {  // 3-D, one-argument f(x - y) lambda vs the built-in 1/r kernel
        const size_t n = 20000;
        auto x = pts<double, 3>(n);
        std::vector<double> w(n);
        for (auto &v : w) v = uni();
        fmma::FMMA<double, 3> f;
        f.set_fn([] __host__ __device__(const std::array<double, 3> &d) {
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            return r2 == 0.0 ? 0.0 : 1.0 / std::sqrt(r2);
        });
        f.set_solver_type("fmm");
        f.set_poly_ord(3);
        f.set_Depth(4);
        std::vector<double> a, b, e;
        f.solve(x, w, x, a);
        fmma::FMMA<double, 3> g;
        g.set_kernel(FMMA_KERNEL_INV_R_REG);
        g.set_solver_type("fmm");
        g.set_poly_ord(3);
        g.set_Depth(4);
        g.set_trans_sym_flag(1);
        g.solve(x, w, x, b);
        f.set_solver_type("exact");
        f.solve(x, w, x, e);
        check("lambda_vs_builtin_3d", rel(a, b), 1e-12);
        check("fmm_vs_exact_3d", rel(a, e), 1e-3);
        check("depth_kept", std::abs(f.get_Depth() - 4), 0);
    }
    {  // 2-D two-argument f(x, y), not translation invariant (test/src/fmma/fmm.cpp:151-161)
        const size_t n = 3000;
        auto t = pts<double, 2>(n), s = pts<double, 2>(n);
        std::vector<double> w(n);
        for (auto &v : w) v = uni();
        fmma::FMMA<double, 2> f;
        f.set_fn([] __host__ __device__(const std::array<double, 2> &x, const std::array<double, 2> &y) {
            double l2 = 0, s2 = 0;
            for (int d = 0; d < 2; ++d) {
                l2 += (x[d] - y[d]) * (x[d] - y[d]);
                s2 += (x[d] + y[d]) * (x[d] + y[d]);
            }
            return s2 / l2;
        });
        f.set_solver_type("fmm");
        f.set_poly_ord(5);
        f.set_Depth(4);
        std::vector<double> a, e, b;
        f.fmm(t, w, s, a);
        f.exact(t, w, s, e);
        fmma::FMMA<double, 2> g;
        g.set_kernel(FMMA_KERNEL_SUM2_R2);
        g.set_poly_ord(5);
        g.set_Depth(4);
        g.fmm(t, w, s, b);
        check("general_fmm_vs_exact_2d", rel(a, e), 1e-3);
        check("general_lambda_vs_builtin_2d", rel(a, b), 1e-12);
    }
    {  // 1-D float, unweighted, the reference's README usage (README.md:75-78)
        const size_t n = 4096;
        auto t = pts<float, 1>(n), s = pts<float, 1>(n);
        fmma::FMMA<float, 1> f;
        f.set_fn([] __host__ __device__(const std::array<float, 1> &x, const std::array<float, 1> &y) {
            return 1.0f / (x[0] - y[0]);
        });
        f.set_trans_sym_flag(1);
        f.set_solver_type("fmm");
        f.set_poly_ord(7);
        std::vector<float> a, e;
        f.solve(t, s, a);
        f.set_solver_type("exact");
        f.solve(t, s, e);
        check("float_1d_fmm_vs_exact", rel(a, e), 1e-3);
    }
    {  // an odd user kernel that declares symmetry -1: the mutual P2P form
       // (targets = sources) against the direct form and the exact solver
        const size_t n = 20000;
        auto x = pts<double, 3>(n);
        std::vector<double> w(n);
        for (auto &v : w) v = 2 * uni() - 1;
        fmma::FMMA<double, 3> f;
        f.set_fn(OddKernel{});
        f.set_solver_type("fmm");
        f.set_poly_ord(4);
        f.set_Depth(4);
        std::vector<double> a, b, e;
        f.solve(x, w, x, a);
        fmma_plan_set_p2p_mutual(f.plan(), 0);
        f.solve(x, w, x, b);
        f.set_solver_type("exact");
        f.solve(x, w, x, e);
        check("odd_mutual_vs_direct_3d", rel(a, b), 1e-13);
        check("odd_fmm_vs_exact_3d", rel(a, e), 1e-2);
    }
    {  // differentiation with a user kernel and its gradient functor vs the built-in Gaussian
        const size_t n = 20000;
        auto t = pts<double, 3>(n), s = pts<double, 3>(n);
        std::vector<double> w(n), cot(n);
        for (auto &v : w) v = 2 * uni() - 1;
        for (auto &v : cot) v = 2 * uni() - 1;
        const double a = 1.0 / (2 * 0.1 * 0.1);
        fmma::FMMA<double, 3> f;
        f.set_fn([a] __host__ __device__(const std::array<double, 3> &d) {
                     return std::exp(-a * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
                 },
                 [a] __host__ __device__(const std::array<double, 3> &d) {
                     const double e = -2 * a * std::exp(-a * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
                     return std::array<double, 3>{e * d[0], e * d[1], e * d[2]};
                 });
        f.set_solver_type("fmm");
        f.set_poly_ord(3);
        f.set_Depth(4);
        f.set_trans_sym_flag(1);
        fmma::FMMA<double, 3> g;
        g.set_kernel(FMMA_KERNEL_GAUSS, {0.1});
        g.set_solver_type("fmm");
        g.set_poly_ord(3);
        g.set_Depth(4);
        g.set_trans_sym_flag(1);
        std::vector<std::array<double, 3>> ga, gb;
        std::vector<double> ta, tb;
        f.gradient(t, w, s, ga);
        g.gradient(t, w, s, gb);
        f.solve_transpose(t, cot, s, ta);  // no symmetry declared: the transposed launcher table
        g.solve_transpose(t, cot, s, tb);
        std::vector<double> fa, fb;
        for (size_t i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) {
                fa.push_back(ga[i][d]);
                fb.push_back(gb[i][d]);
            }
        check("user_grad_vs_builtin_gauss", rel(fa, fb), 1e-12);
        check("user_transpose_vs_builtin_gauss", rel(ta, tb), 1e-12);
    }
    {  // transpose of a kernel that is neither symmetric nor translation invariant, exact solver, vs the host sum
        const size_t nt = 300, ns = 200;
        auto t = pts<double, 2>(nt), s = pts<double, 2>(ns);
        std::vector<double> cot(nt), out;
        for (auto &v : cot) v = 2 * uni() - 1;
        auto fn = [] __host__ __device__(const std::array<double, 2> &x, const std::array<double, 2> &y) {
            return (1.0 + x[0]) * std::exp(-(x[0] - y[0]) * (x[0] - y[0]) - 2.0 * (x[1] - y[1]) * (x[1] - y[1]));
        };
        fmma::FMMA<double, 2> f;
        f.set_fn(fn);
        f.set_solver_type("exact");
        f.solve_transpose(t, cot, s, out);
        std::vector<double> host(ns, 0.0);
        for (size_t j = 0; j < ns; ++j)
            for (size_t i = 0; i < nt; ++i) host[j] += cot[i] * fn(t[i], s[j]);
        check("general_transpose_exact_vs_host", rel(out, host), 1e-13);
    }
    {  // tree and nrnmm with a user lambda vs the built-in kernel, and vs the exact sum
        const size_t n = 8000;
        auto x = pts<double, 3>(n);
        std::vector<double> w(n);
        for (auto &v : w) v = uni();
        fmma::FMMA<double, 3> f, g;
        f.set_fn([] __host__ __device__(const std::array<double, 3> &d) {
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            return r2 == 0.0 ? 0.0 : 1.0 / std::sqrt(r2);
        });
        g.set_kernel(FMMA_KERNEL_INV_R_REG);
        for (auto *p : {&f, &g}) {
            p->set_poly_ord(4);
            p->set_Depth(4);
            p->set_nrn_N(6);
        }
        std::vector<double> ta, tb, na, nb, e;
        f.tree(x, w, x, ta);
        g.tree(x, w, x, tb);
        f.nrnmm(x, w, x, na);
        g.nrnmm(x, w, x, nb);
        f.exact(x, w, x, e);
        check("tree_lambda_vs_builtin", rel(ta, tb), 1e-12);
        check("nrnmm_lambda_vs_builtin", rel(na, nb), 1e-12);
        check("tree_vs_exact", rel(ta, e), 1e-3);
        check("nrnmm_vs_exact", rel(na, e), 1e-3);
    }
    {  // a two-argument kernel written with the point operators of fmma.hpp
       // (include/fmma/fmma.hpp:15-49), in device code, vs the built-in 1/r
        const size_t n = 6000;
        auto x = pts<double, 3>(n);
        std::vector<double> w(n);
        for (auto &v : w) v = uni();
        fmma::FMMA<double, 3> f, g;
        f.set_fn([] __host__ __device__(const std::array<double, 3> &a, const std::array<double, 3> &b) {
            using fmma::operator-;
            using fmma::operator*;
            const std::array<double, 3> d = a - b, q = d * d;
            const double r2 = q[0] + q[1] + q[2];
            return r2 == 0.0 ? 0.0 : 1.0 / std::sqrt(r2);
        });
        g.set_kernel(FMMA_KERNEL_INV_R_REG);
        std::vector<double> a, b;
        for (auto *p : {&f, &g}) {
            p->set_solver_type("fmm");
            p->set_poly_ord(3);
            p->set_Depth(3);
            p->set_trans_sym_flag(1);
        }
        f.solve(x, w, x, a);
        g.solve(x, w, x, b);
        check("point_operators_vs_builtin", rel(a, b), 1e-12);
        using fmma::operator+;
        using fmma::operator/;
        const std::array<double, 2> u{1.0, 6.0}, v{3.0, 2.0}, s = u + v, r = u / v;
        check("point_operators_host", std::abs(s[0] - 4.0) + std::abs(s[1] - 8.0) + std::abs(r[1] - 3.0), 0);
    }
    {  // every point at one position: the cube has size 0 and the solve is rejected
        std::vector<std::array<double, 3>> x(100, std::array<double, 3>{0.5, 0.5, 0.5});
        std::vector<double> w(100, 1.0), a;
        fmma::FMMA<double, 3> f;
        f.set_solver_type("fmm");
        f.set_poly_ord(2);
        bool rejected = false;
        try {
            f.solve(x, w, x, a);
        } catch (const std::runtime_error &) {
            rejected = true;
        }
        check("zero_size_cube_rejected", rejected ? 0 : 1, 0);
    }
    {  // time log keys of the reference (src/fmm.cpp:57-141, src/fmma.cpp:189)
        fmma::FMMA<double, 2> f;
        auto x = pts<double, 2>(5000);
        std::vector<double> a;
        f.set_solver_type("fmm");
        f.set_kernel(FMMA_KERNEL_LOG_R);
        f.set_poly_ord(4);
        f.solve(x, x, a);
        auto log = f.get_time_log();
        int have = 0;
        for (const char *k : {"P2M", "M2M", "M2L", "L2L", "L2P", "P2P", "total time(fmm)"}) have += log.count(k);
        check("time_log_missing_keys", 7 - have, 0);
    }
    printf(fails ? "FAILED\n" : "OK\n");
    return fails ? 1 : 0;
}
