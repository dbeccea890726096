// fmma/fmma.hpp -- drop-in fmma::FMMA<TYPE,DIM> for AMD MI355X.
//
// Same names, overloads, argument order and state as the reference class
// (fockl/FMMA include/fmma/fmma.hpp:52-170): set_fn / set_fn1 / set_fn2,
// set_solver_type, set_nrn_N, set_poly_ord, set_Depth, set_trans_sym_flag,
// check_blas, and solve / exact / exact_matvec / nrnmm / tree / fmm in the
// weighted (target, source_weight, source, ans) and unweighted
// (target, source, ans) forms.  Every solve runs on the GPU through the C-ABI
// of libfmma_hip.so (fmma_hip.h); nothing is computed on the host.
//
// Differences a caller sees:
//  * f is a __host__ __device__ functor, passed as a template parameter
//    (set_fn, from fmma/fmma_device.hpp, compiled with hipcc) or one of the
//    built-in kernels (set_kernel); a host-only std::function cannot run on
//    the GPU and is rejected at compile time.
//  * New setters: set_kernel, set_m2l_precision; new getter get_time_log;
//    new differentiation entry points gradient and solve_transpose.
// Misuse the reference ends with exit(EXIT_FAILURE) ends the same way here:
// an unknown solver type (src/fmma.cpp:183-186) and a source_weight size
// that differs from the source size (src/fmm.cpp:32-35).
#pragma once

#include <array>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <initializer_list>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../fmma_hip.h"

#if defined(__HIPCC__)
#define FMMA_HOST_DEVICE __host__ __device__
#else
#define FMMA_HOST_DEVICE
#endif

namespace fmma {

// Elementwise +, -, *, / of two points, as the reference defines them for
// kernels written over std::array (include/fmma/fmma.hpp:15-49); usable in
// host code and, compiled with hipcc, inside a device functor.
#define FMMA_POINT_OPERATOR(OP)                                                                             \
    template <typename TYPE, std::size_t DIM>                                                               \
    FMMA_HOST_DEVICE inline std::array<TYPE, DIM> operator OP(const std::array<TYPE, DIM> &a,              \
                                                             const std::array<TYPE, DIM> &b)               \
    {                                                                                                       \
        std::array<TYPE, DIM> r{};                                                                          \
        for (std::size_t d = 0; d < DIM; ++d) r[d] = a[d] OP b[d];                                          \
        return r;                                                                                           \
    }
FMMA_POINT_OPERATOR(+)
FMMA_POINT_OPERATOR(-)
FMMA_POINT_OPERATOR(*)
FMMA_POINT_OPERATOR(/)
#undef FMMA_POINT_OPERATOR

template <typename TYPE, std::size_t DIM>
class FMMA {
    static_assert(DIM >= 1 && DIM <= 3, "DIM must be 1, 2 or 3");
    static_assert(std::is_same<TYPE, double>::value || std::is_same<TYPE, float>::value,
                  "TYPE must be double or float");

  public:
    FMMA()
    {
        check(fmma_plan_create(&plan_, (int)DIM, std::is_same<TYPE, double>::value ? FMMA_F64 : FMMA_F32));
    }
    ~FMMA() { fmma_plan_destroy(plan_); }
    FMMA(const FMMA &) = delete;
    FMMA &operator=(const FMMA &) = delete;

    // ---- kernel f(x, y)
    // Built-in kernels by id (FMMA_KERNEL_* in fmma_hip.h).  The default is
    // 1/|x-y| with inf at x = y, the reference's default
    // (include/fmma/fmma.hpp:55-63).
    void set_kernel(int kernel_id, std::initializer_list<double> params = {})
    {
        std::vector<double> p(params);
        check(fmma_plan_set_kernel(plan_, kernel_id, p.empty() ? nullptr : p.data(), (int)p.size()));
        user_.reset();
    }

    // set_fn(f(x-y)) / set_fn(f(x, y)) with a __host__ __device__ functor,
    // defined in fmma/fmma_device.hpp (hipcc).  A one-argument f is taken as
    // f(x - y) (include/fmma/fmma.hpp:65-72), a two-argument f as f(x, y).
    template <class F>
    void set_fn(F f);
    // f with its gradient grad_f(x, y) -> std::array<TYPE, DIM> = grad_x f
    // (or grad_f(x - y) for a one-argument f), for gradient().
    template <class F, class G>
    void set_fn(F f, G grad_f);
    template <class F>
    void set_fn1(F f) { set_fn(f); }
    template <class F>
    void set_fn2(F f) { set_fn(f); }

    void set_solver_type(const std::string new_type)
    {
        solver_type_ = new_type;
        check(fmma_plan_set_solver_type(plan_, new_type.c_str()));
    }
    void set_nrn_N(int n) { check(fmma_plan_set_nrn_N(plan_, n)); }
    void set_poly_ord(int p) { check(fmma_plan_set_poly_ord(plan_, p)); }
    void set_Depth(int d) { check(fmma_plan_set_depth(plan_, d)); }
    void set_trans_sym_flag(int f) { check(fmma_plan_set_trans_sym_flag(plan_, f)); }
    void set_m2l_precision(int precision) { check(fmma_plan_set_m2l_precision(plan_, precision)); }

    // The reference reports whether it was built against CBLAS; this build
    // uses no BLAS (its dense contraction is hand-written MFMA code).
    bool check_blas(void) { return false; }

    int get_Depth() const
    {
        int d = -1;
        fmma_plan_get_depth(plan_, &d);
        return d;
    }

    // Per-stage milliseconds (HIP events), accumulated like the reference's
    // time_log (include/fmma/fmma.hpp:168).
    std::map<std::string, double> get_time_log() const
    {
        std::map<std::string, double> m;
        const int n = fmma_plan_time_log(plan_, nullptr, nullptr, 0);
        std::vector<const char *> names(n > 0 ? n : 1);
        std::vector<double> ms(n > 0 ? n : 1);
        fmma_plan_time_log(plan_, names.data(), ms.data(), n);
        for (int i = 0; i < n; ++i) m[names[i]] = ms[i];
        return m;
    }

    // ---- solvers (src/fmma.cpp:143-207)
    void solve(const std::vector<std::array<TYPE, DIM>> &target, const std::vector<std::array<TYPE, DIM>> &source,
               std::vector<TYPE> &ans)
    {
        run(solver_type_, target, nullptr, source, ans);
    }
    void solve(const std::vector<std::array<TYPE, DIM>> &target, const std::vector<TYPE> &source_weight,
               const std::vector<std::array<TYPE, DIM>> &source, std::vector<TYPE> &ans)
    {
        run(solver_type_, target, &source_weight, source, ans);
    }

#define FMMA_SOLVER_PAIR(NAME)                                                                                   \
    void NAME(const std::vector<std::array<TYPE, DIM>> &target, const std::vector<std::array<TYPE, DIM>> &source, \
              std::vector<TYPE> &ans)                                                                            \
    {                                                                                                            \
        run(#NAME, target, nullptr, source, ans);                                                                \
    }                                                                                                            \
    void NAME(const std::vector<std::array<TYPE, DIM>> &target, const std::vector<TYPE> &source_weight,          \
              const std::vector<std::array<TYPE, DIM>> &source, std::vector<TYPE> &ans)                          \
    {                                                                                                            \
        run(#NAME, target, &source_weight, source, ans);                                                         \
    }
    FMMA_SOLVER_PAIR(exact)
    FMMA_SOLVER_PAIR(exact_matvec)
    FMMA_SOLVER_PAIR(nrnmm)
    FMMA_SOLVER_PAIR(tree)
    FMMA_SOLVER_PAIR(fmm)
#undef FMMA_SOLVER_PAIR

    // ---- differentiation (new; the reference has none)
    // grad[i][d] = d ans_i / d x_{i,d} of the current solver's result
    // (fmm: the derivative of the approximation; exact: of the direct sum).
    // Needs a kernel with a gradient: a built-in, or set_fn(f, grad_f).
    void gradient(const std::vector<std::array<TYPE, DIM>> &target, const std::vector<TYPE> &source_weight,
                  const std::vector<std::array<TYPE, DIM>> &source, std::vector<std::array<TYPE, DIM>> &grad)
    {
        grad.resize(target.size());
        check_sizes(target, &source_weight, source);
        check(fmma_plan_gradient_host(plan_, (long)target.size(), ptr(target), source_weight.data(),
                                      (long)source.size(), ptr(source), grad.empty() ? nullptr : grad.data()->data()));
    }
    // out[j] = sum_i cot[i] f(x_i, y_j): the transpose of the solve in the
    // weights, on the tree and Depth that solve(target, ., source) uses.
    void solve_transpose(const std::vector<std::array<TYPE, DIM>> &target, const std::vector<TYPE> &cot,
                         const std::vector<std::array<TYPE, DIM>> &source, std::vector<TYPE> &out)
    {
        out.resize(source.size());
        if (cot.size() != target.size()) throw std::runtime_error("fmma: cot size != target size");
        check(fmma_plan_solve_transpose_host(plan_, (long)target.size(), ptr(target), cot.data(),
                                             (long)source.size(), ptr(source), out.data()));
    }

    fmma_plan plan() const { return plan_; }

  private:
    static const void *ptr(const std::vector<std::array<TYPE, DIM>> &v)
    {
        return v.empty() ? nullptr : v.data()->data();
    }
    static void check_sizes(const std::vector<std::array<TYPE, DIM>> &target, const std::vector<TYPE> *source_weight,
                            const std::vector<std::array<TYPE, DIM>> &source)
    {
        if (source_weight && !target.empty() && !source.empty() && source_weight->size() != source.size()) {
            fprintf(stderr, "%s:%d ERROR : source_weight size (%zu) != source size (%zu)\n", __FILE__, __LINE__,
                    source_weight->size(), source.size());
            exit(EXIT_FAILURE);
        }
    }

    static void check(int rc)
    {
        if (rc != FMMA_OK) throw std::runtime_error(std::string("fmma: ") + fmma_last_error());
    }

    void run(const std::string &solver, const std::vector<std::array<TYPE, DIM>> &target,
             const std::vector<TYPE> *source_weight, const std::vector<std::array<TYPE, DIM>> &source,
             std::vector<TYPE> &ans)
    {
        ans.resize(target.size());
        check_sizes(target, source_weight, source);
        const bool swap = solver != solver_type_;
        if (swap) check(fmma_plan_set_solver_type(plan_, solver.c_str()));
        // std::array<TYPE, DIM> elements are contiguous: the vectors are the
        // [n][DIM] arrays the C-ABI takes.
        const void *t = target.empty() ? nullptr : target.data()->data();
        const void *s = source.empty() ? nullptr : source.data()->data();
        const int rc = fmma_plan_solve_host(plan_, (long)target.size(), t,
                                            source_weight ? source_weight->data() : nullptr, (long)source.size(), s,
                                            ans.data());
        if (swap) fmma_plan_set_solver_type(plan_, solver_type_.c_str());
        if (rc == FMMA_ERR_SOLVER) {
            fprintf(stderr, "%s:%d ERROR : solve type %s not undefined\n", __FILE__, __LINE__, solver.c_str());
            exit(EXIT_FAILURE);
        }
        check(rc);
#if FMMA_TIME_LOG
        write_time_log();
#endif
    }

#if FMMA_TIME_LOG
    // As the reference with FMMA_TIME_LOG (src/fmma.cpp:191-204): entries
    // sorted by time, to stderr and to fmma.log.
    void write_time_log()
    {
        std::multimap<double, std::string, std::greater<double>> m;
        for (const auto &kv : get_time_log()) m.insert({kv.second, kv.first});
        FILE *fp = fopen("fmma.log", "w");
        if (fp) fprintf(fp, "name time[ms]\n");
        for (const auto &kv : m) {
            fprintf(stderr, "%s %lf\n", kv.second.c_str(), kv.first);
            if (fp) fprintf(fp, "%s %lf\n", kv.second.c_str(), kv.first);
        }
        if (fp) fclose(fp);
    }
#endif

    fmma_plan plan_ = nullptr;
    std::string solver_type_ = "exact";
    std::shared_ptr<void> user_;  // the device functor of set_fn, kept alive with the plan
};

}  // namespace fmma

#if defined(__HIPCC__)
#include "fmma_device.hpp"
#else
namespace fmma {
template <typename TYPE, std::size_t DIM>
template <class F>
void FMMA<TYPE, DIM>::set_fn(F)
{
    static_assert(sizeof(F) == 0,
                  "FMMA::set_fn takes a __host__ __device__ functor and must be compiled with hipcc; "
                  "use set_kernel(FMMA_KERNEL_*) for the built-in kernels");
}
template <typename TYPE, std::size_t DIM>
template <class F, class G>
void FMMA<TYPE, DIM>::set_fn(F, G)
{
    static_assert(sizeof(F) == 0,
                  "FMMA::set_fn takes __host__ __device__ functors and must be compiled with hipcc; "
                  "use set_kernel(FMMA_KERNEL_*) for the built-in kernels");
}
}  // namespace fmma
#endif
