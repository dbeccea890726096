// pyfmma.cpp -- the drop-in for the reference's Python module pyfmma
// (cpython/pyfmma.cpp:140-168): classes fmmad1, fmmaf1, fmmad2, fmmaf2,
// fmmad3, fmmaf3 with solve (weighted and unweighted, ans filled in place),
// check_blas, set_fn1, set_fn2, set_solver_type, set_nrn_N, set_poly_ord,
// set_Depth, set_trans_sym_flag.  Built on fmma::FMMA (include/fmma/fmma.hpp),
// so every solve runs on the GPU.
//
// Shape checks and their RuntimeError messages follow the reference
// (cpython/pyfmma.cpp:22-53).  One deliberate difference: ans must already
// have the class's dtype and be C-contiguous (the reference would silently
// fill a converted temporary otherwise).  set_fn1(f) / set_fn2(f) take the
// reference's Python callables, f(x - y) and f(x, y) over the point's
// coordinates: fmma_amd.jit traces f once into an expression and compiles
// it for the GPU (the reference calls f per pair on the CPU).  Built-in
// kernels are also available by name with set_kernel(name, *params).  The
// GIL is released while the GPU works.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <map>
#include <string>

#include "../../include/fmma/fmma.hpp"

namespace py = pybind11;

namespace {

const std::map<std::string, int> &kernel_ids()
{
    static const std::map<std::string, int> m = {
        {"inv_r", FMMA_KERNEL_INV_R},   {"inv_r_reg", FMMA_KERNEL_INV_R_REG}, {"inv_diff", FMMA_KERNEL_INV_DIFF},
        {"log_r", FMMA_KERNEL_LOG_R},   {"gauss", FMMA_KERNEL_GAUSS},         {"yukawa", FMMA_KERNEL_YUKAWA},
        {"sum_r2", FMMA_KERNEL_SUM_R2}, {"sum2_r2", FMMA_KERNEL_SUM2_R2}};
    return m;
}

template <typename TYPE, std::size_t DIM>
class PyFMMA : public fmma::FMMA<TYPE, DIM> {
  public:
    using Arr = py::array_t<TYPE, py::array::c_style | py::array::forcecast>;

    void pysolve_w(const Arr &target, const Arr &source_weight, const Arr &source, py::array &ans)
    {
        py::buffer_info bt = target.request(), bw = source_weight.request(), bs = source.request();
        if (bt.ndim != 2) throw std::runtime_error("target shape must be 2 but " + std::to_string(bt.ndim));
        if (bw.ndim != 1)
            throw std::runtime_error("source_weight shape must be 2 but " + std::to_string(bw.ndim));
        if (bs.ndim != 2) throw std::runtime_error("source shape must be 2 but " + std::to_string(bs.ndim));
        check_common(bt, bs, ans);
        if (bs.shape[0] != bw.shape[0])
            throw std::runtime_error("source num " + std::to_string(bs.shape[0]) +
                                     " is different from source_weight num " + std::to_string(bw.shape[0]));
        run(bt, &bw, bs, ans);
    }

    void pysolve(const Arr &target, const Arr &source, py::array &ans)
    {
        py::buffer_info bt = target.request(), bs = source.request();
        if (bt.ndim != 2) throw std::runtime_error("target shape must be 2 but " + std::to_string(bt.ndim));
        if (bs.ndim != 2) throw std::runtime_error("source shape must be 2 but " + std::to_string(bs.ndim));
        check_common(bt, bs, ans);
        run(bt, nullptr, bs, ans);
    }

    void set_kernel_py(const std::string &name, py::args params)
    {
        auto it = kernel_ids().find(name);
        if (it == kernel_ids().end()) throw std::runtime_error("unknown kernel " + name);
        std::vector<double> p;
        for (auto h : params) p.push_back(h.cast<double>());
        const int rc = fmma_plan_set_kernel(this->plan(), it->second, p.empty() ? nullptr : p.data(), (int)p.size());
        if (rc != FMMA_OK) throw std::runtime_error(std::string("fmma: ") + fmma_last_error());
    }

    void set_m2l_precision_py(const std::string &name)
    {
        const int v = name == "native" ? FMMA_M2L_NATIVE : name == "fp32" ? FMMA_M2L_FP32
                      : name == "bf16x3" ? FMMA_M2L_BF16X3 : -1;
        if (v < 0) throw std::runtime_error("m2l precision must be native, fp32 or bf16x3");
        this->set_m2l_precision(v);
    }

    // set_fn1 (nargs 1: f(x - y)) and set_fn2 (nargs 2: f(x, y))
    // (cpython/pyfmma.cpp:150-151): the callable traced and compiled for the
    // GPU by fmma_amd.jit, whose launcher table the plan copies.
    void set_fn_py(py::object f, int nargs)
    {
        if (!PyCallable_Check(f.ptr())) throw std::runtime_error("set_fn needs a callable");
        py::object jit = py::module_::import("fmma_amd.jit");
        const uintptr_t addr = jit.attr("kernel_ops_address")(f, (int)DIM, sizeof(TYPE) == 8 ? "float64" : "float32",
                                                              nargs)
                                   .template cast<uintptr_t>();
        const int rc = fmma_plan_set_kernel_ops(this->plan(), reinterpret_cast<const fmma_kernel_ops *>(addr));
        if (rc != FMMA_OK) throw std::runtime_error(std::string("fmma: ") + fmma_last_error());
    }
    void set_fn1_py(py::object f) { set_fn_py(f, 1); }
    void set_fn2_py(py::object f) { set_fn_py(f, 2); }

  private:
    static void check_common(const py::buffer_info &bt, const py::buffer_info &bs, py::array &ans)
    {
        if (ans.ndim() != 1) throw std::runtime_error("ans shape must be 1 but " + std::to_string(ans.ndim()));
        if ((std::size_t)bt.shape[1] != DIM)
            throw std::runtime_error("target dimension must be " + std::to_string(DIM) + " but " +
                                     std::to_string(bt.shape[1]));
        if ((std::size_t)bs.shape[1] != DIM)
            throw std::runtime_error("source dimension must be " + std::to_string(DIM) + " but " +
                                     std::to_string(bs.shape[1]));
        if (ans.shape(0) != bt.shape[0])
            throw std::runtime_error("ans num " + std::to_string(ans.shape(0)) + " is different from target num " +
                                     std::to_string(bt.shape[0]));
        if (!py::isinstance<py::array_t<TYPE>>(ans) || !(ans.flags() & py::array::c_style) || !ans.writeable())
            throw std::runtime_error(std::string("ans must be a writeable C-contiguous ") +
                                     (sizeof(TYPE) == 8 ? "float64" : "float32") + " array");
    }

    void run(const py::buffer_info &bt, const py::buffer_info *bw, const py::buffer_info &bs, py::array &ans)
    {
        const std::size_t nt = bt.shape[0], ns = bs.shape[0];
        std::vector<std::array<TYPE, DIM>> t(nt), s(ns);
        if (nt) std::memcpy(t.data()->data(), bt.ptr, nt * DIM * sizeof(TYPE));
        if (ns) std::memcpy(s.data()->data(), bs.ptr, ns * DIM * sizeof(TYPE));
        std::vector<TYPE> w, a(nt);
        if (bw) w.assign(static_cast<const TYPE *>(bw->ptr), static_cast<const TYPE *>(bw->ptr) + ns);
        TYPE *out = static_cast<TYPE *>(ans.mutable_data());
        // the reference passes a fresh zero vector (cpython/pyfmma.cpp:56,113)
        {
            py::gil_scoped_release nogil;
            if (bw)
                this->solve(t, w, s, a);
            else
                this->solve(t, s, a);
        }
        if (nt) std::memcpy(out, a.data(), nt * sizeof(TYPE));
    }
};

template <typename TYPE, std::size_t DIM>
void make_class(py::module &m, const std::string &type_str)
{
    using C = PyFMMA<TYPE, DIM>;
    const std::string name = "fmma" + type_str + std::to_string(DIM);
    py::class_<C>(m, name.c_str())
        .def(py::init<>())
        .def("solve", &C::pysolve_w, py::arg("target"), py::arg("source_weight"), py::arg("source"), py::arg("ans"))
        .def("solve", &C::pysolve, py::arg("target"), py::arg("source"), py::arg("ans"))
        .def("check_blas", &C::check_blas)
        .def("set_fn1", &C::set_fn1_py, py::arg("fn"))
        .def("set_fn2", &C::set_fn2_py, py::arg("fn"))
        .def("set_kernel", &C::set_kernel_py)
        .def("set_solver_type", &C::set_solver_type)
        .def("set_nrn_N", &C::set_nrn_N)
        .def("set_poly_ord", &C::set_poly_ord)
        .def("set_Depth", &C::set_Depth)
        .def("get_Depth", &C::get_Depth)
        .def("set_trans_sym_flag", &C::set_trans_sym_flag)
        .def("set_m2l_precision", &C::set_m2l_precision_py)
        .def("time_log", &C::get_time_log);
}

}  // namespace

PYBIND11_MODULE(pyfmma, m)
{
    m.doc() = "Black-box FMM on AMD MI355X; drop-in for fockl/FMMA's pyfmma";
    make_class<double, 1>(m, "d");
    make_class<float, 1>(m, "f");
    make_class<double, 2>(m, "d");
    make_class<float, 2>(m, "f");
    make_class<double, 3>(m, "d");
    make_class<float, 3>(m, "f");
}
