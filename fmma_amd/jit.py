"""Python kernels on the GPU.

The reference's pyfmma takes any Python callable as the kernel
(``set_fn1(f)`` for f(x - y), ``set_fn2(f)`` for f(x, y);
cpython/pyfmma.cpp:150-151) and calls it once per pair on the CPU.  Here the
callable is traced once: it is called with symbolic points, and every
arithmetic operation, ``abs`` and numpy function it applies (np.fabs,
np.sqrt, np.exp, np.log, ...) is recorded as an expression.  The expression,
and its derivative with respect to x for the gradient, are written out as a
``__host__ __device__`` functor, compiled once with hipcc for gfx950 into a
small shared library next to the launcher templates of
include/fmma/detail/pair_kernels.hpp (cached on disk under the expression's
hash), and handed to a plan as its launcher table.

The functor computes in double, as the reference's Python kernel does
whatever TYPE is.  Symmetry (f(y, x) = +-f(x, y), which lets targets =
sources evaluate each pair once) is taken from the callable itself on 256
random pairs; translation invariance (f depends on x - y only) from the
expression: every coordinate appears only as x[d] - y[d].  Before a kernel is
used, the compiled functor is evaluated on the host and compared with the
callable on random pairs.

What cannot be traced raises TypeError: branching on a coordinate (if,
comparisons), math.* functions (use the numpy ones), and calls into code
that needs real numbers.
"""
import ctypes
import hashlib
import math
import os
import subprocess
import tempfile
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

__all__ = ["trace", "compile_kernel", "kernel_ops_address", "KernelSpec"]


# ----------------------------------------------------------------- tracing

class Sym:
    """A traced scalar: an expression node (op, args...)."""

    __slots__ = ("node",)
    __array_priority__ = 1000

    def __init__(self, node):
        self.node = node

    # arithmetic
    def __add__(self, o):
        return _bin("add", self, o)

    def __radd__(self, o):
        return _bin("add", o, self)

    def __sub__(self, o):
        return _bin("sub", self, o)

    def __rsub__(self, o):
        return _bin("sub", o, self)

    def __mul__(self, o):
        return _bin("mul", self, o)

    def __rmul__(self, o):
        return _bin("mul", o, self)

    def __truediv__(self, o):
        return _bin("div", self, o)

    def __rtruediv__(self, o):
        return _bin("div", o, self)

    def __pow__(self, o):
        return _bin("pow", self, o)

    def __rpow__(self, o):
        return _bin("pow", o, self)

    def __neg__(self):
        return Sym(("neg", self.node))

    def __pos__(self):
        return self

    def __abs__(self):
        return _call("fabs", self)

    # what cannot be traced
    def __bool__(self):
        raise TypeError("a kernel traced for the GPU cannot branch on a coordinate (if / comparisons); "
                        "write it with arithmetic and numpy functions (np.where is not supported)")

    def __float__(self):
        raise TypeError("a kernel traced for the GPU needs numpy functions (np.sqrt, np.exp, ...), "
                        "not math.* or float()")

    __int__ = __float__
    __index__ = __float__

    def __lt__(self, o):
        return self.__bool__()

    __le__ = __gt__ = __ge__ = __lt__

    # numpy: ufuncs on a Sym, and the methods object arrays call per element
    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        if method != "__call__" or kwargs.get("out") is not None:
            return NotImplemented
        name = _UFUNCS.get(ufunc.__name__)
        if name is None:
            raise TypeError("numpy function %s is not supported in a GPU kernel" % ufunc.__name__)
        return name(*inputs) if callable(name) else _call(name, *inputs)

    def sqrt(self):
        return _call("sqrt", self)

    def exp(self):
        return _call("exp", self)

    def log(self):
        return _call("log", self)

    def sin(self):
        return _call("sin", self)

    def cos(self):
        return _call("cos", self)

    def tanh(self):
        return _call("tanh", self)

    def fabs(self):
        return _call("fabs", self)

    def absolute(self):
        return _call("fabs", self)

    def conjugate(self):
        return self


def _node(v):
    if isinstance(v, Sym):
        return v.node
    if isinstance(v, (bool, np.bool_)):
        raise TypeError("booleans are not supported in a GPU kernel")
    if isinstance(v, (int, float, np.integer, np.floating)):
        return ("const", float(v))
    raise TypeError("a GPU kernel can combine coordinates with numbers only, not %r" % type(v).__name__)


def _bin(op, a, b):
    return Sym((op, _node(a), _node(b)))


def _call(fn, *args):
    return Sym(("call", fn) + tuple(_node(a) for a in args))


_UFUNCS = {
    "add": lambda a, b: _bin("add", a, b), "subtract": lambda a, b: _bin("sub", a, b),
    "multiply": lambda a, b: _bin("mul", a, b), "true_divide": lambda a, b: _bin("div", a, b),
    "divide": lambda a, b: _bin("div", a, b), "power": lambda a, b: _bin("pow", a, b),
    "float_power": lambda a, b: _bin("pow", a, b), "negative": lambda a: -Sym(_node(a)),
    "positive": lambda a: Sym(_node(a)), "square": lambda a: _bin("mul", a, a),
    "reciprocal": lambda a: _bin("div", 1.0, a),
    "absolute": "fabs", "fabs": "fabs", "sqrt": "sqrt", "cbrt": "cbrt", "exp": "exp", "exp2": "exp2",
    "expm1": "expm1", "log": "log", "log2": "log2", "log10": "log10", "log1p": "log1p", "sin": "sin",
    "cos": "cos", "tan": "tan", "arcsin": "asin", "arccos": "acos", "arctan": "atan", "arctan2": "atan2",
    "sinh": "sinh", "cosh": "cosh", "tanh": "tanh", "hypot": "hypot", "maximum": "fmax", "minimum": "fmin",
    "fmax": "fmax", "fmin": "fmin",
}


class SymPoint:
    """A traced point: x[d] is the symbolic coordinate d."""

    def __init__(self, name, dim):
        self.name = name
        self.dim = dim

    def __getitem__(self, k):
        if isinstance(k, slice):
            return [self[i] for i in range(self.dim)[k]]
        k = int(k)
        if k < 0:
            k += self.dim
        if not 0 <= k < self.dim:
            raise IndexError("coordinate %d of a %d-D point" % (k, self.dim))
        return Sym(("var", self.name, k))

    def __len__(self):
        return self.dim

    def __iter__(self):
        return (self[i] for i in range(self.dim))

    def __array__(self, dtype=None, copy=None):
        return np.array([self[i] for i in range(self.dim)], dtype=object)


def trace(fn, dim, nargs):
    """The expression of fn: f(x, y) (nargs 2) or f(d) with d = x - y (1)."""
    if nargs == 2:
        out = fn(SymPoint("x", dim), SymPoint("y", dim))
    else:
        out = fn(SymPoint("d", dim))
    if isinstance(out, np.ndarray) and out.shape == () and out.dtype == object:
        out = out.item()
    return _node(out)


# ---------------------------------------------------------- simplification

def _is_const(n, v=None):
    return n[0] == "const" and (v is None or n[1] == v)


def _mk(op, *args):
    """Node construction with the constant folding the derivative needs."""
    if op in ("add", "sub", "mul", "div") and all(_is_const(a) for a in args):
        a, b = args[0][1], args[1][1]
        if op != "div" or b != 0.0:
            return _c({"add": a + b, "sub": a - b, "mul": a * b, "div": a / b if b else 0.0}[op])
    if op == "add":
        a, b = args
        if _is_const(a, 0.0):
            return b
        if _is_const(b, 0.0):
            return a
    elif op == "sub":
        a, b = args
        if _is_const(b, 0.0):
            return a
        if _is_const(a, 0.0):
            return ("neg", b)
    elif op == "mul":
        a, b = args
        if _is_const(a, 0.0) or _is_const(b, 0.0):
            return ("const", 0.0)
        if _is_const(a, 1.0):
            return b
        if _is_const(b, 1.0):
            return a
    elif op == "div":
        a, b = args
        if _is_const(a, 0.0):
            return ("const", 0.0)
        if _is_const(b, 1.0):
            return a
    elif op == "neg":
        (a,) = args
        if _is_const(a):
            return ("const", -a[1])
        if a[0] == "neg":
            return a[1]
    return (op,) + tuple(args)


def _c(v):
    return ("const", float(v))


def _f(name, *args):
    return ("call", name) + tuple(args)


def derivative(n, var):
    """d n / d var, var = (name, k); a node of the same form."""
    op = n[0]
    if op == "const":
        return _c(0)
    if op == "var":
        return _c(1) if (n[1], n[2]) == var else _c(0)
    if op == "neg":
        return _mk("neg", derivative(n[1], var))
    if op in ("add", "sub"):
        return _mk(op, derivative(n[1], var), derivative(n[2], var))
    if op == "mul":
        a, b = n[1], n[2]
        return _mk("add", _mk("mul", derivative(a, var), b), _mk("mul", a, derivative(b, var)))
    if op == "div":
        a, b = n[1], n[2]
        da, db = derivative(a, var), derivative(b, var)
        return _mk("sub", _mk("div", da, b), _mk("div", _mk("mul", a, db), _mk("mul", b, b)))
    if op == "pow":
        a, b = n[1], n[2]
        da, db = derivative(a, var), derivative(b, var)
        if _is_const(db, 0.0):
            return _mk("mul", _mk("mul", b, ("pow", a, _mk("sub", b, _c(1)))), da)
        return _mk("mul", n, _mk("add", _mk("mul", db, _f("log", a)), _mk("div", _mk("mul", b, da), a)))
    assert op == "call"
    name, args = n[1], n[2:]
    a = args[0]
    da = derivative(a, var)
    if name in ("fmax", "fmin", "atan2", "hypot"):
        b = args[1]
        db = derivative(b, var)
        if name == "fmax":
            return ("select_ge", a, b, da, db)
        if name == "fmin":
            return ("select_ge", b, a, da, db)
        if name == "atan2":
            return _mk("div", _mk("sub", _mk("mul", b, da), _mk("mul", a, db)),
                       _mk("add", _mk("mul", a, a), _mk("mul", b, b)))
        return _mk("div", _mk("add", _mk("mul", a, da), _mk("mul", b, db)), n)
    if _is_const(da, 0.0):
        return _c(0)
    d = {
        "fabs": lambda: _f("copysign", _c(1), a),
        "sqrt": lambda: _mk("div", _c(0.5), n),
        "cbrt": lambda: _mk("div", _c(1), _mk("mul", _c(3), _mk("mul", n, n))),
        "exp": lambda: n,
        "exp2": lambda: _mk("mul", _c(math.log(2.0)), n),
        "expm1": lambda: _f("exp", a),
        "log": lambda: _mk("div", _c(1), a),
        "log2": lambda: _mk("div", _c(1), _mk("mul", _c(math.log(2.0)), a)),
        "log10": lambda: _mk("div", _c(1), _mk("mul", _c(math.log(10.0)), a)),
        "log1p": lambda: _mk("div", _c(1), _mk("add", _c(1), a)),
        "sin": lambda: _f("cos", a),
        "cos": lambda: _mk("neg", _f("sin", a)),
        "tan": lambda: _mk("div", _c(1), _mk("mul", _f("cos", a), _f("cos", a))),
        "asin": lambda: _mk("div", _c(1), _f("sqrt", _mk("sub", _c(1), _mk("mul", a, a)))),
        "acos": lambda: _mk("neg", _mk("div", _c(1), _f("sqrt", _mk("sub", _c(1), _mk("mul", a, a))))),
        "atan": lambda: _mk("div", _c(1), _mk("add", _c(1), _mk("mul", a, a))),
        "sinh": lambda: _f("cosh", a),
        "cosh": lambda: _f("sinh", a),
        "tanh": lambda: _mk("sub", _c(1), _mk("mul", n, n)),
    }[name]()
    return _mk("mul", d, da)


# ------------------------------------------------------------ C++ emission

class Emitter:
    """C++ for a set of expressions, with every subexpression that occurs
    more than once (as the derivative's do) held in a local constant."""

    def __init__(self, roots):
        self.count = {}
        for r in roots:
            self._visit(r)
        self.names = {}
        self.lines = []

    def _visit(self, n):
        if n[0] in ("const", "var"):
            return
        seen = n in self.count
        self.count[n] = self.count.get(n, 0) + 1
        if not seen:
            for c in n[1:]:
                if isinstance(c, tuple):
                    self._visit(c)

    def expr(self, n):
        if n in self.names:
            return self.names[n]
        text = self._text(n)
        if n[0] not in ("const", "var") and self.count.get(n, 0) > 1:
            name = "t%d" % len(self.names)
            self.lines.append("const double %s = %s;" % (name, text))
            self.names[n] = name
            return name
        return text

    def _text(self, n):
        op = n[0]
        if op == "const":
            v = n[1]
            if math.isnan(v):
                return "__builtin_nan(\"\")"
            if math.isinf(v):
                return "(%s__builtin_inf())" % ("-" if v < 0 else "")
            return "(%s)" % v.hex()
        if op == "var":
            return "(double)%s[%d]" % (n[1], n[2])
        if op == "neg":
            return "(-%s)" % self.expr(n[1])
        if op in ("add", "sub", "mul", "div"):
            sym = {"add": "+", "sub": "-", "mul": "*", "div": "/"}[op]
            return "(%s %s %s)" % (self.expr(n[1]), sym, self.expr(n[2]))
        if op == "pow":
            b = n[2]
            if b[0] == "const" and b[1] in (2.0, 3.0, 4.0):
                # small integer powers as products: the value of pow(a, k) to
                # an ulp, at a fraction of its code in every inlined pair loop
                a = self.expr(n[1])
                return "(" + " * ".join([a] * int(b[1])) + ")"
            if b[0] == "const" and b[1] == 1.0:
                return self.expr(n[1])
            return "std::pow(%s, %s)" % (self.expr(n[1]), self.expr(b))
        if op == "select_ge":
            return "(%s >= %s ? %s : %s)" % tuple(self.expr(a) for a in n[1:])
        assert op == "call"
        return "std::%s(%s)" % (n[1], ", ".join(self.expr(a) for a in n[2:]))


def emit_body(roots, ret):
    """Statements computing `roots`, then `ret` formatted with their values."""
    em = Emitter(roots)
    vals = [em.expr(r) for r in roots]
    return em.lines + [ret % tuple(vals)]


def _uses_only_differences(n, dim):
    """Every coordinate occurs as x[d] - y[d] or y[d] - x[d] (same d)."""
    op = n[0]
    if op == "var":
        return False
    if op == "const":
        return True
    if op == "sub" and n[1][0] == "var" and n[2][0] == "var":
        a, b = n[1], n[2]
        return a[2] == b[2] and {a[1], b[1]} == {"x", "y"}
    return all(_uses_only_differences(c, dim) for c in n[1:] if isinstance(c, tuple))


# ------------------------------------------------------ symmetry and checks

def _eval_py(fn, nargs, x, y):
    try:
        with np.errstate(all="ignore"):
            v = fn(list(x), list(y)) if nargs == 2 else fn(list(np.asarray(x) - np.asarray(y)))
        return float(v)
    except (ZeroDivisionError, OverflowError, ValueError):
        return float("nan")


def _sample_pairs(dim, k=256, seed=12345):
    rng = np.random.default_rng(seed)
    x = rng.random((k, dim))
    y = rng.random((k, dim))
    return x, y


def symmetry_of(fn, dim, nargs):
    """+1 if f(y,x) == f(x,y), -1 if == -f(x,y) on 256 random pairs (exactly,
    in double), else 0."""
    x, y = _sample_pairs(dim)
    sym = anti = True
    seen = 0
    for a, b in zip(x, y):
        u, v = _eval_py(fn, nargs, a, b), _eval_py(fn, nargs, b, a)
        if not (math.isfinite(u) and math.isfinite(v)):
            continue
        seen += 1
        sym = sym and u == v
        anti = anti and u == -v
    if seen == 0:
        return 0
    return 1 if sym else (-1 if anti and not sym else 0)


class KernelSpec:
    """A traced kernel: the expressions, flags and the source to compile."""

    def __init__(self, fn, dim, nargs, dtype):
        self.dim, self.nargs = int(dim), int(nargs)
        self.dtype = np.dtype(dtype)
        self.expr = trace(fn, self.dim, self.nargs)
        pv = "x" if self.nargs == 2 else "d"
        self.grad = [derivative(self.expr, (pv, k)) for k in range(self.dim)]
        self.symmetry = symmetry_of(fn, self.dim, self.nargs)
        self.translation_invariant = self.nargs == 1 or _uses_only_differences(self.expr, self.dim)
        self.sources = self._sources()
        h = hashlib.sha256()
        for name in sorted(self.sources):
            h.update(name.encode() + b"\0" + self.sources[name].encode() + b"\0")
        self.key = h.hexdigest()[:24]

    def _sources(self):
        """The library's translation units (compiled in parallel): the
        functor and the launcher table, then the launcher groups."""
        T = "double" if self.dtype == np.float64 else "float"
        D = self.dim
        if self.nargs == 2:
            args = "const std::array<%s, %d> &x, const std::array<%s, %d> &y" % (T, D, T, D)
            hx = "std::array<%s, %d> x, y;\n    for (int k = 0; k < %d; ++k) { x[k] = (%s)px[k]; y[k] = (%s)py[k]; }" % (
                T, D, D, T, T)
            call = "PyKernel{}(x, y)"
        else:
            args = "const std::array<%s, %d> &d" % (T, D)
            hx = "std::array<%s, %d> d;\n    for (int k = 0; k < %d; ++k) d[k] = (%s)px[k] - (%s)py[k];" % (
                T, D, D, T, T)
            call = "PyKernel{}(d)"
        prelude = [
            "// generated by fmma_amd.jit from a traced Python kernel",
            "#include <cmath>",
            "#include <fmma/fmma.hpp>",
            "#include <fmma/fmma_device.hpp>",
            "",
            "struct PyKernel {",
            "    static constexpr int symmetry = %d;" % self.symmetry,
            "    __host__ __device__ double operator()(%s) const" % args,
            "    {",
        ] + ["        " + l for l in emit_body([self.expr], "return %s;")] + [
            "    }",
            "};",
            "",
            "struct PyGrad {",
            "    __host__ __device__ std::array<%s, %d> operator()(%s) const" % (T, D, args),
            "    {",
        ] + ["        " + l for l in emit_body(self.grad, "return std::array<%s, %d>{%s};" % (
            T, D, ", ".join(["(%s)%%s" % T] * D)))] + [
            "    }",
            "};",
            "",
            "using K = fmma::detail::UserKernel<%s, %d, PyKernel, %s, PyGrad>;" % (
                T, D, "true" if self.nargs == 2 else "false"),
            "using KT = fmma::detail::Transposed<K>;",
            "#define HIDDEN __attribute__((visibility(\"hidden\")))",
            "#define EXPORT __attribute__((visibility(\"default\")))",
            "",
        ]
        ops = "fmma::detail::ops_for<%s, %d, %%s>::%%s" % (T, D)
        sig = {"p2p": "fmma_p2p_args", "exact": "fmma_exact_args", "m2l_op": "fmma_m2l_op_args",
               "m2l_general": "fmma_m2l_general_args", "m2p": "fmma_m2p_args", "p2p_grad": "fmma_p2p_grad_args",
               "exact_grad": "fmma_exact_grad_args"}

        def launchers(kt, names):
            out = []
            for nm in names:
                out += ['extern "C" HIDDEN int jit_%s%s(const %s *a, void *ctx, void *s)' % (
                            "t_" if kt else "", nm, sig[nm]),
                        "{",
                        "    return %s(a, ctx, s);" % (ops % ("KT" if kt else "K", nm)),
                        "}", ""]
            return out

        groups = {"p2p": (False, ["p2p"]), "grad": (False, ["p2p_grad", "exact_grad"]),
                  "far": (False, ["exact", "m2l_op", "m2l_general"]), "m2p": (False, ["m2p"])}
        if self.symmetry == 0:
            groups["t_p2p"] = (True, ["p2p"])
            groups["t_far"] = (True, ["exact", "m2l_op", "m2l_general", "m2p"])
        tus = {g: "\n".join(prelude + launchers(kt, names)) for g, (kt, names) in groups.items()}

        decl = []
        for g, (kt, names) in groups.items():
            for nm in names:
                decl.append('extern "C" int jit_%s%s(const %s *, void *, void *);' % ("t_" if kt else "", nm, sig[nm]))
        fill = ["    o.%s = &jit_%s;" % (nm, nm) for nm in ("p2p", "exact", "m2l_op", "m2l_general", "m2p",
                                                           "p2p_grad", "exact_grad")]
        main = prelude + decl + [
            "",
            "namespace {",
            "K g_k{PyKernel{}, PyGrad{}};",
            "KT g_kt{K{PyKernel{}, PyGrad{}}};",
            "fmma_kernel_ops g_ops, g_ops_t;",
            "bool g_ready = false;",
            "}  // namespace",
            "",
            "// The launcher table (kept alive with this library): the reference's",
            "// kernel form, its symmetry, the translation invariance found in the",
            "// expression, and the transposed kernel's table when f has no symmetry.",
            'extern "C" EXPORT const fmma_kernel_ops *fmma_jit_ops()',
            "{",
            "    if (!g_ready) {",
            "        fmma_kernel_ops &o = g_ops;",
            "        o = fmma_kernel_ops{};",
            "        o.dim = %d;" % D,
            "        o.dtype = %s;" % ("FMMA_F64" if T == "double" else "FMMA_F32"),
            "        o.translation_invariant = %d;" % (1 if self.translation_invariant else 0),
            "        o.symmetry = %d;" % self.symmetry,
            "        o.ctx = &g_k;",
        ] + fill
        if self.symmetry == 0:
            main += [
                "        fmma_kernel_ops &t = g_ops_t;",
                "        t = fmma_kernel_ops{};",
                "        t.dim = %d;" % D,
                "        t.dtype = o.dtype;",
                "        t.translation_invariant = o.translation_invariant;",
                "        t.ctx = &g_kt;",
            ] + ["        t.%s = &jit_t_%s;" % (nm, nm) for nm in ("p2p", "exact", "m2l_op", "m2l_general", "m2p")] + [
                "        o.transposed = &g_ops_t;",
            ]
        main += [
            "        g_ready = true;",
            "    }",
            "    return &g_ops;",
            "}",
            "",
            "// The functor evaluated on the host, to check the trace against Python.",
            'extern "C" EXPORT double fmma_jit_eval(const double *px, const double *py)',
            "{",
            "    %s" % hx,
            "    return %s;" % call,
            "}",
            "",
            'extern "C" EXPORT void fmma_jit_eval_grad(const double *px, const double *py, double *g)',
            "{",
            "    %s" % hx,
            "    const std::array<%s, %d> r = PyGrad{}(%s);" % (T, D, "x, y" if self.nargs == 2 else "d"),
            "    for (int k = 0; k < %d; ++k) g[k] = (double)r[k];" % D,
            "}",
            "",
        ]
        tus["main"] = "\n".join(main)
        return tus


# -------------------------------------------------------------- compiling

_lock = threading.Lock()
_loaded = {}


def cache_dir():
    d = os.environ.get("FMMA_JIT_CACHE")
    if not d:
        base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
        d = os.path.join(base, "fmma_amd_jit")
    try:
        os.makedirs(d, exist_ok=True)
        probe = os.path.join(d, ".probe")
        open(probe, "w").close()
        os.remove(probe)
    except OSError:
        d = os.path.join(tempfile.gettempdir(), "fmma_amd_jit_%d" % os.getuid())
        os.makedirs(d, exist_ok=True)
    return d


def _hipcc(src, out, extra):
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-w", "-fvisibility=hidden", "-I" + os.path.join(ROOT, "include")] + extra + [src, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on the traced kernel %s:\n%s" % (src, r.stderr[-4000:]))


def compile_kernel(spec):
    """The shared library of a traced kernel (compiled once, its launcher
    groups in parallel, then cached on disk by the sources' hash)."""
    import concurrent.futures
    with _lock:
        if spec.key in _loaded:
            return _loaded[spec.key]
        d = cache_dir()
        so = os.path.join(d, "pykernel_%s.so" % spec.key)
        if not os.path.exists(so):
            work = tempfile.mkdtemp(prefix="pykernel_%s_" % spec.key, dir=d)
            objs = []
            jobs = []
            for name, text in spec.sources.items():
                src = os.path.join(work, name + ".hip")
                with open(src, "w") as f:
                    f.write(text)
                obj = os.path.join(work, name + ".o")
                objs.append(obj)
                jobs.append((src, obj))
            workers = max(1, min(len(jobs), os.cpu_count() or 4, 16))
            with concurrent.futures.ThreadPoolExecutor(workers) as ex:
                list(ex.map(lambda j: _hipcc(j[0], j[1], ["-c"]), jobs))
            tmp = so + ".tmp%d" % os.getpid()
            r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs,
                               capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError("linking the traced kernel failed:\n%s" % r.stderr[-4000:])
            os.replace(tmp, so)
        L = ctypes.CDLL(so)
        L.fmma_jit_ops.restype = ctypes.c_void_p
        L.fmma_jit_eval.restype = ctypes.c_double
        L.fmma_jit_eval.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.fmma_jit_eval_grad.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _loaded[spec.key] = L
        return L


def check_against_python(L, fn, spec):
    """The compiled functor against the callable on random pairs: equal to
    1e-13 relative (both in double; libm and the compiler may round a
    function differently by an ulp)."""
    x, y = _sample_pairs(spec.dim, 64, seed=777)
    for a, b in zip(x, y):
        want = _eval_py(fn, spec.nargs, a, b)
        a64 = np.ascontiguousarray(a, dtype=np.float64)
        b64 = np.ascontiguousarray(b, dtype=np.float64)
        got = L.fmma_jit_eval(a64.ctypes.data, b64.ctypes.data)
        if spec.dtype == np.float32:
            want = float(np.float32(_eval_py(fn, spec.nargs, a.astype(np.float32), b.astype(np.float32))))
            got = float(np.float32(got))
            tol = 1e-6
        else:
            tol = 1e-13
        if math.isfinite(want) and not abs(got - want) <= tol * max(abs(want), 1e-300):
            raise RuntimeError("the traced kernel differs from the Python function at x=%s, y=%s: %r vs %r"
                               % (list(a), list(b), got, want))


def kernel_ops_address(fn, dim, dtype, nargs):
    """Address of the fmma_kernel_ops table for the Python kernel fn (traced,
    compiled and checked on first use), for fmma_plan_set_kernel_ops."""
    if not callable(fn):
        raise TypeError("the kernel must be callable")
    spec = KernelSpec(fn, dim, nargs, dtype)
    L = compile_kernel(spec)
    check_against_python(L, fn, spec)
    return int(L.fmma_jit_ops())
