"""User-defined point kernels: one formula, three evaluators (numpy, torch, run-time compiled HIP).

The reference keeps its ``KernelOperator`` plugin open through PyKeOps formulas
(``skelfmm/operators.py:17-108``, lazy kernels ``skelfmm/util_batched.py:189-224``). Here a
user operator is a real scalar formula of the pair ``(x, y)`` written in a small expression
language; ``operators.custom_operator`` turns it into

* a numpy dense kernel (the host build's proxy matrices and the subset direct check),
* a torch dense block apply (``torch_apply``), and
* the near-field kernel of the device apply: ``kernel_value<>`` specialized for the
  formula and compiled with hiprtc together with the library's own one-sided near-field
  kernel (``csrc/p2p_core.h``), for FP64 and the FP32 precision form and channel blocks of
  1, 2, 4 and 8 (``skf_jit_compile``).

Expression language: numbers; ``+ - * / **``; the names ``r2`` (squared distance), ``r``,
``rinv`` (``1/r``), ``dx dy dz`` (components of ``x - y``; ``dz`` is 0 in 2D), ``param`` (the
operator parameter, ``SkelFMM(..., param=...)``) and ``pi``; the functions ``exp log sqrt
sin cos tan tanh erf abs min max``. Pairs at distance zero (the self pair) contribute 0,
as for the built-in kernels; on the device ``r2`` is clamped to at least 1e-15 (FP64) or
biased by 1e-20 (FP32) as the built-in single layers are.
"""

from __future__ import annotations

import ast
import hashlib
import math
import os

import numpy as np

__all__ = ["UserKernel", "parse_expression", "kernel_source", "compile_kernels"]

HEADER_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "p2p_core.h")
VARIABLES = ("r2", "r", "rinv", "dx", "dy", "dz", "param")
FUNCTIONS = {"exp": ("exp", np.exp), "log": ("log", np.log), "sqrt": ("sqrt", np.sqrt), "sin": ("sin", np.sin),
             "cos": ("cos", np.cos), "tan": ("tan", np.tan), "tanh": ("tanh", np.tanh), "abs": ("fabs", np.abs),
             "min": ("fmin", np.minimum), "max": ("fmax", np.maximum), "erf": ("erf", None)}
_ARITY = {"min": 2, "max": 2}
KID_USER_2D = 16
KID_USER_3D = 17
CHANNEL_BLOCKS = (1, 2, 4, 8)


def _erf_np(x):
    from scipy.special import erf

    return erf(x)


def parse_expression(expr):
    """Validate ``expr`` against the expression language; returns its ``ast.Expression``."""
    if not isinstance(expr, str) or not expr.strip():
        raise ValueError("the kernel expression must be a non-empty string")
    try:
        tree = ast.parse(expr.strip(), mode="eval")
    except SyntaxError as e:
        raise ValueError(f"kernel expression does not parse: {e.msg}") from None
    used = set()
    for node in ast.walk(tree):
        if isinstance(node, (ast.Expression, ast.Load, ast.Add, ast.Sub, ast.Mult, ast.Div, ast.Pow, ast.USub,
                             ast.UAdd)):
            continue
        if isinstance(node, ast.BinOp):
            if not isinstance(node.op, (ast.Add, ast.Sub, ast.Mult, ast.Div, ast.Pow)):
                raise ValueError(f"operator {type(node.op).__name__} is not allowed")
            continue
        if isinstance(node, ast.UnaryOp):
            if not isinstance(node.op, (ast.USub, ast.UAdd)):
                raise ValueError(f"operator {type(node.op).__name__} is not allowed")
            continue
        if isinstance(node, ast.Constant):
            if isinstance(node.value, bool) or not isinstance(node.value, (int, float)):
                raise ValueError(f"constant {node.value!r} is not a real number")
            continue
        if isinstance(node, ast.Name):
            if node.id not in VARIABLES and node.id != "pi" and node.id not in FUNCTIONS:
                raise ValueError(f"unknown name {node.id!r} (allowed: {', '.join(VARIABLES)}, pi)")
            used.add(node.id)
            continue
        if isinstance(node, ast.Call):
            if not isinstance(node.func, ast.Name) or node.func.id not in FUNCTIONS:
                raise ValueError("only the functions " + ", ".join(sorted(FUNCTIONS)) + " may be called")
            if node.keywords or len(node.args) != _ARITY.get(node.func.id, 1):
                raise ValueError(f"{node.func.id} takes {_ARITY.get(node.func.id, 1)} positional argument(s)")
            continue
        raise ValueError(f"{type(node).__name__} is not allowed in a kernel expression")
    for node in ast.walk(tree):  # function names only in call position, variables never called
        if isinstance(node, ast.Name) and node.id in FUNCTIONS:
            if not any(isinstance(c, ast.Call) and c.func is node for c in ast.walk(tree)):
                raise ValueError(f"{node.id} must be called")
    return tree


def _to_c(node):
    """C++ for one validated node; literals become ``R(...)`` so the formula stays in the kernel's precision."""
    if isinstance(node, ast.Expression):
        return _to_c(node.body)
    if isinstance(node, ast.Constant):
        return f"R({float(node.value)!r})"
    if isinstance(node, ast.Name):
        return "R(3.14159265358979323846)" if node.id == "pi" else node.id
    if isinstance(node, ast.UnaryOp):
        return ("-" if isinstance(node.op, ast.USub) else "+") + "(" + _to_c(node.operand) + ")"
    if isinstance(node, ast.BinOp):
        a, b = _to_c(node.left), _to_c(node.right)
        if isinstance(node.op, ast.Pow):
            e = node.right
            if isinstance(e, ast.Constant) and float(e.value).is_integer() and 1 <= abs(int(e.value)) <= 4:
                n = abs(int(e.value))
                prod = "*".join(["(" + a + ")"] * n)
                return f"({prod})" if e.value > 0 else f"(R(1) / ({prod}))"
            return f"pow({a}, {b})"
        op = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/"}[type(node.op)]
        return f"({a} {op} {b})"
    if isinstance(node, ast.Call):
        name = FUNCTIONS[node.func.id][0]
        return f"{name}(" + ", ".join(_to_c(a) for a in node.args) + ")"
    raise ValueError("unexpected node")


def _to_py(node):
    """Python source (numpy ufuncs under the names ``_f_<fn>``) for one validated node."""
    if isinstance(node, ast.Expression):
        return _to_py(node.body)
    if isinstance(node, ast.Constant):
        return repr(float(node.value))
    if isinstance(node, ast.Name):
        return repr(math.pi) if node.id == "pi" else node.id
    if isinstance(node, ast.UnaryOp):
        return ("-" if isinstance(node.op, ast.USub) else "+") + "(" + _to_py(node.operand) + ")"
    if isinstance(node, ast.BinOp):
        op = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/", ast.Pow: "**"}[type(node.op)]
        return f"({_to_py(node.left)} {op} {_to_py(node.right)})"
    if isinstance(node, ast.Call):
        return f"_f_{node.func.id}(" + ", ".join(_to_py(a) for a in node.args) + ")"
    raise ValueError("unexpected node")


class UserKernel:
    """A validated kernel formula with its numpy and torch evaluators and its HIP source."""

    def __init__(self, expression, ndim):
        if ndim not in (2, 3):
            raise ValueError("ndim must be 2 or 3")
        self.tree = parse_expression(expression)
        self.expression = expression.strip()
        self.ndim = int(ndim)
        self.kernel_id = KID_USER_3D if ndim == 3 else KID_USER_2D
        self.names_used = {n.id for n in ast.walk(self.tree) if isinstance(n, ast.Name)}
        self.c_expression = _to_c(self.tree)
        self._py = compile(_to_py(self.tree), "<kernel expression>", "eval")
        self.key = hashlib.sha1(f"{self.ndim}:{self.c_expression}".encode()).hexdigest()[:16]

    def _eval(self, lib, names):
        env = {"__builtins__": {}}
        env.update(names)
        for fn, (_, npf) in FUNCTIONS.items():
            env["_f_" + fn] = lib[fn]
        return eval(self._py, env)  # the code object comes from a validated expression only

    def numpy_values(self, diff, r2, param):
        """Kernel values for difference vectors ``diff`` (..., d) and ``r2`` (> 0 where used)."""
        r = np.sqrt(r2)
        with np.errstate(divide="ignore", invalid="ignore"):
            rinv = 1.0 / r
            lib = {fn: (npf if npf is not None else _erf_np) for fn, (_, npf) in FUNCTIONS.items()}
            dz = diff[..., 2] if self.ndim == 3 else np.zeros_like(r2)
            v = self._eval(lib, {"r2": r2, "r": r, "rinv": rinv, "dx": diff[..., 0], "dy": diff[..., 1],
                                 "dz": dz, "param": float(param)})
        return np.broadcast_to(np.asarray(v, dtype=np.float64), r2.shape)

    def torch_values(self, diff, r2, param):
        import torch

        r = torch.sqrt(r2)
        lib = {"exp": torch.exp, "log": torch.log, "sqrt": torch.sqrt, "sin": torch.sin, "cos": torch.cos,
               "tan": torch.tan, "tanh": torch.tanh, "abs": torch.abs, "min": torch.minimum,
               "max": torch.maximum, "erf": torch.erf}
        dz = diff[..., 2] if self.ndim == 3 else torch.zeros_like(r2)

        def wrap(f):  # numbers in the formula (e.g. exp(1.0)) become tensors
            return lambda *a: f(*[x if torch.is_tensor(x) else torch.as_tensor(x, dtype=r2.dtype,
                                                                               device=r2.device) for x in a])

        v = self._eval({k: wrap(f) for k, f in lib.items()},
                       {"r2": r2, "r": r, "rinv": 1.0 / r, "dx": diff[..., 0], "dy": diff[..., 1], "dz": dz,
                        "param": float(param)})
        v = v if torch.is_tensor(v) else torch.as_tensor(v, dtype=r2.dtype, device=r2.device)
        return v.expand_as(r2)

    # --------------------------------------------------------------- device
    def source(self):
        return kernel_source(self)

    def names(self, precision):
        r, prec = ("double", "false") if precision == 0 else ("float", "true")
        return [f"p2p_kernel<{r}, {self.kernel_id}, {n}, {prec}>" for n in CHANNEL_BLOCKS]


def kernel_source(uk):
    """Translation unit for hiprtc: the formula as ``kernel_value<R, id>`` + the shared near-field kernel."""
    kid = uk.kernel_id
    return f"""// near field of a user-defined operator: K = {uk.expression}
#include "p2p_core.h"

template <> struct OpT<{kid}> {{ static constexpr int dim = {uk.ndim}, cplx = 0, geom = 0; }};

template <typename R>
__device__ inline R skf_user_kernel(R dx, R dy, R dz, R r2, R r, R rinv, R param) {{
  return {uk.c_expression};
}}

template <>
__device__ inline Cplx<double> kernel_value<double, {kid}>(double dx, double dy, double dz, double, double param) {{
  const double r2 = fmax(dx * dx + dy * dy + dz * dz, 1e-15);
  const double rinv = Mth<double>::rsq(r2);
  return Cplx<double>{{skf_user_kernel<double>(dx, dy, dz, r2, r2 * rinv, rinv, param), 0.0}};
}}

template <>
__device__ inline Cplx<float> kernel_value<float, {kid}>(float dx, float dy, float dz, float, float param) {{
  float r2 = fmaf(dx, dx, 1e-20f);
  r2 = fmaf(dy, dy, r2);
  r2 = fmaf(dz, dz, r2);
  const float rinv = Mth<float>::rsq(r2);
  return Cplx<float>{{skf_user_kernel<float>(dx, dy, dz, r2, r2 * rinv, rinv, param), 0.0f}};
}}
"""


_MODULES = {}


def _header():
    with open(HEADER_PATH, encoding="utf-8") as fh:
        return fh.read()


def compile_kernels(uk, precision, load=True):
    """hipFunction_t handles (channel blocks 1, 2, 4, 8) of ``uk``'s near field for ``precision``.

    Compiled once per process and formula; ``load=False`` only checks that the source
    compiles (no device needed) and returns the code size in bytes."""
    import ctypes

    from . import hip

    L = hip.lib()
    names = uk.names(precision)
    arr = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
    size = ctypes.c_longlong(0)
    if not load:
        rc = L.skf_jit_compile(uk.source().encode(), b"p2p_core.h", _header().encode(), arr, len(names), None, None,
                               ctypes.byref(size))
        if rc != 0:
            raise RuntimeError("kernel compilation failed: " + L.skf_jit_log().decode(errors="replace"))
        return int(size.value)
    key = (uk.key, int(precision))
    if key not in _MODULES:
        hip.require_gpu()
        mod = ctypes.c_void_p()
        fns = (ctypes.c_void_p * len(names))()
        rc = L.skf_jit_compile(uk.source().encode(), b"p2p_core.h", _header().encode(), arr, len(names),
                               ctypes.byref(mod), fns, ctypes.byref(size))
        if rc != 0:
            raise RuntimeError("kernel compilation failed: " + L.skf_last_error().decode(errors="replace") + "\n" +
                               L.skf_jit_log().decode(errors="replace"))
        _MODULES[key] = (mod, [fns[i] for i in range(len(names))])
    return _MODULES[key][1]

