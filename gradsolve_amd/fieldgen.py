"""Fields from Python right-hand sides: trace ``f(t, y, p)`` into an expression graph and
emit the HIP C++ ``struct Field`` that ``gsx_field_register`` compiles into the fused kernels.

This is the build's counterpart of the reference's jaxpr-to-Warp translator
(gradsolve/warp/jax_field.py:393-514): JAX is not available here, so the right-hand side is
traced by calling it once on symbolic values. ``y`` and ``p`` arrive as NumPy object arrays of
symbols, so code written against NumPy (``np.sin(y)``, ``K @ y``, slicing, broadcasting with
float arrays) traces as written; arithmetic with Python and NumPy numbers folds the numbers in
as literals. The emitted field keeps the traced arithmetic order (each node rounds once, as the
host evaluation of the same Python code does), and its derivatives -- the vector-Jacobian
product every adjoint needs, and for stiff fields the Jacobian and the Jacobian's own VJP --
are generated by reverse-mode differentiation of the graph, with the primal subexpressions
shared.

Supported: + - * / unary -, ``**`` with a numeric exponent (integer exponents up to 8 in
magnitude become products, as XLA's integer_pow does) or a symbolic one, abs, the functions
sin cos tan exp log sqrt tanh sinh cosh arctan log1p expm1 square (as NumPy ufuncs or the
functions of this module), comparisons < <= > >= (traced 0/1 values, as the reference's
translator emits its lt/le/gt/ge primitives, on scalars and on the y and p arrays), and
select / max / min / sign (the reference's select, max, min, sign primitives) as NumPy's own
``np.where``, ``np.maximum``, ``np.minimum``, ``np.fmax``, ``np.fmin``, ``np.clip`` and
``np.sign`` (y and p are handed in as ``TracedArray``s, whose NumPy hooks trace these) or this
module's ``where``, ``maximum``, ``minimum`` and ``sign``. Anything that needs a symbol's
truth value (``if``, Python's ``max()``/``min()``) raises ``UnsupportedRHS`` naming the
construct, as the reference's translator does for primitives outside its subset (``cond``).
"""
from __future__ import annotations

import hashlib
import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .fields import UnsupportedRHS

class Graph:
    """Hash-consed expression nodes: equal (op, args, value) give the same node."""

    def __init__(self):
        self._nodes: Dict[tuple, "Sym"] = {}
        self.order: List["Sym"] = []  # creation order is a topological order

    def node(self, op: str, args: Tuple["Sym", ...] = (), value=None) -> "Sym":
        key = (op, tuple(a.id for a in args), value)
        s = self._nodes.get(key)
        if s is None:
            s = Sym(self, op, args, value, len(self.order))
            self._nodes[key] = s
            self.order.append(s)
        return s

    def const(self, x: float) -> "Sym":
        x = float(x)
        # keep -0.0 and 0.0 apart; NaN payloads are not expected in literals
        return self.node("const", (), (x, math.copysign(1.0, x)))

    def var(self, name: str) -> "Sym":
        return self.node("var", (), name)


class Sym:
    """A node of the traced right-hand side (supports Python and NumPy arithmetic)."""

    __slots__ = ("g", "op", "args", "value", "id")
    __array_priority__ = 1000

    def __init__(self, g, op, args, value, id_):
        self.g, self.op, self.args, self.value, self.id = g, op, args, value, id_

    # -- construction helpers
    def _lift(self, x) -> "Sym":
        if isinstance(x, Sym):
            if x.g is not self.g:
                raise UnsupportedRHS("trace", "symbols from two different traces were mixed")
            return x
        if isinstance(x, (bool, np.bool_)):
            raise UnsupportedRHS("bool", "boolean values cannot enter the right-hand side")
        if isinstance(x, (int, float, np.integer, np.floating)):
            return self.g.const(float(x))
        if isinstance(x, np.ndarray) and x.ndim == 0:
            return self._lift(x.item())
        return NotImplemented

    _NP_OPS = {"add": np.add, "sub": np.subtract, "mul": np.multiply, "div": np.true_divide,
               "pow": np.power}

    def _bin(self, op, other, swap=False):
        if isinstance(other, np.ndarray) and other.ndim > 0:
            # array (op) symbol: broadcast the symbol and let NumPy apply the op elementwise
            mine = np.empty(other.shape, dtype=object)
            mine[...] = self
            fn = Sym._NP_OPS[op]
            return fn(other, mine) if swap else fn(mine, other)
        o = self._lift(other)
        if o is NotImplemented:
            return NotImplemented
        a, b = (o, self) if swap else (self, o)
        return self.g.node(op, (a, b))

    def __add__(self, o): return self._bin("add", o)
    def __radd__(self, o): return self._bin("add", o, True)
    def __sub__(self, o): return self._bin("sub", o)
    def __rsub__(self, o): return self._bin("sub", o, True)
    def __mul__(self, o): return self._bin("mul", o)
    def __rmul__(self, o): return self._bin("mul", o, True)
    def __truediv__(self, o): return self._bin("div", o)
    def __rtruediv__(self, o): return self._bin("div", o, True)
    def __neg__(self): return self.g.node("neg", (self,))
    def __pos__(self): return self
    def __abs__(self): return self.g.node("abs", (self,))

    def __pow__(self, o):
        if isinstance(o, np.ndarray) and o.ndim > 0:
            return self._bin("pow", o)
        if isinstance(o, (int, np.integer)) or (isinstance(o, (float, np.floating))
                                                and float(o).is_integer()
                                                and abs(float(o)) <= 8):
            n = int(o)
            if abs(n) <= 8:
                return self.g.node("ipow", (self,), n)
        return self._bin("pow", o)

    def __rpow__(self, o):
        return self._bin("pow", o, True)

    # NumPy object-array ufuncs call these methods elementwise.
    def sin(self): return self.g.node("sin", (self,))
    def cos(self): return self.g.node("cos", (self,))
    def tan(self): return self.g.node("tan", (self,))
    def exp(self): return self.g.node("exp", (self,))
    def log(self): return self.g.node("log", (self,))
    def sqrt(self): return self.g.node("sqrt", (self,))
    def tanh(self): return self.g.node("tanh", (self,))
    def sinh(self): return self.g.node("sinh", (self,))
    def cosh(self): return self.g.node("cosh", (self,))
    def arctan(self): return self.g.node("atan", (self,))
    def log1p(self): return self.g.node("log1p", (self,))
    def expm1(self): return self.g.node("expm1", (self,))
    def square(self): return self.g.node("ipow", (self,), 2)
    def conjugate(self): return self
    def absolute(self): return abs(self)

    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        # NumPy's max/min/sign/clip on a traced scalar trace (see TracedArray); any other
        # ufunc runs NumPy's object loop on 0-d object arrays, which calls this class's
        # operators and methods as it does without this hook.
        if kwargs.get("out", 0) is None:
            kwargs = {k: v for k, v in kwargs.items() if k != "out"}
        if method == "__call__" and not kwargs and ufunc.__name__ in _TRACED_UFUNCS:
            return _TRACED_UFUNCS[ufunc.__name__](*[_plain(x) for x in inputs])
        args = []
        for x in inputs:
            if isinstance(x, Sym):
                box = np.empty((), dtype=object)
                box[()] = x
                x = box
            args.append(_plain(x))
        out = getattr(ufunc, method)(*args, **kwargs)
        if isinstance(out, np.ndarray) and out.ndim == 0 and out.dtype == object:
            return out[()]
        return _traced(out)

    def __array_function__(self, func, types, args, kwargs):
        # np.where / np.clip with a traced scalar among their arguments trace; other NumPy
        # functions run their own implementation.
        return _array_function(func, args, kwargs)

    def _no(self, what):
        raise UnsupportedRHS(what, f"{what} on a traced value: the right-hand side must be "
                                   "straight-line arithmetic; branch with "
                                   "np.where / np.maximum / np.minimum / np.clip "
                                   "instead of if, max() or min()")

    def __bool__(self): self._no("cond")

    # Comparisons are traced values (1.0 where true, 0.0 where false), usable in arithmetic
    # and as the condition of where(); their truth value cannot steer Python control flow.
    def _cmp(self, op, other):
        if isinstance(other, np.ndarray) and other.ndim > 0:
            mine = np.empty(other.shape, dtype=object)
            mine[...] = self
            return np.frompyfunc(lambda a, b: a._cmp(op, b), 2, 1)(mine, other)
        o = self._lift(other)
        if o is NotImplemented:
            return NotImplemented
        return self.g.node(op, (self, o))

    def __lt__(self, o): return self._cmp("lt", o)
    def __le__(self, o): return self._cmp("le", o)
    def __gt__(self, o): return self._cmp("gt", o)
    def __ge__(self, o): return self._cmp("ge", o)
    def __float__(self): self._no("float")
    def __int__(self): self._no("int")
    def __index__(self): self._no("index")
    def __round__(self, *a): self._no("round")
    def __floor__(self): self._no("floor")
    def __ceil__(self): self._no("ceil")

    def __eq__(self, o):  # identity only; hash-consing keeps equal nodes identical
        return self is o

    def __hash__(self):
        return self.id

    def __repr__(self):
        return f"Sym({self.op}#{self.id})"


def _fn(name):
    def f(x):
        if isinstance(x, Sym):
            return getattr(x, name)()
        return getattr(np, name)(x)  # numbers and arrays (object arrays call the methods)
    f.__name__ = name
    return f


def _lift_pair(a, b):
    g = a.g if isinstance(a, Sym) else b.g
    la = a if isinstance(a, Sym) else g.const(float(a))
    lb = b if isinstance(b, Sym) else g.const(float(b))
    return g, la, lb


def where(cond, a, b):
    """Elementwise ``a`` where ``cond`` is nonzero, else ``b`` (traced select; NumPy values
    pass through to ``np.where``)."""
    if isinstance(cond, np.ndarray) and cond.dtype == object:
        return np.frompyfunc(where, 3, 1)(cond, a, b)
    if not isinstance(cond, Sym):
        return np.where(cond, a, b)
    g = cond.g
    la = a if isinstance(a, Sym) else g.const(float(a))
    lb = b if isinstance(b, Sym) else g.const(float(b))
    return g.node("select", (cond, la, lb))


def maximum(a, b):
    """Elementwise maximum (traced; NumPy values pass through to ``np.maximum``)."""
    if any(isinstance(x, np.ndarray) and x.dtype == object for x in (a, b)):
        return np.frompyfunc(maximum, 2, 1)(a, b)
    if not (isinstance(a, Sym) or isinstance(b, Sym)):
        return np.maximum(a, b)
    g, la, lb = _lift_pair(a, b)
    return g.node("max", (la, lb))


def minimum(a, b):
    """Elementwise minimum (traced; NumPy values pass through to ``np.minimum``)."""
    if any(isinstance(x, np.ndarray) and x.dtype == object for x in (a, b)):
        return np.frompyfunc(minimum, 2, 1)(a, b)
    if not (isinstance(a, Sym) or isinstance(b, Sym)):
        return np.minimum(a, b)
    g, la, lb = _lift_pair(a, b)
    return g.node("min", (la, lb))


def sign(x):
    """sign(x) in {-1, 0, 1} (traced; NumPy values pass through to ``np.sign``)."""
    if isinstance(x, np.ndarray) and x.dtype == object:
        return np.frompyfunc(sign, 1, 1)(x)
    if isinstance(x, Sym):
        return x.g.node("sign", (x,))
    return np.sign(x)


def _clip(a, lo, hi):
    """minimum(maximum(a, lo), hi), as jnp.clip computes it."""
    return minimum(maximum(a, lo), hi)


# NumPy ufuncs whose object loops would ask a symbol for its truth value, traced instead (the
# reference translator's max, min, sign and select primitives).
_TRACED_UFUNCS = {"maximum": maximum, "fmax": maximum, "minimum": minimum, "fmin": minimum,
                  "sign": sign, "clip": _clip}


_COMPARISONS = ("greater", "greater_equal", "less", "less_equal")


def _plain(x):
    if isinstance(x, TracedArray):
        return x.view(np.ndarray)
    if isinstance(x, (list, tuple)):
        return type(x)(_plain(v) for v in x)
    return x


def _traced(x):
    if isinstance(x, np.ndarray) and not isinstance(x, TracedArray) and x.dtype == object:
        return x.view(TracedArray)
    if isinstance(x, tuple):
        return tuple(_traced(v) for v in x)
    return x


class TracedArray(np.ndarray):
    """The object arrays a traced right-hand side receives (y and p). NumPy's maximum,
    minimum, fmax, fmin, sign, clip and where on them trace as this module's maximum,
    minimum, sign and where; every other operation is NumPy's object-array arithmetic on the
    symbols."""

    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        args = [_plain(x) for x in inputs]
        if kwargs.get("out", 0) is None:
            kwargs = {k: v for k, v in kwargs.items() if k != "out"}
        if method == "__call__" and not kwargs and ufunc.__name__ in _TRACED_UFUNCS:
            return _traced(_TRACED_UFUNCS[ufunc.__name__](*args))
        if method == "__call__" and not kwargs and ufunc.__name__ in _COMPARISONS:
            # the object-result loop keeps each comparison a traced 0/1 value
            return _traced(ufunc(*args, dtype=object))
        if "out" in kwargs:
            kwargs["out"] = _plain(kwargs["out"])
        return _traced(getattr(ufunc, method)(*args, **kwargs))

    def __array_function__(self, func, types, args, kwargs):
        return _array_function(func, args, kwargs)


def _array_function(func, args, kwargs):
    args, kwargs = _plain(args), {k: _plain(v) for k, v in kwargs.items()}
    kwargs = {k: v for k, v in kwargs.items() if not (k == "out" and v is None)}
    if func is np.where and len(args) == 3 and not kwargs:
        return _traced(where(*args))
    if func is np.clip and len(args) == 3 and not kwargs:
        return _traced(_clip(*args))
    impl = getattr(func, "_implementation", func)
    return _traced(impl(*args, **kwargs))



sin, cos, tan, exp, log, sqrt = (_fn(n) for n in ("sin", "cos", "tan", "exp", "log", "sqrt"))
tanh, sinh, cosh, arctan, log1p, expm1, square = (
    _fn(n) for n in ("tanh", "sinh", "cosh", "arctan", "log1p", "expm1", "square"))


# ---------------------------------------------------------------------------------------
# Tracing
# ---------------------------------------------------------------------------------------

class Traced:
    """The traced field: outputs over inputs t, y[0..D-1], p[0..P-1]."""

    def __init__(self, g: Graph, out: List[Sym], dim: int, n_params: int):
        self.g, self.out, self.dim, self.n_params = g, out, dim, n_params
        self.t = g.var("t")
        self.y = [g.var(f"y{i}") for i in range(dim)]
        self.p = [g.var(f"p{q}") for q in range(n_params)]

    def reaches(self, roots: Sequence[Sym], target: Sym) -> bool:
        seen, stack = set(), list(roots)
        while stack:
            s = stack.pop()
            if s is target:
                return True
            if s.id in seen:
                continue
            seen.add(s.id)
            stack.extend(s.args)
        return False

    @property
    def time_dependent(self) -> bool:
        return self.reaches(self.out, self.t)


def trace(f, dim: int, n_params: int) -> Traced:
    """Call ``f(t, y, p)`` once on symbols and collect its ``dim`` outputs."""
    g = Graph()
    t = g.var("t")
    y = np.empty(dim, dtype=object)
    for i in range(dim):
        y[i] = g.var(f"y{i}")
    p = np.empty(n_params, dtype=object)
    for q in range(n_params):
        p[q] = g.var(f"p{q}")
    y, p = y.view(TracedArray), p.view(TracedArray)
    try:
        out = f(t, y, p)
    except UnsupportedRHS:
        raise
    except TypeError as exc:
        raise UnsupportedRHS("trace", f"the right-hand side could not be traced: {exc}") from exc
    arr = np.asarray(_plain(out), dtype=object).reshape(-1)
    if arr.shape[0] != dim:
        raise UnsupportedRHS("shape", f"f returns {arr.shape[0]} components, expected "
                                      f"dim={dim}")
    outs = []
    for v in arr:
        if isinstance(v, Sym):
            outs.append(v)
        elif isinstance(v, (int, float, np.integer, np.floating)):
            outs.append(g.const(float(v)))
        else:
            raise UnsupportedRHS("shape", f"f returned a component of type {type(v).__name__}")
    return Traced(g, outs, dim, n_params)


# ---------------------------------------------------------------------------------------
# Reverse-mode differentiation on the graph (derivative nodes join the same graph)
# ---------------------------------------------------------------------------------------

def _is_const(s: Sym, v: Optional[float] = None) -> bool:
    return s.op == "const" and (v is None or s.value[0] == v)


def _add(g, a, b):
    if a is None:
        return b
    if b is None:
        return a
    if _is_const(a, 0.0):
        return b
    if _is_const(b, 0.0):
        return a
    return g.node("add", (a, b))


def _mul(g, a, b):
    if _is_const(a, 1.0):
        return b
    if _is_const(b, 1.0):
        return a
    if _is_const(a, -1.0):
        return _neg(g, b)
    if _is_const(b, -1.0):
        return _neg(g, a)
    if a.op == "const" and b.op == "const":
        return g.const(a.value[0] * b.value[0])
    if a.op == "neg" and b.op == "neg":  # (-x)(-y) = xy exactly
        return _mul(g, a.args[0], b.args[0])
    return g.node("mul", (a, b))


def _neg(g, a):
    if a.op == "neg":
        return a.args[0]
    if a.op == "const":
        return g.const(-a.value[0])
    return g.node("neg", (a,))


def _local(g: Graph, s: Sym, gbar: Sym) -> List[Tuple[Sym, Sym]]:
    """(argument, contribution) pairs of the adjoint gbar of node s."""
    op, a = s.op, s.args
    c = g.const
    if op == "add":
        return [(a[0], gbar), (a[1], gbar)]
    if op == "sub":
        return [(a[0], gbar), (a[1], _neg(g, gbar))]
    if op == "mul":
        return [(a[0], _mul(g, gbar, a[1])), (a[1], _mul(g, gbar, a[0]))]
    if op == "div":  # s = a0 / a1: d/da0 = 1/a1, d/da1 = -s/a1
        q = g.node("div", (gbar, a[1]))
        return [(a[0], q), (a[1], _neg(g, _mul(g, q, s)))]
    if op == "neg":
        return [(a[0], _neg(g, gbar))]
    x = a[0] if a else None
    if op == "sin":
        return [(x, _mul(g, gbar, g.node("cos", (x,))))]
    if op == "cos":
        return [(x, _neg(g, _mul(g, gbar, g.node("sin", (x,)))))]
    if op == "tan":  # 1 + tan^2
        return [(x, _mul(g, gbar, _add(g, c(1.0), g.node("mul", (s, s)))))]
    if op == "exp":
        return [(x, _mul(g, gbar, s))]
    if op == "expm1":
        return [(x, _mul(g, gbar, _add(g, s, c(1.0))))]
    if op == "log":
        return [(x, g.node("div", (gbar, x)))]
    if op == "log1p":
        return [(x, g.node("div", (gbar, _add(g, c(1.0), x))))]
    if op == "sqrt":  # 0.5 / sqrt(x)
        return [(x, g.node("div", (_mul(g, c(0.5), gbar), s)))]
    if op == "tanh":  # 1 - tanh^2
        return [(x, _mul(g, gbar, g.node("sub", (c(1.0), g.node("mul", (s, s))))))]
    if op == "sinh":
        return [(x, _mul(g, gbar, g.node("cosh", (x,))))]
    if op == "cosh":
        return [(x, _mul(g, gbar, g.node("sinh", (x,))))]
    if op == "atan":
        return [(x, g.node("div", (gbar, _add(g, c(1.0), g.node("mul", (x, x))))))]
    if op == "abs":
        return [(x, _mul(g, gbar, g.node("sign", (x,))))]
    if op in ("sign", "lt", "le", "gt", "ge"):
        return []
    if op == "select":  # cond carries no gradient
        z = c(0.0)
        return [(a[1], g.node("select", (a[0], gbar, z))),
                (a[2], g.node("select", (a[0], z, gbar)))]
    if op in ("max", "min"):  # the larger (smaller) argument takes the cotangent; ties split
        u, w = a
        u_wins = g.node("gt" if op == "max" else "lt", (u, w))
        w_wins = g.node("lt" if op == "max" else "gt", (u, w))
        share_u = g.node("select", (u_wins, c(1.0), g.node("select", (w_wins, c(0.0), c(0.5)))))
        share_w = g.node("select", (w_wins, c(1.0), g.node("select", (u_wins, c(0.0), c(0.5)))))
        return [(u, _mul(g, gbar, share_u)), (w, _mul(g, gbar, share_w))]
    if op == "ipow":
        n = s.value
        if n == 0:
            return []
        if n == 1:
            return [(x, gbar)]
        lower = x if n == 2 else g.node("ipow", (x,), n - 1)
        return [(x, _mul(g, gbar, _mul(g, c(float(n)), lower)))]
    if op == "pow":
        base, ex = a
        out = []
        if ex.op == "const":
            e = ex.value[0]
            d = _mul(g, c(e), g.node("pow", (base, c(e - 1.0))))
            out.append((base, _mul(g, gbar, d)))
        else:  # d/dbase = ex * base^(ex-1) = ex * s / base; d/dex = s * log(base)
            out.append((base, _mul(g, gbar, g.node("div", (_mul(g, ex, s), base)))))
            out.append((ex, _mul(g, gbar, _mul(g, s, g.node("log", (base,))))))
        return out
    raise UnsupportedRHS(op, f"no derivative rule for {op}")


def vjp_nodes(g: Graph, roots: Sequence[Sym], seeds: Sequence[Sym],
              wrt: Sequence[Sym]) -> List[Optional[Sym]]:
    """sum_i seeds[i] * d roots[i] / d wrt[k] for every k (None where identically zero)."""
    need = set()
    stack = list(roots)
    while stack:
        s = stack.pop()
        if s.id in need:
            continue
        need.add(s.id)
        stack.extend(s.args)
    adj: Dict[int, Sym] = {}
    for r, sd in zip(roots, seeds):
        adj[r.id] = _add(g, adj.get(r.id), sd)
    # graph.order is topological; walk it backwards over the nodes the roots depend on
    for s in reversed(list(g.order)):
        if s.id not in need or s.id not in adj or not s.args:
            continue
        gbar = adj[s.id]
        for arg, contrib in _local(g, s, gbar):
            if arg.op == "const":
                continue
            adj[arg.id] = _add(g, adj.get(arg.id), contrib)
    return [adj.get(w.id) for w in wrt]


# ---------------------------------------------------------------------------------------
# C++ emission
# ---------------------------------------------------------------------------------------

_CFUN = {"sin": "std::sin", "cos": "std::cos", "tan": "std::tan", "exp": "std::exp",
         "log": "std::log", "sqrt": "std::sqrt", "tanh": "std::tanh", "sinh": "std::sinh",
         "cosh": "std::cosh", "atan": "std::atan", "log1p": "std::log1p",
         "expm1": "std::expm1", "abs": "std::fabs"}


def _literal(x: float) -> str:
    if math.isinf(x):
        return "(T(1.0) / T(0.0))" if x > 0 else "(-T(1.0) / T(0.0))"
    if math.isnan(x):
        raise UnsupportedRHS("const", "NaN literal in the right-hand side")
    return f"T({float.hex(x)})"


class _Emitter:
    def __init__(self, names: Dict[int, str]):
        self.names = dict(names)
        self.lines: List[str] = []

    def ref(self, s: Sym) -> str:
        if s.op == "const":
            return _literal(s.value[0])
        return self.names[s.id]

    def emit(self, roots: Sequence[Sym]):
        need = set()
        stack = [r for r in roots if r is not None]
        while stack:
            s = stack.pop()
            if s.id in need or s.id in self.names or s.op == "const":
                continue
            need.add(s.id)
            stack.extend(s.args)
        for s in s_order(roots, need):
            if s.op == "var":
                raise UnsupportedRHS("var", f"unbound input {s.value}")
            self.names[s.id] = f"t{s.id}"
            self.lines.append(f"    const T t{s.id} = {self._expr(s)};")

    def _expr(self, s: Sym) -> str:
        a = [self.ref(x) for x in s.args]
        op = s.op
        if op == "add":
            return f"{a[0]} + {a[1]}"
        if op == "sub":
            return f"{a[0]} - {a[1]}"
        if op == "mul":
            return f"{a[0]} * {a[1]}"
        if op == "div":
            return f"{a[0]} / {a[1]}"
        if op == "neg":
            return f"-{a[0]}"
        if op == "pow":
            return f"std::pow({a[0]}, {a[1]})"
        if op == "ipow":
            n = s.value
            if n == 0:
                return "T(1.0)"
            prod = " * ".join([a[0]] * abs(n))
            return prod if n > 0 else f"T(1.0) / ({prod})"
        if op == "sign":  # sign(0) = 0, the derivative JAX takes for |x| at 0
            return f"T(({a[0]} > T(0.0)) - ({a[0]} < T(0.0)))"
        if op in ("lt", "le", "gt", "ge"):
            sym = {"lt": "<", "le": "<=", "gt": ">", "ge": ">="}[op]
            return f"({a[0]} {sym} {a[1]} ? T(1.0) : T(0.0))"
        if op == "select":
            return f"({a[0]} != T(0.0) ? {a[1]} : {a[2]})"
        if op == "max":  # NaN-propagating, as jnp.maximum
            return f"({a[0]} != {a[0]} || {a[0]} > {a[1]} ? {a[0]} : {a[1]})"
        if op == "min":
            return f"({a[0]} != {a[0]} || {a[0]} < {a[1]} ? {a[0]} : {a[1]})"
        if op in _CFUN:
            return f"{_CFUN[op]}({a[0]})"
        raise UnsupportedRHS(op, f"cannot emit {op}")


def s_order(roots, need):
    g = next(r.g for r in roots if r is not None)
    return [s for s in g.order if s.id in need]


def _inputs(tr: Traced, extra: Dict[str, str]) -> Dict[int, str]:
    names = {tr.t.id: "t"}
    for i, s in enumerate(tr.y):
        names[s.id] = f"y[{i}]"
    for q, s in enumerate(tr.p):
        names[s.id] = f"p[{q}]"
    for k, v in extra.items():
        names[tr.g.var(k).id] = v
    return names


def _emit_time_derivative(tr: Traced, wrt) -> list:
    """``dfdt`` (df/dt at (t, y)) and ``dfdt_vjp`` (yb += d/dy ftb . df/dt, pb likewise) of
    a stiff field that reads t: the df/dt term of the Rodas5P stages
    (gradsolve/solvers/rodas5p_step.py:119-146) and its adjoint."""
    g, D = tr.g, tr.dim
    one = g.const(1.0)
    ft = [vjp_nodes(g, [tr.out[i]], [one], [tr.t])[0] for i in range(D)]
    em = _Emitter(_inputs(tr, {}))
    em.emit([x for x in ft if x is not None])
    lines = ["  template <class T>",
             "  GSX_HD static void dfdt(T t, const T* y, const T* p, T* ft, const double*) {",
             "    (void)t; (void)y; (void)p;"]
    lines += em.lines
    lines += [f"    ft[{i}] = {em.ref(x) if x is not None else 'T(0.0)'};"
              for i, x in enumerate(ft)]
    lines += ["  }"]
    roots, seeds = [], []
    for i, x in enumerate(ft):
        if x is not None and x.op != "const":
            roots.append(x)
            seeds.append(g.var(f"ftb{i}"))
    fg = vjp_nodes(g, roots, seeds, wrt) if roots else [None] * len(wrt)
    em = _Emitter(_inputs(tr, {f"ftb{i}": f"ftb[{i}]" for i in range(D)}))
    em.emit([x for x in fg if x is not None])
    lines += ["  template <class T>",
              "  GSX_HD static void dfdt_vjp(T t, const T* y, const T* p, const T* ftb, T* yb, "
              "T* pb,",
              "                              const double*) {",
              "    (void)t; (void)y; (void)p; (void)ftb; (void)yb; (void)pb;"]
    lines += em.lines
    for k, gk in enumerate(fg):
        if gk is None:
            continue
        tgt = f"yb[{k}]" if k < D else f"pb[{k - D}]"
        lines.append(f"    {tgt} += {em.ref(gk)};")
    lines += ["  }"]
    return lines


def emit_field(tr: Traced, stiff: bool = False) -> str:
    """The ``struct Field`` source for ``gsx_field_register``."""
    g, D = tr.g, tr.dim
    P = max(tr.n_params, 1)
    autonomous = not tr.time_dependent
    lines = ["struct Field {",
             f"  static constexpr int D = {D}, P = {P};",
             f"  static constexpr bool autonomous = {'true' if autonomous else 'false'};"]
    # f
    em = _Emitter(_inputs(tr, {}))
    em.emit(tr.out)
    # y / dy (and v / yb below) are any indexable views: plain arrays in the register lanes,
    # strided workspace rows in the wide lanes
    lines += ["  template <class T, class YP, class DP>",
              "  GSX_HD static void f(T t, YP y, const T* p, DP dy, const double*) {",
              "    (void)t; (void)y; (void)p;"]
    lines += em.lines
    lines += [f"    dy[{i}] = {em.ref(o)};" for i, o in enumerate(tr.out)]
    lines += ["  }"]
    # vjp: yb += v^T df/dy, pb += v^T df/dp
    v = [g.var(f"v{i}") for i in range(D)]
    wrt = list(tr.y) + list(tr.p)
    grads = vjp_nodes(g, tr.out, v, wrt)
    em = _Emitter(_inputs(tr, {f"v{i}": f"v[{i}]" for i in range(D)}))
    em.emit([x for x in grads if x is not None])
    lines += ["  template <class T, class YP, class VP, class YBP>",
              "  GSX_HD static void vjp(T t, YP y, const T* p, VP v, YBP yb, T* pb,",
              "                         const double*) {",
              "    (void)t; (void)y; (void)p; (void)v; (void)yb; (void)pb;"]
    lines += em.lines
    for k, gk in enumerate(grads):
        if gk is None:
            continue
        tgt = f"yb[{k}]" if k < D else f"pb[{k - D}]"
        lines.append(f"    {tgt} += {em.ref(gk)};")
    lines += ["  }"]
    if stiff:
        one = g.const(1.0)
        J = []
        for i in range(D):
            row = vjp_nodes(g, [tr.out[i]], [one], list(tr.y))
            J.append(row)
        em = _Emitter(_inputs(tr, {}))
        em.emit([x for row in J for x in row if x is not None])
        lines += ["  template <class T>",
                  "  GSX_HD static void jac(T t, const T* y, const T* p, T* J, const double*) {",
                  "    (void)t; (void)y; (void)p;"]
        lines += em.lines
        for i in range(D):
            for j in range(D):
                x = J[i][j]
                lines.append(f"    J[{i * D + j}] = {em.ref(x) if x is not None else 'T(0.0)'};")
        lines += ["  }"]
        # jac_vjp: yb += d/dy sum Jb_ij J_ij, pb += d/dp sum Jb_ij J_ij
        roots, seeds = [], []
        for i in range(D):
            for j in range(D):
                if J[i][j] is not None and J[i][j].op != "const":
                    roots.append(J[i][j])
                    seeds.append(g.var(f"Jb{i * D + j}"))
        jg = vjp_nodes(g, roots, seeds, wrt) if roots else [None] * len(wrt)
        em = _Emitter(_inputs(tr, {f"Jb{k}": f"Jb[{k}]" for k in range(D * D)}))
        em.emit([x for x in jg if x is not None])
        lines += ["  template <class T>",
                  "  GSX_HD static void jac_vjp(T t, const T* y, const T* p, const T* Jb, "
                  "T* yb, T* pb,",
                  "                             const double*) {",
                  "    (void)t; (void)y; (void)p; (void)Jb; (void)yb; (void)pb;"]
        lines += em.lines
        for k, gk in enumerate(jg):
            if gk is None:
                continue
            tgt = f"yb[{k}]" if k < D else f"pb[{k - D}]"
            lines.append(f"    {tgt} += {em.ref(gk)};")
        lines += ["  }"]
        if not autonomous:
            lines += _emit_time_derivative(tr, wrt)
    lines += ["};", ""]
    return "\n".join(lines)


def source_key(src: str) -> str:
    return hashlib.sha256(src.encode()).hexdigest()[:16]
