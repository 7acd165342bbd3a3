"""User-written field models on the device: the reference's plugin contract
(subclass ``_Field``, override ``B`` and/or ``E``; rapt/fields.py:5-74 and
examples/Creating new fields.ipynb) traced into a small program.

``trace(field)`` calls the subclass's own ``B(tpos)`` and ``E(tpos)`` once
with symbolic ``tpos = [t, x, y, z]``.  Arithmetic (``+ - * / **``, unary
minus, ``abs``, Python's ``pow``) and numpy's elementwise functions
(``np.sqrt``, ``np.exp``, ``np.tanh``, ``np.arctan2``, ``np.maximum``,
``np.minimum``, ``np.clip``, ``np.sign``, ...) on those symbols, and on numpy
arrays built from them (``np.dot``, ``np.cross``, ``np.linalg.norm``,
``np.sum``), record an expression graph; Python
numbers are constants, combined as Python combines them.  The graph becomes a
register program (four int32 per instruction: opcode, destination, operand a,
operand b) that the device kernels and the oracle evaluate in double
precision in numpy's operation order, so a field written for the reference
runs unchanged on the GPU.  What cannot be traced: branches on field values
(``if z > 0``), conversions of field values to Python floats, and calls
outside numpy's elementwise math; those raise ``NotImplementedError`` naming
the limitation.
"""
import math

import numpy as np

# opcodes shared with rapt_fields.hpp (expr_eval) and the oracle
OPS = {
    "const": 0, "add": 1, "sub": 2, "mul": 3, "div": 4, "neg": 5, "pow": 6, "sqrt": 7,
    "exp": 8, "log": 9, "sin": 10, "cos": 11, "tan": 12, "tanh": 13, "sinh": 14, "cosh": 15,
    "arctan": 16, "arcsin": 17, "arccos": 18, "abs": 19, "arctan2": 20, "mov": 21,
    "log10": 22, "hypot": 23, "maximum": 24, "minimum": 25, "sign": 26,
}
NREG = 32          # registers of the device interpreter (0-3 hold t, x, y, z; 4-6 results)
MAXCODE = 256      # instructions of both programs together
MAXCONST = 128

_NUMPY = {"add": np.add, "sub": np.subtract, "mul": np.multiply, "div": np.divide,
          "pow": np.power, "arctan2": np.arctan2, "hypot": np.hypot, "neg": np.negative,
          "sqrt": np.sqrt, "exp": np.exp, "log": np.log, "log10": np.log10, "sin": np.sin,
          "cos": np.cos, "tan": np.tan, "tanh": np.tanh, "sinh": np.sinh, "cosh": np.cosh,
          "arctan": np.arctan, "arcsin": np.arcsin, "arccos": np.arccos, "abs": np.abs,
          "maximum": np.maximum, "minimum": np.minimum, "sign": np.sign}
_UFUNC = {"true_divide": "div", "divide": "div", "add": "add", "subtract": "sub",
          "multiply": "mul", "power": "pow", "negative": "neg", "absolute": "abs",
          "fabs": "abs", "clip": "clip"}


def _num(v):
    return isinstance(v, (int, float, np.floating, np.integer)) and not isinstance(v, bool)


def _apply(name, *args):
    """One operation: on numbers it is evaluated as numpy would (float64);
    otherwise it becomes a node of the graph."""
    if name == "clip":  # np.clip(a, lo, hi) is minimum(maximum(a, lo), hi) elementwise
        return _apply("minimum", _apply("maximum", args[0], args[1]), args[2])
    if all(_num(a) for a in args):
        return float(_NUMPY[name](*[np.float64(a) for a in args]))
    for a in args:
        if not (_num(a) or isinstance(a, Sym)):
            raise NotImplementedError("unsupported value in a traced field: %r" % (type(a),))
    return Sym(name, args)


class Sym:
    """A traced scalar: an operation on other Syms and constants."""
    __array_priority__ = 1000

    def __init__(self, op, args=(), value=None):
        self.op, self.args, self.value = op, tuple(args), value

    def __add__(self, o): return _binop("add", self, o)
    def __radd__(self, o): return _binop("add", o, self)
    def __sub__(self, o): return _binop("sub", self, o)
    def __rsub__(self, o): return _binop("sub", o, self)
    def __mul__(self, o): return _binop("mul", self, o)
    def __rmul__(self, o): return _binop("mul", o, self)
    def __truediv__(self, o): return _binop("div", self, o)
    def __rtruediv__(self, o): return _binop("div", o, self)
    def __pow__(self, o): return _binop("pow", self, o)
    def __rpow__(self, o): return _binop("pow", o, self)
    def __neg__(self): return Sym("neg", (self,))
    def __pos__(self): return self
    def __abs__(self): return Sym("abs", (self,))

    # numpy's object-array loops call these methods of the elements
    def sqrt(self): return Sym("sqrt", (self,))
    def exp(self): return Sym("exp", (self,))
    def log(self): return Sym("log", (self,))
    def log10(self): return Sym("log10", (self,))
    def sin(self): return Sym("sin", (self,))
    def cos(self): return Sym("cos", (self,))
    def tan(self): return Sym("tan", (self,))
    def tanh(self): return Sym("tanh", (self,))
    def sinh(self): return Sym("sinh", (self,))
    def cosh(self): return Sym("cosh", (self,))
    def arctan(self): return Sym("arctan", (self,))
    def arcsin(self): return Sym("arcsin", (self,))
    def arccos(self): return Sym("arccos", (self,))
    def conjugate(self): return self

    def clip(self, a_min=None, a_max=None, out=None, **kw):
        """np.clip of one traced value (numpy calls the element's method)."""
        v = self if a_min is None else _apply("maximum", self, a_min)
        return v if a_max is None else _apply("minimum", v, a_max)

    def __bool__(self):
        raise NotImplementedError("a branch on a field value cannot run on the device: "
                                  "write the field without value-dependent if/else, "
                                  "max() or min() (np.maximum, np.minimum, np.clip and "
                                  "np.sign are traced)")

    def __float__(self):
        raise NotImplementedError("a field value converted to a Python float cannot be traced")

    def __lt__(self, o): return self.__bool__()
    __le__ = __gt__ = __ge__ = __lt__

    def __array_ufunc__(self, ufunc, method, *inputs, **kw):
        if method != "__call__" or kw:
            raise NotImplementedError("numpy %s.%s cannot be traced" % (ufunc.__name__, method))
        name = _UFUNC.get(ufunc.__name__, ufunc.__name__)
        if name not in _NUMPY and name != "clip":
            raise NotImplementedError("numpy function %s cannot be traced" % ufunc.__name__)
        if any(isinstance(v, np.ndarray) for v in inputs):
            f = np.frompyfunc(lambda *a: _apply(name, *a), len(inputs), 1)
            args = []
            for v in inputs:
                a = np.empty((), dtype=object) if not isinstance(v, np.ndarray) else None
                if a is not None:
                    a[()] = v
                args.append(a if a is not None else v.astype(object))
            return f(*args)
        return _apply(name, *inputs)


def _binop(name, a, b):
    if isinstance(a, np.ndarray) or isinstance(b, np.ndarray):
        return NotImplemented  # numpy applies the operation elementwise
    return _apply(name, a, b)


def _node(v):
    if isinstance(v, Sym):
        return v
    if _num(v):
        return Sym("const", value=float(v))
    raise NotImplementedError("unsupported value in a traced field: %r" % (type(v),))


def _components(out, what):
    if _num(out) and float(out) == 0.0:
        return [_node(0.0)] * 3  # the base class's scalar 0
    arr = np.asarray(out, dtype=object).reshape(-1)
    if arr.size != 3:
        raise NotImplementedError("%s must return three components" % what)
    return [_node(v) for v in arr]


class Program:
    """Register programs for B and E (device and oracle format): code
    [nb + ne][4] int32 (B's program, then E's), constants, and the two lengths;
    each program leaves its three components in registers 4-6."""

    def __init__(self, code, consts, nb, ne):
        self.code = np.ascontiguousarray(np.reshape(code, (-1, 4)), dtype=np.int32)
        self.consts = np.ascontiguousarray(consts, dtype=np.float64)
        self.nb, self.ne = int(nb), int(ne)

    def key(self):
        return (self.code.tobytes(), self.consts.tobytes(), self.nb, self.ne)


def _compile(outputs, consts, const_idx):
    """Instructions for three output nodes, in dependency order, registers
    reused once a value's last use has been read."""
    order, done = [], set()
    for o in outputs:
        stack = [(o, False)]
        while stack:
            node, ready = stack.pop()
            if id(node) in done:
                continue
            if ready:
                done.add(id(node))
                order.append(node)
                continue
            stack.append((node, True))
            for a in reversed(node.args):
                if isinstance(a, Sym) and id(a) not in done:
                    stack.append((a, False))
    # constant operands become const nodes with stable identity
    order2, cnode = [], {}
    for n in order:
        n.args = tuple(a if isinstance(a, Sym) else cnode.setdefault(float(a),
                                                                     Sym("const", value=float(a)))
                       for a in n.args)
    seen = set()
    for n in order:
        for a in n.args:
            if a.op == "const" and id(a) not in seen:
                seen.add(id(a))
                order2.append(a)
        if id(n) not in seen:
            seen.add(id(n))
            order2.append(n)
    order = order2
    last = {}
    for i, n in enumerate(order):
        for a in n.args:
            last[id(a)] = i
    for o in outputs:
        last[id(o)] = len(order)  # outputs stay live until the moves
    reg, free, code = {}, list(range(NREG - 1, 6, -1)), []
    for i, n in enumerate(order):
        if n.op == "input":
            reg[id(n)] = n.value
            continue
        args = [reg[id(a)] for a in n.args]
        for a in {id(a): a for a in n.args}.values():
            if last.get(id(a)) == i and reg[id(a)] >= 7:
                free.append(reg[id(a)])
        if not free:
            raise NotImplementedError("field expression needs more than %d registers" % NREG)
        d = free.pop()
        if n.op == "const":
            if n.value not in const_idx:
                const_idx[n.value] = len(consts)
                consts.append(n.value)
            code.append((OPS["const"], d, const_idx[n.value], 0))
        else:
            code.append((OPS[n.op], d, args[0], args[1] if len(args) > 1 else 0))
        reg[id(n)] = d
    for k, o in enumerate(outputs):
        code.append((OPS["mov"], 4 + k, reg[id(o)], 0))
    return code


def trace(field):
    """Program of a field whose B/E are Python code (the reference's override
    contract).  Raises NotImplementedError when they cannot be traced."""
    tpos = np.array([Sym("input", value=k) for k in range(4)], dtype=object)
    try:
        with np.errstate(all="ignore"):
            B = _components(type(field).B(field, tpos.copy()), "B")
            E = _components(type(field).E(field, tpos.copy()), "E")
    except NotImplementedError:
        raise
    except (TypeError, AttributeError, ValueError) as e:
        # code that needs concrete numbers (a lookup, a library call
        # outside numpy's elementwise math) fails on the symbols
        raise NotImplementedError("the field's code needs concrete values (%s: %s)"
                                  % (type(e).__name__, e)) from None
    consts, cidx = [], {}
    cb = _compile(B, consts, cidx)
    e_zero = all(n.op == "const" and n.value == 0.0 for n in E)
    ce = [] if e_zero else _compile(E, consts, cidx)
    if len(cb) + len(ce) > MAXCODE or len(consts) > MAXCONST:
        raise NotImplementedError("field expression too long for the device (%d instructions, "
                                  "%d constants)" % (len(cb) + len(ce), len(consts)))
    return Program(cb + ce, consts, len(cb), len(ce))


_MATH = {"add": lambda a, b: a + b, "sub": lambda a, b: a - b, "mul": lambda a, b: a * b,
         "div": lambda a, b: a / b if b != 0.0 else float(np.divide(a, b)),
         "pow": math.pow, "arctan2": math.atan2, "hypot": math.hypot, "neg": lambda a: -a,
         "sqrt": lambda a: math.sqrt(a) if a >= 0.0 else math.nan, "exp": math.exp,
         "log": math.log, "log10": math.log10, "sin": math.sin, "cos": math.cos,
         "tan": math.tan, "tanh": math.tanh, "sinh": math.sinh, "cosh": math.cosh,
         "arctan": math.atan, "arcsin": math.asin, "arccos": math.acos, "abs": abs,
         "maximum": lambda a, b: a if (a > b or a != a) else b,
         "minimum": lambda a, b: a if (a < b or a != a) else b,
         "sign": lambda a: 1.0 if a > 0 else (-1.0 if a < 0 else (0.0 if a == 0 else a))}


def evaluate(prog, tpos):
    """Host evaluation of a program with the C library's scalar math (the
    arithmetic numpy applies to scalars, and the oracle's): (B, E) at one
    tpos."""
    inv = {v: k for k, v in OPS.items()}
    out = []
    for lo, hi in ((0, prog.nb), (prog.nb, prog.nb + prog.ne)):
        r = [0.0] * NREG
        r[:4] = [float(v) for v in tpos]
        for op, d, a, b in prog.code[lo:hi]:
            name = inv[int(op)]
            if name == "const":
                r[d] = float(prog.consts[a])
            elif name == "mov":
                r[d] = r[a]
            elif name in ("add", "sub", "mul", "div", "pow", "arctan2", "hypot", "maximum",
                          "minimum"):
                r[d] = _MATH[name](r[a], r[b])
            else:
                r[d] = _MATH[name](r[a])
        out.append(np.array(r[4:7]) if hi > lo else np.zeros(3))
    return out[0], out[1]
