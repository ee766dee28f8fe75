"""A small expression IR for weak forms: the neutral front end of the HIP code generator.

The reference builds its element expressions with GiNaC and prints C per residual
(src/codegen.cpp:70-97 print_simplest_form, :3912-4109 write_generic_RJM). Here an
expression is an immutable, hash-consed DAG of

    sym, const, add, sub, mul, div, neg, exp, pow (integer exponent)

built with Python operators. Construction applies only identities that do not change a
floating-point result (x + 0, x - 0, 0 - x = -x, x * 1, x * 0 = 0, x / 1, --x); nothing is
reordered, so the printed code evaluates exactly the tree that was written, operand by
operand -- which is what lets a generated kernel reproduce a hand-written one bit for bit.
Derivatives follow the textbook rules on the tree (``diff``); ``CPrinter`` prints fully
parenthesised C with every shared subexpression in a temporary.
"""

__all__ = ["Expr", "sym", "const", "exp", "diff", "is_zero", "is_one", "CPrinter", "evaluate"]


class Expr:
    __slots__ = ("op", "args", "val", "_hash")
    _table = {}

    def __new__(cls, op, args=(), val=None):
        key = (op, args, val)
        e = Expr._table.get(key)
        if e is None:
            e = object.__new__(cls)
            e.op, e.args, e.val = op, args, val
            e._hash = hash(key)
            Expr._table[key] = e
        return e

    def __hash__(self):
        return self._hash

    def __eq__(self, other):
        return self is other

    # -- construction with result-preserving identities only ------------------------------
    def __add__(self, o):
        o = _wrap(o)
        if is_zero(o):
            return self
        if is_zero(self):
            return o
        return Expr("add", (self, o))

    def __radd__(self, o):
        return _wrap(o) + self

    def __sub__(self, o):
        o = _wrap(o)
        if is_zero(o):
            return self
        if is_zero(self):
            return -o
        return Expr("sub", (self, o))

    def __rsub__(self, o):
        return _wrap(o) - self

    def __mul__(self, o):
        o = _wrap(o)
        if is_zero(self) or is_zero(o):
            return ZERO
        if is_one(o):
            return self
        if is_one(self):
            return o
        return Expr("mul", (self, o))

    def __rmul__(self, o):
        return _wrap(o) * self

    def __truediv__(self, o):
        o = _wrap(o)
        if is_one(o):
            return self
        if is_zero(self):
            return ZERO
        return Expr("div", (self, o))

    def __rtruediv__(self, o):
        return _wrap(o) / self

    def __neg__(self):
        if is_zero(self):
            return ZERO
        if self.op == "neg":
            return self.args[0]
        return Expr("neg", (self,))

    def __pow__(self, n):
        if not isinstance(n, int) or n < 0:
            raise ValueError("only non-negative integer powers")
        if n == 0:
            return ONE
        if n == 1:
            return self
        return Expr("pow", (self,), n)

    def __repr__(self):
        return CPrinter("double").expr_text(self)


def sym(name):
    """A symbol; ``name`` is the C expression it prints as (a variable of the kernel)."""
    return Expr("sym", (), name)


def const(v):
    return Expr("const", (), float(v))


def _wrap(v):
    if isinstance(v, Expr):
        return v
    if isinstance(v, (int, float)):
        return const(v)
    raise TypeError("cannot use %r in an expression" % (v,))


ZERO = const(0.0)
ONE = const(1.0)


def is_zero(e):
    return e.op == "const" and e.val == 0.0


def is_one(e):
    return e.op == "const" and e.val == 1.0


def exp(a):
    return Expr("exp", (_wrap(a),))


def diff(e, s):
    """d e / d s for a symbol s (memoised per call)."""
    memo = {}

    def d(x):
        r = memo.get(x)
        if r is not None:
            return r
        op, a = x.op, x.args
        if op == "sym":
            r = ONE if x is s else ZERO
        elif op == "const":
            r = ZERO
        elif op == "add":
            r = d(a[0]) + d(a[1])
        elif op == "sub":
            r = d(a[0]) - d(a[1])
        elif op == "neg":
            r = -d(a[0])
        elif op == "mul":
            da, db = d(a[0]), d(a[1])
            r = da * a[1] + a[0] * db
        elif op == "div":
            da, db = d(a[0]), d(a[1])
            if is_zero(db):
                r = da / a[1]
            else:
                r = (da * a[1] - a[0] * db) / (a[1] * a[1])
        elif op == "exp":
            r = x * d(a[0])
        elif op == "pow":
            n = x.val
            r = (const(n) * (a[0] ** (n - 1))) * d(a[0])
        else:
            raise ValueError(op)
        memo[x] = r
        return r

    return d(e)


def evaluate(e, env):
    """Value of e with symbols looked up by name in env (Python floats / numpy arrays)."""
    import numpy as np
    memo = {}

    def ev(x):
        if x in memo:
            return memo[x]
        op, a = x.op, x.args
        if op == "sym":
            r = env[x.val]
        elif op == "const":
            r = x.val
        elif op == "add":
            r = ev(a[0]) + ev(a[1])
        elif op == "sub":
            r = ev(a[0]) - ev(a[1])
        elif op == "neg":
            r = -ev(a[0])
        elif op == "mul":
            r = ev(a[0]) * ev(a[1])
        elif op == "div":
            r = ev(a[0]) / ev(a[1])
        elif op == "exp":
            r = np.exp(ev(a[0]))
        elif op == "pow":
            b = ev(a[0])
            r = b
            for _ in range(x.val - 1):
                r = r * b
        memo[x] = r
        return r

    return ev(e)


class CPrinter:
    """Prints expressions as C for element type ``ctype`` ("T" in the kernel template).

    ``assign(outputs)`` returns the statements that compute every output: subexpressions
    used more than once (and not a symbol or constant) go into temporaries ``t<k>``, in an
    order where each is defined before use; every operation is parenthesised, so the C
    evaluates the tree as built.
    """

    def __init__(self, ctype="T", prefix="t"):
        self.ctype = ctype
        self.prefix = prefix
        self.names = {}
        self.lines = []

    def _const(self, v):
        r = repr(float(v))
        if "e" not in r and "." not in r and "inf" not in r and "nan" not in r:
            r += ".0"
        return "%s(%s)" % (self.ctype, r)

    def expr_text(self, x):
        n = self.names.get(x)
        if n is not None:
            return n
        op, a = x.op, x.args
        t = self.expr_text
        if op == "sym":
            return x.val
        if op == "const":
            return self._const(x.val)
        if op == "add":
            return "(%s + %s)" % (t(a[0]), t(a[1]))
        if op == "sub":
            return "(%s - %s)" % (t(a[0]), t(a[1]))
        if op == "mul":
            return "(%s * %s)" % (t(a[0]), t(a[1]))
        if op == "div":
            return "(%s / %s)" % (t(a[0]), t(a[1]))
        if op == "neg":
            return "(-%s)" % t(a[0])
        if op == "exp":
            return "texp(%s)" % t(a[0])
        if op == "pow":
            b = t(a[0])
            s = b
            for _ in range(x.val - 1):
                s = "(%s * %s)" % (s, b)
            return s
        raise ValueError(op)

    def assign(self, outputs):
        """outputs: list of (C lvalue, Expr). Returns the C statements (list of str)."""
        count = {}
        order = []

        def visit(x):
            count[x] = count.get(x, 0) + 1
            if count[x] > 1:
                return
            for c in x.args:
                visit(c)
            order.append(x)

        for _, e in outputs:
            visit(e)
        out = []
        for x in order:
            if count[x] > 1 and x.op not in ("sym", "const") and x not in self.names:
                name = "%s%d" % (self.prefix, len(self.names))
                out.append("const %s %s = %s;" % (self.ctype, name, self.expr_text(x)))
                self.names[x] = name
        for lv, e in outputs:
            out.append("%s = %s;" % (lv, self.expr_text(e)))
        return out
