"""OpenQASM 2.0 reader: program text -> qpp_amd.QCircuit, lowered the way
qpp's reader lowers it (qasm/qasm.hpp:41-823, parser from qasmtools):

* qreg / creg declarations take consecutive qubit / dit indices in program
  order (Context::alloc_qubit / alloc_bit, qasm.hpp:293-321);
* ``U(theta, phi, lambda)`` is the Qiskit-convention matrix (qasm.hpp:557-581,
  the default build without QASMTOOLS_QASM2_SPECS), one gate step per qubit;
* ``CX`` is qpp's CNOT as a two-qubit gate step, mapped over registers
  (qasm.hpp:588-616);
* a gate with an entry in qpp's table of known matrices (qasm.hpp:41-198)
  becomes one gate step with that matrix, mapped over registers; any other
  declared gate (the standard library's u3, u2, u1, u0, r, cswap, cu1, cu3,
  or the program's own) is expanded from its body (qasm.hpp:624-690);
* ``measure a -> c`` is a non-destructive measurement per qubit
  (qasm.hpp:507-516); ``reset`` is the circuit's reset step
  (qasm.hpp:518-523); ``barrier`` is dropped (qasm.hpp:622);
* ``if (c == n) qop`` is a runtime conditional block whose condition is
  "dit c[i] equals bit i of n" (qasm.hpp:525-544).

The standard library is the Qiskit-flavoured qelib1.inc that qasmtools
supplies for ``include "qelib1.inc"``
(qasmtools/include/qasmtools/parser/preprocessor.hpp:83-132): its
definitions for the gates that are not in the known-matrix table are kept
below as bodies, so that every program expands to the same steps as in qpp.
"""
import math
import os
import re

import numpy as np

from . import QCircuit

__all__ = ["QasmError", "read", "read_from_string", "read_from_file"]


class QasmError(ValueError):
    """A malformed program, or a construct qpp's reader rejects."""


# ---------------------------------------------------------------------------
# matrices
# ---------------------------------------------------------------------------
_S2 = 1 / math.sqrt(2)


# numpy's IEEE functions, as qpp's std::cos / std::exp: a non-finite
# parameter gives a NaN matrix, not an exception
def _u(theta, phi, lam):
    with np.errstate(all="ignore"):
        c, s = np.cos(theta / 2), np.sin(theta / 2)
        return np.array([[c, -np.exp(1j * lam) * s],
                         [np.exp(1j * phi) * s, np.exp(1j * (phi + lam)) * c]])


def _ctrl(U):
    M = np.eye(4, dtype=complex)
    M[2:, 2:] = U
    return M


_X = np.array([[0, 1], [1, 0]], dtype=complex)
_Y = np.array([[0, -1j], [1j, 0]])
_Z = np.diag([1, -1]).astype(complex)
_H = np.array([[_S2, _S2], [_S2, -_S2]], dtype=complex)
_S = np.diag([1, 1j])
_T = np.diag([1, np.exp(1j * math.pi / 4)])


def _rx(t):
    with np.errstate(all="ignore"):
        c, s = np.cos(t / 2), np.sin(t / 2)
        return np.array([[c, -1j * s], [-1j * s, c]])


def _ry(t):
    with np.errstate(all="ignore"):
        c, s = np.cos(t / 2), np.sin(t / 2)
        return np.array([[c, -s], [s, c]], dtype=complex)


def _rz(t):
    with np.errstate(all="ignore"):
        return np.diag([np.exp(-0.5j * t), np.exp(0.5j * t)])


_CNOT = np.eye(4, dtype=complex)[[0, 1, 3, 2]]
_SWAP = np.eye(4, dtype=complex)[[0, 2, 1, 3]]
_TOF = np.eye(8, dtype=complex)[[0, 1, 2, 3, 4, 5, 7, 6]]

# qpp's known matrices (qasm.hpp:41-198): name -> (parameter count, builder)
KNOWN = {
    "cx": (0, lambda a: _CNOT), "id": (0, lambda a: np.eye(2, dtype=complex)),
    "x": (0, lambda a: _X), "y": (0, lambda a: _Y), "z": (0, lambda a: _Z),
    "h": (0, lambda a: _H), "s": (0, lambda a: _S), "sdg": (0, lambda a: _S.conj()),
    "t": (0, lambda a: _T), "tdg": (0, lambda a: _T.conj()),
    "rx": (1, lambda a: _rx(a[0])), "ry": (1, lambda a: _ry(a[0])),
    "rz": (1, lambda a: _rz(a[0])),
    "cz": (0, lambda a: np.diag([1, 1, 1, -1]).astype(complex)),
    "cy": (0, lambda a: _ctrl(_Y)), "swap": (0, lambda a: _SWAP),
    "ch": (0, lambda a: _ctrl(_H)), "ccx": (0, lambda a: _TOF),
    "crz": (1, lambda a: _ctrl(_rz(a[0]))),
}

# qelib1.inc gates without a known matrix, as (parameters, qubits, body);
# every name in KNOWN is also declared by the library with its own
# parameter and qubit counts
_STDLIB = """
gate u3(theta,phi,lambda) q { U(theta,phi,lambda) q; }
gate u2(phi,lambda) q { U(pi/2,phi,lambda) q; }
gate u1(lambda) q { U(0,0,lambda) q; }
gate u0(gamma) q { U(0,0,0) q; }
gate r(theta,phi) a { u3(theta,phi-pi/2,-phi+pi/2) a; }
gate cswap a,b,c { cx c,b; ccx a,b,c; cx c,b; }
gate cu1(lambda) a,b { u1(lambda/2) a; cx a,b; u1(-lambda/2) b; cx a,b; u1(lambda/2) b; }
gate cu3(theta,phi,lambda) c,t { u1((lambda+phi)/2) c; u1((lambda-phi)/2) t; cx c,t;
  u3(-theta/2,0,-(phi+lambda)/2) t; cx c,t; u3(theta/2,phi,0) t; }
"""
_KNOWN_ARITY = {"cx": 2, "cz": 2, "cy": 2, "swap": 2, "ch": 2, "crz": 2, "ccx": 3}


# ---------------------------------------------------------------------------
# tokens
# ---------------------------------------------------------------------------
_TOKEN = re.compile(r"""
    (?P<ws>\s+|//[^\n]*|/\*.*?\*/)
  | (?P<real>(?:\d+\.\d*|\.\d+)(?:[eE][+-]?\d+)?|\d+[eE][+-]?\d+)
  | (?P<int>\d+)
  | (?P<id>[A-Za-z_][A-Za-z0-9_]*)
  | (?P<str>"[^"]*")
  | (?P<sym>->|==|[;,()\[\]{}+\-*/^])
""", re.VERBOSE | re.DOTALL)


def _tokens(text, src):
    out, pos, line = [], 0, 1
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise QasmError(f"{src}:{line}: unexpected character {text[pos]!r}")
        kind = m.lastgroup
        if kind != "ws":
            out.append((kind, m.group(), line))
        line += m.group().count("\n")
        pos = m.end()
    out.append(("eof", "", line))
    return out


# ---------------------------------------------------------------------------
# parser -> statement tuples
# ---------------------------------------------------------------------------
class _Parser:
    def __init__(self, text, src, base_dir):
        self.t = _tokens(text, src)
        self.i = 0
        self.src = src
        self.base_dir = base_dir

    def peek(self, k=0):
        return self.t[self.i + k]

    def next(self):
        tok = self.t[self.i]
        self.i += 1
        return tok

    def err(self, msg):
        raise QasmError(f"{self.src}:{self.peek()[2]}: {msg}")

    def expect(self, value):
        tok = self.next()
        if tok[1] != value:
            self.i -= 1
            self.err(f"expected {value!r}, got {tok[1]!r}")
        return tok

    def end(self):
        # a statement ends at ';' or at the end of the input (qasmtools'
        # consume_until(semicolon) stops at eof without an error,
        # parser.hpp:149-166)
        if self.peek()[0] != "eof":
            self.expect(";")

    def ident(self):
        tok = self.next()
        if tok[0] != "id":
            self.i -= 1
            self.err(f"expected an identifier, got {tok[1]!r}")
        return tok[1]

    def integer(self):
        tok = self.next()
        if tok[0] != "int":
            self.i -= 1
            self.err(f"expected an integer, got {tok[1]!r}")
        return int(tok[1])

    # program ---------------------------------------------------------------
    def program(self):
        stmts = []
        if self.peek()[1] == "OPENQASM":
            self.next()
            v = self.next()
            if v[1] not in ("2.0", "2"):
                self.err(f"unsupported OPENQASM version {v[1]}")
            self.end()
        while self.peek()[0] != "eof":
            stmts.extend(self.statement())
        return stmts

    def statement(self):
        tok = self.peek()
        w = tok[1]
        if w == "include":
            self.next()
            name = self.next()
            if name[0] != "str":
                self.err("expected a file name")
            self.end()
            fname = name[1][1:-1]
            if fname == "qelib1.inc":
                return [("include_std",)]
            path = os.path.join(self.base_dir, fname)
            try:
                with open(path) as f:
                    text = f.read()
            except OSError:
                raise QasmError(f"{self.src}:{tok[2]}: cannot open {fname}") from None
            return _Parser(text, path, os.path.dirname(path)).program()
        if w in ("qreg", "creg"):
            self.next()
            name = self.ident()
            self.expect("[")
            n = self.integer()
            self.expect("]")
            self.end()
            return [("reg", w == "qreg", name, n)]
        if w in ("gate", "opaque"):
            self.next()
            name = self.ident()
            params = []
            if self.peek()[1] == "(":
                self.next()
                if self.peek()[1] != ")":
                    params = self.idlist()
                self.expect(")")
            qargs = self.idlist()
            if w == "opaque":
                self.end()
                return [("opaque", name, params, qargs)]
            self.expect("{")
            body = []
            while self.peek()[1] != "}":
                if self.peek()[1] == "barrier":
                    self.next()
                    self.idlist()
                    self.end()
                    continue
                body.append(self.qop(in_gate=True))
            self.expect("}")
            return [("gate", name, params, qargs, body)]
        if w == "barrier":
            self.next()
            self.arglist()
            self.end()
            return []
        if w == "if":
            self.next()
            self.expect("(")
            creg = self.ident()
            self.expect("==")
            val = self.integer()
            self.expect(")")
            return [("if", creg, val, self.qop(in_gate=False), tok[2])]
        return [self.qop(in_gate=False)]

    def idlist(self):
        out = [self.ident()]
        while self.peek()[1] == ",":
            self.next()
            out.append(self.ident())
        return out

    def arg(self):
        line = self.peek()[2]
        name = self.ident()
        off = None
        if self.peek()[1] == "[":
            self.next()
            off = self.integer()
            self.expect("]")
        return (name, off, line)

    def arglist(self):
        out = [self.arg()]
        while self.peek()[1] == ",":
            self.next()
            out.append(self.arg())
        return out

    def qop(self, in_gate):
        tok = self.peek()
        w = tok[1]
        if w == "measure" and not in_gate:
            self.next()
            q = self.arg()
            self.expect("->")
            c = self.arg()
            self.end()
            return ("measure", q, c)
        if w == "reset" and not in_gate:
            self.next()
            a = self.arg()
            self.end()
            return ("reset", a)
        if w == "U":
            self.next()
            self.expect("(")
            e = [self.expr()]
            for _ in range(2):
                self.expect(",")
                e.append(self.expr())
            self.expect(")")
            a = self.arg()
            self.end()
            return ("U", e, a)
        if w == "CX":
            self.next()
            a = self.arg()
            self.expect(",")
            b = self.arg()
            self.end()
            return ("CX", a, b)
        name = self.ident()
        cargs = []
        if self.peek()[1] == "(":
            self.next()
            if self.peek()[1] != ")":
                cargs.append(self.expr())
                while self.peek()[1] == ",":
                    self.next()
                    cargs.append(self.expr())
            self.expect(")")
        qargs = self.arglist()
        self.end()
        return ("call", name, cargs, qargs, tok[2])

    # expressions: + - < * / < unary - < ^ (right) < atoms -------------------
    def expr(self):
        e = self.term()
        while self.peek()[1] in ("+", "-"):
            op = self.next()[1]
            e = (op, e, self.term())
        return e

    def term(self):
        e = self.unary()
        while self.peek()[1] in ("*", "/"):
            op = self.next()[1]
            e = (op, e, self.unary())
        return e

    def unary(self):
        if self.peek()[1] == "-":
            self.next()
            return ("neg", self.unary())
        if self.peek()[1] == "+":
            self.next()
            return self.unary()
        return self.power()

    def power(self):
        base = self.atom()
        if self.peek()[1] == "^":
            self.next()
            return ("^", base, self.unary())
        return base

    def atom(self):
        kind, val, line = self.next()
        if kind in ("int", "real"):
            return ("num", float(val))
        if val == "(":
            e = self.expr()
            self.expect(")")
            return e
        if kind == "id":
            if val == "pi":
                return ("num", math.pi)
            if val in ("sin", "cos", "tan", "exp", "ln", "sqrt"):
                self.expect("(")
                e = self.expr()
                self.expect(")")
                return ("fn", val, e)
            return ("var", val, line)
        self.i -= 1
        self.err(f"unexpected {val!r} in an expression")


# evaluated in IEEE doubles like qpp's visitor (qasm.hpp:437-505): 1/0 is
# inf, ln(-1) and (-8)^(1/3) are nan
_FN = {"sin": np.sin, "cos": np.cos, "tan": np.tan, "exp": np.exp, "ln": np.log,
       "sqrt": np.sqrt}
_OP = {"+": np.add, "-": np.subtract, "*": np.multiply, "/": np.divide, "^": np.power}


def _eval(e, env):
    k = e[0]
    if k == "num":
        return e[1]
    if k == "var":
        if e[1] not in env:
            raise QasmError(f"line {e[2]}: unknown parameter {e[1]!r}")
        return env[e[1]]
    with np.errstate(all="ignore"):
        if k == "neg":
            return -_eval(e[1], env)
        if k == "fn":
            return float(_FN[e[1]](np.float64(_eval(e[2], env))))
        return float(_OP[k](np.float64(_eval(e[1], env)), np.float64(_eval(e[2], env))))


# ---------------------------------------------------------------------------
# builder
# ---------------------------------------------------------------------------
class _Builder:
    def __init__(self, stmts):
        self.stmts = stmts
        self.qregs, self.cregs, self.gates, self.opaque = {}, {}, {}, set()
        self.declared = set()
        nq = nc = 0
        for s in stmts:
            if s[0] == "reg":
                if s[1]:
                    self.qregs[s[2]] = list(range(nq, nq + s[3]))
                    nq += s[3]
                else:
                    self.cregs[s[2]] = list(range(nc, nc + s[3]))
                    nc += s[3]
        self.qc = QCircuit(nq, nc)

    def build(self):
        for s in self.stmts:
            k = s[0]
            if k == "include_std":
                for g in _Parser(_STDLIB, "qelib1.inc", ".").program():
                    self.gates[g[1]] = g
                self.declared |= set(self.gates) | set(KNOWN)
            elif k == "gate":
                self.gates[s[1]] = s
                self.declared.add(s[1])
            elif k == "opaque":
                self.opaque.add(s[1])
                self.declared.add(s[1])
            elif k == "reg":
                pass
            elif k == "if":
                _, creg, val, op, line = s
                if creg not in self.cregs:
                    raise QasmError(f"line {line}: unknown creg {creg!r}")
                dits = self.cregs[creg]
                self.qc.cond_if(dits, [(val >> i) & 1 for i in range(len(dits))])
                self.op(op, {}, {})
                self.qc.cond_end()
            else:
                self.op(s, {}, {})
        return self.qc

    def qubits(self, a, qenv):
        name, off, line = a
        if off is None and name in qenv:
            return [qenv[name]]
        if name not in self.qregs:
            raise QasmError(f"line {line}: unknown qubit or qreg {name!r}")
        reg = self.qregs[name]
        if off is None:
            return list(reg)
        if off >= len(reg):
            raise QasmError(f"line {line}: index {off} out of range for {name}")
        return [reg[off]]

    def bits(self, a):
        name, off, line = a
        if name not in self.cregs:
            raise QasmError(f"line {line}: unknown creg {name!r}")
        reg = self.cregs[name]
        if off is None:
            return list(reg)
        if off >= len(reg):
            raise QasmError(f"line {line}: index {off} out of range for {name}")
        return [reg[off]]

    def op(self, s, env, qenv):
        k = s[0]
        if k == "measure":
            q, c = self.qubits(s[1], qenv), self.bits(s[2])
            if len(q) != len(c):
                raise QasmError(f"line {s[1][2]}: measure of {len(q)} qubits into {len(c)} bits")
            for qi, ci in zip(q, c):
                self.qc.measure(qi, ci, destructive=False)
        elif k == "reset":
            q = self.qubits(s[1], qenv)
            self.qc.reset(q[0] if len(q) == 1 else q)
        elif k == "U":
            M = _u(*(_eval(e, env) for e in s[1]))
            for q in self.qubits(s[2], qenv):
                self.qc.gate(M, q)
        elif k == "CX":
            self.mapped([self.qubits(s[1], qenv), self.qubits(s[2], qenv)],
                        lambda args: self.qc.gate(_CNOT, *args))
        elif k == "call":
            self.call(s, env, qenv)
        else:
            raise QasmError(f"unexpected statement {k}")

    @staticmethod
    def mapped(qargs, emit):
        # qasm.hpp:639-647: a register argument maps the gate over its qubits
        size, is_reg = 1, [len(a) > 1 for a in qargs]
        for a in qargs:
            if len(a) > 1:
                size = len(a)
        for j in range(size):
            emit([a[j] if r else a[0] for a, r in zip(qargs, is_reg)])

    def call(self, s, env, qenv):
        _, name, cexprs, qexprs, line = s
        if name not in self.declared:
            raise QasmError(f"line {line}: unknown gate {name!r}")
        if name in self.opaque and name not in self.gates and name not in KNOWN:
            raise QasmError(f"line {line}: opaque gate {name!r} has no definition")
        cargs = [_eval(e, env) for e in cexprs]
        qargs = [self.qubits(a, qenv) for a in qexprs]
        if name in KNOWN:
            npar, make = KNOWN[name]
            nq = _KNOWN_ARITY.get(name, 1)
            if len(cargs) != npar or len(qargs) != nq:
                raise QasmError(f"line {line}: {name} takes {npar} parameters and {nq} qubits")
            M = make(cargs)
            self.mapped(qargs, lambda args: self.qc.gate(M, *args))
            return
        _, _, params, qparams, body = self.gates[name]
        if len(cargs) != len(params) or len(qargs) != len(qparams):
            raise QasmError(f"line {line}: {name} takes {len(params)} parameters and "
                            f"{len(qparams)} qubits")
        inner_env = dict(zip(params, cargs))

        def expand(args):
            inner_q = dict(zip(qparams, args))
            for g in body:
                self.op(g, inner_env, inner_q)

        self.mapped(qargs, expand)


def read_from_string(text, base_dir="."):
    """qpp::qasm::read_from_string (qasm.hpp:805-809)."""
    return _Builder(_Parser(text, "<string>", base_dir).program()).build()


def read_from_file(path):
    """qpp::qasm::read_from_file (qasm.hpp:818-821)."""
    with open(path) as f:
        text = f.read()
    return _Builder(_Parser(text, str(path), os.path.dirname(os.path.abspath(path))).program()).build()


def read(stream):
    """qpp::qasm::read(std::istream&) (qasm.hpp:794-797)."""
    return read_from_string(stream.read())
