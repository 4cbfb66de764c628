"""OpenQASM 2 and 3 reader for the state-vector engine (src/qasm.rs).

Reads what the reference's reader reads on the state-vector path:
register declarations (``qreg``/``creg``, OpenQASM 3's ``qubit``/``bit``,
laid out in declaration order), the gates of ``qelib1.inc`` and
``stdgates.inc`` plus the ``U``/``CX`` primitives (src/qasm.rs:1377-1435),
``gate`` definitions expanded where they are called, parameter expressions
(``+ - * / ^ **``, ``pi tau euler``, ``sin cos tan exp ln sqrt asin acos
atan``), broadcast over registers and OpenQASM 3 ranges ``r[a:b]`` /
``r[a:step:b]``, ``measure`` in both syntaxes, ``reset``, ``barrier``,
``gphase`` (ignored), the ``inv @``, ``ctrl(n) @``, ``negctrl(n) @`` and
integer ``pow(k) @`` modifiers (a modified gate becomes an exact unitary),
and ``if`` on one classical bit whose bodies apply Pauli gates only, read as
classical feedback (src/qasm.rs:636-684). The relative-phase Toffolis
``rccx`` and ``rc3x`` come from their qelib1.inc definitions and may be
redefined. Everything else that needs classical state is rejected with the
line and column of the statement.
"""
from __future__ import annotations

import math

import numpy as np

from . import _native
from .circuit import Circuit


class QasmError(ValueError):
    """A syntax or semantic error at a 1-based line and column."""

    def __init__(self, line: int, column: int, message: str):
        super().__init__(f"line {line}, column {column}: {message}")
        self.line, self.column, self.message = line, column, message


# The qelib1.inc definitions of the relative-phase Toffolis (OpenQASM 2's
# standard library; src/qasm.rs:84-97). A program may redefine them.
PRELUDE = """
gate rccx a,b,c
{
  u2(0,pi) c; u1(pi/4) c; cx b, c; u1(-pi/4) c; cx a, c;
  u1(pi/4) c; cx b, c; u1(-pi/4) c; u2(0,pi) c;
}
gate rc3x a,b,c,d
{
  u2(0,pi) d; u1(pi/4) d; cx c,d; u1(-pi/4) d; u2(0,pi) d;
  cx a,d; u1(pi/4) d; cx b,d; u1(-pi/4) d; cx a,d;
  u1(pi/4) d; cx b,d; u1(-pi/4) d; u2(0,pi) d;
  u1(pi/4) d; cx c,d; u1(-pi/4) d; u2(0,pi) d;
}
"""

# Statements of OpenQASM 3 that need classical state (src/qasm.rs:391-421).
UNSUPPORTED = {"else", "for", "while", "def", "defcal", "defcalgrammar", "cal", "let", "const", "input", "output",
               "int", "uint", "float", "angle", "bool", "complex", "duration", "stretch", "delay", "box", "return",
               "break", "continue", "end", "extern", "array", "opaque", "pragma"}

CONSTANTS = {"pi": math.pi, "π": math.pi, "tau": math.tau, "τ": math.tau, "euler": math.e, "ℇ": math.e}
FUNCTIONS = {"sin": math.sin, "cos": math.cos, "tan": math.tan, "exp": math.exp, "ln": math.log, "sqrt": math.sqrt,
             "asin": math.asin, "arcsin": math.asin, "acos": math.acos, "arccos": math.acos, "atan": math.atan,
             "arctan": math.atan}

SYMBOLS2 = ("->", "**", "==", "!=")
SYMBOLS1 = set(";,()[]{}=+-*/^@:<>!&|%~$")

# (kind, value, line, column); kind: "id", "num", "str", "sym", "end"
Token = tuple


def lex(source: str) -> list:
    tokens = []
    i, line, col = 0, 1, 1
    n = len(source)

    def advance(k):
        nonlocal i, line, col
        for _ in range(k):
            if source[i] == "\n":
                line, col = line + 1, 1
            else:
                col += 1
            i += 1

    while i < n:
        c = source[i]
        if c.isspace():
            advance(1)
        elif source.startswith("//", i):
            while i < n and source[i] != "\n":
                advance(1)
        elif source.startswith("/*", i):
            at = (line, col)
            advance(2)
            while i < n and not source.startswith("*/", i):
                advance(1)
            if i >= n:
                raise QasmError(*at, "unterminated comment")
            advance(2)
        elif c.isalpha() or c == "_":
            j = i
            while j < n and (source[j].isalnum() or source[j] == "_"):
                j += 1
            tokens.append(("id", source[i:j], line, col))
            advance(j - i)
        elif c.isdigit() or (c == "." and i + 1 < n and source[i + 1].isdigit()):
            j = i
            while j < n and (source[j].isdigit() or source[j] in "._"):
                j += 1
            if j < n and source[j] in "eE":
                k = j + 1
                if k < n and source[k] in "+-":
                    k += 1
                if k < n and source[k].isdigit():
                    j = k
                    while j < n and source[j].isdigit():
                        j += 1
            text = source[i:j].replace("_", "")
            try:
                value = float(text)
            except ValueError:
                raise QasmError(line, col, f"invalid number `{text}`") from None
            tokens.append(("num", value, line, col))
            advance(j - i)
        elif c in "\"'":
            j = i + 1
            while j < n and source[j] != c and source[j] != "\n":
                j += 1
            if j >= n or source[j] != c:
                raise QasmError(line, col, "unterminated string")
            tokens.append(("str", source[i + 1:j], line, col))
            advance(j + 1 - i)
        elif source[i:i + 2] in SYMBOLS2:
            tokens.append(("sym", source[i:i + 2], line, col))
            advance(2)
        elif c in SYMBOLS1:
            tokens.append(("sym", c, line, col))
            advance(1)
        else:
            raise QasmError(line, col, f"unexpected character `{c}`")
    tokens.append(("end", None, line, col))
    return tokens


def describe(tok) -> str:
    kind, value = tok[0], tok[1]
    if kind == "end":
        return "the end of the program"
    if kind == "str":
        return f'"{value}"'
    if kind == "num":
        return f"`{value:g}`"
    return f"`{value}`"


# -- expressions: trees of tuples evaluated against a parameter environment --

def evaluate(expr, env: dict) -> float:
    kind = expr[0]
    if kind == "num":
        return expr[1]
    if kind == "name":
        _, name, line, col = expr
        if name in CONSTANTS:
            return CONSTANTS[name]
        if name in env:
            return env[name]
        raise QasmError(line, col, f"unknown parameter `{name}`")
    if kind == "neg":
        return -evaluate(expr[1], env)
    if kind == "call":
        _, name, arg, line, col = expr
        if name not in FUNCTIONS:
            raise QasmError(line, col, f"unknown function `{name}`")
        return FUNCTIONS[name](evaluate(arg, env))
    _, op, a, b = expr
    x, y = evaluate(a, env), evaluate(b, env)
    if op == "+":
        return x + y
    if op == "-":
        return x - y
    if op == "*":
        return x * y
    if op == "/":
        return x / y
    return x ** y


# -- the engine's gate table, probed for arity and parameter counts ----------

_BUILTIN: dict = {}


def builtin(name: str):
    """(parameter count, qubit count) of a gate the engine's table knows, or
    None."""
    if name not in _BUILTIN:
        found = None
        for k in range(5):
            try:
                controls, base = _native.gate_lower(name, [0.0] * k)
            except _native.NativeError as e:
                if "unknown gate" in str(e):
                    break
                continue
            found = (k, controls + int(round(math.log2(base.shape[0]))))
            break
        _BUILTIN[name] = found
    return _BUILTIN[name]


def controlled(m: np.ndarray, count: int, state: int) -> np.ndarray:
    """m with `count` controls prepended on the low index bits, active in
    `state` (src/matrix.rs:177-195)."""
    dim = m.shape[0]
    out = np.eye(dim << count, dtype=np.complex128)
    rows = [(r << count) | state for r in range(dim)]
    out[np.ix_(rows, rows)] = m
    return out


def embed(m: np.ndarray, qubits, k: int) -> np.ndarray:
    """The 2^k x 2^k operator of `m` acting on `qubits` of k qubits."""
    dim = 1 << k
    kk = len(qubits)
    out = np.zeros((dim, dim), dtype=np.complex128)
    mask = sum(1 << q for q in qubits)
    for col in range(dim):
        rest = col & ~mask
        c = sum(((col >> q) & 1) << j for j, q in enumerate(qubits))
        for r in range(1 << kk):
            row = rest | sum(((r >> j) & 1) << q for j, q in enumerate(qubits))
            out[row, col] += m[r, c]
    return out


class Parser:
    def __init__(self, tokens, definitions=None):
        self.tokens = tokens
        self.at = 0
        self.qubits = self.clbits = 0
        self.qregs: dict = {}
        self.cregs: dict = {}
        self.gates: dict = dict(definitions or {})
        # ("gate", name, qubits, params) | ("unitary", matrix, qubits) |
        # ("measure", qubit, clbit) | ("reset", qubit) | ("barrier", qubits) |
        # ("feedback", pauli, qubit, clbit)
        self.out: list = []

    # -- tokens --
    def peek(self):
        return self.tokens[self.at]

    def next(self):
        tok = self.tokens[self.at]
        if tok[0] != "end":
            self.at += 1
        return tok

    def fail(self, tok, message: str):
        raise QasmError(tok[2], tok[3], message)

    def is_sym(self, s: str) -> bool:
        tok = self.peek()
        return tok[0] == "sym" and tok[1] == s

    def expect(self, s: str):
        tok = self.next()
        if tok[0] != "sym" or tok[1] != s:
            self.fail(tok, f"expected `{s}`, found {describe(tok)}")
        return tok

    def ident(self):
        tok = self.next()
        if tok[0] != "id":
            self.fail(tok, f"expected a name, found {describe(tok)}")
        return tok[1], tok

    def integer(self) -> int:
        tok = self.next()
        if tok[0] != "num" or tok[1] != int(tok[1]) or tok[1] < 0:
            self.fail(tok, f"expected a non-negative integer, found {describe(tok)}")
        return int(tok[1])

    # -- statements --
    def program(self):
        while self.peek()[0] != "end":
            self.statement()

    def statement(self):
        tok = self.peek()
        if tok[0] != "id":
            self.fail(tok, f"expected a statement, found {describe(tok)}")
        word = tok[1]
        if word == "OPENQASM":
            self.next()
            v = self.next()
            if v[0] != "num" or not 2.0 <= v[1] < 4.0:
                self.fail(v, "only OpenQASM 2 and 3 are supported")
            self.expect(";")
        elif word == "include":
            self.next()
            f = self.next()
            if f[0] != "str":
                self.fail(f, "expected a file name")
            if f[1] not in ("qelib1.inc", "stdgates.inc"):
                self.fail(f, f"cannot include `{f[1]}`: only qelib1.inc and stdgates.inc are built in")
            self.expect(";")
        elif word in ("qreg", "creg"):
            self.next()
            name, at = self.ident()
            self.expect("[")
            size = self.integer()
            self.expect("]")
            self.expect(";")
            self.declare(word == "qreg", name, size, at)
        elif word in ("qubit", "bit"):
            self.next()
            size = 1
            if self.is_sym("["):
                self.next()
                size = self.integer()
                self.expect("]")
            name, at = self.ident()
            if self.is_sym("="):
                self.fail(at, "initialising classical bits is not supported")
            self.expect(";")
            self.declare(word == "qubit", name, size, at)
        elif word == "gate":
            self.definition()
        elif word == "if":
            self.conditional()
        elif word == "measure":
            self.next()
            qubits = self.operand(True)
            if not self.is_sym("->"):
                self.fail(tok, "a measurement must store its result (`measure q -> c;` or `c = measure q;`)")
            self.next()
            clbits = self.operand(False)
            self.expect(";")
            self.measure(qubits, clbits, tok)
        elif word == "reset":
            self.next()
            qubits = self.operand(True)
            self.expect(";")
            self.out += [("reset", q) for q in qubits]
        elif word == "barrier":
            self.next()
            qubits = []
            while not self.is_sym(";"):
                qubits += self.operand(True)
                if self.is_sym(","):
                    self.next()
            self.expect(";")
            self.out.append(("barrier", qubits or list(range(self.qubits))))
        elif word == "gphase":
            # a global phase changes nothing measurable
            while not self.is_sym(";") and self.peek()[0] != "end":
                self.next()
            self.expect(";")
        elif word in UNSUPPORTED:
            self.fail(tok, f"`{word}` is not supported (the engine runs gates, measurements and resets only)")
        else:
            save = self.at
            self.next()
            if self.is_sym("[") or self.is_sym("="):
                # c = measure q;  c[0] = measure q[0];
                self.at = save
                clbits = self.operand(False)
                self.expect("=")
                keyword, at = self.ident()
                if keyword != "measure":
                    self.fail(at, "assigning classical values other than measurements is not supported")
                qubits = self.operand(True)
                self.expect(";")
                self.measure(qubits, clbits, tok)
            else:
                self.at = save
                call = self.call()
                self.expect(";")
                self.top_level_call(call)

    def declare(self, quantum: bool, name: str, size: int, at):
        if name in self.qregs or name in self.cregs:
            self.fail(at, f"`{name}` is already declared")
        if quantum:
            self.qregs[name] = (self.qubits, size)
            self.qubits += size
        else:
            self.cregs[name] = (self.clbits, size)
            self.clbits += size

    def measure(self, qubits, clbits, at):
        if len(qubits) != len(clbits):
            self.fail(at, f"measuring {len(qubits)} qubits into {len(clbits)} bits")
        self.out += [("measure", q, c) for q, c in zip(qubits, clbits)]

    def operand(self, quantum: bool) -> list:
        """name, name[i] or name[a:b] / name[a:step:b] as global indices."""
        name, at = self.ident()
        regs = self.qregs if quantum else self.cregs
        if name not in regs:
            self.fail(at, f"unknown {'qubit' if quantum else 'classical bit'} register `{name}`")
        first, size = regs[name]
        if not self.is_sym("["):
            return list(range(first, first + size))
        self.next()
        start = self.integer()
        idx = [start]
        if self.is_sym(":"):
            self.next()
            second = self.integer()
            step, end = (1, second)
            if self.is_sym(":"):
                self.next()
                step, end = second, self.integer()
            if step == 0:
                self.fail(at, "a range step must be positive")
            idx = list(range(start, end + 1, step))
        self.expect("]")
        for i in idx:
            if i >= size:
                self.fail(at, f"index {i} is out of range for `{name}` of size {size}")
        return [first + i for i in idx]

    # -- if: classical feedback of Pauli gates --
    def conditional(self):
        at = self.next()
        self.expect("(")
        clbit, wanted = self.condition()
        self.expect(")")
        then = self.pauli_body(at)
        other = {}
        if self.peek()[0] == "id" and self.peek()[1] == "else":
            self.next()
            other = self.pauli_body(at)
        on_one, on_zero = (then, other) if wanted else (other, then)
        # the body for 0 always, then the product of both bodies where the
        # bit reads 1 (a Pauli phase conditioned on a measured bit is not
        # observable; src/qasm.rs:630-642)
        for q, bits in sorted(on_zero.items()):
            if bits:
                self.out.append(("gate", "xzy"[bits - 1], [q], []))
        for q in sorted(set(on_one) | set(on_zero)):
            flip = on_one.get(q, 0) ^ on_zero.get(q, 0)
            if flip:
                self.out.append(("feedback", "XZY"[flip - 1], q, clbit))

    def condition(self):
        negated = False
        while self.is_sym("!"):
            self.next()
            negated = not negated
        name, at = self.ident()
        if name not in self.cregs:
            self.fail(at, f"unknown classical bit register `{name}` in a condition")
        first, size = self.cregs[name]
        if self.is_sym("["):
            self.next()
            index = self.integer()
            self.expect("]")
            if index >= size:
                self.fail(at, f"index {index} is out of range for `{name}` of size {size}")
            clbit = first + index
        elif size == 1:
            clbit = first
        else:
            self.fail(at, f"a condition on `{name}`, a register of {size} bits, is not supported: only one bit can "
                          f"be tested (`if ({name}[k])`)")
        wanted = True
        if self.is_sym("==") or self.is_sym("!="):
            equal = self.next()[1] == "=="
            tok = self.next()
            if tok[0] == "num" and tok[1] in (0.0, 1.0):
                value = tok[1] == 1.0
            elif tok[0] == "id" and tok[1] in ("true", "false"):
                value = tok[1] == "true"
            else:
                self.fail(tok, f"a bit can only be compared with 0, 1, true or false, not {describe(tok)}")
            wanted = value == equal
        return clbit, wanted != negated

    def pauli_body(self, at) -> dict:
        """One gate call or a { } block of them, as the Pauli each qubit gets
        (bit 0: X, bit 1: Z, up to a phase)."""
        block = self.is_sym("{")
        if block:
            self.next()
        start = len(self.out)
        while True:
            if block and self.is_sym("}"):
                self.next()
                break
            tok = self.peek()
            if tok[0] == "id" and tok[1] == "barrier":
                while not self.is_sym(";") and self.peek()[0] != "end":
                    self.next()
                self.expect(";")
            elif tok[0] == "id" and (tok[1] in ("measure", "reset", "if", "gate") or tok[1] in UNSUPPORTED):
                self.fail(tok, f"`{tok[1]}` inside `if` is not supported: an `if` body can only apply Pauli gates "
                               f"(x, y, z, id)")
            elif tok[0] == "id":
                call = self.call()
                self.expect(";")
                before = len(self.out)
                self.top_level_call(call)
                if not all(i[0] == "gate" and i[1] in ("x", "y", "z", "id", "i") for i in self.out[before:]):
                    self.fail(tok, f"`{call['name']}` inside `if` is not supported: an `if` body can only apply "
                                   f"Pauli gates (x, y, z, id), which run as classical feedback")
            elif tok[0] == "end":
                self.fail(at, "unterminated `if` block")
            else:
                self.fail(tok, f"expected a gate call, found {describe(tok)}")
            if not block:
                break
        body: dict = {}
        for ins in self.out[start:]:
            bits = {"x": 1, "z": 2, "y": 3}.get(ins[1], 0)
            body[ins[2][0]] = body.get(ins[2][0], 0) ^ bits
        del self.out[start:]
        return body

    # -- gate calls and definitions --
    def call(self) -> dict:
        """modifiers @ name(params) args, the arguments as (name, token
        position of an index or range or None)."""
        modifiers = []
        while self.peek()[0] == "id" and self.peek()[1] in ("inv", "ctrl", "negctrl", "pow"):
            tok = self.peek()
            save = self.at
            self.next()
            arg = None
            if self.is_sym("("):
                self.next()
                arg = self.expression()
                self.expect(")")
            if not self.is_sym("@"):
                self.at = save
                break
            self.next()
            kind = tok[1]
            if kind in ("ctrl", "negctrl"):
                count = 1
                if arg is not None:
                    x = evaluate(arg, {})
                    if x < 1 or x != int(x):
                        self.fail(tok, "a control count must be a positive integer")
                    count = int(x)
                modifiers.append((kind, count))
            elif kind == "pow":
                if arg is None:
                    self.fail(tok, "`pow` needs an exponent")
                modifiers.append(("pow", arg))
            else:
                modifiers.append(("inv", None))
        name, at = self.ident()
        params = []
        if self.is_sym("("):
            self.next()
            if not self.is_sym(")"):
                while True:
                    params.append(self.expression())
                    if self.is_sym(","):
                        self.next()
                    else:
                        break
            self.expect(")")
        args = []
        while not self.is_sym(";") and not self.is_sym("}") and self.peek()[0] != "end":
            start = self.at
            arg, _ = self.ident()
            index_at = None
            if self.is_sym("["):
                while not self.is_sym("]") and self.peek()[0] != "end":
                    self.next()
                self.expect("]")
                index_at = start
            args.append((arg, index_at))
            if self.is_sym(","):
                self.next()
            else:
                break
        return {"modifiers": modifiers, "name": name, "params": params, "args": args, "at": at}

    def top_level_call(self, call: dict):
        at = call["at"]
        operands = []
        for arg, index_at in call["args"]:
            if index_at is not None:
                save = self.at
                self.at = index_at
                operands.append(self.operand(True))
                self.at = save
            elif arg in self.qregs:
                first, size = self.qregs[arg]
                operands.append(list(range(first, first + size)))
            else:
                self.fail(at, f"unknown qubit register `{arg}`")
        width = max([len(o) for o in operands if len(o) != 1] or [1])
        if any(len(o) not in (1, width) for o in operands):
            self.fail(at, "registers of different sizes cannot be broadcast together")
        params = [evaluate(e, {}) for e in call["params"]]
        for i in range(width):
            qubits = [o[0] if len(o) == 1 else o[i] for o in operands]
            if len(set(qubits)) != len(qubits):
                self.fail(at, f"`{call['name']}` is applied to the same qubit twice")
            self.expand(call["modifiers"], call["name"], params, qubits, at, self.out, 0)

    def expand(self, modifiers, name, params, qubits, at, out, depth):
        if depth > 64:
            self.fail(at, f"gate `{name}` is defined recursively")
        if modifiers:
            controls = sum(m[1] for m in modifiers if m[0] in ("ctrl", "negctrl"))
            if len(qubits) <= controls:
                self.fail(at, f"`{name}` with {controls} controls needs more qubits")
            k = len(qubits) - controls
            body = []
            self.expand([], name, params, list(range(k)), at, body, depth + 1)
            m = np.eye(1 << k, dtype=np.complex128)
            for ins in body:
                if ins[0] == "gate":
                    g = _native.gate_matrix(ins[1], ins[3])
                else:
                    g = ins[1]
                m = embed(g, ins[2], k) @ m
            # right to left: the modifier nearest the gate first
            for kind, arg in reversed(modifiers):
                if kind == "inv":
                    m = m.conj().T
                elif kind == "pow":
                    p = evaluate(arg, {})
                    if p != int(p):
                        self.fail(at, "only integer powers are supported")
                    base = m.conj().T if p < 0 else m
                    r = np.eye(m.shape[0], dtype=np.complex128)
                    for _ in range(abs(int(p))):
                        r = base @ r
                    m = r
                else:
                    m = controlled(m, arg, (1 << arg) - 1 if kind == "ctrl" else 0)
            # the controls nearest the gate come last among the operands
            out.append(("unitary", m, list(qubits)))
            return
        if name in self.gates:
            d = self.gates[name]
            if len(d["params"]) != len(params) or len(d["qubits"]) != len(qubits):
                self.fail(at, f"`{name}` takes {len(d['params'])} parameters and {len(d['qubits'])} qubits, given "
                              f"{len(params)} and {len(qubits)}")
            env = dict(zip(d["params"], params))
            where = dict(zip(d["qubits"], qubits))
            for c in d["body"]:
                values = [evaluate(e, env) for e in c["params"]]
                args = []
                for a, _ in c["args"]:
                    if a not in where:
                        self.fail(c["at"], f"unknown qubit `{a}` in the definition of `{name}`")
                    args.append(where[a])
                self.expand(c["modifiers"], c["name"], values, args, c["at"], out, depth + 1)
            return
        b = builtin(name)
        if b is None:
            self.fail(at, f"unknown gate `{name}`")
        nparams, arity = b
        if len(params) != nparams:
            self.fail(at, f"`{name}` takes {nparams} parameters, given {len(params)}")
        if arity != len(qubits):
            self.fail(at, f"`{name}` acts on {arity} qubits, given {len(qubits)}")
        out.append(("gate", name, list(qubits), list(params)))

    def definition(self):
        self.next()
        name, at = self.ident()
        params, qubits = [], []
        if self.is_sym("("):
            self.next()
            while not self.is_sym(")"):
                params.append(self.ident()[0])
                if self.is_sym(","):
                    self.next()
            self.expect(")")
        while not self.is_sym("{"):
            qubits.append(self.ident()[0])
            if self.is_sym(","):
                self.next()
        self.expect("{")
        body = []
        while not self.is_sym("}"):
            if self.peek()[0] == "end":
                self.fail(at, f"unterminated definition of `{name}`")
            if self.peek()[0] == "id" and self.peek()[1] == "barrier":
                while not self.is_sym(";"):
                    self.next()
                self.expect(";")
                continue
            c = self.call()
            if any(index_at is not None for _, index_at in c["args"]):
                self.fail(at, f"the body of `{name}` can only use its own qubit names")
            body.append(c)
            self.expect(";")
        self.expect("}")
        self.gates[name] = {"params": params, "qubits": qubits, "body": body}

    # -- expressions: + - over * / over ^ ** (right associative) over unary --
    def expression(self):
        left = self.term()
        while self.is_sym("+") or self.is_sym("-"):
            op = self.next()[1]
            left = ("bin", op, left, self.term())
        return left

    def term(self):
        left = self.power()
        while self.is_sym("*") or self.is_sym("/"):
            op = self.next()[1]
            left = ("bin", op, left, self.power())
        return left

    def power(self):
        base = self.unary()
        if self.is_sym("^") or self.is_sym("**"):
            self.next()
            return ("bin", "^", base, self.power())
        return base

    def unary(self):
        if self.is_sym("-"):
            self.next()
            return ("neg", self.unary())
        if self.is_sym("+"):
            self.next()
            return self.unary()
        tok = self.next()
        if tok[0] == "num":
            return ("num", tok[1])
        if tok[0] == "sym" and tok[1] == "(":
            e = self.expression()
            self.expect(")")
            return e
        if tok[0] == "id":
            if self.is_sym("("):
                self.next()
                arg = self.expression()
                self.expect(")")
                return ("call", tok[1], arg, tok[2], tok[3])
            return ("name", tok[1], tok[2], tok[3])
        self.fail(tok, f"expected an expression, found {describe(tok)}")


_PRELUDE_GATES = None


def parse(source: str) -> Circuit:
    """An OpenQASM 2 or 3 program as a Circuit; qubits and classical bits are
    numbered in declaration order across registers (src/qasm.rs:55-77)."""
    global _PRELUDE_GATES
    if _PRELUDE_GATES is None:
        pre = Parser(lex(PRELUDE))
        pre.program()
        _PRELUDE_GATES = pre.gates
    p = Parser(lex(source), _PRELUDE_GATES)
    p.program()
    c = Circuit(p.qubits, p.clbits)
    for ins in p.out:
        kind = ins[0]
        if kind == "gate":
            c.gate(ins[1], ins[2], ins[3])
        elif kind == "unitary":
            c.unitary(ins[1], ins[2])
        elif kind == "measure":
            c.measure(ins[1], ins[2])
        elif kind == "reset":
            c.reset(ins[1])
        elif kind == "barrier":
            c.barrier()
        else:
            c.feedback(ins[1], ins[2], ins[3])
    return c
