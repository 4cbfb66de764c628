"""A reader for the OpenQASM 2 subset the engine's circuits use.

Covers what ``Circuit.to_qasm`` writes and common hand-written programs:
``qreg``/``creg`` declarations (several registers are laid out in order),
qelib1/stdgates gate applications with parameter expressions over numbers
and ``pi``, register broadcast (``h q;``), ``measure``, ``reset`` and
``barrier``. Gate definitions, ``if`` and OpenQASM 3 syntax are rejected
with a ValueError naming the line (the reference's full reader,
src/qasm.rs, is not part of the state-vector path).
"""
from __future__ import annotations

import ast
import math
import operator
import re

from .circuit import Circuit

_OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv,
        ast.Pow: operator.pow}


def _eval(expr: str, line: int) -> float:
    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            return float(node.value)
        if isinstance(node, ast.Name) and node.id in ("pi", "PI", "π"):
            return math.pi
        if isinstance(node, ast.BinOp) and type(node.op) in _OPS:
            return _OPS[type(node.op)](ev(node.left), ev(node.right))
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
            v = ev(node.operand)
            return -v if isinstance(node.op, ast.USub) else v
        raise ValueError(f"line {line}: unsupported parameter expression `{expr}`")

    try:
        return ev(ast.parse(expr.strip().replace("π", "pi"), mode="eval"))
    except SyntaxError:
        raise ValueError(f"line {line}: unsupported parameter expression `{expr}`") from None


def parse(source: str) -> Circuit:
    text = re.sub(r"//[^\n]*", "", source)
    qregs, cregs = {}, {}
    nq = nc = 0
    stmts = []
    line = 1
    for chunk in text.split(";"):
        stmt = chunk.strip()
        start_line = line + chunk[: len(chunk) - len(chunk.lstrip())].count("\n")
        line += chunk.count("\n")
        if stmt:
            stmts.append((start_line, " ".join(stmt.split())))
    body = []
    for ln, s in stmts:
        if s.startswith("OPENQASM"):
            if not s.split()[1].startswith("2"):
                raise ValueError(f"line {ln}: only OpenQASM 2 is read here")
            continue
        if s.startswith("include"):
            continue
        m = re.fullmatch(r"(qreg|creg)\s+(\w+)\s*\[\s*(\d+)\s*\]", s)
        if m:
            kind, name, size = m.group(1), m.group(2), int(m.group(3))
            if kind == "qreg":
                qregs[name] = (nq, size)
                nq += size
            else:
                cregs[name] = (nc, size)
                nc += size
            continue
        body.append((ln, s))
    circuit = Circuit(nq, nc)

    def operands(text_ops, regs, ln):
        out = []
        for op in text_ops.split(","):
            op = op.strip()
            m = re.fullmatch(r"(\w+)\s*\[\s*(\d+)\s*\]", op)
            if m:
                if m.group(1) not in regs:
                    raise ValueError(f"line {ln}: unknown register `{m.group(1)}`")
                base, size = regs[m.group(1)]
                idx = int(m.group(2))
                if idx >= size:
                    raise ValueError(f"line {ln}: index {idx} out of range for `{m.group(1)}`")
                out.append([base + idx])
            elif op in regs:
                base, size = regs[op]
                out.append(list(range(base, base + size)))
            else:
                raise ValueError(f"line {ln}: bad operand `{op}`")
        return out

    for ln, s in body:
        if s.startswith("gate ") or s.startswith("if") or s.startswith("opaque"):
            raise ValueError(f"line {ln}: `{s.split()[0]}` is not supported by this reader")
        if s.startswith("measure"):
            m = re.fullmatch(r"measure\s+(.+?)\s*->\s*(.+)", s)
            if not m:
                raise ValueError(f"line {ln}: bad measure")
            qs, cs = operands(m.group(1), qregs, ln)[0], operands(m.group(2), cregs, ln)[0]
            if len(qs) != len(cs):
                raise ValueError(f"line {ln}: register sizes differ")
            for q, c in zip(qs, cs):
                circuit.measure(q, c)
            continue
        if s.startswith("reset"):
            for q in operands(s[len("reset"):], qregs, ln)[0]:
                circuit.reset(q)
            continue
        if s.startswith("barrier"):
            circuit.barrier()
            continue
        m = re.fullmatch(r"(\w+)\s*(?:\((.*)\))?\s+(.+)", s)
        if not m:
            raise ValueError(f"line {ln}: cannot parse `{s}`")
        name, params_text, ops_text = m.group(1), m.group(2), m.group(3)
        params = [_eval(p, ln) for p in params_text.split(",")] if params_text and params_text.strip() else []
        groups = operands(ops_text, qregs, ln)
        width = max(len(g) for g in groups)
        for i in range(width):
            qubits = [g[i] if len(g) > 1 else g[0] for g in groups]
            try:
                circuit.gate(name, qubits, params)
            except ValueError as e:
                raise ValueError(f"line {ln}: {e}") from None
    return circuit
