"""C/HIP emitter for staged pointwise algebra.

Takes a :class:`~sympygr_amd.codegen.bssn_eqs.Stages` object, runs one
common-subexpression pass over every stage and output together, orders the
resulting statements so each temporary is computed right before its first
consumer (short live ranges, which is what keeps the GPU kernel's VGPR count
down), and prints straight-line C++ templated on the arithmetic type ``R``.

Inputs are not function arguments: the printed code reads them through macros
that the including translation unit defines, at the first statement that uses
them:

    BSSN_U(f)          field f at this point
    BSSN_D1(f, d)      first derivative of field f along d
    BSSN_D2(f, s)      second derivative of field f, symmetric slot s
    BSSN_ADV(f)        advection of field f: beta^d times its advective derivative
    BSSN_LAP(f)        igt^{lm} d_l d_m of field f (conformal metric, shift)
    BSSN_GDB(j)        d_j (d_k beta^k)
    BSSN_OUT(f, v)     store RHS of field f
    BSSN_OUT_ALT(f, v) its value under the alternative (Rochester) gauge

Parameters (lambda0..3, lambda_f0..1, eta) are plain variables of type R
that must be in scope.  This is the project's own printer; it does not use or
re-target the reference's code printer (GR/dendro.py:567-622).
"""
from __future__ import annotations

import re

import sympy as sp
from sympy.printing.c import C99CodePrinter

from .bssn_eqs import FIELD_INDEX as _BSSN_FIELD_INDEX, SYM

# Field indices of the unzipped arrays: bssn::VAR order, and Theta after
# them in the CCZ4 system (field 24).
FIELD_INDEX = dict(_BSSN_FIELD_INDEX, Theta=24)

_D1 = re.compile(r"^grad_(\d)_(\w+)$")
_D2 = re.compile(r"^grad2_(\d)_(\d)_(\w+)$")
_ADV = re.compile(r"^adv_(\w+)$")
_LAP = re.compile(r"^lap_(\w+)$")
_GDB = re.compile(r"^gdivb(\d)$")


def classify_input(name: str):
    """Return (kind, field, index) for an input symbol name, or None."""
    if name in FIELD_INDEX:
        return ("U", name, None)
    m = _D2.match(name)
    if m:
        return ("D2", m.group(3), SYM[(int(m.group(1)), int(m.group(2)))])
    m = _D1.match(name)
    if m:
        return ("D1", m.group(2), int(m.group(1)))
    m = _ADV.match(name)
    if m and m.group(1) in FIELD_INDEX:
        return ("ADV", m.group(1), None)
    m = _LAP.match(name)
    if m and m.group(1) in FIELD_INDEX:
        return ("LAP", m.group(1), None)
    m = _GDB.match(name)
    if m:
        return ("GDB", "beta0", int(m.group(1)))
    return None


class _Printer(C99CodePrinter):
    """Prints numbers as R(...) so the same text works for double and float."""

    def _print_Integer(self, expr):
        return f"R({int(expr)})"

    def _print_Rational(self, expr):
        return f"R({expr.p}.0/{expr.q}.0)"

    def _print_Float(self, expr):
        return f"R({float(expr)!r})"

    def _print_Pow(self, expr):
        b, e = expr.as_base_exp()
        if e == sp.Rational(1, 2) or e == sp.Rational(-1, 2):  # the Psi4 tetrad's norms
            root = f"sqrt({self._print(b)})"
            return root if e > 0 else f"(R(1)/{root})"
        if not e.is_Integer:
            raise ValueError(f"non-integer power in algebra: {expr}")
        n = int(e)
        bs = self.parenthesize(b, 100)
        if n == 0:
            return "R(1)"
        mul = "*".join([bs] * abs(n))
        if n > 0:
            return f"({mul})"
        return f"(R(1)/({mul}))"


def count_flops(expr: sp.Expr) -> int:
    """Arithmetic operations of one printed expression (add/sub/mul/div);
    a multiplication by -1 is folded into the surrounding add and not
    counted."""
    if expr.is_Atom:
        return 0
    n = 0
    if expr.is_Add:
        n += len(expr.args) - 1
    elif expr.is_Mul:
        args = [a for a in expr.args if not (a.is_Number and abs(a) == 1)]
        n += max(len(args) - 1, 0)
    elif expr.is_Pow:
        k = int(expr.exp)
        n += abs(k) - 1 + (1 if k < 0 else 0)
    return n + sum(count_flops(a) for a in expr.args)


class Emitted:
    def __init__(self, text, flops, inputs, ntemps):
        self.text = text
        self.flops = flops
        self.inputs = inputs  # set of (kind, field, index)
        self.ntemps = ntemps


def emit_stages(S, tmp_prefix="t_", indent="    ", output_order=None, barrier=False) -> Emitted:
    names = []
    exprs = []
    for _, pairs in S.stages:
        for s, e in pairs:
            names.append(s)
            exprs.append(sp.expand(e) if False else e)
    out_names = [n for n, _ in S.outputs]
    out_exprs = [e for _, e in S.outputs]
    alt = getattr(S, "alt_outputs", {})
    alt_idx = sorted(alt)
    repl, reduced = sp.cse(exprs + out_exprs + [alt[i] for i in alt_idx], symbols=sp.numbered_symbols(tmp_prefix))
    defs = {}
    for s, e in repl:
        defs[s] = e
    for s, e in zip(names, reduced[: len(names)]):
        defs[s] = e
    out_reduced = reduced[len(names): len(names) + len(out_names)]
    alt_reduced = dict(zip(alt_idx, reduced[len(names) + len(out_names):]))

    printer = _Printer()
    lines = []
    emitted = set()
    loaded = set()
    inputs = set()
    flops = 0

    def load_inputs(expr):
        for fs in sorted(expr.free_symbols, key=lambda x: x.name):
            if fs in defs or fs in loaded:
                continue
            c = classify_input(fs.name)
            if c is None:
                continue  # a parameter in scope
            loaded.add(fs)
            inputs.add(c)
            kind, f, idx = c
            fi = FIELD_INDEX[f]
            if kind == "U":
                lines.append(f"{indent}const R {fs.name} = BSSN_U({fi});")
            elif kind in ("ADV", "LAP"):
                lines.append(f"{indent}const R {fs.name} = BSSN_{kind}({fi});")
            elif kind == "GDB":
                lines.append(f"{indent}const R {fs.name} = BSSN_GDB({idx});")
            else:
                lines.append(f"{indent}const R {fs.name} = BSSN_{kind}({fi}, {idx});")

    def emit(sym):
        nonlocal flops
        if sym in emitted:
            return
        expr = defs[sym]
        for fs in sorted(expr.free_symbols, key=lambda x: x.name):
            if fs in defs:
                emit(fs)
        load_inputs(expr)
        flops += count_flops(expr)
        lines.append(f"{indent}const R {sym.name} = {printer.doprint(expr)};")
        emitted.add(sym)

    if not output_order:
        for s in names:
            emit(s)
    order = output_order or list(range(len(out_names)))
    for oi in order:
        n, e = out_names[oi], out_reduced[oi]
        for fs in sorted(e.free_symbols, key=lambda x: x.name):
            if fs in defs:
                emit(fs)
        load_inputs(e)
        flops += count_flops(e)
        lines.append(f"{indent}BSSN_OUT({oi}, {printer.doprint(e)});")
        if barrier:
            lines.append(f"{indent}BSSN_BARRIER();")
    # alternative-gauge values of this group's outputs, in one block at the
    # end: their extra inputs are loaded only when that gauge is on
    alts = [oi for oi in order if oi in alt_reduced]
    if alts:
        outer = indent
        indent = outer + "    "
        saved_loaded, saved_emitted = set(loaded), set(emitted)
        lines.append(f"{outer}if (BSSN_ALT_ON) {{")
        for oi in alts:
            ea = alt_reduced[oi]
            for fs in sorted(ea.free_symbols, key=lambda x: x.name):
                if fs in defs:
                    emit(fs)
            load_inputs(ea)
            flops += count_flops(ea)
            lines.append(f"{indent}BSSN_OUT_ALT({oi}, {printer.doprint(ea)});")
        lines.append(f"{outer}}}")
        loaded.clear()
        loaded.update(saved_loaded)
        emitted.clear()
        emitted.update(saved_emitted)
    return Emitted("\n".join(lines) + "\n", flops, inputs, len(emitted))
