"""Field-staged emission of the pointwise algebra for the fused kernel.

The fused kernel visits the 24 fields one after another; while field f's
tile is in LDS, every thread evaluates the derivatives of f at its point.
This emitter orders the CSE'd algebra so that each statement is computed in
the first stage at which all of its inputs exist (ASAP), and each derivative
value is read right when its field is staged.  Derivative registers therefore
die within their own stage, and what stays live across stages are the
contracted intermediates (Christoffels, Ricci pieces, partial RHS terms).

Macros the including kernel defines (all inside one function body):

    BSSN_FIRST_FIELD(f)          field staged first (loaded by the prologue)
    BSSN_STAGE_BEGIN(f, s, d2, fn)  field f becomes current (stage s; d2: it
                                 has second/mixed derivatives; fn: the field
                                 staged next, -1 after the last)
    BSSN_STAGE_END(f, s, fn)     end of stage s
    BSSN_U(f)                    value of field f at the point (any time)
    BSSN_D1(f, d) BSSN_D2(f, s) BSSN_ADV(f) BSSN_LAP(f)   derivatives of the current field
    BSSN_FIELD_DONE(f)           per-field work that needs only field f
                                 (Kreiss-Oliger, boundary value)
    BSSN_OUT(f, v)               final value of RHS f
"""
from __future__ import annotations

import sympy as sp

from .bssn_eqs import FIELD_INDEX, FIELDS
from .emit import _Printer, classify_input, count_flops

# beta first (its derivatives feed every Lie derivative), then the metric
# (Christoffels, Ricci), chi, alpha, Gamma, K, B and At.
DEFAULT_FIELD_ORDER = ["beta0", "beta1", "beta2", "gt0", "gt1", "gt2", "gt3", "gt4", "gt5", "chi", "alpha",
                       "Gt0", "Gt1", "Gt2", "K", "B0", "B1", "B2", "At0", "At1", "At2", "At3", "At4", "At5"]

D2_FIELDS = {"alpha", "chi", "beta0", "beta1", "beta2", "gt0", "gt1", "gt2", "gt3", "gt4", "gt5"}


class StagedEmitted:
    def __init__(self, text, flops, max_live, live_profile):
        self.text = text
        self.flops = flops
        self.max_live = max_live
        self.live_profile = live_profile


def emit_staged(S, field_order=None, indent="    ", outputs=None) -> StagedEmitted:
    """outputs: indices of S.outputs to emit (default all).  Only fields whose
    derivatives the selected outputs need, or which are selected outputs
    themselves (Kreiss-Oliger, boundary value), are staged."""
    order = list(field_order or DEFAULT_FIELD_ORDER)
    assert sorted(order) == sorted(FIELDS)

    names, exprs = [], []
    for _, pairs in S.stages:
        for s, e in pairs:
            names.append(s)
            exprs.append(e)
    sel = list(range(len(S.outputs))) if outputs is None else list(outputs)
    out_names = [S.outputs[i][0] for i in sel]
    out_exprs = [S.outputs[i][1] for i in sel]
    out_fields = [FIELDS[i] for i in sel]
    # fields that must be staged: those with a derivative input reachable
    # from the selected outputs, plus the output fields themselves
    defs0 = dict(zip(names, exprs))
    need = set(out_fields)
    seen_sym = set()
    stack = list(out_exprs)
    while stack:
        e = stack.pop()
        for x in e.free_symbols:
            if x in seen_sym:
                continue
            seen_sym.add(x)
            if x in defs0:
                stack.append(defs0[x])
            else:
                c = classify_input(x.name)
                if c is not None and c[0] != "U":
                    need.add(c[1])
    order = [f for f in order if f in need]
    stage_of_field = {f: i for i, f in enumerate(order)}
    repl, reduced = sp.cse(exprs + out_exprs, symbols=sp.numbered_symbols("t_"))
    defs = {}
    for s, e in repl:
        defs[s] = e
    for s, e in zip(names, reduced[: len(names)]):
        defs[s] = e
    out_reduced = reduced[len(names):]
    # keep only what the selected outputs reach
    reach, stack = set(), list(out_reduced)
    while stack:
        e = stack.pop()
        for x in e.free_symbols:
            if x in defs and x not in reach:
                reach.add(x)
                stack.append(defs[x])
    defs = {k: v for k, v in defs.items() if k in reach}

    # ready stage of every symbol: field values -1, derivatives of f at stage(f)
    ready = {}

    def rdy(sym):
        if sym in ready:
            return ready[sym]
        if sym in defs:
            r = max([rdy(x) for x in defs[sym].free_symbols] + [-1])
        else:
            c = classify_input(sym.name)
            if c is None or c[0] == "U":
                r = -1
            else:
                r = stage_of_field[c[1]]
        ready[sym] = r
        return r

    # Split every sum whose terms become available at different stages into
    # running partial sums, one per stage, so early terms (e.g. the second
    # derivatives of the metric in the Ricci tensor) are folded in as soon as
    # they exist instead of staying live until the last term is ready.
    counter = [0]

    def split(expr):
        if expr.is_Atom:
            return expr
        args = [split(a) for a in expr.args]
        if expr.is_Add:
            groups = {}
            for a in args:
                groups.setdefault(max([rdy(x) for x in a.free_symbols] + [-1]), []).append(a)
            if len(groups) > 1:
                acc = None
                for st in sorted(groups):
                    part = sp.Add(*groups[st])
                    if acc is not None:
                        part = sp.Add(acc, part, evaluate=False)
                    counter[0] += 1
                    q = sp.Symbol(f"p_{counter[0]}")
                    defs[q] = part
                    ready[q] = st
                    acc = q
                return acc
            return sp.Add(*args)
        return expr.func(*args)

    for sym in list(defs.keys()):
        if sym.name.startswith("p_"):
            continue
        defs[sym] = split(defs[sym])
    out_reduced = [split(e) for e in out_reduced]

    out_ready = [max([rdy(x) for x in e.free_symbols] + [-1]) for e in out_reduced]
    for i, n in enumerate(out_names):
        f = out_fields[i]
        assert out_ready[i] >= stage_of_field[f], (n, "output ready before its own field is staged")

    # derivative inputs used, grouped by stage
    deriv_inputs = {st: [] for st in range(len(order))}
    seen = set()
    for e in list(defs.values()) + out_reduced:
        for x in e.free_symbols:
            if x in defs or x in seen:
                continue
            c = classify_input(x.name)
            if c is not None and c[0] != "U":
                seen.add(x)
                deriv_inputs[stage_of_field[c[1]]].append((x, c))

    printer = _Printer()
    lines = []
    emitted = set()
    loaded_u = set()
    flops = 0
    # liveness bookkeeping: definition position and last use position
    pos_def, pos_last = {}, {}
    step = [0]

    def touch_uses(expr):
        for x in expr.free_symbols:
            pos_last[x] = step[0]

    def load_u(expr):
        for x in sorted(expr.free_symbols, key=lambda q: q.name):
            if x in defs or x in loaded_u:
                continue
            c = classify_input(x.name)
            if c is not None and c[0] == "U":
                loaded_u.add(x)
                lines.append(f"{indent}const R {x.name} = BSSN_U({FIELD_INDEX[c[1]]});")
                pos_def[x] = step[0]

    def emit(sym):
        nonlocal flops
        if sym in emitted:
            return
        e = defs[sym]
        for x in sorted(e.free_symbols, key=lambda q: q.name):
            if x in defs:
                emit(x)
        load_u(e)
        flops += count_flops(e)
        step[0] += 1
        touch_uses(e)
        lines.append(f"{indent}const R {sym.name} = {printer.doprint(e)};")
        pos_def[sym] = step[0]
        emitted.add(sym)

    # statements that read a derivative of stage s directly are emitted in
    # stage s (so the derivative registers die there); everything else is
    # emitted just in time, right before its first consumer.
    direct = {}
    for sym, e in defs.items():
        st = [stage_of_field[c[1]] for x in e.free_symbols if x not in defs
              for c in [classify_input(x.name)] if c is not None and c[0] != "U"]
        if (st and max(st) == rdy(sym)) or sym.name.startswith("p_"):
            direct.setdefault(rdy(sym), []).append(sym)
    rank = {s: i for i, s in enumerate(list(defs.keys()))}

    def nonlocal_flops(e):
        nonlocal flops
        flops += count_flops(e)

    def emit_ready(stage):
        for sym in sorted(direct.get(stage, []), key=lambda q: rank[q]):
            emit(sym)
        for i in range(len(out_names)):
            if out_ready[i] == stage:
                e = out_reduced[i]
                for x in sorted(e.free_symbols, key=lambda q: q.name):
                    if x in defs:
                        emit(x)
                load_u(e)
                nonlocal_flops(e)
                step[0] += 1
                touch_uses(e)
                lines.append(f"{indent}BSSN_OUT({FIELD_INDEX[out_fields[i]]}, {printer.doprint(e)});")

    lines.append(f"{indent}BSSN_FIRST_FIELD({FIELD_INDEX[order[0]]});")
    emit_ready(-1)
    for s, f in enumerate(order):
        fi = FIELD_INDEX[f]
        fnext = FIELD_INDEX[order[s + 1]] if s + 1 < len(order) else -1
        lines.append(f"{indent}BSSN_STAGE_BEGIN({fi}, {s}, {1 if f in D2_FIELDS else 0}, {fnext});")
        for x, (kind, fld, idx) in sorted(deriv_inputs[s], key=lambda t: t[0].name):
            args = f"{fi}" if kind in ("ADV", "LAP") else f"{fi}, {idx}"
            lines.append(f"{indent}const R {x.name} = BSSN_{kind}({args});")
            step[0] += 1
            pos_def[x] = step[0]
        if f in out_fields:
            lines.append(f"{indent}BSSN_FIELD_DONE({fi});")
        emit_ready(s)
        lines.append(f"{indent}BSSN_STAGE_END({fi}, {s}, {fnext});")

    # live-value profile at every step (values of type R)
    events = []
    for x, d in pos_def.items():
        last = pos_last.get(x, d)
        events.append((d, 1))
        events.append((last + 1, -1))
    events.sort()
    live, mx, prof = 0, 0, []
    for p, dlt in events:
        live += dlt
        mx = max(mx, live)
        prof.append((p, live))
    return StagedEmitted("\n".join(lines) + "\n", flops, mx, prof)
