"""The generated algebra is reproducible from the SymPy statement, and the
statement is the reference's system (checked numerically in test_oracle.py
through the fixtures; here: structure and exact symbolic identities)."""
import os

import sympy as sp

from sympygr_amd.codegen import __main__ as gen
from sympygr_amd.codegen.bssn_eqs import FIELDS, bssn_rhs_stages

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_committed_algebra_is_up_to_date():
    E = gen.generate(write=False)
    text = gen.HEADER.format(flops=E.flops, ntemps=E.ntemps, ninputs=len(E.inputs)) + E.text
    with open(os.path.join(ROOT, "oracle/gen/bssn_algebra.inc")) as f:
        assert f.read() == text, "oracle/gen/bssn_algebra.inc is stale: run python -m sympygr_amd.codegen"
    for g, (_, gtext) in gen.generate_groups(write=False).items():
        rel = f"sympygr_amd/csrc/gen/bssn_algebra_{g}.inc"
        with open(os.path.join(ROOT, rel)) as f:
            assert f.read() == gtext, f"{rel} is stale: run python -m sympygr_amd.codegen"


def test_input_set_matches_reference_derivative_lists():
    """45 first derivatives and the second derivatives of the 11 fields that
    bssn/scripts/bssnrhs_*.h allocate them for: alpha and chi in full (12),
    beta0..2 and gt0..5 only through igt^{lm} d_l d_m f (9) and, for the
    shift, d_j d_k beta^k (3); the 72 advective derivatives enter only as
    one contraction beta^d agrad_d f per field (24); the 24 field values
    (the Rochester gauge's shift equation reads beta and Gt themselves)."""
    E = gen.generate(write=False)
    kinds = {}
    for k, f, i in E.inputs:
        kinds.setdefault(k, set()).add((f, i))
    assert len(kinds["D1"]) == 45
    assert len(kinds["D2"]) == 12
    assert {f for f, _ in kinds["LAP"]} == {f"gt{s}" for s in range(6)} | {"beta0", "beta1", "beta2"}
    assert {i for _, i in kinds["GDB"]} == {0, 1, 2}
    assert len(kinds["ADV"]) == 24
    assert len(kinds["U"]) == 24
    assert {f for f, _ in kinds["ADV"]} == set(FIELDS)
    d2fields = {f for f, _ in kinds["D2"]}
    assert d2fields == {"alpha", "chi"}


def _substituted_outputs():
    S = bssn_rhs_stages()
    sub = {}
    for _, pairs in S.stages:
        for s, e in pairs:
            sub[s] = e.xreplace(sub)
    return {n: e.xreplace(sub) for n, e in S.outputs}


def test_factorised_ricci_equals_textbook_contraction():
    """The C1u factorisation of the quadratic Ricci terms equals the direct
    triple sum of GR/dendro.py:495-499, symbolically (random rational point)."""
    import random
    from sympygr_amd.codegen.bssn_eqs import Inputs, SYM, E3, E33, sym_slot
    random.seed(3)
    I = Inputs()
    vals = {}
    gt = [[I.u[f"gt{sym_slot(i, j)}"] for j in E3] for i in E3]
    for s in range(6):
        vals[I.u[f"gt{s}"]] = sp.Rational(random.randint(5, 9), 4) if s in (0, 3, 5) else sp.Rational(random.randint(-3, 3), 7)
    for k in E3:
        for s in range(6):
            vals[I.d1(f"gt{s}", k)] = sp.Rational(random.randint(-9, 9), 5)
    g = sp.Matrix(3, 3, lambda i, j: gt[i][j]).xreplace(vals)
    ig = g.inv()
    dg = lambda k, i, j: I.d1(f"gt{sym_slot(i, j)}", k).xreplace(vals)
    C1 = [[[sp.Rational(1, 2) * (dg(j, k, i) + dg(i, k, j) - dg(k, i, j)) for j in E3] for i in E3] for k in E3]
    C2 = [[[sum(ig[i, l] * C1[l][j][k] for l in E3) for k in E3] for j in E3] for i in E3]
    direct = {}
    for (i, j) in SYM:
        direct[(i, j)] = sum(ig[l, m] * (C2[k][l][i] * C1[j][k][m] + C2[k][l][j] * C1[i][k][m] + C2[k][i][m] * C1[k][l][j])
                             for k in E3 for (l, m) in E33)
    S = bssn_rhs_stages()
    sub = dict(vals)
    quad = {}
    for name, pairs in S.stages:
        for s, e in pairs:
            sub[s] = e.xreplace(sub)
    C1u = lambda a, b, l: sub[sp.Symbol(f"C1u_{a}_{b}_{l}")]
    C2s = lambda i, j, k: sub[sp.Symbol(f"C2_{i}_{min(j, k)}{max(j, k)}")]
    for (i, j) in SYM:
        q = sum(C2s(k, l, i) * C1u(j, k, l) + C2s(k, l, j) * C1u(i, k, l) for k in E3 for l in E3)
        q += sum(C2s(k, i, m) * C1u(k, j, m) for k in E3 for m in E3)
        assert sp.simplify(q - direct[(i, j)]) == 0


def test_tail_variant_differs_from_full_algebra_by_the_tails():
    """The GPU kernels run the algebra with every advection left out and add
    per-output tails.  Symbolically, full - tails equals exactly the tail's
    advection part: c_f adv_f (c = lambda0 for alpha, lambda1 for the shift,
    1 otherwise), and (1 - lambda3) adv_Gt^i + lambda2 adv_B^i for B^i; in the
    Rochester alternatives xi1 adv_beta^i for the shift and 0 for B."""
    from sympygr_amd.codegen.bssn_eqs import RHS_NAMES, Inputs
    full, tl = bssn_rhs_stages(), bssn_rhs_stages(tails=True)
    fs = dict(p for _, pairs in full.stages for p in pairs)
    ts = dict(p for _, pairs in tl.stages for p in pairs)
    I = Inputs()
    P = I.params
    adv = I.adv
    # stages differ only where an advection enters (the Gt equation)
    for s, e in fs.items():
        d = sp.expand(e - ts[s])
        if s.name.startswith("Gt_rhs_s"):
            assert d == sp.expand(adv(f"Gt{s.name[-1]}")), s
        else:
            assert d == 0, s
    for k, ((n, ef), (_, et)) in enumerate(zip(full.outputs, tl.outputs)):
        f = FIELDS[k]
        d = ef - et
        if f.startswith("B"):
            i = int(f[1])
            # both read the stage symbol Gt_rhs_s<i>, which in the full
            # algebra carries + adv Gt^i (checked above)
            want = P["lambda2"] * adv(f) - P["lambda3"] * adv(f"Gt{i}")
            assert sp.expand(d - want) == 0, (n, d)  # + adv Gt^i arrives through the stage
        elif f.startswith("Gt"):
            assert d == 0, n  # its advection arrives through the Gt_rhs_s stage (checked above)
        elif f == "alpha":
            assert sp.expand(d - P["lambda0"] * adv(f)) == 0, n
        elif f.startswith("beta"):
            assert sp.expand(d - P["lambda1"] * adv(f)) == 0, n
        else:
            assert sp.expand(d - adv(f)) == 0, (n, d)
    for k, e in full.alt_outputs.items():
        d = sp.expand(e - tl.alt_outputs[k])
        f = FIELDS[k]
        assert d == (P["xi1"] * adv(f) if f.startswith("beta") else 0), f
    assert len(RHS_NAMES) == 24
