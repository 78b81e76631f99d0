"""Symbolic statement of the BSSN right-hand side (standard puncture gauge,
constant eta), written as an ordered list of staged assignments.

This is this project's own statement of the equations the reference states in
GR/bssn.py:57-104 (``bssn_puncture_gauge(eta, isStaged=False)``) with the
tensor helpers of GR/dendro.py:170-527.  The result is mathematically equal
to the reference's generated pointwise code
(bssn/src/bssneqs_eta_const_standard_gauge.cpp); it differs only in how the
arithmetic is grouped:

* intermediate tensors (inverse metric, Christoffel symbols, Ricci pieces,
  raised A~) are kept as named stages instead of being substituted into one
  expression and CSE'd as a whole (GR/dendro.py:541-563);
* the quadratic Christoffel terms of the conformal Ricci tensor are contracted
  through a once-raised first-kind Christoffel ``C1u[a][b][l] =
  sum_m igt[l][m] C1[a][b][m]``, which removes about a third of the flops.

Variable names follow the reference: fields ``alpha chi K Gt0..2 beta0..2
B0..2 gt0..5 At0..5`` (bssn/include/grDef.h:31, symmetric order
xx,xy,xz,yy,yz,zz as GR/dendro.py:86-98), first derivatives
``grad_<d>_<f>``, second derivatives ``grad2_<i>_<j>_<f>`` with i<=j
(bssn/scripts/bssnrhs_derivs*.h).  The advective derivatives ``agrad_<d>_<f>``
appear in the reference system only as beta^d agrad_d f; that contraction is
the input ``adv_<f>``, one per field.  The second derivatives of gt_ij appear
only as igt^{lm} grad2_l_m_gt_ij, the input ``lap_<gt_ij>``, and those of the
shift only as ``lap_beta<i>`` and ``gdivb<j>`` = d_j d_k beta^k.
"""
from __future__ import annotations

import sympy as sp

# Field order of the unzipped arrays, bssn/include/grDef.h:31.
FIELDS = [
    "alpha", "chi", "K", "Gt0", "Gt1", "Gt2", "beta0", "beta1", "beta2",
    "B0", "B1", "B2", "gt0", "gt1", "gt2", "gt3", "gt4", "gt5",
    "At0", "At1", "At2", "At3", "At4", "At5",
]
FIELD_INDEX = {n: i for i, n in enumerate(FIELDS)}

# Output names in the same order (bssn/src/rhs.cpp:114-137).
RHS_NAMES = [
    "a_rhs", "chi_rhs", "K_rhs", "Gt_rhs0", "Gt_rhs1", "Gt_rhs2",
    "b_rhs0", "b_rhs1", "b_rhs2", "B_rhs0", "B_rhs1", "B_rhs2",
    "gt_rhs00", "gt_rhs01", "gt_rhs02", "gt_rhs11", "gt_rhs12", "gt_rhs22",
    "At_rhs00", "At_rhs01", "At_rhs02", "At_rhs11", "At_rhs12", "At_rhs22",
]

# symmetric 3x3 storage slot of (i, j)
SYM = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}


def sym_slot(i: int, j: int) -> int:
    return SYM[(min(i, j), max(i, j))]


E3 = (0, 1, 2)
E33 = [(i, j) for i in E3 for j in E3]


class Inputs:
    """Symbols for everything the pointwise algebra reads at one point."""

    def __init__(self):
        self.u = {n: sp.Symbol(n) for n in FIELDS}
        self.params = {
            "lambda0": sp.Symbol("lambda0"), "lambda1": sp.Symbol("lambda1"),
            "lambda2": sp.Symbol("lambda2"), "lambda3": sp.Symbol("lambda3"),
            "lambda_f0": sp.Symbol("lambda_f0"), "lambda_f1": sp.Symbol("lambda_f1"),
            "eta": sp.Symbol("eta"),
            "xi1": sp.Symbol("xi1"), "xi2": sp.Symbol("xi2"),
        }

    # d_i f  (centred, with bflag closures)
    @staticmethod
    def d1(f: str, i: int) -> sp.Symbol:
        return sp.Symbol(f"grad_{i}_{f}")

    # d_i d_j f, symmetric
    @staticmethod
    def d2(f: str, i: int, j: int) -> sp.Symbol:
        a, b = min(i, j), max(i, j)
        return sp.Symbol(f"grad2_{a}_{b}_{f}")

    # advection of field f: beta^i ad_i f, the only form in which the
    # advective derivatives agrad_i_f (upwinded by sign(beta^i), or = d1 in
    # centred mode) enter the system; evaluated beside the derivatives as
    # (beta0 ad_0 f + beta1 ad_1 f) + beta2 ad_2 f
    @staticmethod
    def adv(f: str) -> sp.Symbol:
        return sp.Symbol(f"adv_{f}")

    # igt^{lm} d_l d_m gt_ij: the only form in which the second derivatives of
    # the conformal metric enter (the first term of the conformal Ricci
    # tensor), evaluated beside the derivatives with the inverse metric of the
    # point as sum over slots s of w_s igt_s d2_s gt_ij, w = (1,2,2,1,2,1)
    @staticmethod
    def lap(f: str) -> sp.Symbol:
        return sp.Symbol(f"lap_{f}")

    # d_j (d_k beta^k), j = 0..2: the other form in which the shift's second
    # derivatives enter (the Gamma-driver), evaluated beside the derivatives
    # as (d_j d_0 beta^0 + d_j d_1 beta^1) + d_j d_2 beta^2
    @staticmethod
    def gdivb(j: int) -> sp.Symbol:
        return sp.Symbol(f"gdivb{j}")


def _sym3(prefix: str):
    return [[sp.Symbol(f"{prefix}{sym_slot(i, j)}") for j in E3] for i in E3]


class Stages:
    """Ordered list of named stages; each stage is a list of (symbol, expr)
    whose expressions only reference inputs and symbols of earlier stages."""

    def __init__(self):
        self.stages: list[tuple[str, list[tuple[sp.Symbol, sp.Expr]]]] = []
        self.outputs: list[tuple[str, sp.Expr]] = []
        # alternative value of an output under another gauge: output index ->
        # expression (stored instead of the output when that gauge is on)
        self.alt_outputs: dict[int, sp.Expr] = {}

    def stage(self, name: str, pairs):
        self.stages.append((name, list(pairs)))


# Fields whose advection the GPU pointwise kernel does not read: the stencil
# pass folds it, with the Kreiss-Oliger term, into a per-output tail that the
# kernel adds after the algebra (every field).
TAIL_FIELDS = set(FIELDS)


def bssn_rhs_stages(contract_beta: bool = True, tails: bool = False) -> Stages:
    """BSSN RHS, GR/bssn.py:57-104, as staged assignments.  contract_beta:
    the shift's second derivatives enter as igt^{jk} d_j d_k beta^i
    (lap_beta<i>) and d_j (d_k beta^k) (gdivb<j>), formed beside the
    derivatives, instead of as 18 separate inputs.  tails: the advection of
    the TAIL_FIELDS is left out (zero); the caller adds, per output, the tail
    c_f adv f + sigma KO f (c_f the advection's coefficient in the equation;
    for B^i: (1 - lambda3) adv Gt^i + lambda2 adv B^i + sigma KO B^i), which
    makes the outputs equal to the full ones."""
    I = Inputs()
    u = I.u
    P = I.params
    d1, d2 = I.d1, I.d2

    def adv(f):
        return sp.Integer(0) if tails and f in TAIL_FIELDS else I.adv(f)
    S = Stages()

    alpha, chi, K = u["alpha"], u["chi"], u["K"]
    b = [u["beta0"], u["beta1"], u["beta2"]]
    bn = ["beta0", "beta1", "beta2"]
    B = [u["B0"], u["B1"], u["B2"]]
    Bn = ["B0", "B1", "B2"]
    Gt = [u["Gt0"], u["Gt1"], u["Gt2"]]
    Gtn = ["Gt0", "Gt1", "Gt2"]
    gtn = [[f"gt{sym_slot(i, j)}" for j in E3] for i in E3]
    Atn = [[f"At{sym_slot(i, j)}" for j in E3] for i in E3]
    gt = [[u[gtn[i][j]] for j in E3] for i in E3]
    At = [[u[Atn[i][j]] for j in E3] for i in E3]
    half = sp.Rational(1, 2)

    # ---- inverse conformal metric: adjugate / det (GR/dendro.py:374-389,
    # inv('ADJ')).  The determinant is not assumed to be one.
    cof = [[None] * 3 for _ in E3]
    cof[0][0] = gt[1][1] * gt[2][2] - gt[1][2] ** 2
    cof[0][1] = gt[0][2] * gt[1][2] - gt[0][1] * gt[2][2]
    cof[0][2] = gt[0][1] * gt[1][2] - gt[0][2] * gt[1][1]
    cof[1][1] = gt[0][0] * gt[2][2] - gt[0][2] ** 2
    cof[1][2] = gt[0][1] * gt[0][2] - gt[0][0] * gt[1][2]
    cof[2][2] = gt[0][0] * gt[1][1] - gt[0][1] ** 2
    detg = gt[0][0] * cof[0][0] + gt[0][1] * cof[0][1] + gt[0][2] * cof[0][2]
    idet = sp.Symbol("idet")
    S.stage("idet", [(idet, 1 / detg)])
    igt = _sym3("igt")
    S.stage("igt", [(igt[i][j], cof[i][j] * idet) for (i, j) in SYM])

    inv_chi = sp.Symbol("inv_chi")
    S.stage("inv_chi", [(inv_chi, 1 / chi)])

    # ---- first-kind Christoffel, GR/dendro.py:392-415:
    # C1[k,i,j] = 1/2 (d_j g_ki + d_i g_kj - d_k g_ij)
    C1 = [[[None] * 3 for _ in E3] for _ in E3]
    pairs = []
    for k in E3:
        for (i, j) in SYM:
            s = sp.Symbol(f"C1_{k}_{i}{j}")
            pairs.append((s, half * (d1(gtn[k][i], j) + d1(gtn[k][j], i) - d1(gtn[i][j], k))))
            C1[k][i][j] = C1[k][j][i] = s
    S.stage("C1", pairs)

    # ---- second-kind Christoffel, GR/dendro.py:418-437: C2[i,j,k] = igt[i,l] C1[l,j,k]
    C2 = [[[None] * 3 for _ in E3] for _ in E3]
    pairs = []
    for i in E3:
        for (j, k) in SYM:
            s = sp.Symbol(f"C2_{i}_{j}{k}")
            pairs.append((s, sum(igt[i][l] * C1[l][j][k] for l in E3)))
            C2[i][j][k] = C2[i][k][j] = s
    S.stage("C2", pairs)

    # igt^{im} d_m chi and the raised first-kind symbol used by Ricci
    dchi = [d1("chi", i) for i in E3]
    igt_dchi = [sp.Symbol(f"igt_dchi{i}") for i in E3]
    S.stage("igt_dchi", [(igt_dchi[i], sum(igt[i][m] * dchi[m] for m in E3)) for i in E3])

    # ---- physical-metric Christoffel, GR/dendro.py:440-465:
    # C3[i,j,k] = C2[i,j,k] - 1/(2 chi) (delta_ij d_k chi + delta_ik d_j chi - gt_jk igt^{im} d_m chi)
    C3 = [[[None] * 3 for _ in E3] for _ in E3]
    pairs = []
    for i in E3:
        for (j, k) in SYM:
            s = sp.Symbol(f"C3_{i}_{j}{k}")
            dij = 1 if i == j else 0
            dik = 1 if i == k else 0
            pairs.append((s, C2[i][j][k] - half * inv_chi * (dij * dchi[k] + dik * dchi[j] - gt[j][k] * igt_dchi[i])))
            C3[i][j][k] = C3[i][k][j] = s
    S.stage("C3", pairs)

    # ---- conformal connection from the metric, GR/dendro.py:493: CalGt^i = igt^{kl} C2[i,k,l]
    CalGt = [sp.Symbol(f"CalGt{i}") for i in E3]
    S.stage("CalGt", [(CalGt[i], sum(igt[k][l] * C2[i][k][l] for (k, l) in E33)) for i in E3])

    # C1u[a][b][l] = sum_m igt[l][m] C1[a][b][m]
    C1u = [[[sp.Symbol(f"C1u_{a}_{bb}_{l}") for l in E3] for bb in E3] for a in E3]
    S.stage("C1u", [(C1u[a][bb][l], sum(igt[l][m] * C1[a][bb][m] for m in E3))
                    for a in E3 for bb in E3 for l in E3])

    # ---- conformal Ricci, GR/dendro.py:495-499:
    # Rt_ij = -1/2 igt^{lm} d_l d_m gt_ij + 1/2 (gt_ki d_j Gt^k + gt_kj d_i Gt^k)
    #       + 1/2 CalGt^k (C1[i,j,k] + C1[j,i,k])
    #       + igt^{lm} (C2[k,l,i] C1[j,k,m] + C2[k,l,j] C1[i,k,m] + C2[k,i,m] C1[k,l,j])
    # where igt^{lm} C1[j,k,m] = C1u[j][k][l] and igt^{lm} C1[k,l,j] = C1u[k][j][m].
    Rt = [[None] * 3 for _ in E3]
    pairs = []
    for (i, j) in SYM:
        s = sp.Symbol(f"Rt{sym_slot(i, j)}")
        e = -half * I.lap(gtn[i][j])
        e += half * sum(gt[k][i] * d1(Gtn[k], j) + gt[k][j] * d1(Gtn[k], i) for k in E3)
        e += half * sum(CalGt[k] * (C1[i][j][k] + C1[j][i][k]) for k in E3)
        e += sum(C2[k][l][i] * C1u[j][k][l] + C2[k][l][j] * C1u[i][k][l] for k in E3 for l in E3)
        e += sum(C2[k][i][m] * C1u[k][j][m] for k in E3 for m in E3)
        pairs.append((s, e))
        Rt[i][j] = Rt[j][i] = s
    S.stage("Rt", pairs)

    # ---- conformal-factor part of Ricci, GR/dendro.py:509-518
    trchi = sp.Symbol("Rphi_tr")
    S.stage("Rphi_tr", [(trchi, sum(igt[k][l] * (d2("chi", k, l) - sp.Rational(3, 2) * inv_chi * dchi[k] * dchi[l]) for (k, l) in E33)
                         - sum(CalGt[m] * dchi[m] for m in E3))])
    Rphi = [[None] * 3 for _ in E3]
    pairs = []
    for (i, j) in SYM:
        s = sp.Symbol(f"Rphi{sym_slot(i, j)}")
        e = half * inv_chi * (d2("chi", i, j) - sum(C2[k][j][i] * dchi[k] for k in E3)) \
            - sp.Rational(1, 4) * inv_chi ** 2 * dchi[i] * dchi[j] \
            + half * inv_chi * gt[i][j] * trchi
        pairs.append((s, e))
        Rphi[i][j] = Rphi[j][i] = s
    S.stage("Rphi", pairs)

    # ---- D_i D_j alpha with the physical metric, GR/dendro.py:177-188
    DDa = [[None] * 3 for _ in E3]
    pairs = []
    for (i, j) in SYM:
        s = sp.Symbol(f"DDa{sym_slot(i, j)}")
        pairs.append((s, d2("alpha", i, j) - sum(C3[l][i][j] * d1("alpha", l) for l in E3)))
        DDa[i][j] = DDa[j][i] = s
    S.stage("DDa", pairs)

    # ---- raised A~: At^{ij} (GR/dendro.py:206-214) and A_ik igt^kl A_lj
    AtU = [[sp.Symbol(f"AtU{sym_slot(i, j)}") for j in E3] for i in E3]
    Amix = [[sp.Symbol(f"Amix_{i}{j}") for j in E3] for i in E3]  # igt^{ik} A_kj
    S.stage("Amix", [(Amix[i][j], sum(igt[i][k] * At[k][j] for k in E3)) for i in E3 for j in E3])
    S.stage("AtU", [(AtU[i][j], sum(Amix[i][k] * igt[k][j] for k in E3)) for (i, j) in SYM])
    AA = [[sp.Symbol(f"AA{sym_slot(i, j)}") for j in E3] for i in E3]
    S.stage("AA", [(AA[i][j], sum(At[i][k] * Amix[k][j] for k in E3)) for (i, j) in SYM])

    divb = sp.Symbol("div_beta")
    S.stage("div_beta", [(divb, sum(d1(bn[k], k) for k in E3))])

    # ---- the RHS, GR/bssn.py:67-100
    out = {}
    out["a_rhs"] = P["lambda0"] * adv("alpha") - 2 * alpha * K
    for i in E3:
        out[f"b_rhs{i}"] = sp.Rational(3, 4) * (P["lambda_f0"] + P["lambda_f1"] * alpha) * B[i] \
            + P["lambda1"] * adv(bn[i])
    w = -sp.Rational(2, 3)
    for (i, j) in SYM:
        lie = adv(gtn[i][j]) + sum(gt[i][k] * d1(bn[k], j) + gt[k][j] * d1(bn[k], i) for k in E3) \
            + w * gt[i][j] * divb
        out[f"gt_rhs{i}{j}"] = lie - 2 * alpha * At[i][j]
    out["chi_rhs"] = adv("chi") + w * chi * divb + sp.Rational(2, 3) * chi * alpha * K

    # trace-free part of (alpha R - DDa), GR/dendro.py:321-336
    X = [[None] * 3 for _ in E3]
    for (i, j) in SYM:
        X[i][j] = X[j][i] = alpha * (Rt[i][j] + Rphi[i][j]) - DDa[i][j]
    trX = sp.Symbol("trX")
    S.stage("trX", [(trX, sum(igt[k][l] * X[k][l] for (k, l) in E33))])
    for (i, j) in SYM:
        lie = adv(Atn[i][j]) + sum(At[i][k] * d1(bn[k], j) + At[k][j] * d1(bn[k], i) for k in E3) \
            + w * At[i][j] * divb
        out[f"At_rhs{i}{j}"] = lie + chi * (X[i][j] - gt[i][j] * trX * sp.Rational(1, 3)) \
            + alpha * (K * At[i][j] - 2 * AA[i][j])

    # K: lie - laplacian(alpha) + alpha (K^2/3 + A_ij A^ij); laplacian uses
    # the inverse physical metric chi igt (GR/dendro.py:259-275)
    out["K_rhs"] = adv("K") \
        - chi * sum(igt[i][j] * DDa[i][j] for (i, j) in E33) \
        + alpha * (K * K * sp.Rational(1, 3) + sum(At[i][j] * AtU[i][j] for (i, j) in E33))

    Gt_rhs = []
    for i in E3:
        e = adv(Gtn[i])
        e -= sum(CalGt[j] * d1(bn[i], j) for j in E3)
        e += sp.Rational(2, 3) * CalGt[i] * divb
        if contract_beta:
            e += I.lap(bn[i]) + sum(igt[i][j] * I.gdivb(j) for j in E3) * sp.Rational(1, 3)
        else:
            e += sum(igt[j][k] * d2(bn[i], j, k) + igt[i][j] * d2(bn[k], j, k) * sp.Rational(1, 3) for (j, k) in E33)
        e -= sum(2 * AtU[i][j] * d1("alpha", j) for j in E3)
        e += sum(2 * alpha * C2[i][j][k] * AtU[j][k] for (j, k) in E33)
        e -= sum(alpha * (3 * inv_chi * AtU[i][j] * dchi[j] + sp.Rational(4, 3) * igt[i][j] * d1("K", j)) for j in E3)
        Gt_rhs.append(e)
    Gts = [sp.Symbol(f"Gt_rhs_s{i}") for i in E3]
    S.stage("Gt_rhs", [(Gts[i], Gt_rhs[i]) for i in E3])
    for i in E3:
        out[f"Gt_rhs{i}"] = Gts[i]
        out[f"B_rhs{i}"] = Gts[i] - P["eta"] * B[i] + P["lambda2"] * adv(Bn[i]) \
            - P["lambda3"] * adv(Gtn[i])

    name_map = {"gt_rhs00": "gt_rhs00", "gt_rhs01": "gt_rhs01", "gt_rhs02": "gt_rhs02",
                "gt_rhs11": "gt_rhs11", "gt_rhs12": "gt_rhs12", "gt_rhs22": "gt_rhs22"}
    for n in RHS_NAMES:
        S.outputs.append((name_map.get(n, n), out[n]))

    # Rochester puncture gauge, GR/bssn.py:194-209: the shift follows
    #   b_rhs^i = xi_1 beta^j ad_j beta^i + 3/4 xi_2 Gt^i - eta beta^i
    # (xi_k = BSSN_XI[k]) and B is not evolved (B_rhs = 0 before dissipation).
    # Every other equation is the standard one.
    for i in E3:
        S.alt_outputs[RHS_NAMES.index(f"b_rhs{i}")] = P["xi1"] * adv(bn[i]) \
            + sp.Rational(3, 4) * P["xi2"] * Gt[i] - P["eta"] * b[i]
        S.alt_outputs[RHS_NAMES.index(f"B_rhs{i}")] = sp.Integer(0)
    return S


def constraint_stages(psi4: bool = False, parts: bool = False) -> Stages:
    """Hamiltonian and momentum constraints, GR/constraints.py:161-173, with the
    Ricci tensor of GR/dendro.py:468-527 (the same staged pieces as the RHS):

        H   = chi igt^{jk} R_jk - A_ij A^ij + 2/3 K^2
        M_i = igt^{jk} (d_k A_ij - C2[m,k,i] A_jm) - Gt^j A_ij
              - 3/2 igt^{jk} A_ki d_j chi / chi - 2/3 d_i K

    Outputs are named ham, mom0..2 and stored with BSSN_OUT(0..3, v).  With
    psi4, also the real and imaginary parts of the Weyl scalar Psi4
    (GR/constraints.py:46-155; psi4_stages), BSSN_OUT(4..5, v)."""
    R = bssn_rhs_stages()
    keep = {"idet", "igt", "inv_chi", "C1", "C2", "igt_dchi", "CalGt", "C1u", "Rt", "Rphi_tr", "Rphi",
            "Amix", "AtU"}
    S = Stages()
    for name, pairs in R.stages:
        if name in keep:
            S.stage(name, pairs)
    I = Inputs()
    u = I.u
    d1 = I.d1
    chi, K = u["chi"], u["K"]
    Gt = [u["Gt0"], u["Gt1"], u["Gt2"]]
    Atn = [[f"At{sym_slot(i, j)}" for j in E3] for i in E3]
    At = [[u[Atn[i][j]] for j in E3] for i in E3]
    igt = _sym3("igt")
    C2 = [[[sp.Symbol(f"C2_{i}_{min(j, k)}{max(j, k)}") for k in E3] for j in E3] for i in E3]
    Rt = _sym3("Rt")
    Rphi = _sym3("Rphi")
    AtU = _sym3("AtU")
    inv_chi = sp.Symbol("inv_chi")
    ham = chi * sum(igt[j][k] * (Rt[j][k] + Rphi[j][k]) for (j, k) in E33) \
        - sum(At[i][j] * AtU[i][j] for (i, j) in E33) + sp.Rational(2, 3) * K * K
    S.outputs.append(("ham", ham))
    for i in E3:
        m = sum(igt[j][k] * (d1(Atn[i][j], k) - sum(C2[mm][k][i] * At[j][mm] for mm in E3)) for (j, k) in E33)
        m -= sum(Gt[j] * At[i][j] for j in E3)
        m -= sp.Rational(3, 2) * sum(igt[j][k] * At[k][i] * d1("chi", j) * inv_chi for (j, k) in E33)
        m -= sp.Rational(2, 3) * d1("K", i)
        S.outputs.append((f"mom{i}", m))
    if psi4:
        psi4_stages(S, parts)
    return S


# 1/sqrt(2) as the reference writes it (GR/constraints.py:55)
INV_SQRT2 = sp.Float("0.7071067811865475244", 30)


def psi4_stages(S: Stages, parts: bool = False) -> None:
    """The Weyl scalar Psi4 for wave extraction, GR/constraints.py:46-155,
    appended to a constraint Stages (whose igt, inv_chi, C2 and Ricci stages
    it reads) as outputs psi4_real, psi4_img.

    Tetrad: the coordinate vectors r = (x, y, z), theta = (x z, y z,
    -(x^2 + y^2)), phi = (-y, x, 0) at the point (px, py, pz, plain variables
    in scope), made orthonormal in that order by Gram-Schmidt with the
    physical metric gt / chi; m = (theta + i phi) / sqrt(2).  With
    A_i = A~_ij r^j, MM = Re(m m), NN = Im(m m), MR / NR the antisymmetric
    products of Re m / Im m with r, and U_ij (V_ij) = (Re m)^k ((Im m)^k)
    (d_j A~_ki + C2^m_ki A~_mj):

        Re Psi4 =   R.MM + a A~.MM - (MR.U - NR.V)/chi - (...)/chi^2
        Im Psi4 = -(R.NN + a A~.NN - (NR.U + MR.V)/chi - (...)/chi^2)

    with a = (A_i r^i + K chi/3 + 1/2 r.d chi) / chi^2, R the Ricci tensor of
    the physical metric and the last terms as GR/constraints.py:149-150."""
    I = Inputs()
    u = I.u
    d1 = I.d1
    chi, K = u["chi"], u["K"]
    Atn = [[f"At{sym_slot(i, j)}" for j in E3] for i in E3]
    At = [[u[Atn[i][j]] for j in E3] for i in E3]
    gt = [[u[f"gt{sym_slot(i, j)}"] for j in E3] for i in E3]
    C2 = [[[sp.Symbol(f"C2_{i}_{min(j, k)}{max(j, k)}") for k in E3] for j in E3] for i in E3]
    Rt = _sym3("Rt")
    Rphi = _sym3("Rphi")
    inv_chi = sp.Symbol("inv_chi")
    px, py, pz = sp.Symbol("px"), sp.Symbol("py"), sp.Symbol("pz")
    half = sp.Rational(1, 2)

    def dot(a, b):  # physical inner product gt_ij / chi a^i b^j
        return inv_chi * sum(gt[i][j] * a[i] * b[j] for (i, j) in E33)

    # Gram-Schmidt, r first, then theta, then phi (GR/constraints.py:57-92)
    rv = [px, py, pz]
    n_r = sp.Symbol("psi4_nr")
    S.stage("psi4_nr", [(n_r, 1 / sp.sqrt(dot(rv, rv)))])
    rn = [sp.Symbol(f"psi4_r{i}") for i in E3]
    S.stage("psi4_r", [(rn[i], rv[i] * n_r) for i in E3])
    th = [px * pz, py * pz, -(px * px + py * py)]
    t1, t2 = sp.Symbol("psi4_tt"), sp.Symbol("psi4_rt")
    S.stage("psi4_th_ip", [(t1, dot(th, th)), (t2, dot(rn, th))])
    n_t = sp.Symbol("psi4_nt")
    S.stage("psi4_nt", [(n_t, 1 / sp.sqrt(t1 - t2 * t2))])
    tn = [sp.Symbol(f"psi4_t{i}") for i in E3]
    S.stage("psi4_t", [(tn[i], (th[i] - t2 * rn[i]) * n_t) for i in E3])
    ph = [-py, px, sp.Integer(0)]
    q1, q2, q3 = sp.Symbol("psi4_pp"), sp.Symbol("psi4_rp"), sp.Symbol("psi4_tp")
    S.stage("psi4_ph_ip", [(q1, dot(ph, ph)), (q2, dot(rn, ph)), (q3, dot(tn, ph))])
    n_p = sp.Symbol("psi4_np")
    S.stage("psi4_np", [(n_p, 1 / sp.sqrt(q1 - q2 * q2 - q3 * q3))])
    pn = [sp.Symbol(f"psi4_p{i}") for i in E3]
    S.stage("psi4_p", [(pn[i], (ph[i] - q2 * rn[i] - q3 * tn[i]) * n_p) for i in E3])

    # the complex null vector m = (theta + i phi)/sqrt(2) and its products
    mr = [tn[i] * INV_SQRT2 for i in E3]
    mi = [pn[i] * INV_SQRT2 for i in E3]
    MM = [[sp.Symbol(f"psi4_MM{sym_slot(i, j)}") for j in E3] for i in E3]
    NN = [[sp.Symbol(f"psi4_NN{sym_slot(i, j)}") for j in E3] for i in E3]
    S.stage("psi4_MMNN", [(MM[i][j], mr[i] * mr[j] - mi[i] * mi[j]) for (i, j) in SYM] +
            [(NN[i][j], mr[i] * mi[j] + mr[j] * mi[i]) for (i, j) in SYM])

    def contract_sym(T, W):  # T_ij W^ij over a symmetric pair, diagonal + 2 x upper
        return sum(T[i][i] * W[i][i] for i in E3) + 2 * sum(T[i][j] * W[i][j] for i in E3 for j in E3 if j > i)

    Avec = [sp.Symbol(f"psi4_A{i}") for i in E3]
    S.stage("psi4_A", [(Avec[i], sum(At[i][j] * rn[j] for j in E3)) for i in E3])
    # U_ij, V_ij (GR/constraints.py:125-128)
    wpairs = []
    Wt = [[[None] * 3 for _ in E3] for _ in E3]
    for k in E3:
        for i in E3:
            for j in E3:
                s = sp.Symbol(f"psi4_W{k}{i}{j}")
                wpairs.append((s, d1(Atn[k][i], j) + sum(C2[m][k][i] * At[m][j] for m in E3)))
                Wt[k][i][j] = s
    S.stage("psi4_W", wpairs)
    U = [[sp.Symbol(f"psi4_U{i}{j}") for j in E3] for i in E3]
    V = [[sp.Symbol(f"psi4_V{i}{j}") for j in E3] for i in E3]
    S.stage("psi4_UV", [(U[i][j], sum(mr[k] * Wt[k][i][j] for k in E3)) for i in E3 for j in E3] +
            [(V[i][j], sum(mi[k] * Wt[k][i][j] for k in E3)) for i in E3 for j in E3])
    rdchi = sum(rn[i] * d1("chi", i) for i in E3)
    atmp = sp.Symbol("psi4_a")
    S.stage("psi4_a", [(atmp, inv_chi * inv_chi * (sum(Avec[i] * rn[i] for i in E3) + K * chi / 3 + half * rdchi))])
    mr_dchi = sum(mr[i] * d1("chi", i) for i in E3)
    mi_dchi = sum(mi[i] * d1("chi", i) for i in E3)
    mr_A = sum(mr[i] * Avec[i] for i in E3)
    mi_A = sum(mi[i] * Avec[i] for i in E3)
    R = [[Rt[i][j] + Rphi[i][j] for j in E3] for i in E3]
    MR = [[mr[i] * rn[j] - mr[j] * rn[i] for j in E3] for i in E3]
    NR = [[mi[i] * rn[j] - mi[j] * rn[i] for j in E3] for i in E3]
    p1r, p1i = contract_sym(R, MM), contract_sym(R, NN)
    p2r, p2i = atmp * contract_sym(At, MM), atmp * contract_sym(At, NN)
    p3r = inv_chi * sum(MR[i][j] * U[i][j] - NR[i][j] * V[i][j] for (i, j) in E33)
    p3i = inv_chi * sum(NR[i][j] * U[i][j] + MR[i][j] * V[i][j] for (i, j) in E33)
    p4r = inv_chi * inv_chi * (mr_A * (mr_A + half * mr_dchi) - mi_A * (mi_A + half * mi_dchi))
    p4i = inv_chi * inv_chi * (mr_A * (mi_A - half * mi_dchi) + mi_A * (mr_A - half * mr_dchi))
    S.outputs.append(("psi4_real", p1r + p2r - p3r - p4r))
    S.outputs.append(("psi4_img", -(p1i + p2i - p3i - p4i)))
    if parts:
        for n, e in (("p1r", p1r), ("p2r", p2r), ("p3r", p3r), ("p4r", p4r),
                     ("p1i", p1i), ("p2i", p2i), ("p3i", p3i), ("p4i", p4i),
                     ("mrA", mr_A), ("miA", mi_A), ("mrd", mr_dchi), ("mid", mi_dchi), ("ic", inv_chi)):
            S.outputs.append((f"psi4_{n}", e))
