"""Symbolic statement of the CCZ4 right-hand side (conformal and covariant
Z4 with constraint damping), the system the reference states in
ccz4_eqns.py:195-431 with dendrosym's helpers (dendrosym/nr.py), following
its source paper (Alic, Bona-Casas, Bona, Rezzolla, Palenzuela 2012,
arXiv:1106.2254, eqs. 14-22).

Fields: the 24 BSSN fields in bssn::VAR order, with the Gt slots holding the
evolved Gamma-hat^i of the paper, and Theta appended as field 24.

The reference file leaves three things open; this statement fixes them:

* Conformal factor.  dendrosym's compute_ricci, get_complete_christoffel and
  laplacian (dendrosym/nr.py:582-618, 966-1016, 1018-) read their scalar as
  chi with gamma_ij = gt_ij / chi, the convention of the BSSN code, while
  ccz4_eqns.py:260-262 and :279-283 write the At and phi equations in the
  paper's phi = chi^(1/2).  Here chi is used throughout: the chi equation
  and the chi prefactor of the At equation are BSSN's, so that the system
  reduces to the BSSN one when Theta = Z = 0.
* Z.  ccz4_eqns.py:140,167-173 keep Z as a free symbol with
  Gammat^i = Gammahat^i - 2 gt^ij Z_j.  Taking Gammat^i as the connection of
  the conformal metric (CalGt^i = gt^kl C2^i_kl) closes it:
  Z_i = 1/2 gt_ij (Gammahat^j - CalGt^j).
* R_ij + 2 D_(i Z_j).  Written out, the d_i Z_j of ccz4_eqns.py:230-235
  needs second derivatives of gt that no stencil provides.  The paper's own
  device makes it first-order in Z: the conformal Ricci tensor built with
  Gammahat in place of CalGt (in gt_k(i d_j) Gamma^k and Gamma^k C1_(ij)k)
  equals R~_ij + 2 D~_(i Z_j); the physical covariant derivative adds
  (Z_i d_j chi + Z_j d_i chi - gt_ij gt^kl Z_k d_l chi) / chi.
* kappa1..3 are real (ccz4_eqns.py:42-45 declares them unsigned int, which
  would truncate kappa1 = 0.02 to 0); vacuum: S_ij = S_i = tau = 0.

Outputs are named as the BSSN outputs plus Theta_rhs.  With ``tails`` the
advection of every field is left out, as for the BSSN kernels.
"""
from __future__ import annotations

import sympy as sp

from .bssn_eqs import E3, E33, FIELDS as BSSN_FIELDS, RHS_NAMES as BSSN_RHS_NAMES, SYM, Inputs, Stages, _sym3, \
    bssn_rhs_stages, sym_slot

FIELDS = BSSN_FIELDS + ["Theta"]
RHS_NAMES = BSSN_RHS_NAMES + ["Theta_rhs"]
TAIL_FIELDS = set(FIELDS)


def ccz4_rhs_stages(tails: bool = False) -> Stages:
    """CCZ4 right-hand side as staged assignments (see the module text).
    Geometry stages (inverse metric, Christoffels, Ricci pieces, D_i D_j
    alpha, raised A~) are the BSSN statement's; the Ricci tensor's product
    term uses Gammahat, and Z terms, Theta and the damping terms are added."""
    B = bssn_rhs_stages(tails=tails)
    keep = ["idet", "igt", "inv_chi", "C1", "C2", "igt_dchi", "C3", "CalGt", "C1u", "Rphi_tr", "Rphi", "DDa", "Amix",
            "AtU", "AA", "div_beta"]
    stages = dict(B.stages)
    S = Stages()
    I = Inputs()
    u = dict(I.u)
    u["Theta"] = sp.Symbol("Theta")
    P = dict(I.params)
    kappa1, kappa2, kappa3 = sp.symbols("kappa1 kappa2 kappa3")
    d1, d2 = I.d1, I.d2

    def adv(f):
        return sp.Integer(0) if tails and f in TAIL_FIELDS else I.adv(f)

    for name in keep[:9]:
        S.stage(name, stages[name])
    half = sp.Rational(1, 2)
    alpha, chi, K, Th = u["alpha"], u["chi"], u["K"], u["Theta"]
    bn = ["beta0", "beta1", "beta2"]
    Bv = [u["B0"], u["B1"], u["B2"]]
    Bn = ["B0", "B1", "B2"]
    Gh = [u["Gt0"], u["Gt1"], u["Gt2"]]   # Gamma-hat^i, evolved
    Gtn = ["Gt0", "Gt1", "Gt2"]
    gtn = [[f"gt{sym_slot(i, j)}" for j in E3] for i in E3]
    Atn = [[f"At{sym_slot(i, j)}" for j in E3] for i in E3]
    gt = [[u[gtn[i][j]] for j in E3] for i in E3]
    At = [[u[Atn[i][j]] for j in E3] for i in E3]
    igt = _sym3("igt")
    inv_chi = sp.Symbol("inv_chi")
    C1 = [[[sp.Symbol(f"C1_{k}_{min(i, j)}{max(i, j)}") for j in E3] for i in E3] for k in E3]
    C2 = [[[sp.Symbol(f"C2_{i}_{min(j, k)}{max(j, k)}") for k in E3] for j in E3] for i in E3]
    C1u = [[[sp.Symbol(f"C1u_{a}_{b}_{l}") for l in E3] for b in E3] for a in E3]
    CalGt = [sp.Symbol(f"CalGt{i}") for i in E3]
    dchi = [d1("chi", i) for i in E3]

    # Z_i = 1/2 gt_ij (Gammahat^j - CalGt^j); Zu^i = gt^ij Z_j = 1/2 (Gammahat^i - CalGt^i)
    Zu = [sp.Symbol(f"Zu{i}") for i in E3]
    S.stage("Zu", [(Zu[i], half * (Gh[i] - CalGt[i])) for i in E3])
    Zl = [sp.Symbol(f"Zl{i}") for i in E3]
    S.stage("Zl", [(Zl[i], sum(gt[i][j] * Zu[j] for j in E3)) for i in E3])

    # conformal Ricci with Gammahat in the connection terms: R~_ij + 2 D~_(i Z_j)
    Rh = [[None] * 3 for _ in E3]
    pairs = []
    for (i, j) in SYM:
        s = sp.Symbol(f"Rh{sym_slot(i, j)}")
        e = -half * I.lap(gtn[i][j])
        e += half * sum(gt[k][i] * d1(Gtn[k], j) + gt[k][j] * d1(Gtn[k], i) for k in E3)
        e += half * sum(Gh[k] * (C1[i][j][k] + C1[j][i][k]) for k in E3)
        e += sum(C2[k][l][i] * C1u[j][k][l] + C2[k][l][j] * C1u[i][k][l] for k in E3 for l in E3)
        e += sum(C2[k][i][m] * C1u[k][j][m] for k in E3 for m in E3)
        pairs.append((s, e))
        Rh[i][j] = Rh[j][i] = s
    S.stage("Rh", pairs)
    for name in ("Rphi_tr", "Rphi"):
        S.stage(name, stages[name])
    Rphi = _sym3("Rphi")
    # R_ij + 2 D_(i Z_j): the physical covariant derivative adds the chi terms
    igt_dchi = [sp.Symbol(f"igt_dchi{i}") for i in E3]
    zdchi = sp.Symbol("Z_dchi")  # gt^kl Z_k d_l chi = Zu^l d_l chi
    S.stage("Z_dchi", [(zdchi, sum(Zu[l] * dchi[l] for l in E3))])
    RZ = [[None] * 3 for _ in E3]
    pairs = []
    for (i, j) in SYM:
        s = sp.Symbol(f"RZ{sym_slot(i, j)}")
        pairs.append((s, Rh[i][j] + Rphi[i][j] + inv_chi * (Zl[i] * dchi[j] + Zl[j] * dchi[i] - gt[i][j] * zdchi)))
        RZ[i][j] = RZ[j][i] = s
    S.stage("RZ", pairs)
    trRZ = sp.Symbol("trRZ")  # R + 2 D_i Z^i, physical trace chi gt^ij (R_ij + 2 D_(i Z_j))
    S.stage("trRZ", [(trRZ, chi * sum(igt[i][j] * RZ[i][j] for (i, j) in E33))])
    for name in keep[11:]:
        S.stage(name, stages[name])
    DDa = _sym3("DDa")
    AtU = _sym3("AtU")
    AA = _sym3("AA")
    divb = sp.Symbol("div_beta")
    del igt_dchi

    out = {}
    # lapse (ccz4_eqns.py:388-389): lambda1 beta.d alpha - 2 alpha (K - 2 Theta - K0), K0 = 0
    out["a_rhs"] = P["lambda0"] * adv("alpha") - 2 * alpha * (K - 2 * Th)
    # shift and B (ccz4_eqns.py:395-421): as BSSN, f(alpha) = lambda_f0 + lambda_f1 alpha
    for i in E3:
        out[f"b_rhs{i}"] = sp.Rational(3, 4) * (P["lambda_f0"] + P["lambda_f1"] * alpha) * Bv[i] \
            + P["lambda1"] * adv(bn[i])
    w = -sp.Rational(2, 3)
    # conformal metric (ccz4_eqns.py:197-210) with the trace-free part of At
    trA = sp.Symbol("trA")
    S.stage("trA", [(trA, sum(igt[k][l] * At[k][l] for (k, l) in E33))])
    for (i, j) in SYM:
        lie = adv(gtn[i][j]) + sum(gt[i][k] * d1(bn[k], j) + gt[k][j] * d1(bn[k], i) for k in E3) \
            + w * gt[i][j] * divb
        out[f"gt_rhs{i}{j}"] = lie - 2 * alpha * (At[i][j] - gt[i][j] * trA * sp.Rational(1, 3))
    # chi (ccz4_eqns.py:277-283, in chi): beta.d chi + 2/3 chi (alpha K - d_k beta^k)
    out["chi_rhs"] = adv("chi") + w * chi * divb + sp.Rational(2, 3) * chi * alpha * K
    # At (ccz4_eqns.py:240-270): chi [-D_i D_j alpha + alpha (R_ij + 2 D_(i Z_j))]^TF
    #   + alpha At_ij (K - 2 Theta) - 2 alpha At_il At^l_j + Lie_beta At
    X = [[None] * 3 for _ in E3]
    for (i, j) in SYM:
        X[i][j] = X[j][i] = alpha * RZ[i][j] - DDa[i][j]
    trX = sp.Symbol("trX")
    S.stage("trX", [(trX, sum(igt[k][l] * X[k][l] for (k, l) in E33))])
    for (i, j) in SYM:
        lie = adv(Atn[i][j]) + sum(At[i][k] * d1(bn[k], j) + At[k][j] * d1(bn[k], i) for k in E3) \
            + w * At[i][j] * divb
        out[f"At_rhs{i}{j}"] = lie + chi * (X[i][j] - gt[i][j] * trX * sp.Rational(1, 3)) \
            + alpha * (At[i][j] * (K - 2 * Th) - 2 * AA[i][j])
    AtAt = sp.Symbol("AtAt")
    S.stage("AtAt", [(AtAt, sum(At[i][j] * AtU[i][j] for (i, j) in E33))])
    # K (ccz4_eqns.py:288-296): -D^i D_i alpha + alpha (R + 2 D_i Z^i + K^2 - 2 Theta K)
    #   + beta.d K - 3 alpha kappa1 (1 + kappa2) Theta
    out["K_rhs"] = adv("K") - chi * sum(igt[i][j] * DDa[i][j] for (i, j) in E33) \
        + alpha * (trRZ + K * K - 2 * Th * K) - 3 * alpha * kappa1 * (1 + kappa2) * Th
    # Theta (ccz4_eqns.py:301-310): alpha/2 (R + 2 D_i Z^i - At_ij At^ij + 2/3 K^2 - 2 Theta K)
    #   - Z^i d_i alpha + beta.d Theta - alpha kappa1 (2 + kappa2) Theta, Z^i = chi Zu^i
    out["Theta_rhs"] = half * alpha * (trRZ - AtAt + sp.Rational(2, 3) * K * K - 2 * Th * K) \
        - chi * sum(Zu[i] * d1("alpha", i) for i in E3) + adv("Theta") - alpha * kappa1 * (2 + kappa2) * Th
    # Gammahat (ccz4_eqns.py:318-378)
    Gh_rhs = []
    for i in E3:
        e = adv(Gtn[i])
        e += sum(2 * alpha * C2[i][j][k] * AtU[j][k] for (j, k) in E33)
        e -= sum(3 * alpha * inv_chi * AtU[i][j] * dchi[j] for j in E3)
        e -= sum(sp.Rational(4, 3) * alpha * igt[i][j] * d1("K", j) for j in E3)
        e += sum(2 * igt[k][i] * (alpha * d1("Theta", k) - Th * d1("alpha", k)) for k in E3)
        e -= sum(sp.Rational(4, 3) * alpha * K * igt[k][i] * Zl[k] for k in E3)
        e -= sum(2 * AtU[i][j] * d1("alpha", j) for j in E3)
        e += I.lap(bn[i]) + sum(igt[i][j] * I.gdivb(j) for j in E3) * sp.Rational(1, 3)
        e += sp.Rational(2, 3) * CalGt[i] * divb
        e -= sum(CalGt[j] * d1(bn[i], j) for j in E3)
        e += 2 * kappa3 * (sp.Rational(2, 3) * Zu[i] * divb - sum(Zu[j] * d1(bn[i], j) for j in E3))
        e -= 2 * alpha * kappa1 * Zu[i]
        Gh_rhs.append(e)
    Ghs = [sp.Symbol(f"Gt_rhs_s{i}") for i in E3]
    S.stage("Gt_rhs", [(Ghs[i], Gh_rhs[i]) for i in E3])
    for i in E3:
        out[f"Gt_rhs{i}"] = Ghs[i]
        out[f"B_rhs{i}"] = Ghs[i] - P["eta"] * Bv[i] + P["lambda2"] * adv(Bn[i]) - P["lambda3"] * adv(Gtn[i])
    for n in RHS_NAMES:
        S.outputs.append((n, out[n]))
    return S
