"""The FP32 path evaluates three groups of pressure coefficients through
identities that are exact in real arithmetic but do not cancel in FP32
(dycore.hip: phillips_prs, layer_coef). These tests check each identity in FP64
against the reference's own expressions (pressure_data.f90:76-88,
dyn_run.f90:393-410) on the 64-level hybrid coordinate over a range of surface
pressures, to the series' truncation bound, and that a float32 evaluation of
the identity form is far closer to the FP64 value than a float32 evaluation of
the reference form (the reason the forms exist).
"""
import numpy as np

import levels

KAPPA = 287.05 / 1004.6
AK, BK = levels.hybrid_levels(64)
PS = np.linspace(5.0e4, 1.06e5, 301)


def reference_forms(K, ps):
    """the reference's expressions, layer K (1-based, top -> bottom), FP64"""
    p1, p2 = AK[K - 1] + BK[K - 1] * ps, AK[K] + BK[K] * ps
    dp = p2 - p1
    dbk = BK[K] - BK[K - 1]
    rlnp = np.log(p2 / p1)
    alfa = 1 - (p1 / dp) * rlnp
    cofa = -(1 / dp) * (BK[K] * p1 / p2 - BK[K - 1] + rlnp * (BK[K - 1] - p1 * dbk / dp))
    dbp = BK[K] / p2 - BK[K - 1] / p1
    prs = ((p2 ** KAPPA * p2 - p1 ** KAPPA * p1) / ((KAPPA + 1) * dp)) ** (1 / KAPPA)
    return dict(rlnp=rlnp, alfa=alfa, cofa=cofa, dbp=dbp, prs=prs)


def identity_forms(K, ps, dt=np.float64):
    """the FP32 path's forms (same operation order as the kernel)"""
    c = lambda v: dt(v)
    ak1, bk1, ak2, bk2 = c(AK[K - 1]), c(BK[K - 1]), c(AK[K]), c(BK[K])
    ck = c(AK[K] * BK[K - 1] - AK[K - 1] * BK[K])
    dak, dbk = c(AK[K] - AK[K - 1]), c(BK[K] - BK[K - 1])
    ps = ps.astype(dt)
    p1, p2 = ak1 + bk1 * ps, ak2 + bk2 * ps
    dp = dak + dbk * ps
    x = dp / p1
    u = x / (c(2) + x)
    u2 = u * u
    y = c(2) * u * (c(1) + u2 * (c(1 / 3) + u2 * (c(.2) + u2 * (c(1 / 7) + u2 * (c(1 / 9) + u2 * (
        c(1 / 11) + u2 * (c(1 / 13) + u2 * c(1 / 15))))))))
    y = np.where(x <= 1, y, np.log1p(x))
    y2 = y * y
    alfa = y * c(.5) - y2 * (c(1 / 12) - y2 * (c(1 / 720) - y2 * (c(1 / 30240) - y2 * (
        c(1 / 1209600) - y2 * c(1 / 47900160)))))
    alfa = np.where(y <= 1, alfa, c(1) - (p1 / dp) * y)
    cofa = ck / dp * (alfa - dp / p2) / p1
    dbp = -ck / (p1 * p2)
    a = c(KAPPA) + c(1)
    a2 = a * a
    a4 = a2 * a2
    d = [(a2 - 1) / (c(24) * c(KAPPA)), -(a4 - 1) / (c(2880) * c(KAPPA)), (a4 * a2 - 1) / (c(181440) * c(KAPPA)),
         -(a4 * a4 - 1) / (c(9676800) * c(KAPPA)), (a4 * a4 * a2 - 1) / (c(479001600) * c(KAPPA))]
    S = y2 * (d[0] + y2 * (d[1] + y2 * (d[2] + y2 * (d[3] + y2 * d[4]))))
    prs = np.sqrt(p1 * p2) * np.exp(S)
    return dict(rlnp=y, alfa=alfa, cofa=cofa, dbp=dbp, prs=prs)


def rel(a, b):
    scale = np.abs(b).max()
    return np.abs(a - b).max() / scale if scale > 0 else np.abs(a).max()


def test_identities_equal_reference_expressions_in_fp64():
    for K in range(2, 65):
        ref, idf = reference_forms(K, PS), identity_forms(K, PS)
        for k in ref:
            # the series truncations are <= 5e-9 relative; FP64 rounding of
            # the reference forms (cancellation) is ~1e-13
            assert rel(idf[k], ref[k]) < 1e-8, (K, k, rel(idf[k], ref[k]))


def test_identities_in_fp32_beat_the_reference_expressions_in_fp32():
    worst_id, worst_ref = {}, {}
    f = np.float32
    for K in range(2, 65):
        ref = reference_forms(K, PS)
        idf = identity_forms(K, PS, f)
        # the reference expressions evaluated in float32 from float32 inputs
        p1 = f(AK[K - 1]) + f(BK[K - 1]) * PS.astype(f)
        p2 = f(AK[K]) + f(BK[K]) * PS.astype(f)
        dp = p2 - p1
        dbk = f(BK[K] - BK[K - 1])
        rl = np.log(p2 / p1)
        al = f(1) - (p1 / dp) * rl
        naive = dict(rlnp=rl, alfa=al,
                     cofa=-(f(1) / dp) * (f(BK[K]) * p1 / p2 - f(BK[K - 1]) + rl * (f(BK[K - 1]) - p1 * dbk / dp)),
                     dbp=f(BK[K]) / p2 - f(BK[K - 1]) / p1,
                     prs=((p2 ** f(KAPPA) * p2 - p1 ** f(KAPPA) * p1) / ((f(KAPPA) + 1) * dp)) ** (f(1) / f(KAPPA)))
        for k in ref:
            worst_id[k] = max(worst_id.get(k, 0), rel(idf[k].astype(float), ref[k]))
            worst_ref[k] = max(worst_ref.get(k, 0), rel(naive[k].astype(float), ref[k]))
    print("float32 identity forms:", worst_id)
    print("float32 reference forms:", worst_ref)
    for k in worst_id:
        assert worst_id[k] < 3e-6, k
        assert worst_id[k] < 0.5 * worst_ref[k] or worst_ref[k] < 3e-6, k
    # the two that motivated the change
    assert worst_id["alfa"] < 1e-2 * worst_ref["alfa"]
    assert worst_id["prs"] < 1e-1 * worst_ref["prs"]
