"""GPU Fisher matrices (FP64 stencil -> matrix-core Gram -> FP64 inversion) against the CPU oracle.

Bar: every element within 1e-6 of the oracle, measured relative to sqrt(F_ii F_jj) (the element's
natural scale: off-diagonal elements pass through zero as correlations change sign), and the
diagonal within 1e-6 relative.  The FP32 Gram (GWH_MIXED) is compared with the FP64 one at 1e-4.

Conditioning.  With the reference's absolute step 1e-8 the central-difference numerators keep only
~5-8 significant digits (the luminosity distance, ~400 Mpc, moves by 2.5e-11 relative), so the
reference algorithm itself is reproducible only to 3e-6..4e-5: changing the step by one part in 1e7
moves the oracle's own result that much (its rounding jitter).  Two correct FP64 implementations
with different libm rounding therefore differ by that jitter.  The 1e-6 parity bar is checked at an
absolute step of 1e-5 (same stencil, same formula; the jitter there is below 1e-7); at the
reference's 1e-8 step the engine must agree with the oracle within 10x the oracle's own jitter.
"""
import numpy as np
import pytest

import gw_analysis_tools_amd as G
import oracle as O
from gw_analysis_tools_amd import workloads as W

pytestmark = pytest.mark.gpu

FISHER_TOL = 1e-6
MIXED_FISHER_TOL = 1e-4
EPS_PARITY = 1e-5   # well-conditioned step for the 1e-6 bar
EPS_REF = 1e-8      # the reference's step (fisher.cpp:158)

INJ = dict(mass1=36., mass2=29., spin1=[0, 0, 0.3], spin2=[0, 0, -0.1], DL=410., tc=0., phiRef=0., f_ref=20.,
           RA=1.37, DEC=-1.26, psi=2.66, incl=2.9, gmst=W.GMST, shift_time=1, shift_phase=1)
PREC = dict(INJ, spin1=[0.3, 0.1, 0.3], spin2=[-0.1, 0.2, -0.1], chip=-1.)
# the 13-parameter basis differentiates (chip, phip), so its points carry them (reduced parametrisation)
REDUCED = dict(chip=0.35, phip=0.8)

# (engine basis, oracle method string, dimension, sky_average, model)
BASES = [
    (G.BASIS_PHENOMD_SKYAVG, "IMRPhenomD", 7, 1, G.IMRPHENOMD),
    (G.BASIS_PHENOMD_FULL, "IMRPhenomD", 11, 0, G.IMRPHENOMD),
    (G.BASIS_PHENOMD_MCMC, "MCMC_IMRPhenomD", 11, 0, G.IMRPHENOMD),
    (G.BASIS_PV2_FULL, "IMRPhenomPv2", 13, 0, G.IMRPHENOMPV2),
    (G.BASIS_PV2_MCMC, "MCMC_IMRPhenomPv2", 15, 0, G.IMRPHENOMPV2),
]


def norm_err(a, b):
    s = np.sqrt(np.abs(np.diag(b)))
    return np.max(np.abs(a - b) / np.outer(s, s))


def diag_err(a, b):
    return np.max(np.abs(np.diag(a) - np.diag(b)) / np.abs(np.diag(b)))


def _grid():
    f = W.uniform_grid(20., 1024., 0.25)
    return f, W.analytic_psd(f)


def _kw(model, basis=None, **over):
    kw = dict(PREC if model == G.IMRPHENOMPV2 else INJ)
    if basis == G.BASIS_PV2_FULL:
        kw.update(REDUCED)
    kw.update(over)
    return kw


@pytest.mark.parametrize("basis,method,dim,sky,model", BASES)
@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("dets", [(0, 0), (1, 0)])
def test_fisher_numerical_vs_oracle(basis, method, dim, sky, model, order, dets):
    if sky and dets != (0, 0):
        pytest.skip("the sky-averaged basis has no detector")
    f, psd = _grid()
    kw = _kw(model, basis, sky_average=sky)
    det, ref = dets
    gp, op = G.make_params(**kw), O.make_params(**kw)
    a = G.fisher_numerical(f, psd, basis, det, ref, dim, gp, order=order, epsilon=EPS_PARITY)
    b = O.fisher_numerical(f, psd, method, det, ref, dim, op, order=order, epsilon=EPS_PARITY)
    e, ed = norm_err(a, b), diag_err(a, b)
    assert np.allclose(a, a.T, rtol=0, atol=0)
    # the reference's own step: within 10x the oracle's rounding jitter there
    a8 = G.fisher_numerical(f, psd, basis, det, ref, dim, gp, order=order, epsilon=EPS_REF)
    b8 = O.fisher_numerical(f, psd, method, det, ref, dim, op, order=order, epsilon=EPS_REF)
    jitter = max(norm_err(O.fisher_numerical(f, psd, method, det, ref, dim, op, order=order,
                                             epsilon=EPS_REF * (1 + j * 1e-7)), b8) for j in (1, 2, 3, 4, 5))
    e8 = norm_err(a8, b8)
    print("fisher %s sky=%d order=%d det=%s: eps 1e-5 max normalised %.3e diag rel %.3e | eps 1e-8 %.3e "
          "(oracle jitter %.3e)" % (method, sky, order, dets, e, ed, e8, jitter))
    assert e <= FISHER_TOL and ed <= FISHER_TOL, (e, ed)
    assert e8 <= 10 * jitter, (e8, jitter)


@pytest.mark.parametrize("basis,method,dim,sky,model", BASES)
@pytest.mark.parametrize("order", [2, 4])
def test_fisher_derivatives_vs_oracle(basis, method, dim, sky, model, order):
    """bare stencil derivatives (calculate_derivatives) per parameter, relative to the largest |dh/dtheta_i|"""
    f, psd = _grid()
    kw = _kw(model, basis, sky_average=sky)
    det, ref = (0, 0) if sky else (1, 0)
    a = G.fisher_derivatives(f, basis, det, ref, dim, G.make_params(**kw), order=order, epsilon=EPS_PARITY)
    b = O.calculate_derivatives(f, method, det, ref, dim, O.make_params(**kw), order=order, epsilon=EPS_PARITY)
    err = [np.max(np.abs(a[i] - b[i])) / max(np.max(np.abs(b[i])), 1e-300) for i in range(dim)]
    print("derivatives %s order %d: worst parameter %d rel %.3e" % (method, order, int(np.argmax(err)), max(err)))
    assert max(err) <= 1e-6, err


def test_fisher_gram_matches_weighted_sum():
    """the matrix-core Gram equals 4 Simpson(Re(d_i d_j*)/S) of the engine's own derivatives"""
    f, psd = _grid()
    kw = _kw(G.IMRPHENOMD, G.BASIS_PHENOMD_MCMC)
    p = G.make_params(**kw)
    d = G.fisher_derivatives(f, G.BASIS_PHENOMD_MCMC, 1, 0, 11, p)
    F = G.fisher_numerical(f, psd, G.BASIS_PHENOMD_MCMC, 1, 0, 11, p)
    n = len(f)
    w = np.ones(n)
    w[1:n - 1:2] = 4
    w[2:n - 1:2] = 2  # endpoints 1 whatever the parity of n (util.h:841-856)
    w = w * (f[1] - f[0]) / 3 * 4 / psd
    ref = np.real((d * w) @ np.conj(d).T)
    print("Gram vs numpy: max normalised %.3e" % norm_err(F, ref))
    assert norm_err(F, ref) <= 1e-12


def test_fisher_eta_edge():
    """eta within epsilon of 1/4: the eta row uses the one-sided difference of fisher.cpp:417-425."""
    f, psd = _grid()
    kw = _kw(G.IMRPHENOMD, mass1=30., mass2=30.)
    a = G.fisher_numerical(f, psd, G.BASIS_PHENOMD_MCMC, 0, 0, 11, G.make_params(**kw))
    b = O.fisher_numerical(f, psd, "MCMC_IMRPhenomD", 0, 0, 11, O.make_params(**kw))
    assert norm_err(a, b) <= FISHER_TOL, norm_err(a, b)


def _network():
    f = W.uniform_grid(20., 2048., 0.125)
    psd = W.analytic_psd(f)
    dets = [0, 1, 2]
    zeros = [np.zeros(len(f), complex)] * 3
    return f, psd, dets, G.Event(f, zeros, [psd] * 3, dets)


def _points(model, n):
    if model == G.IMRPHENOMPV2:
        P = W.precessing_params(n, seed_stream=7)
        g = W.rng(3, 0, stream=8)
        P["chip"] = g.uniform(0.05, 0.9, n)
        P["phip"] = g.uniform(0, 2 * np.pi, n)
        return P
    return W.perturbed_params(W.injection_params(1)[0], n, step=11, config=3)


def _oracle_params(rec):
    return O.make_params(**{k: (list(rec[k]) if rec[k].shape else rec[k].item()) for k in rec.dtype.names})


@pytest.mark.parametrize("basis,method,dim,model", [(G.BASIS_PHENOMD_MCMC, "MCMC_IMRPhenomD", 11, G.IMRPHENOMD),
                                                    (G.BASIS_PV2_FULL, "IMRPhenomPv2", 13, G.IMRPHENOMPV2)])
def test_fisher_batch_network_vs_oracle(basis, method, dim, model):
    """network Fisher = sum over detectors of the single-detector Fisher with the first detector as the
    arrival-time reference (fisher.cpp:81-145 called per detector, as the MCMC Fisher wrapper does)."""
    f, psd, dets, ev = _network()
    P = _points(model, 3)
    F, sig = G.fisher_batch(ev, P, model, basis, dim, precision=G.FP64, with_sigma=True, epsilon=EPS_PARITY)
    worst = worst_sigma = 0.0
    for i in range(len(P)):
        op = _oracle_params(P[i])
        ref = sum(O.fisher_numerical(f, psd, method, d, dets[0], dim, op, epsilon=EPS_PARITY) for d in dets)
        worst = max(worst, norm_err(F[i], ref), diag_err(F[i], ref))
        # the inversion kernel against the oracle's LU inverse of the same matrix
        inv = O.lu_invert(F[i])
        s_ref = np.sqrt(np.diag(inv))
        # (normalised condition numbers reach 1e10, so the two LU orders agree to ~cond * 1e-16)
        ok = np.isfinite(s_ref)
        assert np.array_equal(np.isfinite(sig[i]), ok)
        assert np.allclose(sig[i][ok], s_ref[ok], rtol=1e-5, atol=0), np.max(np.abs(sig[i][ok] / s_ref[ok] - 1))
        s_orc = np.sqrt(np.diag(O.lu_invert(ref)))
        both = ok & np.isfinite(s_orc)
        if both.any():
            worst_sigma = max(worst_sigma, np.max(np.abs(sig[i][both] / s_orc[both] - 1)))
    print("network fisher %s: max normalised / diag rel error %.3e, sigma rel error vs oracle %.3e"
          % (method, worst, worst_sigma))
    assert worst <= FISHER_TOL, worst


@pytest.mark.parametrize("basis,dim,model", [(G.BASIS_PHENOMD_MCMC, 11, G.IMRPHENOMD),
                                             (G.BASIS_PV2_FULL, 13, G.IMRPHENOMPV2)])
def test_fisher_batch_mixed_vs_fp64(basis, dim, model):
    """GWH_MIXED contracts FP64 numerators on the FP64 matrix cores (sigmas need FP64 elements): identical to FP64;
    the FP32 Gram (GWH_FISHER_GRAM_FP32) agrees on the elements to MIXED_FISHER_TOL of sqrt(F_ii F_jj)"""
    f, psd, dets, ev = _network()
    P = _points(model, 16)
    a, sa = G.fisher_batch(ev, P, model, basis, dim, precision=G.FP64, with_sigma=True)
    m, sm = G.fisher_batch(ev, P, model, basis, dim, precision=G.MIXED, with_sigma=True)
    assert np.array_equal(m, a) and np.array_equal(np.isfinite(sm), np.isfinite(sa))
    assert np.array_equal(sm[np.isfinite(sm)], sa[np.isfinite(sa)])
    b, sb = G.fisher_batch(ev, P, model, basis, dim, precision=G.FISHER_GRAM_FP32, with_sigma=True)
    e = max(norm_err(b[i], a[i]) for i in range(len(P)))
    rel = np.abs(b - a) / np.maximum(np.abs(a), 1e-300)
    fin = np.isfinite(sa) & np.isfinite(sb)
    srel = np.abs(sb[fin] / sa[fin] - 1)
    # marginal sigmas of these ill-conditioned matrices (normalised condition numbers 1e8..1e10) are
    # reported, not bounded: an FP32 Gram cannot pin them
    print("mixed Gram vs FP64: max normalised %.3e, max element rel %.3e (median %.3e); sigma rel max %.3e "
          "median %.3e over %d finite of %d" % (e, rel.max(), np.median(rel), srel.max(), np.median(srel),
                                                 fin.sum(), fin.size))
    assert e <= MIXED_FISHER_TOL, e


@pytest.mark.parametrize("model,basis,dim", [(G.IMRPHENOMD, G.BASIS_PHENOMD_MCMC, 11),
                                             (G.IMRPHENOMPV2, G.BASIS_PV2_MCMC, 15)])
def test_fisher_batch_composition_independent(model, basis, dim):
    """a point's Fisher matrix is bitwise the same alone and inside a larger batch, also in the later passes of a
    batch (300 points run as passes of 128, 128 and 44), for the network Fisher of both models"""
    f, psd, dets, ev = _network()
    P = _points(model, 300)
    full, sig = G.fisher_batch(ev, P, model, basis, dim, with_sigma=True)
    for i in (0, 5, 130, 255, 256, 299):
        one, s1 = G.fisher_batch(ev, P[i:i + 1], model, basis, dim, with_sigma=True)
        assert np.array_equal(one[0], full[i]), i
        assert np.array_equal(s1[0], sig[i], equal_nan=True), i
    assert np.all(np.isfinite(full))


@pytest.mark.parametrize("basis,method,dim,sky,model", [BASES[0], BASES[2], BASES[3]])
def test_fisher_relative_step_vs_oracle(basis, method, dim, sky, model):
    """The benchmark configuration's "1e-7 relative" step (epsilon < 0: |epsilon| * |theta_i| per basis
    coordinate, |epsilon| where theta_i is 0) on the 20 + 0.25 k Hz, k < 4017 grid; the bar is the
    parity bar, or 10x the oracle's own rounding jitter at that step where that is larger."""
    f = 20. + 0.25 * np.arange(4017)
    psd = W.analytic_psd(f)
    kw = _kw(model, basis, sky_average=sky)
    gp, op = G.make_params(**kw), O.make_params(**kw)
    a = G.fisher_numerical(f, psd, basis, 1 - sky, 0, dim, gp, epsilon=-1e-7)
    b = O.fisher_numerical(f, psd, method, 1 - sky, 0, dim, op, epsilon=-1e-7)
    jitter = max(norm_err(O.fisher_numerical(f, psd, method, 1 - sky, 0, dim, op, epsilon=-1e-7 * (1 + j * 1e-7)),
                          b) for j in (1, 2, 3))
    e = norm_err(a, b)
    print("relative step %s sky=%d: max normalised %.3e (oracle jitter %.3e)" % (method, sky, e, jitter))
    assert e <= max(FISHER_TOL, 10 * jitter)
    # and it is not the absolute step of the same size
    c = G.fisher_numerical(f, psd, basis, 1 - sky, 0, dim, gp, epsilon=1e-7)
    assert not np.array_equal(a, c)


def test_fisher_errors():
    f, psd = _grid()
    p = G.make_params(**INJ)
    for bad in (0., float("nan"), float("inf")):
        with pytest.raises(G.GWHError):
            G.fisher_numerical(f, psd, G.BASIS_PHENOMD_MCMC, 0, 0, 11, p, epsilon=bad)
    with pytest.raises(G.GWHError):
        G.fisher_numerical(f, psd, G.BASIS_PHENOMD_MCMC, 0, 0, 7, p)  # wrong dimension
    with pytest.raises(G.GWHError):
        G.fisher_numerical(f, psd, G.BASIS_PHENOMD_SKYAVG, 0, 0, 7, p)  # sky basis without sky_average
