"""CPU tests that pin the oracle: known-answer vectors, closed forms,
limits, independent numerical integration and finite differences.  A
plausible restatement error (a dropped term, a wrong sign, a transposed
operand) fails one of these."""
import numpy as np
import pytest
from scipy import integrate, linalg, special, stats

from gacelle_amd import _abi, models
from gacelle_amd.askadam import askadam
from oracle import oracle as orc

from .helpers import FA_JOINT, TE12, TE_JOINT, TR_JOINT, mcmc_cfg, r2s_setup


# ---------------------------------------------------------------------------
# Philox4x32-10 known-answer vectors (Random123 distribution, kat_vectors)
@pytest.mark.parametrize("ctr,key,out", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answers(ctr, key, out):
    assert orc.philox(ctr, key) == out


def test_native_normals_are_standard_normal():
    z = np.array([orc.native_normal(7, it, c, 0, q) for it in range(1, 60) for c in range(40) for q in range(3)])
    assert abs(z.mean()) < 0.05 and abs(z.std() - 1) < 0.05
    from scipy import stats
    assert stats.kstest(z, "norm").pvalue > 1e-3


def test_norm_inv_matches_the_normal_quantile():
    """Giles' single-precision erfinv behind the odd normal of the MH stream:
    sqrt(2) erfinv(2u - 1) against scipy's ndtri, body and tails."""
    from scipy import special

    u = np.concatenate([np.linspace(1e-7, 1 - 1e-7, 20001), np.geomspace(3e-8, 1e-2, 400)])
    u = np.concatenate([u, 1 - u[-400:]])
    got = np.array([orc.norm_inv(x) for x in u])
    ref = special.ndtri(u)
    assert np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref))) < 2e-6


@pytest.mark.parametrize("Q", [3, 4, 5])
def test_native_mh_draws_are_independent_normals_and_uniform(Q):
    from scipy import stats

    n, u = zip(*[orc.native_mh_draws(5, it, c, Q) for it in range(1, 400) for c in range(30)])
    n = np.array(n)
    u = np.array(u)
    for q in range(Q):
        assert stats.kstest(n[:, q], "norm").pvalue > 1e-3, q
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    c = np.corrcoef(np.column_stack([n, u]).T)
    assert np.max(np.abs(c - np.eye(Q + 1))) < 0.05


def test_native_partners_are_derangements():
    for it in range(1, 50):
        p = orc.native_partners(11, 0, it, 16, _abi.GC_ENSEMBLE_SIMULTANEOUS)
        assert sorted(p) == list(range(16)) and np.all(p != np.arange(16))
        rb = orc.native_partners(11, 0, it, 16, _abi.GC_ENSEMBLE_REDBLACK)
        assert np.all(rb[:8] >= 8) and np.all(rb[8:] < 8)


# ---------------------------------------------------------------------------
def test_logp_gaussian_closed_form():
    for sos, sig, m in [(0.3, 0.02, 12), (1e-4, 0.1, 48), (2.0, 0.05, 1)]:
        ref = -sos / (2 * sig ** 2) - m / 2 * np.log(2 * np.pi * sig ** 2)
        assert orc.logp_gaussian(sos, sig, m) == pytest.approx(ref, rel=1e-14)
        # equals the log of the product of m normal densities for i.i.d. residuals with sum sq = sos
        r = np.sqrt(sos / m) * np.ones(m)
        assert orc.logp_gaussian(sos, sig, m) == pytest.approx(np.sum(np.log(
            np.exp(-r ** 2 / (2 * sig ** 2)) / np.sqrt(2 * np.pi * sig ** 2))), rel=1e-10)


# ---------------------------------------------------------------------------
def _r2s_md():
    return r2s_setup({})[2]


def test_r2star_closed_form():
    md = _r2s_md()
    th = np.array([[1.3, 0.7], [25.0, 4.0]])
    s = orc.fwd_value(md, th)
    te = np.array([md.table[t] for t in range(12)], dtype=np.float64)
    np.testing.assert_allclose(s, th[0] * np.exp(-np.outer(te, th[1])), rtol=1e-14)


def _joint(tr=TR_JOINT, fa=FA_JOINT):
    m = models.gpuJointR1R2starMapping(TE_JOINT, tr, fa)
    f = askadam.check_set_default_basic(dict(modelParams=["M0", "R1", "R2star"], lb=[0, 0.1, 0.1], ub=[2, 10, 200]))
    return m.gc_model(f, "askadam", None)


def test_joint_closed_form_and_limits():
    md = _joint()
    th = np.array([[1.5], [0.8], [30.0]])
    b1 = np.array([[0.9]])
    s = orc.fwd_value(md, th, b1)
    te = np.array([md.table[t] for t in range(12)], np.float64)
    a = np.array([md.table[12 + f] for f in range(4)], np.float64) * 0.9
    e1 = np.exp(-float(np.float32(TR_JOINT)) * 0.8)
    ref = 1.5 * np.sin(a) * (1 - e1) / (1 - np.cos(a) * e1)
    ref = (ref[None, :] * np.exp(-30.0 * te)[:, None]).reshape(-1, order="F")
    np.testing.assert_allclose(s[:, 0], ref, rtol=1e-13)
    # FA = 90 deg and TR -> infinity reduce to the mono-exponential decay
    md2 = _joint(tr=1e6, fa=np.array([90.0]))
    s2 = orc.fwd_value(md2, th, np.ones((1, 1)))
    np.testing.assert_allclose(s2[:, 0], 1.5 * np.exp(-30.0 * te), rtol=1e-6)


# ---------------------------------------------------------------------------
def _nexi(quad="gl64", lmax=0, rician=False):
    b = np.repeat(np.arange(1, 7, dtype=float), 2)
    D = np.tile([20.0, 30.0], 6)
    m = models.gpuNEXI(b, D, quad=quad, rician=rician)
    names = ["fa", "Da", "De", "ra"] + (["p2"] if lmax == 2 else [])
    f = askadam.check_set_default_basic(dict(modelParams=names, lb=m.lb[: len(names)], ub=m.ub[: len(names)]))
    return m, m.gc_model(f, "askadam", None)


def _karger_M(x, fa, Da, De, ra, re):
    """Karger two-pool signal at one orientation from the matrix exponential:
    d/dt [Sa, Se] = -[[Da x^2 + ra, -re], [-ra, De + re]] [Sa, Se], with
    initial fractions [fa, 1-fa] and all rates already scaled by time."""
    K = np.array([[Da * x * x + ra, -re], [-ra, De + re]])
    return np.array([1.0, 1.0]) @ linalg.expm(-K) @ np.array([fa, 1 - fa])


def test_nexi_integrand_matches_karger_matrix_exponential():
    """The eigenvalue closed form M(x) (NEXI/gpuNEXI.m:1090-1097) equals
    the exponential of the two-pool Karger exchange matrix."""
    m, md = _nexi("gl64")
    rng = np.random.default_rng(0)
    for _ in range(5):
        fa, Da, De = rng.uniform(0.2, 0.8), rng.uniform(1.5, 3), rng.uniform(0.5, 1.5)
        ra = rng.uniform(0.005, 0.3)
        b, Delta = 3.0, 20.0
        dab, deb, rab = b * Da, b * De, Delta * ra
        reb = rab * fa / (1 - fa)
        f = lambda x: _karger_M(x, fa, dab, deb, rab, reb)
        exact = integrate.quad(f, 0, 1, epsabs=1e-13, epsrel=1e-12)[0]
        k = [j for j in range(12) if md.table[j] == 3.0 and md.table[12 + j] == 20.0][0]
        s = orc.fwd_value(md, np.array([[fa], [Da], [De], [ra]]))
        # the Gauss-Legendre nodes and weights travel as float32 in the model table
        assert s[k, 0] == pytest.approx(exact, rel=2e-8)


def test_nexi_no_exchange_limit_is_standard_model():
    """ra -> 0: S = fa*sqrt(pi)*erf(sqrt(bDa))/(2 sqrt(bDa)) + (1-fa) exp(-b De)."""
    m, md = _nexi("gl64")
    fa, Da, De = 0.55, 2.2, 0.9
    s = orc.fwd_value(md, np.array([[fa], [Da], [De], [1e-9]]))
    for k in range(12):
        b = md.table[k]
        ref = fa * np.sqrt(np.pi) * special.erf(np.sqrt(b * Da)) / (2 * np.sqrt(b * Da)) + (1 - fa) * np.exp(-b * De)
        assert s[k, 0] == pytest.approx(ref, rel=2e-7)


def test_nexi_trapezoid_is_the_reference_rule():
    """trapz14 = composite trapezoid on linspace(0,1,14) of the same
    integrand (gacelle_trapz.m), within the reference's accuracy claim."""
    m14, md14 = _nexi("trapz14")
    m64, md64 = _nexi("gl64")
    th = np.array([[0.5], [2.0], [1.0], [0.05]])
    s14 = orc.fwd_value(md14, th)[:, 0]
    s64 = orc.fwd_value(md64, th)[:, 0]
    x = np.linspace(0, 1, 14)
    for k in range(12):
        b, D = md14.table[k], md14.table[12 + k]
        dab, deb, rab = b * 2.0, b * 1.0, D * 0.05
        reb = rab * 0.5 / 0.5
        y = np.array([_karger_M(xx, 0.5, dab, deb, rab, reb) for xx in x])
        tr = np.sum((y[:-1] + y[1:]) * np.diff(x) / 2)
        assert s14[k] == pytest.approx(tr, rel=1e-6)
    nrmse = np.sqrt(np.mean((s14 - s64) ** 2)) / np.sqrt(np.mean(s64 ** 2))
    assert nrmse < 5e-3


def test_nexi_lmax2_second_moment():
    m, md = _nexi("gl64", lmax=2)
    th = np.array([[0.5], [2.0], [1.0], [0.05], [0.4]])
    s = orc.fwd_value(md, th)[:, 0]
    for k in range(12):
        b, D = md.table[k], md.table[12 + k]
        dab, deb, rab = b * 2.0, b * 1.0, D * 0.05
        f = lambda x: _karger_M(x, 0.5, dab, deb, rab, rab) * (3 * x * x - 1) / 2
        ref = 0.4 * abs(integrate.quad(f, 0, 1, epsabs=1e-13)[0])
        assert s[12 + k] == pytest.approx(min(ref, 1.0), rel=2e-7)


def test_rician_mean_against_bessel_and_limits():
    """E|S + n| = sigma sqrt(pi/2) L_1/2(-S^2/2sigma^2) with the exact
    Laguerre function from scaled Bessel functions; the 34-node trapezoid
    of the reference is accurate to its stated NRMSE < 0.05%."""
    m, md = _nexi("gl64", rician=True)
    th = np.array([[0.5], [2.0], [1.0], [0.05]])
    sig = 0.05
    s_r = orc.fwd_value(md, th, np.full((1, 1), sig))[:, 0]
    s_0 = orc.fwd_value(_nexi("gl64")[1], th)[:, 0]
    x = -s_0 ** 2 / (2 * sig ** 2)
    L = np.exp(x / 2) * ((1 - x) * special.i0(-x / 2) - x * special.i1(-x / 2))
    exact = np.maximum(sig * np.sqrt(np.pi / 2) * L, s_0)
    np.testing.assert_allclose(s_r, np.minimum(exact, 1.0), rtol=5e-4)
    # high SNR: E = S + sigma^2 / (2 S) + O(sigma^4 / S^3)
    assert s_r[0] == pytest.approx(s_0[0] + sig ** 2 / (2 * s_0[0]), rel=1e-3)
    # pure noise: the Rayleigh mean sigma sqrt(pi/2)
    th0 = np.array([[0.5], [2.0], [1.0], [0.05]])
    s_n = orc.fwd_value(md, th0, np.full((1, 1), 50.0))[:, 0]
    np.testing.assert_allclose(s_n, 1.0)  # clamped by min(s, 1) since 50 sqrt(pi/2) > 1
    s_n = orc.fwd_value(md, th0, np.full((1, 1), 0.5))[:, 0]
    x = -s_0 ** 2 / (2 * 0.25)
    L = np.exp(x / 2) * ((1 - x) * special.i0(-x / 2) - x * special.i1(-x / 2))
    np.testing.assert_allclose(s_n, np.minimum(0.5 * np.sqrt(np.pi / 2) * L, 1.0), rtol=5e-4)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kind", ["r2s", "joint", "nexi", "nexi_l2", "nexi_rice", "mcrmwi", "mcrmwi_bloch_mag"])
def test_dual_gradient_matches_central_differences(kind):
    rng = np.random.default_rng(1)
    extra = None
    if kind.startswith("mcrmwi"):
        exch = kind == "mcrmwi"
        _, _, md = _mcrmwi_md(exch, cplx=exch, te=np.arange(2, 25, 2) * 1e-3)
        amp = [1.3, 0.15, 0.6, 1.1] + ([2.5] if exch else [])
        th = np.array(amp + [80.0, 20.0, 25.0, 0.08, -0.02] + ([0.01, -0.02, 0.03, 0.0, 0.1] if exch else []))
        th = th[:, None]
        extra = np.array([[0.95], [0.3], [0.9], [0.05], [0.0], [-0.05], [0.02]])
    elif kind == "r2s":
        md = _r2s_md()
        th = np.array([[1.1], [33.0]])
    elif kind == "joint":
        md = _joint()
        th = np.array([[1.2], [0.9], [25.0]])
        extra = np.array([[1.05]])
    else:
        lmax = 2 if kind == "nexi_l2" else 0
        _, md = _nexi("trapz14", lmax=lmax, rician=kind == "nexi_rice")
        th = np.array([[0.45], [2.1], [0.8], [0.04]] + ([[0.3]] if lmax == 2 else []))
        if kind == "nexi_rice":
            extra = np.array([[0.03]])
    _, jac = orc.fwd(md, th, extra)
    for p in range(th.shape[0]):
        h = 1e-6 * max(abs(th[p, 0]), 1e-3)
        tp, tm = th.copy(), th.copy()
        tp[p] += h
        tm[p] -= h
        fd = (orc.fwd_value(md, tp, extra) - orc.fwd_value(md, tm, extra)) / (2 * h)
        np.testing.assert_allclose(jac[:, p, 0], fd[:, 0], rtol=1e-5, atol=1e-9)


# ---------------------------------------------------------------------------
def test_tv_constant_impulse_and_wrap():
    dims = (5, 4, 3)
    mask = np.ones(dims, np.uint8)
    n = int(mask.sum())
    lam, vs = 0.7, (2.0, 1.0, 0.5)
    loss, g = orc.tv(np.full(dims, 3.0), mask, _abi.GC_TV_3D, vs, lam, n)
    assert loss == 0 and np.all(g == 0)
    img = np.zeros(dims)
    img[2, 1, 1] = 1.0
    loss, _ = orc.tv(img, mask, _abi.GC_TV_3D, vs, lam, n)
    # an impulse has 2 nonzero forward differences per axis: 2/vx + 2/vy + 2/vz
    assert loss == pytest.approx(lam / n * (2 / 2.0 + 2 / 1.0 + 2 / 0.5))
    loss2, _ = orc.tv(img, mask, _abi.GC_TV_2D, vs, lam, n)
    assert loss2 == pytest.approx(lam / n * (2 / 2.0 + 2 / 1.0))
    # periodic wrap (circshift): a ramp along x has one jump back at the edge
    ramp = np.tile(np.arange(5.0).reshape(5, 1, 1), (1, 4, 3))
    loss3, _ = orc.tv(ramp, mask, _abi.GC_TV_2D, (1, 1, 1), 1.0, n)
    assert loss3 == pytest.approx((4 * 1 + 4) * 4 * 3 / n)
    # only the centre voxel's mask gates a term
    m2 = mask.copy()
    m2[2, 1, 1] = 0
    loss4, _ = orc.tv(img, m2, _abi.GC_TV_3D, vs, lam, n)
    assert loss4 == pytest.approx(lam / n * (1 / 2.0 + 1 / 1.0 + 1 / 0.5))


def test_tv_adjoint_matches_finite_differences():
    rng = np.random.default_rng(2)
    dims = (6, 5, 4)
    mask = (rng.random(dims) < 0.7).astype(np.uint8)
    img = rng.standard_normal(dims) * mask
    n = int(mask.sum())
    loss, g = orc.tv(img, mask, _abi.GC_TV_3D, (1.5, 1.0, 2.0), 0.3, n)
    h = 1e-7
    for idx in [(0, 0, 0), (2, 3, 1), (5, 4, 3), (3, 2, 2)]:
        ip, im = img.copy(), img.copy()
        ip[idx] += h
        im[idx] -= h
        fd = (orc.tv(ip, mask, 1, (1.5, 1.0, 2.0), 0.3, n, grad=False) -
              orc.tv(im, mask, 1, (1.5, 1.0, 2.0), 0.3, n, grad=False)) / (2 * h)
        assert g[idx] == pytest.approx(fd, rel=1e-5, abs=1e-9)


# ---------------------------------------------------------------------------
def _adam_two_iterations(form):
    nv = 5
    rng = np.random.default_rng(4)
    md = _r2s_md()
    f = askadam.check_set_default_basic(dict(modelParams=["M0", "R2star"], lb=[0, 0.1], ub=[2, 200], iteration=2,
                                             lossFunction="L2", adamForm=form, tol=0, convergenceValue=-1,
                                             convergenceStepTol=0, convergenceGradTol=0, initialLearnRate=0.05))
    cfg = askadam().make_cfg(f, None)
    y = rng.uniform(0.2, 1.0, (12, nv)).astype(np.float32)
    z0 = rng.uniform(-1, 1, (2, nv)).astype(np.float32)
    return md, cfg, y, z0


@pytest.mark.parametrize("form", ["effective", "kingma"])
def test_adam_first_step_closed_form(form):
    """iteration = 2: one Adam step from z0 with the gradient at z0
    (bias-corrected first step: dz = -lr*g/(|g| + eps') exactly)."""
    md, cfg, y, z0 = _adam_two_iterations(form)
    out = orc.askadam(md, cfg, y, z0)
    te = np.array([md.table[t] for t in range(12)], np.float64)
    z = z0.astype(np.float64)
    sg = 1 / (1 + np.exp(-z))
    lb, ub = np.array([md.lb[0], md.lb[1]])[:, None], np.array([md.ub[0], md.ub[1]])[:, None]
    th = lb + (ub - lb) * sg
    s = th[0] * np.exp(-np.outer(te, th[1]))
    r = s - y
    dl = 2 * r / r.size
    g0 = np.sum(dl * np.exp(-np.outer(te, th[1])), axis=0) * (ub[0] - lb[0]) * sg[0] * (1 - sg[0])
    g1 = np.sum(dl * (-te[:, None] * s), axis=0) * (ub[1] - lb[1]) * sg[1] * (1 - sg[1])
    g = np.stack([g0, g1])
    m = 0.1 * g
    v = 0.001 * g * g
    if form == "effective":
        step = 0.05 * np.sqrt(1 - 0.999) / (1 - 0.9) * m / (np.sqrt(v) + 1e-8)
    else:
        step = 0.05 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    np.testing.assert_allclose(out["z_final"], z - step, rtol=1e-9, atol=1e-12)
    # the final evaluation is at z1, no update on the last iteration
    assert out["n_iteration"] == 2


def test_askadam_stopping_rules_on_known_sequences():
    """loss < tol stops at the first such iteration; an L2 fit of noiseless
    data converges below tol; no update happens on the stop iteration."""
    nv = 50
    md = _r2s_md()
    th = np.stack([np.full(nv, 1.0), np.full(nv, 30.0)])
    y = orc.fwd_value(md, th).astype(np.float32)
    f = askadam.check_set_default_basic(dict(modelParams=["M0", "R2star"], lb=[0, 0.1], ub=[2, 200],
                                             iteration=3000, lossFunction="L2", initialLearnRate=0.05, tol=1e-6,
                                             convergenceValue=-1, convergenceStepTol=0, convergenceGradTol=0))
    cfg = askadam().make_cfg(f, None)
    z0 = np.zeros((2, nv), np.float32)
    out = orc.askadam(md, cfg, y, z0, trace=True)
    assert out["stop_reason"] == _abi.GC_STOP_TOL
    k = out["n_iteration"]
    assert out["trace"][k] < 1e-6 and np.all(out["trace"][1:k] >= 1e-6)
    # gradient-norm patience: with a huge tolerance every iteration counts,
    # so the stop comes at iteration patienceGrad + 1
    f2 = dict(f, tol=0, convergenceGradTol=1e9, patienceGrad=5)
    out2 = orc.askadam(md, askadam().make_cfg(f2, None), y, z0)
    assert (out2["n_iteration"], out2["stop_reason"]) == (6, _abi.GC_STOP_GRAD)
    # step-norm: iteration 1 has a zero step, so patience also ends at 6
    f3 = dict(f, tol=0, convergenceStepTol=1e9, patienceStep=5)
    out3 = orc.askadam(md, askadam().make_cfg(f3, None), y, z0)
    assert (out3["n_iteration"], out3["stop_reason"]) == (6, _abi.GC_STOP_STEP)


def _l1_sgdm(iteration, lr=0.1, **kw):
    f = askadam.check_set_default_basic(dict(modelParams=["M0", "R2star"], lb=[0, 0.1], ub=[2, 200],
                                             iteration=iteration, lossFunction="L1", optimiser="sgdm",
                                             initialLearnRate=lr, tol=0, convergenceValue=-1,
                                             convergenceStepTol=0, convergenceGradTol=0, **kw))
    return askadam().make_cfg(f, None)


def _pack_codes(codes):
    """[iteration, M, nv] codes (0/1/2) -> [iteration, ceil(M/16), nv] words."""
    it, M, nv = codes.shape
    w = np.zeros((it, (M + 15) // 16, nv), np.uint32)
    for m in range(M):
        w[:, m // 16, :] |= codes[:, m, :].astype(np.uint32) << np.uint32(2 * (m % 16))
    return w


def test_askadam_matched_sign_mode():
    """The L1 derivative follows the supplied residual signs (row e-1 for
    epoch e, the start evaluation with row 0): the oracle's own signs give
    the unmatched run exactly; all-zero codes give no gradient; flipped
    signs reverse the first SGDM step, z1 = z0 + lr*g instead of z0 - lr*g,
    with g = sum_m w*sign(r_m)*ds_m/dz / (M*N) from the closed form."""
    nv = 37
    md = _r2s_md()
    rng = np.random.default_rng(4)
    th = np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)])
    y = (orc.fwd_value(md, th) + 0.01 * rng.standard_normal((12, nv))).astype(np.float32)
    z0 = np.stack([np.full(nv, 0.3), np.full(nv, -1.2)]).astype(np.float32)
    th0 = np.stack([md.lb[p] + (md.ub[p] - md.lb[p]) / (1 + np.exp(-np.float64(z0[p]))) for p in range(2)])
    te = np.array([md.table[t] for t in range(12)], np.float64)
    s0 = th0[0] * np.exp(-np.outer(te, th0[1]))
    sg = np.sign(s0 - y.astype(np.float64))
    own = np.where(sg > 0, 1, np.where(sg < 0, 2, 0))
    cfg = _l1_sgdm(2)
    plain = orc.askadam(md, cfg, y, z0)
    codes = np.stack([own, np.ones_like(own)])          # row 1 (the last epoch) has no update
    matched = orc.askadam(md, cfg, y, z0, sign_trace=_pack_codes(codes))
    np.testing.assert_array_equal(matched["z_final"], plain["z_final"])
    # the loss stays the oracle's own w|r|, whatever the signs
    flipped = orc.askadam(md, cfg, y, z0, sign_trace=_pack_codes(np.stack([3 - own, own])), trace=True)
    ptr = orc.askadam(md, cfg, y, z0, trace=True)["trace"]
    assert flipped["trace"][0] == ptr[0]
    np.testing.assert_allclose(ptr[0], np.mean(np.abs(s0 - y.astype(np.float64))), rtol=1e-12)
    # closed-form first step (sgdm: m1 = -lr*g, z1 = z0 + m1)
    sig = [1 / (1 + np.exp(-np.float64(z0[p]))) for p in range(2)]
    dth = [(md.ub[p] - md.lb[p]) * sig[p] * (1 - sig[p]) for p in range(2)]
    e = np.exp(-np.outer(te, th0[1]))
    J = [e * dth[0], -te[:, None] * th0[0] * e * dth[1]]
    g = np.stack([np.sum(sg * J[p], axis=0) / (12 * nv) for p in range(2)])
    np.testing.assert_allclose(plain["z_final"], z0 - 0.1 * g, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(flipped["z_final"], z0 + 0.1 * g, rtol=1e-6, atol=1e-9)
    zero = orc.askadam(md, _l1_sgdm(6), y, z0, sign_trace=np.zeros((6, 1, nv), np.uint32))
    np.testing.assert_array_equal(zero["z_final"], z0.astype(np.float64))
    # a sign on measurement 5 only: the step is that measurement's term
    one = np.zeros((2, 12, nv), np.int64)
    one[0, 5] = 1
    o5 = orc.askadam(md, cfg, y, z0, sign_trace=_pack_codes(one))
    np.testing.assert_allclose(o5["z_final"], z0 - 0.1 * np.stack([J[p][5] for p in range(2)]) / (12 * nv),
                               rtol=1e-6, atol=1e-12)


def test_askadam_n_mean_scales_the_mean():
    """n_mean = N replaces nv in the mean over residuals (utils/askadam.m:96):
    twice the count halves the loss and the first SGDM step."""
    nv = 25
    md = _r2s_md()
    rng = np.random.default_rng(8)
    th = np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)])
    y = (orc.fwd_value(md, th) + 0.01 * rng.standard_normal((12, nv))).astype(np.float32)
    z0 = np.zeros((2, nv), np.float32)
    a = orc.askadam(md, _l1_sgdm(2), y, z0, trace=True)
    b = orc.askadam(md, _l1_sgdm(2), y, z0, trace=True, n_mean=2 * nv)
    # the start and the first epoch (same parameters) halve; later losses
    # follow the smaller step
    np.testing.assert_allclose(b["trace"][:2], a["trace"][:2] / 2, rtol=1e-14)
    np.testing.assert_allclose(b["z_final"] - z0, (a["z_final"] - z0) / 2, rtol=1e-12, atol=1e-15)


def test_askadam_linear_convergence_stop():
    """convergenceModel 'linear' (update_convergence, utils/askadam.m:1185-1190,
    and the patience rules of :573-597) driven through the oracle's loop on a
    constructed loss sequence.  TE ~ 0 makes s = M0; with the linear
    transform, L1 loss, data below lb and SGDM without momentum every step
    lowers z by the same amount, so the loss falls linearly (slope 0.015 per
    iteration, checked in closed form) until the clamp at z = 0 flattens it.
    The stop must land where minus the least-squares slope over the
    20-iteration buffer (primed with ones) stays <= convergenceValue for more
    than patienceConvergence iterations after the window has filled."""
    m = models.gpuR2starMapping(np.full(12, 1e-9))
    nv = 8
    rng = np.random.default_rng(2)
    lb, ub = [0.5, 0.1], [2.0, 200.0]
    c = 1e-3
    f = askadam.check_set_default_basic(dict(
        modelParams=["M0", "R2star"], lb=lb, ub=ub, iteration=400, lossFunction="L1", optimiser="sgdm",
        sgdmupdateMomentum=0.0, initialLearnRate=nv * 0.01 / 1.5, parameterTransform="linear", tol=0,
        convergenceModel="linear", convergenceWindow=20, convergenceValue=c, patienceConvergence=5,
        convergenceStepTol=0, convergenceGradTol=0))
    md = m.gc_model(f, "askadam", 12)
    cfg = askadam().make_cfg(f, None)
    assert cfg.convergence_model == _abi.GC_CONV_LINEAR and cfg.transform == _abi.GC_TRANSFORM_LINEAR
    y = np.repeat(rng.uniform(0.1, 0.4, (1, nv)), 12, axis=0).astype(np.float32)
    z0 = np.stack([np.full(nv, 0.905), np.full(nv, 0.2)]).astype(np.float32)
    out = orc.askadam(md, cfg, y, z0, trace=True)
    tr = out["trace"]
    assert out["n_iteration"] > 100
    # the constructed sequence: linear with slope -0.015 up to the clamp, then flat
    e = np.arange(1, 80)
    np.testing.assert_allclose(np.diff(tr[e]), -0.015, rtol=1e-6)
    # (R2* still moves, by TE ~ 1e-9 times less)
    assert np.ptp(tr[-7:]) < 1e-12
    # the stop rule restated on that sequence
    buf, pc, stop_at, conv_at = np.ones(20), 0, None, None
    for k in range(1, len(tr)):
        buf = np.append(buf[1:], tr[k])
        conv = -np.polyfit(np.arange(1, 21), buf, 1)[0]
        pc = 0 if (conv > c or k <= 20) else pc + 1
        if pc > 5:
            stop_at, conv_at = k, conv
            break
    assert stop_at is not None and stop_at > 100
    assert out["stop_reason"] == _abi.GC_STOP_CONVERGENCE
    assert out["n_iteration"] == stop_at
    assert out["convergence"] == pytest.approx(conv_at, rel=1e-9, abs=1e-15)
    # a loss flat from the start (learning rate 0) below the buffer's priming
    # ones: the signal is positive until the buffer holds 20 losses, zero at
    # iteration 20, which the window guard still excludes, so patience runs
    # from 21 and the stop is at 20 + 5 + 1
    cfg0 = askadam().make_cfg(dict(f, initialLearnRate=0.0), None)
    flat = orc.askadam(md, cfg0, y, z0, trace=True)
    assert (flat["n_iteration"], flat["stop_reason"]) == (26, _abi.GC_STOP_CONVERGENCE)
    assert abs(flat["convergence"]) < 1e-14


# ---------------------------------------------------------------------------
def test_mh_replay_is_deterministic_and_rejects_out_of_bounds():
    nv = 20
    md = r2s_setup({}, solver="mcmc")[2]
    th = np.stack([np.full(nv, 1.0), np.full(nv, 30.0)])
    y = (orc.fwd_value(md, np.vstack([th, np.full((1, nv), 0.01)]))).astype(np.float32)
    cfg = mcmc_cfg(_abi.GC_MCMC_MH, 200, 5, 20, replay=True, steps=(0.01, 1, 0.005))
    rng = np.random.default_rng(0)
    n = rng.standard_normal((1, 200, nv, 3)).astype(np.float32)
    n[0, 10, 3, 0] = 1e6  # forces M0 out of bounds at iteration 11 for voxel 3
    u = rng.uniform(0, 1, (1, 200, nv)).astype(np.float32)
    x0 = np.vstack([th, np.full((1, nv), 0.01)]).astype(np.float32)
    a = orc.mcmc(md, cfg, y, x0, normals=n, uniforms=u, trace=True, margin=True)
    b = orc.mcmc(md, cfg, y, x0, normals=n, uniforms=u, trace=True)
    np.testing.assert_array_equal(a["posterior"], b["posterior"])
    assert a["trace"][0, 10, 0, 3] == 0 and a["margin"][0, 10, 0, 3] == -np.inf


def test_mh_posterior_of_analytic_target():
    """TE ~ 0 makes s = M0; with fixed noise the M0 posterior is
    N(mean(y), sigma^2/M)."""
    te = np.full(12, 1e-9)
    m = models.gpuR2starMapping(te)
    nv, sigma = 60, 0.02
    rng = np.random.default_rng(3)
    y = (rng.uniform(0.8, 1.2, nv) + sigma * rng.standard_normal((12, nv))).astype(np.float32)
    from gacelle_amd.mcmc import mcmc as M
    f = M.check_set_default_basic(dict(modelParams=["M0", "R2star", "noise"], lb=[0, 0.1, sigma],
                                       ub=[2, 200, sigma]))
    md = m.gc_model(f, "mcmc", 12)
    cfg = mcmc_cfg(_abi.GC_MCMC_MH, 20000, 10, 2000, rep=2, steps=(0.01, 1, 0.0))
    x0 = np.stack([y.mean(axis=0), np.full(nv, 30.0), np.full(nv, sigma)]).astype(np.float32)
    o = orc.mcmc(md, cfg, y, x0, store=False, nthreads=4)
    z = (o["mean"][0] - y.astype(np.float64).mean(axis=0)) / (sigma / np.sqrt(12))
    assert abs(np.mean(z)) < 0.1 and np.std(z) < 0.25
    np.testing.assert_allclose(np.median(o["std"][0]), sigma / np.sqrt(12), rtol=0.06)


def test_gw_native_stretch_density():
    """The oracle's native Goodman-Weare stream: zz = ((a-1)u+1)^2/a from the
    stretch uniform has density proportional to 1/sqrt(z) on [1/a, a]
    (utils/mcmc.m:299), and the acceptance uniform of the same draw is U(0,1)
    and uncorrelated with it."""
    a = 2.0
    n = 60000
    uz = np.empty(n)
    ua = np.empty(n)
    for i in range(n):
        uz[i], ua[i] = orc.native_gw_uniforms(48463, 1 + i % 997, i // 997, i % 2)
    zz = ((a - 1) * uz + 1) ** 2 / a
    # CDF of p(z) ~ 1/sqrt(z) on [1/a, a]: (sqrt(z) - sqrt(1/a)) / (sqrt(a) - sqrt(1/a))
    cdf = lambda z: (np.sqrt(z) - np.sqrt(1 / a)) / (np.sqrt(a) - np.sqrt(1 / a))
    assert stats.kstest(zz, cdf).pvalue > 1e-3
    assert stats.kstest(ua, "uniform").pvalue > 1e-3
    assert abs(np.corrcoef(uz, ua)[0, 1]) < 0.02


def _grid_posterior(y, te, lb, ub, sig):
    """Posterior mean and SD of (M0, R2*, noise) under logP_Gaussian with a
    uniform prior on the box, by integration on a grid around the
    least-squares fit (the noise axis spans its whole box)."""
    from scipy.optimize import least_squares
    p = least_squares(lambda q: q[0] * np.exp(-te * q[1]) - y, [1.0, 30.0]).x
    J = np.stack([np.exp(-te * p[1]), -te * p[0] * np.exp(-te * p[1])], 1)
    sd = np.sqrt(np.diag(np.linalg.inv(J.T @ J))) * sig * 3
    axes = [np.linspace(p[0] - 8 * sd[0], p[0] + 8 * sd[0], 161), np.linspace(p[1] - 8 * sd[1], p[1] + 8 * sd[1], 161),
            np.linspace(lb[2], ub[2], 400)]
    S = axes[0][:, None, None] * np.exp(-te[None, None, :] * axes[1][None, :, None])
    SS = np.sum((S - y) ** 2, axis=2)
    lp = -SS[:, :, None] / (2 * axes[2] ** 2) - te.size * np.log(axes[2])
    P = np.exp(lp - lp.max())
    P /= P.sum()
    out = []
    for ax, g in enumerate(axes):
        marg = P.sum(axis=tuple(i for i in range(3) if i != ax))
        mu = np.sum(marg * g)
        out.append((mu, np.sqrt(np.sum(marg * (g - mu) ** 2))))
    return out


@pytest.mark.parametrize("redblack", [False, True])
def test_gw_posterior_of_analytic_target(redblack):
    """The oracle's plain Goodman-Weare (utils/mcmc.m:192-383) samples the
    posterior of (M0, R2*, noise) of mono-exponential data, all three
    variables live, whose moments come from grid integration: posterior means
    within 0.05 posterior SD and SDs within 3%.  Planted in a copy of the
    oracle (on the CPU), z^0 instead of z^(Q-1) gives SDs x0.65, z^Q x1.22,
    and a stretch about the walker instead of its partner x0.46; a walker
    drawn as its own partner only slows mixing and leaves this target
    unchanged (the derangement is pinned by test_native_partners_are_derangements)."""
    from gacelle_amd.mcmc import mcmc as M
    lb, ub = [0, 0.1, 0.003], [2, 200, 0.05]
    md = models.gpuR2starMapping(TE12).gc_model(
        M.check_set_default_basic(dict(modelParams=["M0", "R2star", "noise"], lb=lb, ub=ub)), "mcmc", 12)
    rng = np.random.default_rng(11)
    nd, rep, sig = 6, 40, 0.01
    th = np.stack([rng.uniform(0.7, 1.3, nd), rng.uniform(15, 50, nd)])
    y0 = (th[0] * np.exp(-np.outer(TE12, th[1])) + sig * rng.standard_normal((12, nd))).astype(np.float32)
    te = np.array([md.table[t] for t in range(12)], np.float64)
    ref = [_grid_posterior(y0[:, i].astype(np.float64), te, lb, ub, sig) for i in range(nd)]
    x0 = np.stack([np.repeat(th[0], rep), np.repeat(th[1], rep), np.full(nd * rep, sig)]).astype(np.float32)
    cfg = mcmc_cfg(_abi.GC_MCMC_GW, 4000, 5, 1000, W=16, redblack=redblack, Q=3)
    o = orc.mcmc(md, cfg, np.repeat(y0, rep, axis=1), x0, store=False, nthreads=4)
    for q in range(3):
        mu = o["mean"][q].reshape(nd, rep).mean(axis=1)
        sd = np.sqrt((o["std"][q].reshape(nd, rep) ** 2).mean(axis=1))
        rm = np.array([r[q][0] for r in ref])
        rs = np.array([r[q][1] for r in ref])
        assert np.all(np.abs(mu - rm) < 0.05 * rs), (q, (mu - rm) / rs)
        np.testing.assert_allclose(sd / rs, 1.0, atol=0.03)


# ---------------------------------------------------------------------------
# MCR-MWI: the Bloch-McConnell steady state against a matrix-exponential
# solve of the two-pool exchange equations, and its no-exchange limit
def _mcrmwi_md(exch, cplx, te=(1e-6, 2e-6, 3e-6)):
    m = models.gpuMCRMWI(np.asarray(te), 38e-3, np.array([5.0, 10.0, 20.0, 50.0]))
    f = m.check_set_default(dict(isComplex=cplx, isFitExchange=exch, isEPG=False))
    m.updateProperty(f)
    f["modelParams"], f["lb"], f["ub"] = list(m.modelParams), list(m.lb), list(m.ub)
    return m, f, m.gc_model(f, "askadam", None)


def _bm_expm(TR, alpha, f, R1f, R1r, kfr):
    """Spoiled-GRE steady state of dM/dt = A M + C (pools free f-pool, restricted
    r-pool, exchange kfr and kfr (1-f)/f): Mz- = (I - cos(a) E)^-1 A^-1 (E - I) C,
    signal = sin(a) Mz- per pool."""
    krf = kfr * (1 - f) / f
    A = np.array([[-R1f - kfr, krf], [kfr, -R1r - krf]])
    C = np.array([R1f * (1 - f), R1r * f])
    E = linalg.expm(A * TR)
    Mz = np.linalg.solve(np.eye(2) - np.cos(alpha) * E, np.linalg.solve(A, (E - np.eye(2)) @ C))
    return np.sin(alpha) * Mz  # [IEW, MW]


@pytest.mark.parametrize("exch", [True, False])
def test_mcrmwi_t1_steady_state(exch):
    """At TE -> 0 with zero frequencies the magnitude equals ff (a_MW + a_IW +
    a_EW): the BM analytical solution (MCRMWI/gpuMCRMWI.m:992-1023) matches
    the matrix-exponential steady state; without exchange the Bloch model."""
    m, f, md = _mcrmwi_md(exch, cplx=False)
    S0, mwf, iwf, r1iew, kiewm = 1.3, 0.15, 0.6, 1.1, 2.5
    th = [S0, mwf, iwf, r1iew] + ([kiewm] if exch else []) + [80.0, 20.0, 25.0, 0.0, 0.0]
    b1, ff = 0.95, 1.0
    ex = np.array([[b1], [0.0], [ff], [0.0], [0.0], [0.0], [0.0]])
    s = orc.fwd_value(md, np.array(th)[:, None], ex)[:, 0]
    rho, t1mw, TR = md.f[3], md.f[4], md.f[0]
    S0MW, S0IW, S0EW = S0 * mwf, S0 * (1 - mwf) * iwf, S0 * (1 - mwf) * (1 - iwf)
    tv = S0IW + S0EW + S0MW / rho
    mvf = (S0MW / rho) / tv
    for fi, fa in enumerate([5.0, 10.0, 20.0, 50.0]):
        alpha = np.deg2rad(np.float32(fa)) * b1
        if exch:
            iew, mw = _bm_expm(TR, alpha, mvf, r1iew, 1 / t1mw, kiewm)
            amp = rho * mw * tv + iew * tv
        else:
            e_mw, e_iew = np.exp(-TR / t1mw), np.exp(-TR * r1iew)
            amp = S0MW * np.sin(alpha) * (1 - e_mw) / (1 - np.cos(alpha) * e_mw) + \
                (S0IW + S0EW) * np.sin(alpha) * (1 - e_iew) / (1 - np.cos(alpha) * e_iew)
        decay = np.exp(-1e-6 * np.array([80.0, 20.0, 25.0]))
        assert s[3 * fi] == pytest.approx(amp * ff, rel=1e-3)  # TE = 1 us: decay ~ 1e-4
        assert s[3 * fi] <= amp * ff * decay.max() * (1 + 1e-9)


def test_mcrmwi_no_exchange_limit():
    """kIEWM -> 0 turns the BM steady state into the Bloch model."""
    _, _, md_x = _mcrmwi_md(True, cplx=False)
    _, _, md_b = _mcrmwi_md(False, cplx=False)
    base = [1.3, 0.15, 0.6, 1.1]
    rest = [80.0, 20.0, 25.0, 0.0, 0.0]
    ex = np.array([[0.95], [0.0], [1.0], [0.0], [0.0], [0.0], [0.0]])
    sx = orc.fwd_value(md_x, np.array(base + [1e-7] + rest)[:, None], ex)
    sb = orc.fwd_value(md_b, np.array(base + rest)[:, None], ex)
    np.testing.assert_allclose(sx, sb, rtol=1e-5)


def test_mcrmwi_complex_phase_and_sum():
    """compute_gremwi_signal: re/im of sum_p a_p exp(-TE R2s_p) exp(i(2 pi TE
    (f_p + f_bkg) B0 gamma + pini)) times ff; the magnitude model is |.| of
    the complex one."""
    te = np.array([3e-3, 7e-3, 12e-3])
    _, _, mc = _mcrmwi_md(True, cplx=True, te=te)
    _, _, mm = _mcrmwi_md(True, cplx=False, te=te)
    amp_par = [1.3, 0.15, 0.6, 1.1, 2.5]
    r2 = [80.0, 20.0, 25.0]
    fr = [0.08, -0.02]
    bkg = [0.01, -0.02, 0.03, 0.0]
    ex = np.array([[1.0], [0.3], [0.9], [0.05], [0.0], [-0.05], [0.02]])
    thc = np.array(amp_par + r2 + fr + bkg + [0.1])[:, None]
    thm = np.array(amp_par + r2 + fr)[:, None]
    sc = orc.fwd_value(mc, thc, ex)[:, 0]
    ex_m = ex.copy()
    sm = orc.fwd_value(mm, thm, ex_m)[:, 0]
    nte, nfa = 3, 4
    re, im = sc[: nte * nfa], sc[nte * nfa:]
    # the magnitude model has no background field, pini or dpini
    thz = np.array(amp_par + r2 + fr + [0, 0, 0, 0] + [0.0])[:, None]
    ex0 = ex.copy()
    ex0[1] = 0.0
    ex0[3:] = 0.0
    s0 = orc.fwd_value(mc, thz, ex0)[:, 0]
    np.testing.assert_allclose(np.hypot(s0[: nte * nfa], s0[nte * nfa:]), sm, rtol=1e-12)
    # a common phase shift (dpini) rotates every (re, im) pair rigidly
    thr = thc.copy()
    thr[-1] += 0.7
    sr = orc.fwd_value(mc, thr, ex)[:, 0]
    z = (re + 1j * im) * np.exp(1j * 0.7)
    np.testing.assert_allclose(sr[: nte * nfa], z.real, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(sr[nte * nfa:], z.imag, rtol=1e-10, atol=1e-14)
    # ff scales the signal linearly
    ex2 = ex.copy()
    ex2[2] *= 2
    np.testing.assert_allclose(orc.fwd_value(mc, thc, ex2)[:, 0], 2 * sc, rtol=1e-12)


def test_mcrmwi_pini_per_flip_angle():
    """GC_FLAG_MCRMWI_PINI_FA: flip angle k uses pini row 3+nfa+k.  Each
    flip-angle block equals the one-pini model evaluated with that row's
    value; with equal rows the two forms agree everywhere."""
    te = np.array([3e-3, 7e-3, 12e-3])
    m, f, mc = _mcrmwi_md(True, cplx=True, te=te)
    f2 = dict(f, piniPerFA=True)
    mp = m.gc_model(f2, "askadam", None)
    assert mp.flags & _abi.GC_FLAG_MCRMWI_PINI_FA and mp.n_extra == 3 + 2 * 4
    th = np.array([1.3, 0.15, 0.6, 1.1, 2.5, 80.0, 20.0, 25.0, 0.08, -0.02, 0.01, -0.02, 0.03, 0.0, 0.1])[:, None]
    base = np.array([[1.0], [0.0], [0.9], [0.05], [0.0], [-0.05], [0.02]])
    pinis = [0.3, -1.2, 2.0, 0.7]
    ex_fa = np.vstack([base, np.array(pinis)[:, None]])
    s_fa = orc.fwd_value(mp, th, ex_fa)[:, 0]
    nte, nfa = 3, 4
    for k, pk in enumerate(pinis):
        exk = base.copy()
        exk[1] = pk
        sk = orc.fwd_value(mc, th, exk)[:, 0]
        for half in (0, 1):
            sl = slice(half * nte * nfa + k * nte, half * nte * nfa + (k + 1) * nte)
            np.testing.assert_allclose(s_fa[sl], sk[sl], rtol=1e-12, atol=1e-15)
    ex_eq = np.vstack([base, np.full((4, 1), 0.4)])
    ex_one = base.copy()
    ex_one[1] = 0.4
    np.testing.assert_allclose(orc.fwd_value(mp, th, ex_eq), orc.fwd_value(mc, th, ex_one), rtol=1e-12)


# ---------------------------------------------------------------------------
# robustConvergence (utils/askadam.m:448-510, 541-566)
def _robust_problem(nv=400, n_out=40, seed=9):
    rng = np.random.default_rng(seed)
    md = _r2s_md()
    th = np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)])
    y = orc.fwd_value(md, th) + 0.005 * rng.standard_normal((12, nv))
    # outlier voxels: heavy noise around the starting point's own signal, so
    # the fit cannot improve them while the other voxels improve a lot
    start = orc.fwd_value(md, np.array([[1.0], [100.05]]))
    y[:, :n_out] = start + 5.0 * rng.standard_normal((12, n_out))
    return md, y.astype(np.float32), np.zeros((2, nv), np.float32)


def _robust_cfg(**kw):
    base = dict(modelParams=["M0", "R2star"], lb=[0, 0.1], ub=[2, 200], iteration=120, initialLearnRate=0.02,
                tol=0, convergenceValue=-1, convergenceStepTol=0, convergenceGradTol=0, robustConvergence=True)
    base.update(kw)
    f = askadam.check_set_default_basic(base)
    return askadam().make_cfg(f, None)


def test_robust_without_flags_equals_plain_loop():
    """Thresholds no voxel can fail and no convergence stop: the robust loop
    takes exactly the plain loop's path."""
    md, y, z0 = _robust_problem()
    plain = orc.askadam(md, _robust_cfg(robustConvergence=False), y, z0, trace=True)
    rob = orc.askadam(md, _robust_cfg(outlierVoxelThres=-1e9, outlierInitThres=-1e9), y, z0, trace=True)
    np.testing.assert_array_equal(plain["trace"], rob["trace"])
    np.testing.assert_array_equal(plain["z_final"], rob["z_final"])


def test_robust_modulates_only_the_epoch_after_a_check():
    """update_outlier_mask returns the modulated weights on check epochs only:
    with every voxel flagged and outlierWeight 0, the (weighted) loss is 0
    exactly on epochs k*weightUpdateInterval + 1 and positive elsewhere."""
    md, y, z0 = _robust_problem()
    cfg = _robust_cfg(outlierVoxelThres=1e9, outlierPopThres=-1e9, outlierInitThres=1e9, outlierInitPopThres=-1e9,
                      outlierWeight=0.0, weightUpdateInterval=5, iteration=40)
    o = orc.askadam(md, cfg, y, z0, trace=True)
    tr = o["trace"]
    zero = [e for e in range(1, 41) if tr[e] == 0]
    assert zero == [6, 11, 16, 21, 26, 31, 36]
    assert np.all(tr[[e for e in range(41) if e not in zero]] > 0)


def test_robust_flags_the_unfittable_voxels():
    """The unfittable voxels stagnate while the median improves: they fail both
    criteria and are down-weighted, so the robust fit's final loss (weighted
    on check-following epochs only) matches the plain fit on the good voxels."""
    md, y, z0 = _robust_problem()
    plain = orc.askadam(md, _robust_cfg(robustConvergence=False, iteration=200), y, z0, trace=True)
    rob = orc.askadam(md, _robust_cfg(iteration=200, outlierWeight=0.0), y, z0, trace=True)
    # once the median improvement passes outlierInitPopThres (the check at
    # epoch 20), the epochs following each check see the outliers at weight 0
    r = rob["trace"] / plain["trace"]
    drops = [e for e in range(21, 200) if (e - 1) % 5 == 0]
    assert np.all(r[drops] < 0.5), r[drops].max()
    np.testing.assert_array_equal(r[:21], 1.0)


def test_rescale_unscale_parameters():
    """askadam.rescale_parameters / unscale_parameters (utils/askadam.m:
    1063-1099): sigmoid maps theta inside [lb + eb, ub - eb] (eb = 1e-4 of
    the range) to its logit and back exactly; outside, theta is clamped to
    those points (z = -/+ logit(1 - 1e-4) = -/+ 9.2102...); linear is the
    affine map to [0, 1] and back."""
    md = _r2s_md()
    lb, ub = np.array([md.lb[0], md.lb[1]]), np.array([md.ub[0], md.ub[1]])
    rng = np.random.default_rng(1)
    th = np.stack([rng.uniform(0.01, 1.99, 50), rng.uniform(1, 199, 50)])
    for tr in (_abi.GC_TRANSFORM_SIGMOID, _abi.GC_TRANSFORM_LINEAR):
        z = orc.rescale_parameters(md, tr, th)
        np.testing.assert_allclose(orc.unscale_parameters(md, tr, z), th, rtol=1e-12)
    zl = orc.rescale_parameters(md, _abi.GC_TRANSFORM_LINEAR, th)
    np.testing.assert_allclose(zl, (th - lb[:, None]) / (ub - lb)[:, None], rtol=1e-15)
    edge = np.array([[-1.0, 0.0, 2.0, 5.0], [0.0, 0.1, 200.0, 1e4]])
    ze = orc.rescale_parameters(md, _abi.GC_TRANSFORM_SIGMOID, edge)
    k = np.log((1 - 1e-4) / 1e-4)
    np.testing.assert_allclose(ze, [[-k, -k, k, k], [-k, -k, k, k]], rtol=1e-9)
    np.testing.assert_allclose(orc.unscale_parameters(md, _abi.GC_TRANSFORM_SIGMOID, np.zeros((2, 1)))[:, 0],
                               (lb + ub) / 2, rtol=1e-15)
