"""The oracle's dictionary search (gpuNEXI.likelihood, NEXI/gpuNEXI.m:822-849)
pinned without retyping it: the scale of each entry from numpy's least
squares, planted exact matches, first-index ties, and MATLAB min's NaN
rule."""
import numpy as np

from oracle import oracle as orc


def _problem(m=6, nv=40, nd=300, seed=0):
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.05, 1.0, (m, nd)).astype(np.float32)
    y = rng.uniform(0.05, 1.0, (m, nv)).astype(np.float32)
    return y, d


def _lstsq_best(y, d, w):
    """argmin_j sum_k w_k (y_k - c_j d_kj)^2 with c_j the least-squares
    scale of d_j onto y (unweighted, as the reference projects)."""
    y = y.astype(np.float64)
    d = d.astype(np.float64)
    best, sse = [], []
    for v in range(y.shape[1]):
        s = np.empty(d.shape[1])
        for j in range(d.shape[1]):
            c = np.linalg.lstsq(d[:, j:j + 1], y[:, v], rcond=None)[0][0]
            s[j] = np.sum(w * (y[:, v] - c * d[:, j]) ** 2)
        best.append(int(np.argmin(s)))
        sse.append(s.min())
    return np.array(best), np.array(sse)


def test_matches_least_squares_projection():
    y, d = _problem()
    w = np.array([1, 1, 1, 0.2, 0.2, 0.2])
    b, s = orc.dict_search(y, d, w)
    bl, sl = _lstsq_best(y, d, w)
    np.testing.assert_array_equal(b, bl)
    np.testing.assert_allclose(s, sl, rtol=1e-10)
    # the projection is unweighted: projecting with the weights picks other
    # entries for some voxels of this problem, so the check above tells them apart
    yy, dd = y.astype(np.float64), d.astype(np.float64)
    cw = ((w[:, None] * yy).T @ dd) / (w @ dd ** 2)
    sw = np.array([[np.sum(w * (yy[:, v] - cw[v, j] * dd[:, j]) ** 2) for j in range(d.shape[1])]
                   for v in range(y.shape[1])])
    assert np.any(sw.argmin(axis=1) != b)
    bu, _ = orc.dict_search(y, d, None)
    bul, _ = _lstsq_best(y, d, np.ones(6))
    np.testing.assert_array_equal(bu, bul)


def test_planted_match_ties_and_nan():
    y, d = _problem(m=5, nv=8, nd=200, seed=1)
    k = np.array([3, 150, 199, 0, 77, 77, 10, 11])
    s = np.linspace(0.5, 2.0, 8)
    y = (d[:, k] * s).astype(np.float32)
    b, sse = orc.dict_search(y, d)
    np.testing.assert_array_equal(b, k)
    assert np.all(sse < 1e-12)
    # duplicated entry: the first copy wins
    d2 = np.concatenate([d, d[:, 150:151]], axis=1)
    assert orc.dict_search(y, d2)[0][1] == 150
    d3 = d.copy()
    d3[:, 5] = d3[:, 150]
    assert orc.dict_search(y, d3)[0][1] == 5
    # a zero entry scores NaN and never wins; an all-NaN voxel keeps -1;
    # a zero voxel scores 0 everywhere and takes the first entry
    d4 = d.copy()
    d4[:, 0] = 0
    y4 = y.copy()
    y4[:, 0] = np.nan
    y4[:, 1] = 0
    b4, s4 = orc.dict_search(y4, d4)
    assert b4[0] == -1 and np.isnan(s4[0])
    assert b4[1] == 1 and s4[1] == 0
    assert b4[3] != 0
