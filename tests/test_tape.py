"""The plain per-op GPU baselines of bench.py (gacelle_amd/tape.py), run on
the CPU: the askadam tape follows the oracle's fit of the same problem, and
the Metropolis-Hastings tape samples the analytic posterior of a TE ~ 0
target (s = M0, M0 posterior N(mean(y), sigma^2 / M))."""
import numpy as np
import torch

from gacelle_amd import _abi, models
from gacelle_amd.askadam import askadam
from gacelle_amd.tape import askadam_r2star_tape, mh_r2star_tape
from oracle import oracle as orc

from .helpers import TE12, logit_start, no_stop


def test_askadam_tape_follows_the_oracle():
    nv = 300
    rng = np.random.default_rng(6)
    m = models.gpuR2starMapping(TE12)
    f = askadam.check_set_default_basic(dict(modelParams=["M0", "R2star"], lb=[0, 0.1], ub=[2, 200],
                                             **no_stop(iteration=200, lossFunction="L2", initialLearnRate=1e-2)))
    md = m.gc_model(f, "askadam", 12)
    th = np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)])
    y = (orc.fwd_value(md, th) + 0.01 * rng.standard_normal((12, nv))).astype(np.float32)
    z0 = logit_start(np.stack([np.ones(nv), np.full(nv, 30.0)]), md)
    ref = orc.askadam(md, askadam().make_cfg(f, m), y, z0)
    z, loss = askadam_r2star_tape(torch.from_numpy(y), torch.tensor(TE12, dtype=torch.float32), [0.0, 0.1],
                                  [2.0, 200.0], torch.from_numpy(z0.astype(np.float32)), 200, lr=1e-2, loss="l2")
    np.testing.assert_allclose(z.numpy(), ref["z_final"], rtol=1e-3, atol=1e-3)
    # the oracle's loss is the mean of the squared residual over M * Nv
    np.testing.assert_allclose(loss, ref["loss"], rtol=1e-4)


def test_mh_tape_samples_the_analytic_posterior():
    te = np.full(12, 1e-9)
    nv, sigma = 200, 0.02
    rng = np.random.default_rng(3)
    y = (rng.uniform(0.8, 1.2, nv) + sigma * rng.standard_normal((12, nv))).astype(np.float32)
    x0 = np.stack([y.mean(axis=0), np.full(nv, 30.0), np.full(nv, sigma)]).astype(np.float32)
    gen = torch.Generator()
    gen.manual_seed(5)
    mean = mh_r2star_tape(torch.from_numpy(y), torch.tensor(te, dtype=torch.float32), torch.from_numpy(x0),
                          [0, 0.1, sigma], [2, 200, sigma], [0.01, 1, 0.0], 6000, 2, 10, 600, gen).numpy()
    z = (mean[0] - y.astype(np.float64).mean(axis=0)) / (sigma / np.sqrt(12))
    assert abs(np.mean(z)) < 0.1 and np.std(z) < 0.3
