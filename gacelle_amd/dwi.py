"""Diffusion-data helpers of the NEXI estimate() glue (utils/DWIutility.m).

The shell table (unique_shell_keepb0), the rotationally invariant signal of
each shell (SHrotinv, Sl_no_normalise, get_Sl_all_no_normalise,
compute_rotationally_invariant_signal) and the b0-normalised variant used by
the Rician NEXI model (Sl, get_Sl_all).  Host-side data formatting that runs
once per volume before the fit; the fit itself runs in the HIP library.

Rotational invariants.  The reference fits real spherical harmonics of every
degree 0..lmax by least squares (pinv) to the antipodally symmetrised samples
{(S, g), (S, -g)} (utils/DWIutility.m:402-417, leastSquaresSHT of the
vendored Spherical-Harmonic-Transform library) and returns, per even degree
l, F_l = ||c_l|| / sqrt(4 pi (2l + 1)) with c_l the 2l+1 coefficients of
degree l.  On symmetrised samples the normal equations split into an even
and an odd block, and the odd right-hand side vanishes, so the even
coefficients equal the minimum-norm least-squares fit of the even degrees
alone to the N original samples; and ||c_l|| is the same in every
orthonormal real basis of degree l.  This module therefore fits only the
even degrees, in its own orthonormal real basis.
"""
from __future__ import annotations

import numpy as np
from scipy.special import lpmv, gammaln

from ._abi import GacelleError, GC_ERR_ARG

EPSILON = 1e-8  # askadam.epsilon


def unique_shell_keepb0(bval, ldelta=None, BDELTA=None, te=None, discardB0=False):
    """Shell table (utils/DWIutility.m:647-775): per echo time (ascending), a
    b=0 entry (unless discardB0 with a single echo time), then for each
    little delta and big delta (ascending) the sorted non-zero b-values."""
    bval = np.asarray(bval, np.float64).ravel()
    n = bval.size
    te = np.ones(n) if te is None or np.size(te) == 0 else np.broadcast_to(np.asarray(te, np.float64).ravel(), (n,))
    ld = np.zeros(n) if ldelta is None or np.size(ldelta) == 0 else \
        np.broadcast_to(np.asarray(ldelta, np.float64).ravel(), (n,))
    bd = np.zeros(n) if BDELTA is None or np.size(BDELTA) == 0 else \
        np.broadcast_to(np.asarray(BDELTA, np.float64).ravel(), (n,))
    out = {"bval": [], "ldelta": [], "BDELTA": [], "te": []}
    te_u = np.unique(te)
    multi = te_u.size > 1

    def add(b, l, d, t):
        out["bval"].append(b)
        out["ldelta"].append(l)
        out["BDELTA"].append(d)
        out["te"].append(t)

    for t in te_u:
        it = te == t
        if np.any(bval[it] == 0) and not (discardB0 and not multi):
            l1 = ld[it].min()
            add(0.0, l1, bd[it & (ld == l1)].min(), t)
        for l in np.unique(ld[it]):
            il = it & (ld == l)
            for d in np.unique(bd[il]):
                bu = np.unique(bval[il & (bd == d)])
                for b in bu[bu > 0]:
                    add(b, l, d, t)
    return tuple(np.array(out[k]) for k in ("bval", "ldelta", "BDELTA", "te"))


def real_sh_even(lmax, azi, incl):
    """Orthonormal real spherical harmonics of the even degrees 0..lmax at
    directions (azimuth, inclination): [Ndir, sum_l (2l+1)], degree-major,
    order m = -l..l within a degree.  Y_l^m = N_lm P_l^|m|(cos incl) times
    sqrt(2) cos(m azi) (m > 0), 1 (m = 0), sqrt(2) sin(|m| azi) (m < 0), with
    N_lm = sqrt((2l+1)/(4 pi) (l-|m|)!/(l+|m|)!) (integral of Y^2 over the
    sphere = 1)."""
    azi = np.asarray(azi, np.float64)
    ct = np.cos(np.asarray(incl, np.float64))
    cols = []
    for l in range(0, int(lmax) + 1, 2):
        for m in range(-l, l + 1):
            am = abs(m)
            norm = np.sqrt((2 * l + 1) / (4 * np.pi) * np.exp(gammaln(l - am + 1) - gammaln(l + am + 1)))
            p = lpmv(am, l, ct)
            if m > 0:
                cols.append(norm * p * np.sqrt(2) * np.cos(m * azi))
            elif m < 0:
                cols.append(norm * p * np.sqrt(2) * np.sin(am * azi))
            else:
                cols.append(norm * p)
    return np.stack(cols, axis=1)


def cart2sph_incl(g):
    """[azimuth, inclination] of unit vectors g [N, 3] (utils/DWIutility.m:491-495)."""
    g = np.asarray(g, np.float64)
    azi = np.arctan2(g[:, 1], g[:, 0])
    elev = np.arctan2(g[:, 2], np.hypot(g[:, 0], g[:, 1]))
    return azi, np.pi / 2 - elev


def sh_rotinv(S, g, lmax):
    """Rotational invariants F [lmax/2 + 1, Nvox] of samples S [Ndir, Nvox]
    on directions g [Ndir, 3] (utils/DWIutility.m:402-417; see the module
    docstring for the even-degree reduction)."""
    S = np.asarray(S, np.float64)
    azi, incl = cart2sph_incl(g)
    Y = real_sh_even(lmax, azi, incl)
    c = np.linalg.pinv(Y) @ S
    F = np.zeros((int(lmax) // 2 + 1, S.shape[1]))
    s = 0
    for k, l in enumerate(range(0, int(lmax) + 1, 2)):
        F[k] = np.sqrt((np.abs(c[s:s + 2 * l + 1]) ** 2).sum(axis=0)) / np.sqrt(4 * np.pi * (2 * l + 1))
        s += 2 * l + 1
    return F


def _bvec_rows(bvec):
    """Directions as [N, 3]; a [3, N] input (the reference's layout, also
    for N = 3) is transposed (utils/DWIutility.m:307-309)."""
    g = np.asarray(bvec, np.float64)
    if g.ndim != 2 or 3 not in g.shape:
        raise GacelleError(GC_ERR_ARG, "bvec must be [3, N] or [N, 3]")
    return g.T if g.shape[0] == 3 else g


def sl_no_normalise(dwi, bval, bvec, lmax):
    """Per unique b-value (ascending) the invariants of degrees 0..lmax:
    [x,y,z, Nb * Nsl], b fastest (utils/DWIutility.m:361-400)."""
    dwi = np.asarray(dwi, np.float64)
    dims = dwi.shape[:3]
    g = _bvec_rows(bvec)
    bval = np.asarray(bval, np.float64).ravel()
    bu = np.unique(bval)
    nsl = int(lmax) // 2 + 1
    out = np.zeros(dims + (bu.size, nsl))
    flat = dwi.reshape((-1, dwi.shape[3]), order="F")
    for k, b in enumerate(bu):
        ind = bval == b
        F = sh_rotinv(flat[:, ind].T, g[ind], lmax)       # [nsl, nvox]
        out[..., k, :] = F.T.reshape(dims + (nsl,), order="F")
    return out.reshape(dims + (bu.size * nsl,), order="F"), bu


def sl(dwi, bval, bvec, lmax):
    """b0-normalised invariants of the non-zero shells (utils/DWIutility.m:306-359):
    each volume divided by the mean b=0 volume (ones when there is none)."""
    dwi = np.asarray(dwi, np.float64)
    dims = dwi.shape[:3]
    g = _bvec_rows(bvec)
    bval = np.asarray(bval, np.float64).ravel()
    bu = np.unique(bval)
    b0 = bval == 0
    dwi_b0 = dwi[..., b0].mean(axis=3) if np.any(b0) else np.ones(dims)
    nz = bu[bu != 0]
    nsl = int(lmax) // 2 + 1
    out = np.zeros(dims + (nz.size, nsl))
    flat = dwi.reshape((-1, dwi.shape[3]), order="F")
    fb0 = dwi_b0.reshape(-1, order="F")
    with np.errstate(divide="ignore", invalid="ignore"):
        for k, b in enumerate(nz):
            ind = bval == b
            F = sh_rotinv((flat[:, ind] / fb0[:, None]).T, g[ind], lmax)
            out[..., k, :] = F.T.reshape(dims + (nsl,), order="F")
    return out.reshape(dims + (nz.size * nsl,), order="F"), nz


def _groups(bval, ldelta, BDELTA):
    for l in np.unique(ldelta):
        il = ldelta == l
        for d in np.unique(BDELTA[il]):
            yield il & (BDELTA == d)


def get_Sl_all(dwi, bval, bvec, ldelta, BDELTA, lmax):
    """b0-normalised invariants grouped by (little delta, big delta)
    (utils/DWIutility.m:276-304): [x,y,z, Nshell * Nsl], shell fastest."""
    dwi = np.asarray(dwi, np.float64)
    dims = dwi.shape[:3]
    bval = np.asarray(bval, np.float64).ravel()
    n = bval.size
    ldelta = np.broadcast_to(np.asarray(ldelta, np.float64).ravel(), (n,))
    BDELTA = np.broadcast_to(np.asarray(BDELTA, np.float64).ravel(), (n,))
    g = _bvec_rows(bvec)
    nsl = int(lmax) // 2 + 1
    blocks, ball = [], []
    for grp in _groups(bval, ldelta, BDELTA):
        s, bu = sl(dwi[..., grp], bval[grp], g[grp], lmax)
        blocks.append(s.reshape(dims + (bu.size, nsl), order="F"))
        ball.append(bu)
    t = np.concatenate(blocks, axis=3)
    return t.reshape(dims + (t.shape[3] * nsl,), order="F"), np.concatenate(ball)


def compute_rotationally_invariant_signal(dwi, bval, bvec, ldelta=None, BDELTA=None, te=None, lmax=0):
    """Rotationally invariant signal (utils/DWIutility.m:10-166): per echo
    time a b=0 block and the (little delta, big delta, b) shells, invariants
    of degrees 0..lmax; with one echo time the non-zero shells divided by
    the b=0 Sl0 (floored at epsilon), l-major with shell fastest.  Returns
    (dwi_out, bval_out, ldelta_out, BDELTA_out, te_out)."""
    dwi = np.asarray(dwi, np.float64)
    dims = dwi.shape[:3]
    bval = np.asarray(bval, np.float64).ravel()
    n = bval.size
    ld = np.ones(n) if ldelta is None or np.size(ldelta) == 0 else \
        np.broadcast_to(np.asarray(ldelta, np.float64).ravel(), (n,))
    bd = np.ones(n) if BDELTA is None or np.size(BDELTA) == 0 else \
        np.broadcast_to(np.asarray(BDELTA, np.float64).ravel(), (n,))
    te = np.zeros(n) if te is None or np.size(te) == 0 else np.broadcast_to(np.asarray(te, np.float64).ravel(), (n,))
    g = _bvec_rows(bvec)
    nsl = int(lmax) // 2 + 1
    blocks, b_out, ld_out, bd_out, te_out = [], [], [], [], []
    te_u = np.unique(te)
    for t in te_u:
        it = te == t
        i0 = it & (bval == 0)
        if np.any(i0):
            blk = np.zeros(dims + (1, nsl))
            blk[..., 0, 0] = dwi[..., i0].mean(axis=3)
            blocks.append(blk)
            b_out.append(np.zeros(1))
            l1 = ld[it].min()
            ld_out.append([l1])
            bd_out.append([bd[it & (ld == l1)].min()])
            te_out.append([t])
        for grp in _groups(bval[it], ld[it], bd[it]):
            idx = np.flatnonzero(it)[grp]
            idx = idx[bval[idx] > 0]
            if idx.size == 0:
                continue
            s, bu = sl_no_normalise(dwi[..., idx], bval[idx], g[idx], lmax)
            blocks.append(s.reshape(dims + (bu.size, nsl), order="F"))
            b_out.append(bu)
            ld_out.append(np.full(bu.size, ld[idx[0]]))
            bd_out.append(np.full(bu.size, bd[idx[0]]))
            te_out.append(np.full(bu.size, t))
    raw5 = np.concatenate(blocks, axis=3)                  # [x,y,z,Nshells,Nsl]
    nsh = raw5.shape[3]
    raw = raw5.reshape(dims + (nsh * nsl,), order="F")
    bl = np.concatenate(b_out)
    ldl, bdl, tel = (np.concatenate([np.ravel(a) for a in v]) for v in (ld_out, bd_out, te_out))
    isb0 = bl == 0
    nonb0 = np.flatnonzero(~isb0)
    if te_u.size == 1:
        b0 = raw[..., 0]     # the b=0 block (the first shell when there is no b=0 volume)
        idx = np.concatenate([kl * nsh + nonb0 for kl in range(nsl)])
        out = raw[..., idx] / np.maximum(b0, EPSILON)[..., None]
        rep = lambda a: np.tile(a[~isb0], nsl)
        return out, rep(bl), rep(ldl), rep(bdl), rep(tel)
    idx = np.concatenate([np.arange(nsh)] + [kl * nsh + nonb0 for kl in range(1, nsl)])
    cat = lambda a: np.concatenate([a] + [a[~isb0]] * (nsl - 1))
    return raw[..., idx], cat(bl), cat(ldl), cat(bdl), cat(tel)
