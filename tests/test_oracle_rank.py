"""Pins the rank-deficient branch of the oracle's least-squares projection (CPU only).

The LC2 fit regresses each patch of U on [I, G, 1]. Where the centred I or G columns
vanish (plateaus, constant-|grad I| ramps) or are collinear, the normal equations are
singular and the projection is defined by a rank rule. The oracle's rule
(lc2o_lc2_exact, explained_var) is, per patch, with tau = 1e-10:
  * a column is constant (dropped) when its centred energy <= tau * its raw energy;
  * the two remaining columns are collinear when det <= tau * C_II * C_GG, i.e. when
    1 - rho^2 <= tau; then U is projected onto the single direction
    I/|I| + sign(rho) G/|G| (centred, unit columns).
That is numpy's truncated-SVD least squares on the centred, unit-normalised columns with
rcond = sqrt(tau)/2 = 5e-6 (for unit columns sigma_min/sigma_max ~ sqrt((1 - |rho|)/2)),
which these tests compute independently with numpy.linalg.lstsq and compare per voxel.
A mistake in the branch (a dropped term, the sign of rho, the wrong scale of "zero")
moves the measure by far more than the tolerance on at least one of these cases; the
sign test below shows it for the collinear cases.
"""
import numpy as np
import pytest

from oracle import oracle as O

TAU = 1e-10
RCOND = 0.5 * np.sqrt(TAU)


def _lstsq_rule(U, I, G, r, rcond=RCOND, flip_sign=False):
    """LC2 and per-voxel measure with the rank rule stated in the module docstring."""
    U64, I64, G64 = (a.astype(np.float64) for a in (U, I, G))
    sm = sw = 0.0
    mmap = np.zeros(U.shape)
    for idx in np.ndindex(*U.shape):
        if any(i < r or i + r >= n for i, n in zip(idx, U.shape)):
            continue
        sl = tuple(slice(i - r, i + r + 1) for i in idx)
        u = U64[sl].ravel()
        N = u.size
        if not np.all(u > 1e-6):
            continue
        uc = u - u.mean()
        cyy = uc @ uc
        if not cyy / (N - 1) > 1e-6:
            continue
        w = np.sqrt(cyy / (N - 1))
        cols = []
        for c in (I64[sl].ravel(), G64[sl].ravel()):
            cc = c - c.mean()
            if cc @ cc > TAU * (c @ c):
                cols.append(cc / np.sqrt(cc @ cc))
        e = 0.0
        if cols:
            A = np.stack(cols, 1)
            if flip_sign and A.shape[1] == 2:
                s = np.linalg.svd(A, compute_uv=False)
                if s[-1] <= rcond * s[0]:
                    # project on the other diagonal direction, I' - sign(rho) G'
                    d = A[:, 0] - np.sign(A[:, 0] @ A[:, 1]) * A[:, 1]
                    d /= np.linalg.norm(d)
                    e = (d @ uc) ** 2
                    A = None
            if A is not None:
                x, *_ = np.linalg.lstsq(A, uc, rcond=rcond)
                fit = A @ x
                e = fit @ fit
        e = min(max(e, 0.0), cyy)
        m = w - ((cyy - e) / (N - 1)) / w
        m = m if (np.isfinite(m) and m > 0) else 0.0
        mmap[idx] = m
        sm += m
        sw += w
    return (sm / sw if sw > 0 else 0.0), mmap


def _us(shape, seed, lo=10.0, hi=200.0):
    return np.random.default_rng(seed).uniform(lo, hi, size=shape).astype(np.float32)


def _label_map(shape, seed, values=(-1000.0, 40.0, 1000.0), noise=0.0):
    """CT-like piecewise-constant map: three classes from random half-planes (plateaus)."""
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    lab = np.zeros(shape, np.int32)
    for k in (1, 2):
        nrm = rng.normal(size=len(shape))
        off = rng.uniform(0.3, 0.7) * sum(n * abs(c) for n, c in zip(shape, nrm))
        lab[sum(c * g for c, g in zip(nrm, grids)) + 0.5 * off > off] = k
    I = np.asarray(values, np.float64)[lab]
    if noise:
        I = I + rng.normal(0.0, noise, size=shape)
    return I.astype(np.float32)


def _cases():
    y2, x2 = np.mgrid[0:34, 0:36].astype(np.float32)
    z3, y3, x3 = np.mgrid[0:13, 0:14, 0:15].astype(np.float32)
    rng = np.random.default_rng(7)
    Irnd = rng.uniform(0, 50, size=(34, 36)).astype(np.float32)
    return {
        # plateaus: interior patches have constant I and G == 0 (both columns dropped)
        "plateau_2d": (_us((34, 36), 1), _label_map((34, 36), 1), None, 3),
        "plateau_3d": (_us((13, 14, 15), 2), _label_map((13, 14, 15), 2), None, 2),
        # the same with sigma = 20 HU noise: full rank almost everywhere
        "plateau_noise_2d": (_us((34, 36), 3), _label_map((34, 36), 3, noise=20.0), None, 3),
        # linear ramps: |Sobel| constant in the interior (G dropped, I kept)
        "ramp_2d": (_us((34, 36), 4), 3.0 * x2 + 2.0 * y2, None, 3),
        "ramp_3d": (_us((13, 14, 15), 5), 2.0 * x3 - 1.0 * y3 + 0.5 * z3, None, 2),
        # collinear columns given directly (G is an input of the projection)
        "collinear_pos": (_us((34, 36), 6), Irnd, 2.0 * Irnd + 5.0, 3),
        "collinear_neg": (_us((34, 36), 8), Irnd, -3.0 * Irnd + 100.0, 3),
    }


@pytest.mark.parametrize("name", list(_cases()))
def test_rank_rule_matches_truncated_lstsq(name):
    U, I, G, r = _cases()[name]
    G = O.sobel(I, "sep") if G is None else G.astype(np.float32)
    res, wm, mm = O.lc2_exact(U, I, G, r, maps=True)
    ref, rm = _lstsq_rule(U, I, G, r)
    assert res.n_valid > 50
    assert np.allclose(mm, rm, rtol=1e-9, atol=1e-9), np.abs(mm - rm).max()
    assert abs(res.lc2 - ref) < 1e-12, (res.lc2, ref)


def test_plateau_interior_measure_is_zero():
    """Constant I and G == 0 explain nothing: measure = w - var(U)/w = 0."""
    U, I, _, r = _cases()["plateau_2d"]
    G = O.sobel(I, "sep")
    _, wm, mm = O.lc2_exact(U, I, G, r, maps=True)
    flat = np.zeros(I.shape, bool)
    for idx in np.ndindex(*I.shape):
        if all(r + 1 <= i < n - r - 1 for i, n in zip(idx, I.shape)):
            sl = tuple(slice(i - r - 1, i + r + 2) for i in idx)
            flat[idx] = np.ptp(I[sl]) == 0
    assert flat.sum() > 50 and (wm[flat] > 0).all()
    assert np.all(mm[flat] < 1e-9 * wm[flat])


@pytest.mark.parametrize("name", ["collinear_pos", "collinear_neg"])
def test_collinear_sign_matters(name):
    """Projecting a collinear patch on the other diagonal direction changes the measure
    by far more than the tolerance, so the test above pins the sign rule."""
    U, I, G, r = _cases()[name]
    _, mm = _lstsq_rule(U, I, G, r)
    _, bad = _lstsq_rule(U, I, G, r, flip_sign=True)
    assert np.abs(bad - mm).max() > 1e-3


def test_full_rank_agrees_with_plain_lstsq():
    """Where the columns are well conditioned the cutoff plays no part: the rule equals
    numpy's default lstsq (rcond = None) on the same centred columns."""
    U, I, _, r = _cases()["plateau_noise_2d"]
    G = O.sobel(I, "sep")
    a, ma = _lstsq_rule(U, I, G, r)
    b, mb = _lstsq_rule(U, I, G, r, rcond=None)
    res = O.lc2_exact(U, I, G, r)
    assert abs(a - b) < 1e-12 and abs(res.lc2 - b) < 1e-12


def test_near_collinear_sensitivity_to_cutoff():
    """G = fp32(-3 I + 100): collinear up to the rounding of G. The cutoff decides whether
    that rounding residue is fitted; numpy's default cutoff (machine epsilon) fits it and
    moves LC2 by 2e-2 on this random data (0.0241 -> 0.0447). The oracle and the kernels
    use the stated tau."""
    U, I, G, r = _cases()["collinear_neg"]
    res = O.lc2_exact(U, I, G, r)
    ref, _ = _lstsq_rule(U, I, G, r)
    loose, _ = _lstsq_rule(U, I, G, r, rcond=None)
    assert abs(res.lc2 - ref) < 1e-12
    assert abs(loose - ref) > 1e-5  # documented sensitivity: the rule is a definition
