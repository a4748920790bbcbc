"""Pins the oracle's Sobel and LC2 (CPU only): closed forms, numpy/scipy brute
force on tiny grids, known answers and invariances."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as O


# ---- Sobel (cu_lc2_3d.cu:100-179, cu_lc2_2d.cu:25-86) ----

def _sobel_brute(I):
    h, hp = np.array([1.0, 2.0, 1.0]), np.array([-1.0, 0.0, 1.0])
    I = I.astype(np.float64)
    if I.ndim == 2:
        ks = [np.outer(h, hp), np.outer(hp, h)]
    else:
        ks = [np.einsum("i,j,k->ijk", h, h, hp), np.einsum("i,j,k->ijk", h, hp, h),
              np.einsum("i,j,k->ijk", hp, h, h)]
    g = [ndimage.correlate(I, k, mode="constant", cval=0.0) for k in ks]
    return np.sqrt(sum(x * x for x in g))


@pytest.mark.parametrize("shape", [(21, 17), (9, 11, 13)])
@pytest.mark.parametrize("order", ["ref", "sep"])
def test_sobel_matches_brute_force(shape, order):
    I = np.random.default_rng(1).normal(size=shape).astype(np.float32) * 100
    G = O.sobel(I, order)
    ref = _sobel_brute(I)
    assert np.allclose(G, ref, rtol=1e-5, atol=1e-4)


def test_sobel_orders_agree_to_rounding():
    I = np.random.default_rng(2).uniform(0, 1000, size=(10, 12, 14)).astype(np.float32)
    a, b = O.sobel(I, "ref"), O.sobel(I, "sep")
    assert np.allclose(a, b, rtol=2e-6, atol=1e-3)


def test_sobel_ramp_closed_form():
    z, y, x = np.mgrid[0:8, 0:9, 0:10].astype(np.float32)
    G = O.sobel(0.5 * x + 0.25 * y, "sep")
    assert np.allclose(G[1:-1, 1:-1, 1:-1], 32 * np.hypot(0.5, 0.25), rtol=1e-6)
    G2 = O.sobel((0.5 * x)[0], "ref")
    assert np.allclose(G2[1:-1, 1:-1], 8 * 0.5, rtol=1e-6)


# ---- LC2 (cu_lc2_3d.cu:181-618, :1119-1194; cu_lc2_2d.cu:88-480, :599-653) ----

def _lc2_brute(U, I, G, r, cols=("I", "G")):
    """Per patch: weight from the US std (ddof 1) with the reference's thresholds,
    numpy lstsq of U on the chosen columns + 1, residual variance ddof 1."""
    U64, I64, G64 = (a.astype(np.float64) for a in (U, I, G))
    nd = U.ndim
    sm = sw = 0.0
    wmap = np.zeros(U.shape)
    mmap = np.zeros(U.shape)
    it = np.ndindex(*U.shape)
    for idx in it:
        if any(i < r or i + r >= n for i, n in zip(idx, U.shape)):
            continue
        sl = tuple(slice(i - r, i + r + 1) for i in idx)
        u = U64[sl].ravel()
        if not np.all(u > 1e-6):
            continue
        var = u.var(ddof=1)
        if not var > 1e-6:
            continue
        w = np.sqrt(var)
        A = [np.ones_like(u)]
        if "I" in cols:
            A.append(I64[sl].ravel())
        if "G" in cols:
            A.append(G64[sl].ravel())
        A = np.stack(A, 1)
        x, *_ = np.linalg.lstsq(A, u, rcond=None)
        res = u - A @ x
        m = w - res.var(ddof=1) / w
        m = m if (np.isfinite(m) and m > 0) else 0.0
        wmap[idx], mmap[idx] = w, m
        sm += m
        sw += w
    return sm / sw if sw > 0 else 0.0, wmap, mmap


def _random_case(shape, seed):
    rng = np.random.default_rng(seed)
    I = rng.uniform(0.1, 1.0, size=shape).astype(np.float32)
    I = ndimage.uniform_filter(I, 2).astype(np.float32)
    G = O.sobel(I, "sep")
    U = (40 * I + 3 * G + rng.normal(0, 5, size=shape) + 60).astype(np.float32)
    U[tuple(rng.integers(0, s) for s in shape)] = 0.0  # one dead US sample
    return U, I, G


@pytest.mark.parametrize("shape,r", [((14, 13), 2), ((16, 15), 3), ((9, 10, 11), 1),
                                     ((11, 10, 12), 2)])
def test_lc2_exact_matches_lstsq(shape, r):
    U, I, G = _random_case(shape, seed=sum(shape) + r)
    res, wm, mm = O.lc2_exact(U, I, G, r, maps=True)
    l_ref, w_ref, m_ref = _lc2_brute(U, I, G, r)
    assert res.n_valid > 10
    assert np.allclose(wm, w_ref, rtol=1e-12, atol=1e-12)
    assert np.allclose(mm, m_ref, rtol=1e-9, atol=1e-10)
    assert abs(res.lc2 - l_ref) < 1e-12


def test_lc2_brute_force_detects_dropped_column():
    """Planted mistake: fitting without the gradient column changes LC2 by far more
    than the parity tolerance, so the lstsq comparison would catch such a slip."""
    U, I, G = _random_case((16, 15), 7)
    good = O.lc2_exact(U, I, G, 3).lc2
    dropped, _, _ = _lc2_brute(U, I, G, 3, cols=("I",))
    assert abs(good - dropped) > 1e-3


def test_lc2_known_answers():
    rng = np.random.default_rng(4)
    I = rng.uniform(0.2, 1.0, size=(12, 13, 14)).astype(np.float32)
    G = O.sobel(I, "sep")
    U = (2.0 * I.astype(np.float64) + 0.5 * G + 1.0).astype(np.float32)
    assert abs(O.lc2_exact(U, I, G, 2).lc2 - 1.0) < 1e-6            # perfect fit
    const = np.full_like(U, 5.0)
    r0 = O.lc2_exact(const, I, G, 2)                                 # var 0 -> w = 0
    assert r0.lc2 == 0.0 and r0.sum_w == 0.0 and r0.n_valid == 0
    Un = (U + rng.normal(0, 0.3, size=U.shape)).astype(np.float32)
    base = O.lc2_exact(Un, I, G, 2).lc2
    scaled = O.lc2_exact((3.0 * Un.astype(np.float64) + 7.0).astype(np.float32), I, G, 2).lc2
    assert abs(base - scaled) < 1e-5                                 # invariant to kU + m
    Iz = (I.astype(np.float64) * 1000 - 50).astype(np.float32)       # and to the scale of I
    assert abs(O.lc2_exact(Un, Iz, G, 2).lc2 - base) < 1e-6


def test_lc2_border_and_zero_exclusion():
    U, I, G = _random_case((20, 21), 9)
    U[10, 10] = 0.0
    _, wm, _ = O.lc2_exact(U, I, G, 3, maps=True)
    assert np.all(wm[:3] == 0) and np.all(wm[-3:] == 0) and np.all(wm[:, :3] == 0)
    assert np.all(wm[7:14, 7:14] == 0)
    assert np.count_nonzero(wm) > 0


def test_cg_fixed_point_is_exact_and_fp32_deviates():
    """The reference's CGLS run in FP64 lands on the exact projection; the FP32
    run (the reference's arithmetic) deviates at the 1e-4..1e-2 level."""
    U, I, G = _random_case((30, 31), 11)
    ex = O.lc2_exact(U, I, G, 3).lc2
    cg64 = O.lc2_cg(U, I, G, 3, fp64=True).lc2
    cg32 = O.lc2_cg(U, I, G, 3).lc2
    assert abs(cg64 - ex) < 1e-7
    assert abs(cg32 - ex) < 5e-2


def test_zrange_partition_sums():
    U, I, G = _random_case((12, 11, 13), 5)
    full = O.lc2_exact(U, I, G, 2)
    a = O.lc2_exact(U, I, G, 2, zlo=0, zhi=6)
    b = O.lc2_exact(U, I, G, 2, zlo=6, zhi=12)
    assert abs(a.sum_measure + b.sum_measure - full.sum_measure) < 1e-9
    assert a.n_valid + b.n_valid == full.n_valid
