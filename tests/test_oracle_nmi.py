"""Oracle of the mutual-information path (patchSize 0): pinned by brute-force histograms
and closed forms of the normalized mutual information."""
import numpy as np

from oracle import oracle as O


def _bins(v, bins):
    v = v.astype(np.float32)
    mn, mx = v.min(), v.max()
    t = (v - mn).astype(np.float64) * (bins - 1.0) / np.float64(mx - mn)
    t32 = t.astype(np.float32).astype(np.float64)        # t >= 0
    return np.floor(t32 + 0.5).astype(int)               # roundf: halves away from zero


def test_histograms_match_brute_force():
    rng = np.random.default_rng(0)
    U = rng.integers(0, 256, (20, 30, 40)).astype(np.float32)   # 255 levels on 64 bins: halves
    U.flat[0], U.flat[1] = 0.0, 255.0
    I = (rng.normal(0, 100, (20, 30, 40)) ** 2).astype(np.float32)
    v, hu, hi, hj = O.nmi(U, I, 64, 32, hists=True)
    bu, bi = _bins(U.ravel(), 64), _bins(I.ravel(), 32)
    assert np.array_equal(hu, np.bincount(bu, minlength=64))
    assert np.array_equal(hi, np.bincount(bi, minlength=32))
    ref = np.zeros((32, 64), int)
    np.add.at(ref, (bi, bu), 1)
    assert np.array_equal(hj, ref)
    n = U.size
    H = lambda h: -sum(c / n * np.log(c / n) for c in h.ravel() if c)
    assert abs(O.nmi(U, I, 64, 32, fp32=False) - (H(hu) + H(hi)) / H(hj)) < 1e-12


def test_identical_images_give_two():
    # k equally filled bins: H_U = H_I = H_joint = ln k, so NMI = 2
    U = np.repeat(np.arange(16, dtype=np.float32), 100)
    assert abs(O.nmi(U, U, 16, 16, fp32=False) - 2.0) < 1e-12
    assert abs(O.nmi(U, U, 16, 16, fp32=True) - 2.0) < 1e-6


def test_independent_images_give_one():
    # every (u, i) pair equally often: H_joint = H_U + H_I, so NMI = 1
    u, i = np.meshgrid(np.arange(8), np.arange(5), indexing="ij")
    U = np.tile(u.ravel(), 7).astype(np.float32)
    I = np.tile(i.ravel(), 7).astype(np.float32)
    assert abs(O.nmi(U, I, 8, 5, fp32=False) - 1.0) < 1e-12


def test_nmi_invariant_to_affine_intensity_maps_and_constant_is_degenerate():
    rng = np.random.default_rng(1)
    U = rng.integers(0, 50, 5000).astype(np.float32)
    I = (U + rng.integers(0, 3, 5000)).astype(np.float32)
    a = O.nmi(U, I, 50, 52, fp32=False)
    b = O.nmi(3.0 * U + 7.0, I - 100.0, 50, 52, fp32=False)
    assert abs(a - b) < 1e-12 and 1.0 < a < 2.0
    assert O.nmi(np.ones(100, np.float32), I[:100], 8, 8) == 0.0
