"""CT precision workload (host side): the int16 Hounsfield volume it synthesises, and the
fact that makes half storage of such a volume lossless."""
import numpy as np

from lc2registration_amd import synth


def test_every_integer_in_hu_range_is_exact_in_half():
    v = np.arange(-2048, 2049, dtype=np.int32)
    assert np.array_equal(v.astype(np.float16).astype(np.int32), v)
    # the first integers that are not: half has an 11-bit significand
    assert int(np.float16(2049)) != 2049 and int(np.float16(-2049)) != -2049


def test_ct_workload_shape_and_ranges():
    w = synth.config3(ct_dims=(64, 64, 56), ct_spacing=2.0, us_dims=(40, 40, 36), us_spacing=1.5)
    ct, us = w.moving.data, w.fixed.data
    assert ct.dtype == np.int16 and ct.shape == (56, 64, 64)
    assert us.dtype == np.uint8 and us.shape == (36, 40, 40)
    assert ct[0, 0, 0] < -900                      # air outside the body
    assert np.median(ct[28, 24:40, 24:40]) > -100  # soft tissue inside
    assert ct.max() > 1000                         # bone
    assert -2048 <= ct.min() and ct.max() <= 2048  # lossless in half storage
    assert w.high_res_is_moving and w.ndim == 3 and len(w.true_pose) == 12
    w2 = synth.config3(ct_dims=(64, 64, 56), ct_spacing=2.0, us_dims=(40, 40, 36), us_spacing=1.5)
    assert np.array_equal(w2.moving.data, ct) and np.array_equal(w2.fixed.data, us)  # seeded
