"""The seeded synthetic generator: deterministic, in range, with the promised content."""
import hashlib

import numpy as np

from anime4kcpp_amd import synth


def test_deterministic_and_seed_dependent():
    a = synth.flat_shaded(5, 64, 48)
    b = synth.flat_shaded(5, 64, 48)
    c = synth.flat_shaded(6, 64, 48)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.dtype == np.float32 and a.min() >= 0 and a.max() <= 1


def test_u8_is_rounded_float():
    f = synth.flat_shaded(9, 100, 60)
    u = synth.flat_shaded(9, 100, 60, dtype=np.uint8)
    assert np.array_equal(u, (f * np.float32(255) + np.float32(0.5)).astype(np.uint8))


def test_content_is_flat_shaded():
    img = synth.flat_shaded(1, 256, 256)
    # mostly flat regions with sigma 0.01 noise: the median local gradient is at noise level
    gx = np.abs(np.diff(img, axis=1))
    assert np.median(gx) < 0.03
    # but there are edges (polygons, strokes)
    assert (gx > 0.1).mean() > 0.001
    # distinct levels (background + polygons)
    hist = np.histogram(img, bins=50)[0]
    assert (hist > 100).sum() >= 3


def test_rgb_and_helpers():
    rgb = synth.flat_shaded(3, 40, 30, channels=3)
    assert rgb.shape == (30, 40, 3)
    r = synth.randint(7, 64, 128, 2048)
    assert r.min() >= 128 and r.max() <= 2048 and np.array_equal(r, synth.randint(7, 64, 128, 2048))
    u = synth.uniform(8, 1000)
    assert 0 <= u.min() and u.max() < 1
