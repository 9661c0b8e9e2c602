"""The committed weight blobs: sizes, provenance hashes and layout invariants
(core/include/AC/Core/Model/ACNet.hpp:34-60, tests/core/src/ModelTest.cpp:5-42)."""
import hashlib
import json
import os

import numpy as np
import pytest

from anime4kcpp_amd.hip import WEIGHTS_DIR, VARIANTS, load_weights

# sha256 prefixes of the packed little-endian f32 arrays, computed when the arrays
# were parsed out of core/internal/AC/Core/Internal/Model/Param/ACNet.p
PINNED = {
    "gan": ("1020739c30bff11c", "94c802cf3220023e"),
    "hdn0": ("df5d8415d83b8e81", "51e1b641911203e4"),
    "hdn1": ("2c673f884aed210c", "169646e8594c45b7"),
    "hdn2": ("9581e07f015377b6", "1be528fca18b3b91"),
    "hdn3": ("3646418ad22e71f2", "163016b5985906fa"),
}


@pytest.mark.parametrize("variant", VARIANTS)
def test_blob_hashes(variant):
    k, b = load_weights(variant)
    assert k.size == 8 * 9 + 8 * 8 * 9 * 8 + 8 * 4 == 4712
    assert b.size == 8 + 8 * 8 == 72
    assert hashlib.sha256(k.astype("<f4").tobytes()).hexdigest().startswith(PINNED[variant][0])
    assert hashlib.sha256(b.astype("<f4").tobytes()).hexdigest().startswith(PINNED[variant][1])
    man = json.load(open(os.path.join(WEIGHTS_DIR, "MANIFEST.json")))
    assert man["variants"][variant]["kernels"]["sha256"].startswith(PINNED[variant][0])


def test_variants_differ_and_are_finite():
    ks = [load_weights(v)[0] for v in VARIANTS]
    for i in range(len(ks)):
        assert np.isfinite(ks[i]).all()
        for j in range(i):
            assert not np.array_equal(ks[i], ks[j])
