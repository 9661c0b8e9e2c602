"""The committed golden fixtures (tests/golden/*.npz: seeded inputs and the outputs of the
reference's own CPU kernels in the Generic and FMA orders, written by
tests/golden/make_golden.py) against the oracle, bit for bit, on the CPU; and against the HIP
paths on the device at the parity bar (u8/u16 within 1 LSB on >= 99.9% of samples, max 2;
f32 within 1e-5)."""
import os

import numpy as np
import pytest

import oracle_lib as OL
from anime4kcpp_amd import hip

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FAMILIES = ("acnet_legacy", "acnet8", "artcnn", "fsrcnnx")


def load(family):
    with np.load(os.path.join(GOLDEN, f"{family}.npz"), allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    cases = sorted({k.rsplit("/", 1)[0] for k in data})
    return [(c, data[c + "/in"], data[c + "/generic"], data[c + "/fma"]) for c in cases]


def oracle(family, variant, x, order):
    if family == "acnet_legacy":
        k, b = hip.load_weights(variant)
        return OL.oracle_pass(x, k, b, order=order)
    if family == "acnet8":
        k, b, a, n = hip.load_weights8(variant)
        return OL.oracle_pass8(x, k, b, a, n, order=order)
    if family == "artcnn":
        k, b, f, n = hip.load_weights_artcnn(variant)
        return OL.oracle_artcnn(x, k, b, f, n, order=order)
    k, b, a, f, n = hip.load_weights_fsrcnnx(variant)
    return OL.oracle_fsrcnnx(x, k, b, a, f, n, order=order)


@pytest.mark.parametrize("family", FAMILIES)
def test_oracle_equals_golden(family):
    cases = load(family)
    assert len(cases) >= 12
    orders_differ = False
    for case, x, gen, fma in cases:
        variant = case.split("/")[0]
        assert np.array_equal(oracle(family, variant, x, OL.ORDER_GENERIC), gen), (case, "generic")
        assert np.array_equal(oracle(family, variant, x, OL.ORDER_FMA), fma), (case, "fma")
        orders_differ |= not np.array_equal(gen, fma)
    # the two summation orders are distinguishable on these inputs, so each pin is its own
    assert orders_differ


def test_fixture_shapes():
    for family in FAMILIES:
        for case, x, gen, fma in load(family):
            w, h = (int(v) for v in case.split("/")[1].split("x"))
            assert x.shape == (h, w) and gen.shape == fma.shape == (2 * h, 2 * w)
            assert x.dtype == gen.dtype == fma.dtype == np.dtype(case.split("/")[2])


@pytest.mark.gpu
@pytest.mark.parametrize("family", FAMILIES)
def test_hip_matches_golden(cuda, family):
    from test_gpu_fp32 import check_int, to_dev, to_host
    nets = {}
    for case, x, gen, fma in load(family):
        variant = case.split("/")[0]
        if variant not in nets:
            nets[variant] = {"acnet_legacy": hip.Net, "acnet8": hip.Net8, "artcnn": hip.NetArt,
                             "fsrcnnx": hip.NetFsr}[family](variant, 0)
        got = to_host(nets[variant].upscale(to_dev(x, "cuda")))
        if x.dtype == np.float32:
            assert np.abs(got - fma).max() <= 1e-5, case
        else:
            check_int(got, fma)
    for n in nets.values():
        n.close()
