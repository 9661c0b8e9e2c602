"""The committed golden fixtures (tests/golden/*.npz: seeded inputs and the outputs of the
reference's own CPU kernels in the Generic and FMA orders, written by
tests/golden/make_golden.py) against the oracle, bit for bit, on the CPU; and against the HIP
paths on the device at the parity bar (u8/u16 within 1 LSB on >= 99.9% of samples, max 2;
f32 within 1e-5).  acnet_legacy_c1.npz is the 256x256 luma workload of the plain pass (f32 and
u8 input, u8 output for every weight set; f32 output for HDN0)."""
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


def load_c1():
    with np.load(os.path.join(GOLDEN, "acnet_legacy_c1.npz"), allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    inputs = {"float32": data.pop("input/float32"), "uint8": data.pop("input/uint8")}
    cases = sorted({k.rsplit("/", 1)[0] for k in data})
    out = []
    for c in cases:
        variant, types = c.split("/")
        tin, tout = types.split("-")
        out.append((c, variant, inputs[tin], data[c + "/generic"], data[c + "/fma"]))
        assert data[c + "/fma"].dtype == np.dtype(tout)
    return out


def test_c1_oracle_equals_golden():
    cases = load_c1()
    assert {c[1] for c in cases} == set(hip.VARIANTS) and len(cases) == 11
    for case, variant, x, gen, fma in cases:
        assert x.shape == (256, 256) and gen.shape == fma.shape == (512, 512)
        k, b = hip.load_weights(variant)
        assert np.array_equal(OL.oracle_pass(x, k, b, out_dtype=fma.dtype, order=OL.ORDER_GENERIC), gen), case
        assert np.array_equal(OL.oracle_pass(x, k, b, out_dtype=fma.dtype, order=OL.ORDER_FMA), fma), case


@pytest.mark.gpu
def test_c1_hip_matches_golden(cuda):
    """Every FP32 configuration of the fused kernel on the 256x256 workload, against the
    reference kernels' own output in both summation orders."""
    import torch
    from test_gpu_fp32 import check_int, to_dev, to_host
    nets = {}
    ncfg = len(hip.configs("fp32"))
    for case, variant, x, gen, fma in load_c1():
        net = nets.setdefault(variant, hip.Net(variant, 0))
        out_dtype = {np.dtype(np.uint8): torch.uint8, np.dtype(np.float32): torch.float32}[fma.dtype]
        for cfg in range(ncfg):
            got = to_host(net.upscale(to_dev(x, "cuda"), config=cfg, out_dtype=out_dtype))
            for want in (fma, gen):
                if fma.dtype == np.float32:
                    assert np.abs(got - want).max() <= 1e-5, (case, cfg)
                else:
                    check_int(got, want)
    for n in nets.values():
        n.close()
