"""FP16 / BF16 MFMA forms of the fused pass: layers 1..8 as a 16-bit implicit GEMM with
FP32 accumulation.  They are not bit-compatible with the reference by design; every
case reports max/mean absolute error and PSNR against the FP64 restatement of the same
network and holds them to these bounds (f32 output, values in [0,1]):
  FP16: max abs <= 4e-3, mean abs <= 4e-4
  BF16: max abs <= 3e-2, mean abs <= 4e-3
and u8 output to >= 99 % (FP16) / >= 90 % (BF16) of pixels within 1 LSB of the
reference FP32 order."""
import numpy as np
import pytest

from anime4kcpp_amd import hip, synth
from oracle_lib import oracle_pass, oracle_pass_f64, to_float
from test_gpu_fp32 import gpu_pass

pytestmark = pytest.mark.gpu

BOUNDS = {"fp16": (4e-3, 4e-4, 0.99), "bf16": (3e-2, 4e-3, 0.90)}


@pytest.fixture(scope="module")
def nets(cuda):
    out = {v: hip.Net(v, 0) for v in hip.VARIANTS}
    yield out
    for n in out.values():
        n.close()


def err_report(got, want64):
    e = np.abs(got.astype(np.float64) - want64)
    mse = (e ** 2).mean()
    psnr = 10 * np.log10(1.0 / mse) if mse > 0 else float("inf")
    return e.max(), e.mean(), psnr


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
@pytest.mark.parametrize("variant", hip.VARIANTS)
@pytest.mark.parametrize("w,h", [(1, 1), (3, 2), (17, 13), (64, 64), (100, 37), (130, 131)])
def test_error_vs_fp64(nets, weights, prec, variant, w, h):
    src = synth.flat_shaded(300 + w + h, w, h)
    k, b = weights[variant]
    got = gpu_pass(nets[variant], src, np.float32, precision=prec)
    mx, mean, psnr = err_report(got, oracle_pass_f64(to_float(src), k, b))
    print(f"{prec} {variant} {w}x{h}: max {mx:.3g} mean {mean:.3g} PSNR {psnr:.1f} dB")
    assert mx <= BOUNDS[prec][0] and mean <= BOUNDS[prec][1]


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
@pytest.mark.parametrize("config", [0, 1])
def test_configs_agree(nets, weights, prec, config):
    src = synth.flat_shaded(9, 150, 97)
    k, b = weights["hdn1"]
    got = gpu_pass(nets["hdn1"], src, np.float32, config=config, precision=prec)
    mx, mean, _ = err_report(got, oracle_pass_f64(to_float(src), k, b))
    assert mx <= BOUNDS[prec][0] and mean <= BOUNDS[prec][1]


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_1080p_u8(nets, weights, prec):
    src = synth.flat_shaded(2, 1920, 1080, dtype=np.uint8)
    k, b = weights["hdn0"]
    got = gpu_pass(nets["hdn0"], src, np.uint8, precision=prec)
    want = oracle_pass(src, k, b)
    d = np.abs(got.astype(int) - want.astype(int))
    frac = (d <= 1).mean()
    print(f"{prec} 1080p u8: {frac * 100:.3f}% within 1 LSB, max {d.max()}")
    assert frac >= BOUNDS[prec][2]


def test_fp16_is_closer_than_bf16(nets, weights):
    src = synth.flat_shaded(4, 200, 120)
    k, b = weights["gan"]
    ref = oracle_pass_f64(to_float(src), k, b)
    e16 = err_report(gpu_pass(nets["gan"], src, np.float32, precision="fp16"), ref)[1]
    ebf = err_report(gpu_pass(nets["gan"], src, np.float32, precision="bf16"), ref)[1]
    assert e16 < ebf
