"""Writes the golden fixtures of the 2x passes: seeded inputs and the outputs of the
reference's own CPU kernels (oracle/_ref: the OpImpl conv/dot kernels of
core/internal/AC/Core/Internal/Processor/CPU/{Generic,X86/AVX}.hpp driven through the layer
sequences of CPUProcessor.cpp, compiled from the reference sources by oracle/Makefile) in the
Generic order (arch 1) and the FMA order (arch 4).

    make -C oracle ref && python tests/golden/make_golden.py          # per-family fixtures
    make -C oracle ref && python tests/golden/make_golden.py --c1     # the 256x256 workload

One .npz per family (numpy, no pickles): for every case "<variant>/<w>x<h>/<dtype>" the
arrays "<case>/in", "<case>/generic" and "<case>/fma".

acnet_legacy_c1.npz holds the 256x256 luma workload of the plain 2x pass: one seeded f32
input ("input/float32") and its u8 quantisation ("input/uint8"), and for every ACNetLegacy
weight set the reference outputs "<variant>/<in>-<out>/{generic,fma}" for f32 -> u8 (the
u8-quantised check of that workload) and u8 -> u8, plus f32 -> f32 for HDN0.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as OL  # noqa: E402
from anime4kcpp_amd import hip, synth  # noqa: E402

SIZES = {
    "acnet_legacy": [((1, 1), np.uint8), ((2, 3), np.uint8), ((17, 13), np.uint8), ((64, 64), np.uint8),
                     ((17, 13), np.float32), ((9, 7), np.uint16)],
    "acnet8": [((3, 2), np.uint16), ((17, 13), np.uint8), ((13, 9), np.float32)],
    "artcnn": [((1, 1), np.uint8), ((17, 13), np.uint8), ((11, 7), np.float32)],
    "fsrcnnx": [((2, 5), np.uint8), ((17, 13), np.uint8), ((11, 7), np.float32)],
}


def cases(family):
    if family == "acnet_legacy":
        for v in hip.VARIANTS:
            k, b = hip.load_weights(v)
            yield v, lambda arch, x, k=k, b=b: OL.ref_pass(arch, x, k, b)
    elif family == "acnet8":
        for v in hip.VARIANTS8:
            k, b, a, blocks = hip.load_weights8(v)
            yield v, lambda arch, x, k=k, b=b, a=a, n=blocks: OL.ref_pass8(arch, x, k, b, a, n)
    elif family == "artcnn":
        for v in hip.VARIANTS_ART:
            k, b, f, blocks = hip.load_weights_artcnn(v)
            yield v, lambda arch, x, k=k, b=b, f=f, n=blocks: OL.ref_artcnn(arch, x, k, b, f, n)
    else:
        for v in hip.VARIANTS_FSR:
            k, b, a, f, blocks = hip.load_weights_fsrcnnx(v)
            yield v, lambda arch, x, k=k, b=b, a=a, f=f, n=blocks: OL.ref_fsrcnnx(arch, x, k, b, a, f, n)


C1_SIZE = 256
C1_SEED = 256256


def c1_arrays():
    x = synth.flat_shaded(C1_SEED, C1_SIZE, C1_SIZE, dtype=np.float32)
    x8 = synth.flat_shaded(C1_SEED, C1_SIZE, C1_SIZE, dtype=np.uint8)
    arrays = {"input/float32": x, "input/uint8": x8}
    for v in hip.VARIANTS:
        k, b = hip.load_weights(v)
        combos = [(x, np.uint8), (x8, np.uint8)] + ([(x, np.float32)] if v == "hdn0" else [])
        for src, out in combos:
            key = f"{v}/{src.dtype.name}-{np.dtype(out).name}"
            arrays[key + "/generic"] = OL.ref_pass(1, src, k, b, out_dtype=out)
            arrays[key + "/fma"] = OL.ref_pass(4, src, k, b, out_dtype=out)
    return arrays


def main():
    if not OL.ref_available():
        raise SystemExit("oracle/_ref is not built (make -C oracle ref needs the reference sources)")
    if "--c1" in sys.argv:
        path = os.path.join(HERE, "acnet_legacy_c1.npz")
        np.savez_compressed(path, **c1_arrays())
        print(f"{path}: {os.path.getsize(path)} bytes")
        return
    for family, sizes in SIZES.items():
        arrays = {}
        for variant, run in cases(family):
            for (w, h), dt in sizes:
                x = synth.flat_shaded(1000 + 31 * w + h, w, h, dtype=dt)
                key = f"{variant}/{w}x{h}/{np.dtype(dt).name}"
                arrays[key + "/in"] = x
                arrays[key + "/generic"] = run(1, x)
                arrays[key + "/fma"] = run(4, x)
        path = os.path.join(HERE, f"{family}.npz")
        np.savez_compressed(path, **arrays)
        print(f"{path}: {len(arrays) // 3} cases, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
