"""Extract the ACNet weight tables into little-endian f32 blobs.

Source: core/internal/AC/Core/Internal/Model/Param/ACNet.p: the five
``ACNetLegacy_{GAN,HDN0..3}_NHWC_{kernels,biases}`` constexpr arrays (bound to the
model variants in core/src/Model.cpp:8-34) and the twelve
``ACNet_F8B{4,8,18}{,_HDN,_Box,_Box_HDN}_NHWC_{kernels,biases,alphas}`` arrays of
ACNet<8> (core/src/Model.cpp:36-119); and from Param/ArtCNN.p the six
``ArtCNN_C4F{16,32}{,_DN,_DS}_NHWC_{kernels,biases}`` arrays of ArtCNN<16>/<32>
(core/src/Model.cpp:232-281); and from Param/FSRCNNX.p the four
``FSRCNNX_F{8,16}{,_DistortPlus}_NHWC_{kernels,biases,alphas}`` arrays of FSRCNNX<8>/<16>
(core/src/Model.cpp:283-326).  The text is parsed, never compiled or executed.  Output beside this script: ``acnet_legacy_<variant>.{kernels,biases}.f32``
(4712 and 72 floats), ``acnet8_<variant>.{kernels,biases,alphas}.f32`` (lengths of
core/include/AC/Core/Model/ACNet.hpp:77-82), and ``MANIFEST.json`` with the source
line ranges and sha256 of every blob.

Run once in a tree that has the reference checked out:
    python anime4kcpp_amd/weights/extract_weights.py /root/reference
"""
import hashlib
import json
import os
import re
import sys

import numpy as np

VARIANTS = ("GAN", "HDN0", "HDN1", "HDN2", "HDN3")
KERNEL_LEN = 8 * 9 + 8 * 8 * 9 * 8 + 8 * 4   # ACNet.hpp:34 with blockNum = 8
BIAS_LEN = 8 + 8 * 8                          # ACNet.hpp:36
_HEAD = re.compile(r"constexpr float ((ACNetLegacy|ACNet|ArtCNN|FSRCNNX)_(\w+?)_NHWC_(kernels|biases|alphas))\[\] = \{")
# ArtCNN variants: blob name -> (table infix, channels); 4 body blocks (C4)
ARTCNN = {f"f{f}{s.lower()}": (f"C4F{f}{s}", f) for f in (16, 32) for s in ("", "_DN", "_DS")}


def artcnn_lengths(f, blocks=4):
    """ArtCNN<F>: kernels F*9 + F*F*9*(blocks+1) + F*4*9, biases F + F*(blocks+1) + 4
    (ArtCNN.hpp:28-31)."""
    return {"kernels": f * 9 + f * f * 9 * (blocks + 1) + f * 4 * 9, "biases": f + f * (blocks + 1) + 4}
# FSRCNNX variants: blob name -> (table infix, channels); 4 blocks (B4)
FSRCNNX = {f"f{f}{s}": (f"F{f}{t}", f) for f in (8, 16) for s, t in (("", ""), ("_dp", "_DistortPlus"))}


def fsrcnnx_lengths(f, blocks=4):
    """FSRCNNX<F>: kernels F*25 + F*F*9*blocks + F*F + F*4*9, biases F + F*(blocks+1) + 4,
    alphas F*(blocks+1) (FSRCNNX.hpp:29-35)."""
    return {"kernels": f * 25 + f * f * 9 * blocks + f * f + f * 4 * 9, "biases": f + f * (blocks + 1) + 4,
            "alphas": f * (blocks + 1)}


# ACNet<8> variants: blob name -> (table infix, blocks)
ACNET8 = {f"b{b}{s.lower()}": (f"F8B{b}{s}", b) for b in (4, 8, 18) for s in ("", "_HDN", "_Box", "_Box_HDN")}


def acnet8_lengths(blocks):
    """ACNet<8> (F = 8): kernels F*9 + F*F*9*blocks + F*4*9, biases F + F*blocks + 4,
    alphas F*(blocks + 1) (ACNet.hpp:77-82)."""
    return {"kernels": 8 * 9 + 64 * 9 * blocks + 8 * 4 * 9, "biases": 8 + 8 * blocks + 4, "alphas": 8 * (blocks + 1)}
_NUM = re.compile(r"[-+]?\d+\.\d+(?:[eE][-+]?\d+)?f")


def parse(path):
    tables = {}
    with open(path, "r") as f:
        lines = f.readlines()
    i = 0
    while i < len(lines):
        m = _HEAD.search(lines[i])
        if not m:
            i += 1
            continue
        start = i + 1
        vals = []
        i += 1
        while "};" not in lines[i]:
            vals.extend(_NUM.findall(lines[i]))
            i += 1
        # decimal literal -> nearest float32, as the C++ compiler does for an
        # 'f'-suffixed literal (7 significant digits never hit a double-rounding tie)
        arr = np.array([np.float32(v[:-1]) for v in vals], dtype="<f4")
        tables[(m.group(2), m.group(3), m.group(4))] = (arr, start, i + 1)
        i += 1
    return tables


def main(ref_root):
    src = os.path.join(ref_root, "core/internal/AC/Core/Internal/Model/Param/ACNet.p")
    tables = parse(src)
    here = os.path.dirname(os.path.abspath(__file__))
    manifest = {"source": "core/internal/AC/Core/Internal/Model/Param/ACNet.p",
                "layout": "NHWC: layer l kernels at ACNet.hpp:48-53 offsets; "
                          "conv k[n][3*ky+kx][c], deconv k2[(2*dy+dx)*8+c]",
                "variants": {}}
    for v in VARIANTS:
        entry = {}
        for kind, n in (("kernels", KERNEL_LEN), ("biases", BIAS_LEN)):
            arr, first, last = tables[("ACNetLegacy", v, kind)]
            if arr.size != n:
                raise SystemExit(f"{v} {kind}: expected {n} values, parsed {arr.size}")
            blob = arr.tobytes()
            name = f"acnet_legacy_{v.lower()}.{kind}.f32"
            with open(os.path.join(here, name), "wb") as f:
                f.write(blob)
            entry[kind] = {"file": name, "count": n, "lines": f"{first}-{last}",
                           "sha256": hashlib.sha256(blob).hexdigest()}
        manifest["variants"][v.lower()] = entry
    manifest["acnet8"] = {}
    for name, (infix, blocks) in ACNET8.items():
        entry = {"blocks": blocks}
        for kind, n in acnet8_lengths(blocks).items():
            arr, first, last = tables[("ACNet", infix, kind)]
            if arr.size != n:
                raise SystemExit(f"ACNet {infix} {kind}: expected {n} values, parsed {arr.size}")
            blob = arr.tobytes()
            fname = f"acnet8_{name}.{kind}.f32"
            with open(os.path.join(here, fname), "wb") as f:
                f.write(blob)
            entry[kind] = {"file": fname, "count": n, "lines": f"{first}-{last}",
                           "sha256": hashlib.sha256(blob).hexdigest()}
        manifest["acnet8"][name] = entry
    art = parse(os.path.join(ref_root, "core/internal/AC/Core/Internal/Model/Param/ArtCNN.p"))
    manifest["artcnn"] = {"source": "core/internal/AC/Core/Internal/Model/Param/ArtCNN.p"}
    for name, (infix, f) in ARTCNN.items():
        entry = {"channels": f, "blocks": 4}
        for kind, n in artcnn_lengths(f).items():
            arr, first, last = art[("ArtCNN", infix, kind)]
            if arr.size != n:
                raise SystemExit(f"ArtCNN {infix} {kind}: expected {n} values, parsed {arr.size}")
            blob = arr.tobytes()
            fname = f"artcnn_{name}.{kind}.f32"
            with open(os.path.join(here, fname), "wb") as fh:
                fh.write(blob)
            entry[kind] = {"file": fname, "count": n, "lines": f"{first}-{last}",
                           "sha256": hashlib.sha256(blob).hexdigest()}
        manifest["artcnn"][name] = entry
    fsr = parse(os.path.join(ref_root, "core/internal/AC/Core/Internal/Model/Param/FSRCNNX.p"))
    manifest["fsrcnnx"] = {"source": "core/internal/AC/Core/Internal/Model/Param/FSRCNNX.p"}
    for name, (infix, f) in FSRCNNX.items():
        entry = {"channels": f, "blocks": 4}
        for kind, n in fsrcnnx_lengths(f).items():
            arr, first, last = fsr[("FSRCNNX", infix, kind)]
            if arr.size != n:
                raise SystemExit(f"FSRCNNX {infix} {kind}: expected {n} values, parsed {arr.size}")
            blob = arr.tobytes()
            fname = f"fsrcnnx_{name}.{kind}.f32"
            with open(os.path.join(here, fname), "wb") as fh:
                fh.write(blob)
            entry[kind] = {"file": fname, "count": n, "lines": f"{first}-{last}",
                           "sha256": hashlib.sha256(blob).hexdigest()}
        manifest["fsrcnnx"][name] = entry
    with open(os.path.join(here, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    print(json.dumps(manifest, indent=1))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
