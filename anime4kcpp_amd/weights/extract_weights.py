"""Extract the ACNetLegacy weight tables into little-endian f32 blobs.

Source: core/internal/AC/Core/Internal/Model/Param/ACNet.p (the five
``ACNetLegacy_{GAN,HDN0..3}_NHWC_{kernels,biases}`` constexpr arrays, bound to
the model variants in core/src/Model.cpp:8-34).  The text is parsed, never
compiled or executed.  Output: ``acnet_legacy_<variant>.kernels.f32`` (4712
floats) and ``.biases.f32`` (72 floats) beside this script, plus
``MANIFEST.json`` with the source line ranges and sha256 of every blob.

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
_HEAD = re.compile(r"constexpr float (ACNetLegacy_(\w+?)_NHWC_(kernels|biases))\[\] = \{")
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
        tables[(m.group(2), m.group(3))] = (arr, start, i + 1)
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
            arr, first, last = tables[(v, kind)]
            if arr.size != n:
                raise SystemExit(f"{v} {kind}: expected {n} values, parsed {arr.size}")
            blob = arr.tobytes()
            name = f"acnet_legacy_{v.lower()}.{kind}.f32"
            with open(os.path.join(here, name), "wb") as f:
                f.write(blob)
            entry[kind] = {"file": name, "count": n, "lines": f"{first}-{last}",
                           "sha256": hashlib.sha256(blob).hexdigest()}
        manifest["variants"][v.lower()] = entry
    with open(os.path.join(here, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    print(json.dumps(manifest, indent=1))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
