"""Write tests/golden/sr_hw_vectors.npz: outputs of gfx950's
v_cvt_sr_bf16_f32 / v_cvt_sr_f16_f32 on the pairs of sr_hw_probe.hip (every
special pair and 1 in 64 of the random ones).  Needs an MI355X; run from the
repository root:  python3 tests/golden/gen_sr_hw.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "sr_hw_probe")
        subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-Xclang",
                        "-target-feature", "-Xclang", "-packed-fp32-ops",
                        os.path.join(HERE, "sr_hw_probe.hip"), "-o", exe], check=True)
        out = os.path.join(tmp, "sr_hw.bin")
        subprocess.run([exe, out], check=True)
        d = np.fromfile(out, dtype=np.uint32)
    n = int(d[0])
    x, r = d[1:1 + n], d[1 + n:1 + 2 * n]
    bf16 = (d[1 + 2 * n:1 + 3 * n] & 0xFFFF).astype(np.uint16)
    f16 = (d[1 + 3 * n:1 + 4 * n] & 0xFFFF).astype(np.uint16)
    special = n - (1 << 20)
    idx = np.concatenate([np.arange(special), special + np.arange(0, 1 << 20, 64)])
    np.savez_compressed(os.path.join(HERE, "sr_hw_vectors.npz"), x=x[idx], r=r[idx],
                        bf16=bf16[idx], f16=f16[idx])
    print("wrote %d pairs" % len(idx))


if __name__ == "__main__":
    sys.exit(main())
