"""Writes the golden vectors under tests/golden/ from the FP64 oracle.

The reference ships no tests or expected outputs, so these vectors pin the oracle against
regressions and give the GPU tests fixed inputs/outputs. Run from the repo root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import light_garden_amd as lg  # noqa: E402
from light_garden_amd import abi, configs  # noqa: E402
from oracle import oracle as orc  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def golden_scenes():
    small = configs.plain_small(num_rays=1500)
    small.width = small.height = 128
    large = configs.large_single(num_rays=1200, width=128)
    return {"plain_small_1500": small, "large_single_1200": large}


def main():
    for name, sc in golden_scenes().items():
        pos, col, key = orc.trace_scene(sc)
        acc = orc.accumulate_f64(pos, col, None, sc.view())
        img = orc.resolve_f64(acc)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), pos=pos, col=col, key=key, img=img)
        print(name, len(key), "segments")
    spec = configs.StringModSpec(modulo=2000, mults=(2, 3, 7), width=128, height=128)
    pos, col = configs.string_mod_segments(spec)
    acc = orc.accumulate_f64(pos, col, col, abi.LgView(-1, 1, -1, 1, 128, 128))
    np.savez_compressed(os.path.join(HERE, "string_mod_2000.npz"), pos=pos, col=col,
                        img=orc.resolve_f64(acc))
    print("string_mod_2000", len(pos), "segments")


if __name__ == "__main__":
    main()
