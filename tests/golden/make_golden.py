"""Writes the golden fixtures of tests/golden/ (run in the build container,
where /root/reference exists).

* rng_words.json -- Philox4x32-10 and Threefry4x64-20 stream words and
  single blocks produced by the REFERENCE's own C engines
  (include/vsmc/rngc/{philox,threefry}.h compiled unchanged into
  oracle/_ref/libvsmc_rngc.so by oracle/Makefile), for keys {0, 1, 101, 102,
  0xFFFFFFFF}, the first 256 words, words after a skip, and blocks whose
  counters exercise the carry of include/vsmc/rng/counter.hpp:42-72 and
  include/vsmc/rngc/philox.h:96-106.
* u01.json -- the reference's U01 conversions (include/vsmc/rngc/u01.h:
  214-217, :243-246, and the u32 -> float form) on edge and random words.
* cv_obs_c0.json -- the 100-step observation track of the CPU workload
  (seed 20161), made by vsmc_amd.models.cv_simulate; inputs, not outputs.

Usage: python tests/golden/make_golden.py
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as ol  # noqa: E402

KEYS = [0, 1, 101, 102, 0xFFFFFFFF]


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                   capture_output=True)
    if not ol.ref_available():
        raise SystemExit("oracle/_ref is not built (reference tree absent)")
    r = ol.ref()
    out = {"philox": {}, "threefry": {}, "philox_blocks": [], "threefry_blocks": []}
    for k in KEYS:
        a = np.zeros(256, np.uint32)
        r.ref_philox4x32_words(k, 0, 256, a)
        b = np.zeros(64, np.uint32)
        r.ref_philox4x32_words(k, 1001, 64, b)
        out["philox"][str(k)] = {"first256": a.tolist(), "skip1001_64": b.tolist()}
        a = np.zeros(256, np.uint64)
        r.ref_threefry4x64_words(k, 0, 256, a)
        b = np.zeros(64, np.uint64)
        r.ref_threefry4x64_words(k, 1001, 64, b)
        out["threefry"][str(k)] = {"first256": [str(int(x)) for x in a],
                                   "skip1001_64": [str(int(x)) for x in b]}
    pctrs = [[0, 0, 0, 0], [1, 0, 0, 0], [0xFFFFFFFF, 0, 0, 0], [0, 1, 0, 0],
             [0xFFFFFFFF, 0xFFFFFFFF, 0, 0], [0, 0, 1, 0],
             [0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF],
             [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344]]
    pkeys = [[0, 0], [101, 0], [0xFFFFFFFF, 0xFFFFFFFF], [0xA4093822, 0x299F31D0]]
    for ctr in pctrs:
        for key in pkeys:
            o = np.zeros(4, np.uint32)
            r.ref_philox4x32_block(np.array(ctr, np.uint32), np.array(key, np.uint32), o)
            out["philox_blocks"].append({"ctr": ctr, "key": key, "out": o.tolist()})
    M = (1 << 64) - 1
    tctrs = [[0, 0, 0, 0], [1, 0, 0, 0], [M, 0, 0, 0], [0, 1, 0, 0], [M, M, M, M]]
    tkeys = [[0, 0, 0, 0], [101, 0, 0, 0], [M, M, M, M],
             [0x452821E638D01377, 0xBE5466CF34E90C6C, 0xC0AC29B7C97C50DD,
              0x3F84D5B5B5470917]]
    for ctr in tctrs:
        for key in tkeys:
            o = np.zeros(4, np.uint64)
            r.ref_threefry4x64_block(np.array(ctr, np.uint64), np.array(key, np.uint64), o)
            out["threefry_blocks"].append({"ctr": [str(x) for x in ctr],
                                           "key": [str(x) for x in key],
                                           "out": [str(int(x)) for x in o]})
    with open(os.path.join(HERE, "rng_words.json"), "w") as f:
        json.dump(out, f)

    rs = np.random.Generator(np.random.PCG64(7))
    w32 = np.concatenate([np.array([0, 1, 2, 255, 256, 0x7FFFFFFF, 0x80000000,
                                    0xFFFFFF00, 0xFFFFFFFE, 0xFFFFFFFF], np.uint32),
                          rs.integers(0, 1 << 32, 118, dtype=np.uint64).astype(np.uint32)])
    w64 = np.concatenate([np.array([0, 1, 2047, 2048, (1 << 63), M - 2047, M - 1, M],
                                   np.uint64),
                          rs.integers(0, 1 << 63, 120, dtype=np.uint64) * 2 + 1])
    d32 = np.zeros(len(w32))
    r.ref_u01_co_u32d(len(w32), w32, d32)
    f32 = np.zeros(len(w32), np.float32)
    r.ref_u01_co_u32f(len(w32), w32, f32)
    d64 = np.zeros(len(w64))
    r.ref_u01_co_u64d(len(w64), w64.astype(np.uint64), d64)
    with open(os.path.join(HERE, "u01.json"), "w") as f:
        json.dump({"w32": w32.tolist(), "u32d": [float(x).hex() for x in d32],
                   "u32f": [float(x).hex() for x in f32],
                   "w64": [str(int(x)) for x in w64],
                   "u64d": [float(x).hex() for x in d64]}, f)

    from vsmc_amd.models import cv_simulate
    st, ox, oy = cv_simulate(100, 20161)
    with open(os.path.join(HERE, "cv_obs_c0.json"), "w") as f:
        json.dump({"seed": 20161, "T": 100, "obs_x": [x.hex() for x in ox],
                   "obs_y": [x.hex() for x in oy]}, f)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
