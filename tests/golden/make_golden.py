"""Writes the committed golden fixtures under tests/golden/.

Run in the development container (it reads the reference's data files from
/root/reference; nothing at test time does):

    python tests/golden/make_golden.py

1. score_crop.npz -- the reference's own unary fixtures scoreFg.npy /
   scoreBg.npy (float32, 800x608; the only data its GraphCut test holds,
   test_GraphCut.py:44-47) cropped to rows 180..439, cols 20..399 (both hard-
   constraint blocks inside), mapped to int32: source = rint(scoreBg),
   sink = rint(scoreFg), values >= 2^31 - 1024 (the hard constraints) -> 2^30.
   Pairwise weights are absent from the reference (its input image is not in
   the repository, test_GraphCut.py:17), so two variants are stored:
   constant w = 8 ("potts8") and w = 1 + (x*7 + y*13) % 20 ("ramp").
   Expected outputs come from scipy.sparse.csgraph.maximum_flow (Dinic) --
   flow value and the set reachable from s in scipy's residual graph --
   independent of the oracle's own solvers.
2. workloads.json -- sha256 of the generated 512x512 (seed 0) workload
   arrays and its max-flow value by scipy, pinning the generator.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "cl-util_amd"))

import gcmi_workloads as G  # noqa: E402
import oracle as O  # noqa: E402

REF = "/root/reference"
HARD = 1 << 30


def to_int(a):
    a = np.asarray(a, np.float64)
    out = np.rint(np.minimum(a, HARD)).astype(np.int64)
    out[a >= 2.0 ** 31 - 1024] = HARD
    return out.astype(np.int32)


def main():
    fg = np.load(os.path.join(REF, "scoreFg.npy"), allow_pickle=False).reshape(608, 800)
    bg = np.load(os.path.join(REF, "scoreBg.npy"), allow_pickle=False).reshape(608, 800)
    rs, cs_ = slice(180, 440), slice(20, 400)
    source = to_int(bg[rs, cs_])
    sink = to_int(fg[rs, cs_])
    H, W = source.shape
    out = {"source": source, "sink": sink}
    ys, xs = np.mgrid[0:H, 0:W]
    variants = {"potts8": np.full((H, W), 8, np.int32),
                "ramp": (1 + (xs * 7 + ys * 13) % 20).astype(np.int32)}
    for name, w in variants.items():
        right = w.copy()
        right[:, -1] = 0
        down = w.copy()
        down[-1, :] = 0
        flow, mask = O.scipy_cut(source, sink, right, down, with_mask=True)
        out[name + "_right"] = right
        out[name + "_down"] = down
        out[name + "_flow"] = np.array([flow], np.int64)
        out[name + "_mask"] = np.packbits(mask.reshape(-1))
        print(name, flow, int(mask.sum()))
    np.savez_compressed(os.path.join(HERE, "score_crop.npz"), **out)

    w = G.centred(512, 512, seed=0)
    hashes = {k: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
              for k, a in zip(("source", "sink", "right", "down"), w.arrays())}
    flow, mask = O.scipy_cut(*w.arrays(), with_mask=True)
    info = {"centred_512_seed0": {"sha256": hashes, "flow": flow,
                                  "mask_sha256": hashlib.sha256(mask.tobytes()).hexdigest(),
                                  "source_pixels": int(mask.sum())}}
    with open(os.path.join(HERE, "workloads.json"), "w") as f:
        json.dump(info, f, indent=1)
    print(info)


if __name__ == "__main__":
    main()
