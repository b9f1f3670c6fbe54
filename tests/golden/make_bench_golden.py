"""Writes tests/golden/bench_expected.json: the oracle's answers for the
benchmark's workloads, so that bench.py checks every result it reports
without running the oracle itself.

    python tests/golden/make_bench_golden.py

Every answer is a Boykov-Kolmogorov cut (oracle/gc_oracle.c) of the
workload's int32 capacities: the max-flow value and the first 16 hex digits
of the sha256 of the minimal source set packed to bits (bit p % 32 of
uint32 word p // 32, little-endian; gcmi_serving.mask_sha).

  plain_4096   the 4096x4096 three-ellipse image (seed 1)
  batched_64   the 64 ragged images (seed 2), in batch order
  warm_2048    the 2048x2048 image (seed 3) after each of its 100 strokes
  served       the 256-session mix (seed 4) after the verification trace
               (gcmi_serving.verification_trace): every request's flow in
               trace order, every session's final flow and mask
"""
import concurrent.futures as cf
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "cl-util_amd"))

import gcmi_serving as S  # noqa: E402
import gcmi_workloads as G  # noqa: E402
import oracle as O  # noqa: E402

OUT = os.path.join(HERE, "bench_expected.json")


def answer(arrays):
    f, m = O.bk_cut(*arrays)
    return {"flow": f, "mask_sha": S.mask_sha(S.pack_mask(m))}


def served():
    mix = S.Mix(nsess=256, seed=4)
    items = S.verification_trace(mix)
    per = [[] for _ in range(mix.nsess)]
    for j, it in enumerate(items):
        per[it[1]].append((j, it[5]))

    def session(i):
        n = mix.sizes[i]
        cur = mix.base(i).copy()
        flows = []
        for j, src in per[i]:
            if isinstance(src, G.Workload):
                cur = src.copy()
            else:
                G.apply_stroke(cur, G.rasterise(src, n, n))
            flows.append((j, O.bk_cut(*cur.arrays())[0]))
        return flows, answer(cur.arrays())

    with cf.ThreadPoolExecutor(8) as ex:
        res = list(ex.map(session, range(mix.nsess)))
    req = [None] * len(items)
    for flows, _ in res:
        for j, f in flows:
            req[j] = f
    return {"trace": dict(S.VERIFY), "requests": len(items),
            "request_flows": req, "sessions": [a for _, a in res]}


def main():
    out = {}
    w = G.three_ellipses(4096, 4096, seed=1)
    out["plain_4096"] = answer(w.arrays())
    print("plain", out["plain_4096"], flush=True)
    works = G.ragged_batch(64, 256, 2048, seed=2)
    with cf.ThreadPoolExecutor(8) as ex:
        out["batched_64"] = list(ex.map(lambda x: answer(x.arrays()), works))
    print("batched done", flush=True)
    cur = G.centred(2048, 2048, seed=3)
    warm = []
    for st in G.strokes(2048, 2048, count=100, seed=3):
        G.apply_stroke(cur, st)
        warm.append(answer(cur.arrays()))
    out["warm_2048"] = warm
    print("warm done", flush=True)
    out["served"] = served()
    print("served done: %d requests" % out["served"]["requests"], flush=True)
    with open(OUT, "w") as f:
        json.dump(out, f, separators=(",", ":"))
        f.write("\n")


if __name__ == "__main__":
    main()
