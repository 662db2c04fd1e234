"""Autotuner of the batched Cholesky: measures every variant of each size bin on the GPU and
writes the variant table the library compiles in (csrc/batched_tuning.hpp).

For each element type (float64, float32) and size bin (n <= 128, n <= 256, n <= 512) the
workload is a batch of HPD matrices G G^T / n + I with n uniform over the bin's range (the
batched configuration's distribution restricted to the bin), and the candidates are the kernel
kinds with their block-column widths (16 and 32; 64-column panels for kind 2) and 1 .. (the
bin's maximum) persistent workgroups per CU.
Each candidate factors the batch `--reps` times (A restored from a copy before each run, the
factorization alone timed with HIP events); the smallest median wins.

Usage (on a GPU):  python -m elemental_amd.autotune_batched --out gpurun_out/batched_tuning.hpp
then copy the file to elemental_amd/csrc/batched_tuning.hpp and rebuild.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

BINS = [(16, 128), (129, 256), (257, 512)]
MAX_OCC = [3, 2, 2]
NBS = [16, 32]
# kernel kinds per bin: 0 block column staged in LDS, 1 operands read straight into registers
# (bins 1 and 2; at most three workgroups per CU, as registers allow), 2 64-column panels
# (bins 1 and 2; at most two workgroups per CU)
KINDS = [[0], [0, 1, 2], [0, 1, 2]]
RD_MAX_OCC = 3
WIDE_MAX_OCC = 2


def candidates(b, kind):
    """(nb, occ) pairs of a kernel kind in bin b."""
    if kind == 2:
        return [(64, occ) for occ in range(1, WIDE_MAX_OCC + 1)]
    top = RD_MAX_OCC if kind == 1 else MAX_OCC[b]
    return [(nb, occ) for nb in NBS for occ in range(1, top + 1)]


def make_batch(dtype, lo, hi, count, seed):
    import numpy as np
    import torch
    from elemental_amd import device as dev
    rng = np.random.default_rng(seed)
    ns = rng.integers(lo, hi + 1, size=count)
    mats = []
    for n in ns:
        G = torch.randn(int(n), int(n), dtype=torch.float64, device="cuda")
        A = G @ G.T / float(n) + torch.eye(int(n), dtype=torch.float64, device="cuda")
        t = dev.colmajor_empty(int(n), int(n), dtype=dtype)
        t.copy_(A)
        mats.append(t)
    return mats, [int(n) for n in ns]


def time_variant(dtype, d, b, nb, occ, mats, saved, reps, kind=0):
    import torch
    from elemental_amd import device as dev
    dev.batched_variant(dtype, b, nb, occ, kind)
    batch = dev.Batch(mats, dtype)
    times = []
    for r in range(reps + 1):
        for m, s in zip(mats, saved):
            m.copy_(s)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        info = dev.potrf_vbatched(batch)
        e1.record()
        torch.cuda.synchronize()
        if int(info.abs().sum().item()) != 0:
            raise RuntimeError(f"variant nb={nb} occ={occ} reported a non-HPD matrix")
        if r > 0:  # the first run builds the schedule
            times.append(e0.elapsed_time(e1))
    times.sort()
    return times[len(times) // 2]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=3000)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=os.path.join("gpurun_out", "batched_tuning.hpp"))
    args = ap.parse_args()
    import torch
    from elemental_amd import device as dev
    dev.load()
    torch.cuda.set_device(0)
    table = [[None] * 3, [None] * 3]
    report = []
    for d, dtype in enumerate([torch.float64, torch.float32]):
        for b, (lo, hi) in enumerate(BINS):
            mats, ns = make_batch(dtype, lo, hi, args.count, 1000 + 10 * d + b)
            flops = sum(n ** 3 / 3.0 for n in ns)
            saved = [dev.colmajor_empty(m.shape[0], m.shape[1], dtype=dtype) for m in mats]
            for s, m in zip(saved, mats):
                s.copy_(m)
            best = None
            for kind in KINDS[b]:
                for nb, occ in candidates(b, kind):
                    ms = time_variant(dtype, d, b, nb, occ, mats, saved, args.reps, kind)
                    tf = flops / (ms * 1e-3) / 1e12
                    report.append({"dtype": str(dtype).split(".")[-1], "bin": [lo, hi],
                                   "kind": kind, "nb": nb, "occ": occ, "ms": round(ms, 4),
                                   "tflops": round(tf, 3)})
                    print(json.dumps(report[-1]), flush=True)
                    if best is None or ms < best[0]:
                        best = (ms, nb, occ, kind)
            dev.batched_variant(dtype, b, 0)
            table[d][b] = (best[1], best[2], best[3])
            del mats, saved
            torch.cuda.empty_cache()
    rows = []
    for d in range(2):
        rows.append("    {" + ", ".join("{%d, %d, %d}" % v for v in table[d]) + "},")
    text = """// Variant table of the batched Cholesky factorization, per element type and size bin.
// Written by elemental_amd/autotune_batched.py from timings on one MI355X (see that script for
// the workloads and the candidates); edit by running the script again, not by hand.
// Measured %s, chosen variants (median factor time of the bin's workload):
%s
#pragma once

namespace elx {
namespace tuned {

struct BatchedVariant {
    int nb;    // block-column width of the left-looking factorization (16 or 32; 64 for kind 2)
    int occ;   // persistent workgroups per CU (at most what the bin's LDS and registers allow)
    int kind;  // 0: block column staged in LDS; 1: operands read straight into registers;
               // 2: 64-column panels, diagonal block factored and inverted in LDS
};

// [0: double, 1: float][bin 0: n <= 128, 1: n <= 256, 2: n <= 512]
constexpr BatchedVariant kBatchedPotrf[2][3] = {
%s
};

}  // namespace tuned
}  // namespace elx
""" % (time.strftime("%Y-%m-%d"), "\n".join(
        "//   %s n in [%d, %d]: kind %d nb %d occ %d %.3f ms" % (
            r["dtype"], r["bin"][0], r["bin"][1], r["kind"], r["nb"], r["occ"], r["ms"])
        for r in report if (r["nb"], r["occ"], r["kind"]) == table[
            0 if r["dtype"] == "float64" else 1][BINS.index(tuple(r["bin"]))]),
        "\n".join(rows))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        f.write(text)
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(report, f, indent=1)
    print("wrote", args.out, flush=True)


if __name__ == "__main__":
    main()
