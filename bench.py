"""Benchmark: chunks meshed per second on one MI355X (batched persistent path vs per-chunk launches).

Workload (the configuration the metric is quoted on): the 4096-chunk cave-terrain batch -- a
16x16x16-chunk region (256^3 voxels) plus a one-chunk rim, seeded 3D fBm caves under a surface
cap, ~15% all-air and ~10% all-stone chunks, spatially correlated random sun/torch light. Inputs
are resident in HBM before timing. One step = one batched meshing pass over all 4096 chunks
(HIP-graph replay of counter reset + persistent kernel). Output is checked bit-exact against
the digests the reference mesher produced for the same chunks.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "chunks meshed/sec (16³, 1 GPU) batched vs plain, bit-exact; p50/p99 remesh ms"
HBM_PEAK = 8.0e12
# Input bytes of the region: 4096 chunks x 16^3 x 3 B plus the one-voxel shell (258^3 - 256^3) x 3 B
REGION_IN_BYTES = 4096 * 4096 * 3 + (258 ** 3 - 256 ** 3) * 3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(blocks, light, nbr, threads):
    """The reference mesher (oracle/_ref, if it was built) else the C oracle, on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cpu_oracle

    if cpu_oracle.Reference.available():
        impl, kind = cpu_oracle.Reference(), "reference"
    else:
        impl, kind = cpu_oracle.Oracle(os.path.join(ROOT, "tests", "golden", "registry.bin")), "port"
    t0 = time.perf_counter()
    q = impl.mesh_many(blocks, light, nbr, threads)
    dt = time.perf_counter() - t0
    return {"value": round(nbr.shape[0] / dt, 1), "unit": "chunks/s", "cores": threads, "kind": kind,
            "sample": f"all {nbr.shape[0]} chunks of the batch once ({int(q.sum())} quads), {dt:.2f} s wall",
            "quads_per_s": round(float(q.sum()) / dt, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build", file=sys.stderr)
        sys.exit(2)

    from hypercraft_amd import hcm, workloads

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import meshdigest

    hcm.require_device()
    log("generating the 4096-chunk cave batch")
    blocks, light, nbr = workloads.cave_region()
    golden = np.load(os.path.join(ROOT, "tests", "golden", "golden.npz"))
    ctx = hcm.Context(pool_quads=1 << 23)
    store = hcm.Store(ctx, blocks.shape[0])
    store.upload(blocks, light)
    batch = hcm.Batch(ctx, store, nbr)
    n = nbr.shape[0]

    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        batch.run()
    batch.sync()

    log(f"timing {args.steps} steps")
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run()
        kernel_ms.append(batch.sync())
    dt = time.perf_counter() - t0
    quads, meshes = batch.totals()
    out = batch.download()
    bad = 0
    for j in range(n):
        d = meshdigest.digest(hcm.job_meshes(j, *out))
        bad += d != golden["cave_region_digest"][j].tobytes()
    bit_exact = bad == 0 and (out[0]["status"] == 0).all()

    plain = None
    if not args.no_plain:
        log("plain per-chunk launches")
        batch.run(hcm.RUN_PER_CHUNK)
        batch.sync()
        t1 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            batch.run(hcm.RUN_PER_CHUNK)
            batch.sync()
        plain = n * reps / (time.perf_counter() - t1)

    value = n * args.steps / dt
    k_ms = float(np.mean(kernel_ms))
    alg_bytes = REGION_IN_BYTES + 44 * quads + 40 * meshes
    achieved = alg_bytes / (k_ms * 1e-3)
    res = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "chunks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "cave-terrain batch: 4096 chunks of 16^3 (16x16x16-chunk region, 256^3 voxels "
                               "+ rim), seed 2, one persistent launch per step",
                   "chunks": n, "quads_per_step": int(quads), "meshes_per_step": int(meshes)},
        "quads_per_s": round(quads * args.steps / dt, 1),
        "bit_exact": bool(bit_exact),
        "plain_per_chunk_chunks_per_s": round(plain, 1) if plain else None,
        "batched_vs_plain": round(value / plain, 2) if plain else None,
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": None,
                     "kernel_ms": round(k_ms, 4), "algorithmic_bytes": int(alg_bytes)},
    }
    if not args.no_cpu:
        log("CPU baseline")
        res["cpu_baseline"] = cpu_baseline(blocks, light, nbr, min(16, os.cpu_count() or 1))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
