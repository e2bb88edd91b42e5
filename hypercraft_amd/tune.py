"""Autotuner for the batched meshing kernel: workgroup size x workgroups per CU.

The persistent kernel's cost per chunk depends on the workgroup size (how many lanes share the
102 visibility tasks, the light-word pass and the emit loop) and on how many workgroups each CU
holds (LDS occupancy vs latency hiding). `autotune` times every candidate on the caller's batch with
HIP events, keeps only settings whose output is bit-identical to the first candidate's, applies the
fastest to the context, and can cache the choice in a JSON file keyed by device and workload.
"""
import hashlib
import json
import os

import numpy as np

THREADS = (128, 256, 512)
WG_PER_CU = (0, 1, 2, 3, 4)  # 0: as many as the LDS footprint allows


def _output_digest(batch):
    """Digest of every job's mesh list in job order (pool positions vary from run to run)."""
    res, recs, verts, idx = batch.download()
    h = hashlib.sha256()
    for f in ("status", "mesh_count", "quad_count"):
        h.update(np.ascontiguousarray(res[f]).tobytes())
    for r in res:
        for m in recs[r["first_mesh"]:r["first_mesh"] + r["mesh_count"]]:
            q0, n = int(m["first_quad"]), int(m["quad_count"])
            h.update(m["aabb"].tobytes())
            h.update(bytes([int(m["transparent"])]))
            h.update(verts[4 * q0:4 * (q0 + n)].tobytes())
            h.update(idx[6 * q0:6 * (q0 + n)].tobytes())
    return h.hexdigest()


def _cache_key(device, workload):
    return f"{device}|{workload}"


def autotune(ctx, batch, workload="", reps=5, cache_path=None, candidates=None, log=None):
    """Times the candidates on `batch` and applies the fastest bit-identical one to `ctx`.

    Returns {"threads", "wg_per_cu", "kernel_ms", "table": [(threads, wg_per_cu, ms), ...]}.
    """
    device = "gfx950"  # the library runs on no other target
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cached = json.load(f).get(_cache_key(device, workload))
        if cached:
            ctx.set_tuning(cached["wg_per_cu"], cached["threads"])
            return dict(cached, cached=True)
    candidates = candidates or [(t, w) for t in THREADS for w in WG_PER_CU]
    table, best, ref = [], None, None
    for threads, wg in candidates:
        ctx.set_tuning(wg, threads)
        batch.run()
        batch.sync()
        ms = []
        for _ in range(reps):
            batch.run()
            ms.append(batch.sync())
        d = _output_digest(batch)
        if ref is None:
            ref = d
        t = float(np.median(ms))
        table.append((threads, wg, round(t, 4), d == ref))
        if log:
            log(f"tune: threads {threads} wg/CU {wg or 'auto'}: {t:.4f} ms" + ("" if d == ref else " (output differs!)"))
        if d == ref and (best is None or t < best[2]):
            best = (threads, wg, t)
    if best is None:
        raise RuntimeError("autotune: no candidate produced the reference output")
    ctx.set_tuning(best[1], best[0])
    out = {"threads": best[0], "wg_per_cu": best[1], "kernel_ms": round(best[2], 4), "table": table}
    if cache_path:
        data = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                data = json.load(f)
        data[_cache_key(device, workload)] = {k: out[k] for k in ("threads", "wg_per_cu", "kernel_ms")}
        with open(cache_path, "w") as f:
            json.dump(data, f, indent=1)
    return out
