"""Autotuner for the batched meshing kernel: workgroup size x workgroups per CU x job order.

The persistent kernel's cost per chunk depends on the workgroup size (how many lanes share the
102 visibility tasks, the face-word pass and the emission), on how many workgroups each CU holds
(LDS occupancy vs latency hiding), and on the order of the job list: the kernel cuts the list into
eight contiguous parts, one per XCD, so the order decides which chunks share an XCD's L2 (a chunk's
halo is read by its 26 neighbours' jobs too) and, within a part, which jobs run last: a persistent launch
ends when its slowest workgroup finishes its last job, so jobs known to be costly (many quads in an
earlier run of the same chunks) are better started first. `autotune` times every candidate on the caller's jobs
with HIP events, keeps only settings whose meshes equal the wanted digests (the reference's, when the
caller has them; else the first candidate's), applies the fastest to the context and returns it with
the job order to build batches with. The choice can be cached in a JSON file keyed by device, the
library build and the workload; a cached choice is checked against the digests before it is used.
"""
import hashlib
import json
import os

import numpy as np

THREADS = (128, 256, 512)
WG_PER_CU = (0, 1, 2, 3, 4)  # 0: as many as the LDS footprint allows
ORDERS = ("linear", "morton", "linear_cost", "morton_cost", "morton_tailcost")
QUEUES = 8  # the kernel's job-list parts, one per XCD (hcm_internal.h kQueues)


def job_order(name, positions, cost=None):
    """Permutation of the jobs: "linear" keeps the caller's order; "morton" sorts the jobs along the
    3-D Z-order curve of their chunk positions, so each eighth of the list (one XCD's part) is a compact
    block of space and consecutive jobs are neighbours. "<order>_cost" keeps each eighth of <order> and
    sorts the jobs inside it by descending cost (per job, in the caller's order; ties keep <order>), so
    that each part ends with its cheapest jobs. "<order>_tailcost" sorts only the second half of each
    part that way: the first half keeps <order>'s locality, and the part still ends with the cheapest of
    its last half."""
    p = np.asarray(positions, np.int64)
    base, _, by_cost = name.partition("_")
    if base not in ("linear", "morton") or by_cost not in ("", "cost", "tailcost"):
        raise ValueError(f"unknown job order {name!r} (orders: {', '.join(ORDERS)})")
    if by_cost and cost is None:
        raise ValueError(f"job order {name!r} needs the jobs' costs")
    if base == "linear":
        perm = np.arange(p.shape[0])
    else:
        q = p - p.min(0)
        code = np.zeros(q.shape[0], np.int64)
        for bit in range(21):
            for a in range(3):
                code |= ((q[:, a] >> bit) & 1) << (3 * bit + a)
        perm = np.argsort(code, kind="stable")
    if by_cost:
        n = perm.shape[0]
        c = np.asarray(cost, np.float64)
        parts = []
        for q in range(QUEUES):
            part = perm[q * n // QUEUES:(q + 1) * n // QUEUES]
            k = part.shape[0] // 2 if by_cost == "tailcost" else 0  # jobs kept in the base order
            parts += [part[:k], part[k:][np.argsort(-c[part[k:]], kind="stable")]]
        perm = np.concatenate(parts)
    return perm


def job_costs(batch):
    """Per job of a run batch (its job order): the quads it produced, a proxy for its meshing time."""
    res = batch.download()[0]
    return res["quad_count"].astype(np.float64)


def job_digests(batch):
    """Per job, the digest of its mesh list (tests/meshdigest.py's layout), in the batch's job order."""
    from . import hcm

    res, recs, verts, idx = batch.download()
    out = np.zeros((res.shape[0], 32), np.uint8)
    for j in range(res.shape[0]):
        h = hashlib.sha256()
        for m in hcm.job_meshes(j, res, recs, verts, idx):
            hdr = np.array([m.vertices.shape[0], m.indices.shape[0], *m.aabb, m.transparent], np.uint32)
            h.update(hdr.tobytes())
            h.update(np.ascontiguousarray(m.vertices, np.uint32).tobytes())
            h.update(np.ascontiguousarray(m.indices, np.uint16).tobytes())
        out[j] = np.frombuffer(h.digest(), np.uint8)
    return out


def library_id():
    """A hash of the loaded library build: cached choices are valid for the build that made them."""
    from . import hcm

    with open(hcm.LIB_PATH, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def _cache_key(device, workload):
    return f"{device}|{library_id()}|{workload}"


def _time(ctx, batch, threads, wg, reps):
    ctx.set_tuning(wg, threads)
    batch.run()
    batch.sync()
    ms = []
    for _ in range(reps):
        batch.run()
        ms.append(batch.sync())
    return float(np.median(ms))


def autotune(ctx, store, nbr, workload="", positions=None, want=None, reps=5, cache_path=None, candidates=None,
             log=None):
    """Times the candidates (threads, wg_per_cu, order) on the jobs `nbr` over `store` and applies the
    fastest whose meshes equal `want` (per-job digests in the caller's job order) to `ctx`. Without a
    candidate list: every workgroup shape in the caller's order, then the other orders (which need the
    chunk positions) with the best shape; the cost orders use each job's quads from a first run.

    Returns {"threads", "wg_per_cu", "order", "perm", "kernel_ms", "table": [(threads, wg, order, ms, ok)]};
    build the batch as hcm.Batch(ctx, store, nbr[perm]) (job k of it is the caller's job perm[k])."""
    from . import hcm

    device = "gfx950"  # the library runs on no other target
    if positions is None:
        orders = {"linear": np.arange(nbr.shape[0])}
    else:
        probe = hcm.Batch(ctx, store, nbr)  # one run in the caller's order gives the jobs' costs
        try:
            probe.run()
            probe.sync()
            cost = job_costs(probe)
        finally:
            probe.close()
        orders = {o: job_order(o, positions, cost) for o in ORDERS}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cached = json.load(f).get(_cache_key(device, workload))
        if cached and cached["order"] in orders:
            perm = orders[cached["order"]]
            batch = hcm.Batch(ctx, store, nbr[perm])
            try:
                ctx.set_tuning(cached["wg_per_cu"], cached["threads"])
                batch.run()
                batch.sync()
                ok = want is None or np.array_equal(job_digests(batch), want[perm])
            finally:
                batch.close()
            if ok:
                return dict(cached, perm=perm, cached=True)
            if log:
                log("tune: the cached choice no longer gives the wanted meshes; tuning again")
    # two stages unless the caller lists the candidates: every workgroup shape in the caller's order, then
    # every other job order with the best shape
    stages = [candidates] if candidates else [[(t, w, "linear") for t in THREADS for w in WG_PER_CU], None]
    table, best, ref = [], None, None
    batches = {}
    try:
        for stage in stages:
            if stage is None:
                if best is None:
                    break
                stage = [(best[0], best[1], o) for o in orders if o != "linear"]
            for threads, wg, order in stage:
                if order not in batches:
                    batches[order] = hcm.Batch(ctx, store, nbr[orders[order]])
                batch = batches[order]
                t = _time(ctx, batch, threads, wg, reps)
                d = job_digests(batch)[np.argsort(orders[order])]  # back to the caller's job order
                if want is None and ref is None:
                    ref = d
                ok = bool(np.array_equal(d, want if want is not None else ref))
                table.append((threads, wg, order, round(t, 4), ok))
                if log:
                    log(f"tune: threads {threads} wg/CU {wg or 'auto'} order {order}: {t:.4f} ms" +
                        ("" if ok else " (output differs!)"))
                if ok and (best is None or t < best[3]):
                    best = (threads, wg, order, t)
    finally:
        for b in batches.values():
            b.close()
    if best is None:
        raise RuntimeError("autotune: no candidate produced the wanted meshes")
    ctx.set_tuning(best[1], best[0])
    out = {"threads": best[0], "wg_per_cu": best[1], "order": best[2], "kernel_ms": round(best[3], 4),
           "table": table, "perm": orders[best[2]]}
    if cache_path:
        data = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                data = json.load(f)
        data[_cache_key(device, workload)] = {k: out[k] for k in ("threads", "wg_per_cu", "order", "kernel_ms")}
        with open(cache_path, "w") as f:
            json.dump(data, f, indent=1)
    return out
