"""Benchmark: chunks meshed per second on one MI355X (batched persistent path vs per-chunk launches),
plus the serving loop's p50/p99 remesh latency.

Headline workload: the 4096-chunk cave-terrain batch -- a 16x16x16-chunk region (256^3 voxels) plus
a one-chunk rim, seeded 3D fBm caves under a surface cap, ~15% all-air and ~10% all-stone chunks,
spatially correlated random sun/torch light. Inputs are resident in HBM before timing. One step =
one batched meshing pass over all 4096 chunks (HIP-graph replay of counter reset + persistent
kernel) with the autotuned workgroup shape. Output is checked bit-exact against the digests the
reference mesher produced for the same chunks.

More legs go into the same JSON line:
  mixed_batches  200 batches of log-uniform size in [1, 4096] drawn from the cave region and a
                 plant-covered terrain region: captured-graph runs vs per-chunk launches, bit-exact;
  serving        a spiral load trace (radius 24 columns x 16 chunks, ~29k terrain chunks, Poisson
                 arrivals at 20k chunks/s) with concurrent single-voxel edits at 2k/s through the
                 continuous-batching server (16 GB vertex pool): p50/p99 request latency; also with the
                 light rebuild, and with loads arriving ten times as fast;
  relight        the whole mesh task on a lit cave region: light rebuilt from the blocks' light levels
                 and the chunks' sunlight heights, then meshing, bit-exact with the reference;
  dropin         the one-chunk call from host memory (hcm_mesh_chunk, hcm_mesh_halo): per-call p50/p99
                 with the PCIe transfers, beside the reference's Generate on one host thread.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-plain] [--no-mixed] [--no-serve]
                    [--no-relight] [--no-dropin]
"""
import argparse
import gc
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


# The GPU box gives one GPU's job 16 host cores, whatever the machine has.
CPU_SHARE = 16


def host_cpus():
    """(cores this process may run on, CPU model name)."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return len(os.sched_getaffinity(0)), model


def cpu_impl():
    """The reference mesher (oracle/_ref, where it was built) else the C oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cpu_oracle

    if cpu_oracle.Reference.available():
        return cpu_oracle.Reference(), "reference"
    return cpu_oracle.Oracle(os.path.join(ROOT, "tests", "golden", "registry.bin")), "port"


def cpu_passes(impl, blocks, light, nbr, threads, min_seconds):
    """Whole passes over the batch with `threads` host threads until min_seconds have run: (chunks/s, quads/s,
    passes, seconds)."""
    passes, quads = 0, 0
    t0 = time.perf_counter()
    while True:
        quads += int(impl.mesh_many(blocks, light, nbr, threads).sum())
        passes += 1
        dt = time.perf_counter() - t0
        if dt >= min_seconds:
            break
    return passes * nbr.shape[0] / dt, quads / dt, passes, dt


def cpu_baseline(blocks, light, nbr, seconds=(8.0, 6.0)):
    """The reference mesher on the host cores: N threads, N the cores this process may use up to the box's
    share for one GPU, and the reference's default of 3/4 of them (client/src/Application.cpp:248)."""
    impl, kind = cpu_impl()
    host, model = host_cpus()
    n = min(host, CPU_SHARE)
    n34 = max(1, n * 3 // 4)
    v, q, passes, dt = cpu_passes(impl, blocks, light, nbr, n, seconds[0])
    v34, q34, passes34, dt34 = cpu_passes(impl, blocks, light, nbr, n34, seconds[1])
    return {"value": round(v, 1), "unit": "chunks/s", "cores": n, "kind": kind,
            "sample": f"{passes} passes over the {nbr.shape[0]}-chunk batch ({passes * nbr.shape[0]} chunks), "
                      f"{dt:.2f} s wall, {n} threads; then {passes34} passes with {n34} threads in {dt34:.2f} s",
            "quads_per_s": round(q, 1), "threads_3_4": n34, "value_3_4": round(v34, 1),
            "quads_per_s_3_4": round(q34, 1), "host_cores": host, "cpu_model": model,
            "threads_policy": f"min(cores available = {host}, {CPU_SHARE} per GPU on this box); 3/4 of that is "
                              "the reference's default worker count"}


def dropin_leg(ctx, hcm, blocks, light, nbr, golden_digest, n=512, reps=3):
    """The drop-in call as the reference's ChunkTaskRunner<kMesh>::Run makes it (ChunkMeshTask.cpp:113-122):
    one chunk per call from host memory, meshes back in host memory, PCIe both ways. hcm_mesh_chunk takes
    the 27 chunks (the library packs the 20 KB halo), hcm_mesh_halo a halo packed by the caller; per-call
    latency over the first n cave jobs. Beside it the reference's Generate on one host thread, same chunks."""
    import ctypes

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import meshdigest

    lib = hcm.lib()
    out = hcm._HostOut()
    jobs = range(min(n, nbr.shape[0]))
    arrays = [workloads_slice(blocks, light, nbr, j) for j in jobs]
    ptrs = []
    for b27, l27 in arrays:
        pb = (ctypes.c_void_p * 27)(*[x.ctypes.data for x in b27])
        pl = (ctypes.c_void_p * 27)(*[x.ctypes.data for x in l27])
        ptrs.append((pb, pl))
    halos = [tuple(np.ascontiguousarray(a) for a in hcm.halo_of_chunks(b27, l27)) for b27, l27 in arrays]
    bad = 0
    for j in range(0, len(ptrs), 64):  # the calls give the reference's meshes
        hcm.check("hcm_mesh_chunk", lib.hcm_mesh_chunk(ctx.ptr, ptrs[j][0], ptrs[j][1], ctypes.byref(out.c)))
        bad += meshdigest.digest(out.meshes()) != golden_digest[j].tobytes()
        hcm.check("hcm_mesh_halo", lib.hcm_mesh_halo(ctx.ptr, halos[j][0].ctypes.data, halos[j][1].ctypes.data,
                                                     ctypes.byref(out.c)))
        bad += meshdigest.digest(out.meshes()) != golden_digest[j].tobytes()
    lat = {"chunk": [], "halo": []}
    for _ in range(reps):
        for (pb, pl), (hb, hl) in zip(ptrs, halos):
            t0 = time.perf_counter()
            hcm.check("hcm_mesh_chunk", lib.hcm_mesh_chunk(ctx.ptr, pb, pl, ctypes.byref(out.c)))
            t1 = time.perf_counter()
            hcm.check("hcm_mesh_halo", lib.hcm_mesh_halo(ctx.ptr, hb.ctypes.data, hl.ctypes.data, ctypes.byref(out.c)))
            t2 = time.perf_counter()
            lat["chunk"].append((t1 - t0) * 1e3)
            lat["halo"].append((t2 - t1) * 1e3)
    impl, kind = cpu_impl()
    sub = np.ascontiguousarray(nbr[:len(ptrs)])
    t0 = time.perf_counter()
    impl.mesh_many(blocks, light, sub, 1)
    ref_ms = (time.perf_counter() - t0) * 1e3 / len(ptrs)
    return {"chunks": len(ptrs), "calls_per_form": len(ptrs) * reps, "bit_exact": bad == 0,
            "mesh_chunk_p50_ms": percentile(lat["chunk"], 50), "mesh_chunk_p99_ms": percentile(lat["chunk"], 99),
            "mesh_halo_p50_ms": percentile(lat["halo"], 50), "mesh_halo_p99_ms": percentile(lat["halo"], 99),
            "upload_bytes_halo": int(hcm.HALO_BLOCKS * 2 + hcm.HALO_LIGHT),
            "cpu_generate_1_thread_ms": round(ref_ms, 4), "cpu_kind": kind,
            "note": "wall time per call from Python through ctypes (a few microseconds of call overhead), host "
                    "buffers in and out"}


def workloads_slice(blocks, light, nbr, j):
    return [blocks[s] for s in nbr[j]], [light[s] for s in nbr[j]]


def relight_leg(ctx, hcm, workloads, meshdigest, steps, cpu_threads, with_cpu, order="linear"):
    """The whole mesh task on the lit cave region (the cave region with glowstone in 0.05% of its air and
    sunlight heights from its columns): light rebuilt from blocks and heights (ChunkMeshTask.cpp:48-111),
    then meshing, one captured graph per step, with the job order the tuner chose for the cave batch
    (both kernels cut the job list into per-XCD parts). Bit-exact against the reference's mesh digests."""
    import hashlib

    from hypercraft_amd import tune

    blocks, heights, nbr = workloads.lit_cave_region()
    golden = np.load(os.path.join(ROOT, "tests", "golden", "light.npz"))
    n = nbr.shape[0]
    store = hcm.Store(ctx, blocks.shape[0])
    store.upload(blocks, np.zeros(blocks.shape, np.uint8))
    store.upload_sunlight(heights)
    perm = np.arange(n)
    if order != "linear":
        cost = None
        if "_" in order:  # the cost orders: the jobs' quads from one run in the region's order
            b0 = hcm.Batch(ctx, store, nbr)
            b0.set_light(hcm.LIGHT_REBUILD)
            b0.run()
            b0.sync()
            cost = tune.job_costs(b0)
            b0.close()
        perm = tune.job_order(order, workloads.region_job_positions(), cost)
    batch = hcm.Batch(ctx, store, nbr[perm])  # job k is the region's job perm[k]
    batch.set_light(hcm.LIGHT_REBUILD)
    for _ in range(3):
        batch.run()
    batch.sync()
    t0 = time.perf_counter()  # steps back to back
    batch.run_repeat(steps)
    ms = [batch.sync() / steps]
    dt = time.perf_counter() - t0
    quads, meshes = batch.totals()
    out = batch.download()
    bad = sum(meshdigest.digest(hcm.job_meshes(j, *out)) != golden["lit_cave_region_mesh_digest"][perm[j]].tobytes()
              for j in range(n))
    bad_light = sum(hashlib.sha256(batch.fetch_light(j).tobytes()).digest() !=
                    golden["lit_cave_region_halo_digest"][perm[j]].tobytes() for j in range(0, n, 16))
    batch.close()
    store.close()
    res = {"workload": "lit cave region: 4096 chunks, glowstone in 0.05% of the air, sunlight heights from the "
                       "columns; light rebuild + meshing per step",
           "chunks_per_s": round(n * steps / dt, 1), "ms_per_step": round(dt * 1e3 / steps, 4),
           "kernel_ms": round(float(np.mean(ms)), 4), "quads_per_step": int(quads), "job_order": order,
           "bit_exact": bool(bad == 0 and bad_light == 0 and (out[0]["status"] == 0).all())}
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import cpu_oracle

        if cpu_oracle.Reference.available():
            impl, kind = cpu_oracle.Reference(), "reference"
        else:
            impl, kind = cpu_oracle.Oracle(os.path.join(ROOT, "tests", "golden", "registry.bin")), "port"
        sample = nbr[:1024]
        passes = 0
        t0 = time.perf_counter()
        while True:
            impl.relight_many(golden["light_levels"], blocks, heights, sample, cpu_threads)
            passes += 1
            cdt = time.perf_counter() - t0
            if cdt >= 5.0:
                break
        res["cpu_baseline"] = {"value": round(passes * sample.shape[0] / cdt, 1), "unit": "chunks/s",
                               "cores": cpu_threads, "kind": kind,
                               "sample": f"{passes} passes over the first {sample.shape[0]} jobs, {cdt:.2f} s wall, "
                                         f"light rebuild + meshing per job"}
    return res


def draw_leg(hcm, workloads, batch, n_meshes, pos):
    """The draw list after meshing: mesh infos of the cave batch's meshes (WorldRenderer::PushChunkMesh),
    frustum culling into indirect draws (chunk_culling.comp) for 20 seeded cameras, and the depth hierarchy
    of a 1920 x 1080 depth image plus occlusion culling for the same cameras; wall time per call, host
    copies of the lists (and of the depth image) included."""
    batch.mesh_infos(pos)
    t0 = time.perf_counter()
    for _ in range(5):
        infos = batch.mesh_infos(pos)
    t_info = (time.perf_counter() - t0) / 5
    rng = np.random.default_rng(7)
    cams = []
    for _ in range(20):
        eye = rng.uniform(-40, 300, 3).astype(np.float32)
        target = rng.uniform(0, 256, 3)
        cams.append((eye, target, workloads.look_at_projection(eye, target)))
    kept = 0
    t0 = time.perf_counter()
    for eye, _, vp in cams:
        o, t = batch.cull(eye, vp)
        kept += len(o) + len(t)
    t_cull = (time.perf_counter() - t0) / len(cams)
    # occlusion: per camera a depth image near the depth of the point 60 blocks ahead of the eye, rippled
    yy, xx = np.mgrid[0:1080, 0:1920]
    wave = (np.sin(xx / 37.0) * np.cos(yy / 23.0)).astype(np.float32)
    kept_occ, t_depth, t_occ = 0, 0.0, 0.0
    for eye, target, vp in cams:
        ahead = 60.0 * (target - eye) / np.linalg.norm(target - eye)
        clip = vp.astype(np.float64).reshape(4, 4).T @ np.append(ahead, 1.0)
        c = np.float32(clip[2] / clip[3])
        img = np.clip(c + np.float32(0.5) * (np.float32(1) - c) * wave, 0, 1).astype(np.float32)
        t0 = time.perf_counter()
        batch.ctx.set_depth(img)
        t1 = time.perf_counter()
        o, t = batch.cull(eye, vp, occlusion=True)
        t_occ += time.perf_counter() - t1
        t_depth += t1 - t0
        kept_occ += len(o) + len(t)
    return {"meshes": int(n_meshes), "infos": int(len(infos)), "mesh_infos_ms": round(t_info * 1e3, 4),
            "cull_ms": round(t_cull * 1e3, 4), "kept_per_camera": round(kept / len(cams), 1),
            "depth_hierarchy_1080p_ms": round(t_depth / len(cams) * 1e3, 4),
            "cull_occlusion_ms": round(t_occ / len(cams) * 1e3, 4),
            "kept_per_camera_occlusion": round(kept_occ / len(cams), 1)}


def measured_traffic(order):
    """HBM bytes per persistent launch of the cave batch with its job list in `order`, from the committed
    rocprofv3 counter summary (None for an order it does not cover)."""
    path = os.path.join(ROOT, "profiles", "mesh_kernel_cave_region.md")
    key = f"traffic_bytes_per_launch[{order}]:"
    if os.path.exists(path):
        for line in open(path):
            if line.startswith(key):
                return int(line.split(":")[1])
    return None


def percentile(a, q):
    return round(float(np.percentile(np.asarray(a, np.float64), q)), 4) if len(a) else None


def mixed_leg(ctx, hcm, workloads, meshdigest, golden):
    """Uneven batches: a new batch's first run (cold: launched directly, wall time), its graph replay (warm,
    captured on the second run, HIP-event time), and per-chunk launches of the same jobs."""
    blocks, light, nbr, _ = workloads.mixed_pool()
    want = np.concatenate([golden["cave_region_digest"], golden["terrain_plants_region_digest"]])
    store = hcm.Store(ctx, blocks.shape[0])
    store.upload(blocks, light)
    batches = workloads.mixed_batches(nbr.shape[0])
    chunks = graph_ms = plain_ms = cold_s = 0.0
    bad = 0
    for k, jobs in enumerate(batches):
        b = hcm.Batch(ctx, store, nbr[jobs])
        t0 = time.perf_counter()
        b.run()
        b.sync()
        cold_s += time.perf_counter() - t0
        b.run()
        graph_ms += b.sync()
        out = b.download()
        for i, j in enumerate(jobs):
            bad += meshdigest.digest(hcm.job_meshes(i, *out)) != want[j].tobytes()
        b.run(hcm.RUN_PER_CHUNK)
        plain_ms += b.sync()
        chunks += jobs.size
        b.close()
        if k % 50 == 49:
            log(f"mixed: {k + 1} batches")
    store.close()
    return {"batches": len(batches), "chunks": int(chunks), "bit_exact": bad == 0,
            "graph_chunks_per_s": round(chunks / (graph_ms * 1e-3), 1),
            "cold_chunks_per_s": round(chunks / cold_s, 1),
            "per_chunk_launch_chunks_per_s": round(chunks / (plain_ms * 1e-3), 1),
            "graph_vs_per_chunk": round(plain_ms / graph_ms, 2)}


def serve_leg(ctx, hcm, workloads, tr, relight=False):
    """Replays the spiral load + edit trace against the server in real time; latency per request kind.
    With relight the server rebuilds light before meshing: loads carry sunlight heights (the first lit
    voxel of each terrain column) and edits remesh the chunks within light distance 14."""
    coords, blocks, light, events = tr["coords"], tr["blocks"], tr["light"], tr["events"]
    pool_quads = (16 << 30) // 44  # a 16 GB vertex pool (44 B per quad)
    srv = hcm.Server(ctx, max_chunks=coords.shape[0], max_batch=4096, max_inflight=2, pool_quads=pool_quads,
                     light_mode=hcm.LIGHT_REBUILD if relight else hcm.LIGHT_STORE)
    if relight:  # the terrain's light is sun 15 at and above the surface: heights = 16 - lit voxels per column
        lit = ((light.reshape(-1, 16, 16, 16) & 15) == 15).sum(1)  # [chunk, z, x]
        heights = (16 - lit).astype(np.int8).reshape(-1, 256)
    kinds = {}
    lag = 0.0
    log(f"serving: {len(events)} requests over {events[-1][0]:.2f} s")
    # the replay follows the trace's request times: a garbage-collection pause of this loop would hold
    # back thousands of requests and release them as one burst (the server's latency is measured from
    # each request's submission)
    gc_was = gc.isenabled()
    gc.disable()
    t0 = time.perf_counter()
    for tag, (t, kind, payload) in enumerate(events):
        now = time.perf_counter() - t0
        while now < t:
            now = time.perf_counter() - t0
        lag = max(lag, now - t)
        if kind == "load":
            c = coords[payload]
            if relight:
                srv.load_sunlit(int(c[0]), int(c[1]), int(c[2]), blocks[payload], heights[payload], tag)
            else:
                srv.load(int(c[0]), int(c[1]), int(c[2]), blocks[payload], light[payload], tag)
        else:
            srv.edit(*payload, tag)
        kinds[tag] = kind
    srv.drain()
    wall = time.perf_counter() - t0
    if gc_was:
        gc.enable()
    done = srv.poll()
    st = srv.stats()
    srv.close()
    lat = {"load": [], "edit": []}
    for c in done:
        lat[kinds[c["tag"]]].append(c["latency_ms"])
    res = {"requests": len(events), "completed": len(done), "loads": len(lat["load"]), "edits": len(lat["edit"]),
           "wall_s": round(wall, 3), "submit_lag_max_ms": round(lag * 1e3, 3),
           "chunks_meshed": int(st["chunks_meshed"]), "quads_meshed": int(st["quads_meshed"]),
           "chunks_meshed_per_s": round(st["chunks_meshed"] / wall, 1), "batches": int(st["batches"]),
           "gpu_busy_frac": round(st["gpu_ms"] * 1e-3 / wall, 4), "retries": int(st["retries"])}
    for k in ("edit", "load"):
        res[f"{k}_p50_ms"] = percentile(lat[k], 50)
        res[f"{k}_p99_ms"] = percentile(lat[k], 99)
        res[f"{k}_max_ms"] = round(max(lat[k]), 4) if lat[k] else None
    return res


def copy_calibration(nbytes, reps):
    """Device-to-device hipMemcpy of nbytes, reps times: the known-size copy that scales the HBM
    counters (FETCH_SIZE / WRITE_SIZE) read for the meshing kernel."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--tuned", default="", metavar="THREADS,WG_PER_CU,ORDER",
                    help="use this tuning instead of sweeping (e.g. 256,3,morton_cost), for profiling runs")
    ap.add_argument("--no-relight", action="store_true")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (for counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration, 5)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build", file=sys.stderr)
        sys.exit(2)

    from hypercraft_amd import hcm, tune, workloads

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import meshdigest

    hcm.require_device()
    log("generating the 4096-chunk cave batch")
    blocks, light, nbr = workloads.cave_region()
    golden = np.load(os.path.join(ROOT, "tests", "golden", "golden.npz"))
    ctx = hcm.Context(pool_quads=1 << 24)
    store = hcm.Store(ctx, blocks.shape[0])
    store.upload(blocks, light)
    n = nbr.shape[0]
    positions = workloads.region_job_positions()

    tuning = ctx.get_tuning()
    tuning["order"] = "linear"
    perm = np.arange(n)
    if args.tuned:
        t, w, o = args.tuned.split(",")
        log(f"tuning given: {t} threads, {w} workgroups per CU, order {o}")
        tuned = tune.autotune(ctx, store, nbr, workload="cave_region", positions=positions,
                              want=golden["cave_region_digest"], reps=1, candidates=[(int(t), int(w), o)], log=log)
    elif not args.no_tune:
        log("autotune")
        tuned = tune.autotune(ctx, store, nbr, workload="cave_region", positions=positions,
                              want=golden["cave_region_digest"], reps=3, log=log)
    if args.tuned or not args.no_tune:
        tuning = {"threads": tuned["threads"], "wg_per_cu": tuned["wg_per_cu"], "order": tuned["order"]}
        perm = tuned["perm"]
    batch = hcm.Batch(ctx, store, nbr[perm])  # job k is the region's job perm[k]

    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        batch.run()
    batch.sync()

    log(f"timing {args.steps} steps")
    # the timed region: K graph replays back to back, bracketed by device synchronisation on the host
    # and by HIP events on the context's stream (the kernel time per launch is their time over K)
    batch.sync()
    t0 = time.perf_counter()
    batch.run_repeat(args.steps)
    kernel_ms = [batch.sync() / args.steps]
    dt = time.perf_counter() - t0
    quads, meshes = batch.totals()
    out = batch.download()
    bad = 0
    for k in range(n):
        d = meshdigest.digest(hcm.job_meshes(k, *out))
        bad += d != golden["cave_region_digest"][perm[k]].tobytes()
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
                   "chunks": n, "quads_per_step": int(quads), "meshes_per_step": int(meshes),
                   "threads": tuning["threads"], "wg_per_cu": tuning["wg_per_cu"], "job_order": tuning["order"]},
        "quads_per_s": round(quads * args.steps / dt, 1),
        "bit_exact": bool(bit_exact),
        "plain_per_chunk_chunks_per_s": round(plain, 1) if plain else None,
        "batched_vs_plain": round(value / plain, 2) if plain else None,
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": measured_traffic(tuning["order"]),
                     "traffic_source": f"profiles/mesh_kernel_cave_region.md, job order {tuning['order']} "
                                       "(FETCH_SIZE + WRITE_SIZE, copy-scaled)",
                     "kernel_ms": round(k_ms, 4), "algorithmic_bytes": int(alg_bytes)},
    }
    res["draw_lists"] = draw_leg(hcm, workloads, batch, int(meshes), positions[perm])
    batch.close()
    store.close()
    if not args.no_mixed:
        log("mixed uneven batches")
        res["mixed_batches"] = mixed_leg(ctx, hcm, workloads, meshdigest, golden)
    if not args.no_serve:
        log("serving trace")
        log("serving: generating the spiral trace")
        trace = workloads.spiral_trace()
        res["serving"] = serve_leg(ctx, hcm, workloads, trace)
        log("serving trace with light rebuild")
        res["serving_relight"] = serve_leg(ctx, hcm, workloads, trace, relight=True)
        log("serving trace at ten times the load rate (200k chunks/s)")
        res["serving_saturated"] = serve_leg(ctx, hcm, workloads, workloads.spiral_trace(load_rate=200000.0))
    if not args.no_relight:
        log("light rebuild + meshing")
        res["relight"] = relight_leg(ctx, hcm, workloads, meshdigest, 20, min(host_cpus()[0], CPU_SHARE),
                                     not args.no_cpu, tuning.get("order", "linear"))
    if not args.no_dropin:
        log("drop-in call latency")
        res["dropin"] = dropin_leg(ctx, hcm, blocks, light, nbr, golden["cave_region_digest"])
    if not args.no_cpu:
        log("CPU baseline")
        res["cpu_baseline"] = cpu_baseline(blocks, light, nbr)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
