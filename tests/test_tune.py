"""The autotuner's job orders (no GPU): each is a permutation; the Morton order makes each eighth of the
list (one XCD's part of the kernel's job list) a compact block of the region; a cost order keeps the
parts of its base order and sorts each part by descending cost, ties in base order; a tail-cost order
sorts only the second half of each part."""
import numpy as np

from hypercraft_amd import tune, workloads


def test_orders_are_permutations_with_the_base_parts():
    pos = workloads.region_job_positions()
    cost = np.random.default_rng(1).integers(0, 50, pos.shape[0])
    n = pos.shape[0]
    for name in tune.ORDERS:
        perm = tune.job_order(name, pos, cost)
        assert sorted(perm.tolist()) == list(range(n)), name
        base = tune.job_order(name.split("_")[0], pos)
        for q in range(tune.QUEUES):
            part = slice(q * n // tune.QUEUES, (q + 1) * n // tune.QUEUES)
            assert set(perm[part].tolist()) == set(base[part].tolist()), (name, q)
            if name.endswith("cost"):
                k = (part.stop - part.start) // 2 if name.endswith("_tailcost") else 0
                assert perm[part][:k].tolist() == base[part][:k].tolist(), (name, q)  # the kept head
                sub = perm[part][k:]
                c = cost[sub]
                assert set(sub.tolist()) == set(base[part][k:].tolist()), (name, q)
                assert (np.diff(c) <= 0).all(), (name, q)
                where = {j: i for i, j in enumerate(base[part].tolist())}
                for v in np.unique(c):  # equal costs keep the base order
                    idx = sub[c == v]
                    assert [where[j] for j in idx.tolist()] == sorted(where[j] for j in idx.tolist())


def test_morton_parts_are_compact_blocks():
    pos = workloads.region_job_positions()  # 16^3 region: each eighth is an 8^3 block
    perm = tune.job_order("morton", pos)
    n = pos.shape[0]
    for q in range(tune.QUEUES):
        p = pos[perm[q * n // 8:(q + 1) * n // 8]]
        assert (p.max(0) - p.min(0) == 7).all()
    assert tune.job_order("linear", pos).tolist() == list(range(n))


def test_unknown_orders_are_refused():
    pos = workloads.region_job_positions()
    for bad in ("hilbert", "morton_foo", "linear_tail"):
        try:
            tune.job_order(bad, pos, np.zeros(pos.shape[0]))
        except ValueError:
            continue
        raise AssertionError(bad)
    try:
        tune.job_order("morton_cost", pos)  # a cost order without costs
    except ValueError:
        return
    raise AssertionError("morton_cost without costs")
