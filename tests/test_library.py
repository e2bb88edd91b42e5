"""libhcm.so loads and exports every entry point declared in include/*.h (no GPU needed);
the workload generators are deterministic."""
import os
import re

import numpy as np

from conftest import ROOT


def declared_functions():
    names = set()
    for h in ("hcm.h", "hcm_workload.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(hcm_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    from hypercraft_amd import hcm

    lib = hcm.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    bound = {s[0] for s in hcm.SIGNATURES}
    assert names == bound, names ^ bound


def test_generators_are_deterministic_and_match_golden_inputs():
    import meshdigest
    from hypercraft_amd import workloads

    golden = np.load(os.path.join(ROOT, "tests", "golden", "golden.npz"))
    b, l, n = workloads.terrain_chunk()
    assert meshdigest.input_digest(b, l, n).encode() == golden["terrain_chunk_input"].tobytes()
    b, l, n = workloads.random_dense(n_jobs=1024)
    assert meshdigest.input_digest(b, l, n).encode() == golden["random_dense_input"].tobytes()


def test_random_dense_mix():
    from hypercraft_amd import hcm, workloads

    b, l, n = workloads.random_dense(n_jobs=16)
    opaque = set(hcm.opaque_values().tolist())
    assert len(opaque) == 33
    v = b.ravel()
    air = (v == 0).mean()
    op = np.isin(v, list(opaque)).mean()
    water = (v == 14).mean()
    assert abs(air - 0.50) < 0.01 and abs(op - 0.35) < 0.01 and abs(water - 0.05) < 0.005
