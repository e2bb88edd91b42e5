"""The block registry: the library's built-in table against the one exported from the reference,
plus the registry invariants of the reference's own tests (block/test/test_block.cpp:10-67)."""
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN, REGISTRY


def load_fixture():
    d = open(REGISTRY, "rb").read()
    assert d[:8] == b"HCREG01\0"
    n_rows, n_meshes, n_tex, _ = struct.unpack_from("<4I", d, 8)
    off = 24
    row_of = np.frombuffer(d, np.uint16, 65536, off)
    off += 131072
    rows = np.frombuffer(d, np.uint8, 16 * n_rows, off).reshape(n_rows, 16)
    off += 16 * n_rows
    meshes = np.frombuffer(d, np.uint8, 400 * n_meshes, off).reshape(n_meshes, 400)
    off += 400 * n_meshes
    tex_flags = np.frombuffer(d, np.uint8, n_tex, off)
    off += n_tex
    src, flow, ftex = struct.unpack_from("<BBH", d, off)
    return {"row_of": row_of, "rows": rows, "meshes": meshes, "tex_flags": tex_flags, "liquid": (src, flow, ftex)}


@pytest.fixture(scope="module")
def fixture():
    return load_fixture()


@pytest.fixture(scope="module")
def builtin():
    from hypercraft_amd import hcm

    return hcm.registry_arrays()


def test_builtin_registry_equals_reference_export(fixture, builtin):
    # the same distinct rows, in the same first-use order, for every one of the 65,536 Block values
    assert np.array_equal(builtin["row_of"], fixture["row_of"])
    assert np.array_equal(builtin["rows"], fixture["rows"])
    assert np.array_equal(builtin["meshes"], fixture["meshes"])
    assert np.array_equal(builtin["tex_flags"], fixture["tex_flags"])
    assert tuple(builtin["liquid"]) == tuple(fixture["liquid"])


def test_builtin_light_levels_equal_reference_export(builtin):
    # Block::GetLightLevel for every Block value, as the reference resolves it (tests/golden/light.npz)
    want = np.load(os.path.join(GOLDEN, "light.npz"))["light_levels"]
    assert np.array_equal(builtin["light_level"], want)


def _row(reg, data):
    r = reg["rows"][reg["row_of"][data]]
    tex = np.frombuffer(r[:12].tobytes(), np.uint16)
    return tex, int(r[12]), int(r[13]), int(np.frombuffer(r[14:16].tobytes(), np.uint16)[0])


def test_air_invariants(fixture):
    # test_block.cpp:10-23: id 0 never collides, passes light both ways and has no custom mesh
    for meta in range(256):
        tex, transparency, collision, mesh = _row(fixture, meta << 8)
        assert collision == 0 and transparency == 2 and mesh == 0xFFFF and not tex.any()


def test_custom_mesh_invariants(fixture):
    # test_block.cpp:25-48: face count in [1, 16], non-empty textures, axis == render_face >> 1
    for m in fixture["meshes"]:
        face_count, kind = struct.unpack_from("<2I", m.tobytes(), 0)
        if kind == 1:  # liquid flow meshes are generated per cell (face_count 0 in the table)
            continue
        assert 1 <= face_count <= 16
        for f in range(face_count):
            axis, lf, rf, _, tex = struct.unpack_from("<4BH", m.tobytes(), 16 + 24 * f)
            assert tex & 0x1FFF != 0
            assert axis == rf >> 1


def test_vertical_light_implies_indirect(fixture):
    # test_block.cpp:58-67: transparency 2 (vertical pass) implies != 0 (indirect pass)
    for r in fixture["rows"]:
        assert r[12] in (0, 1, 2)


def test_texture_flag_table(fixture):
    f = fixture["tex_flags"]
    assert len(f) == 53
    assert f[0] & 1  # texture 0 counts as transparent
    assert f[23] == 7  # water: transparent, transparent pass, liquid
    assert (f & 4).nonzero()[0].tolist() == [23]
