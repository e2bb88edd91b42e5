"""Digests of mesh lists for comparing meshers (test helper).

A job's digest is SHA-256 over, for each mesh in order: vertex count, index count, AABB (6 words)
and transparent flag as little-endian u32, then the vertex words, then the u16 indices. Equal
digests mean byte-identical outputs in the reference's mesh order.
"""
import hashlib

import numpy as np


def digest(meshes):
    h = hashlib.sha256()
    for m in meshes:
        hdr = np.array([m.vertices.shape[0], m.indices.shape[0], *m.aabb, m.transparent], np.uint32)
        h.update(hdr.tobytes())
        h.update(np.ascontiguousarray(m.vertices, np.uint32).tobytes())
        h.update(np.ascontiguousarray(m.indices, np.uint16).tobytes())
    return h.digest()


def quads(meshes):
    return sum(m.vertices.shape[0] // 4 for m in meshes)


def input_digest(blocks, light, nbr):
    h = hashlib.sha256()
    for a in (blocks, light, nbr):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()
