"""MI355X (gfx950) chunk mesher for HyperCraft: a drop-in for BlockMeshAlgo::Generate.

The product is libhcm.so (C ABI: include/hcm.h); this package builds it (build.py), binds it
(hcm.py) and generates seeded workloads (workloads.py).
"""
