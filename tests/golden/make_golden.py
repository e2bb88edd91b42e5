"""Writes the golden fixtures under tests/golden/ from the REAL reference mesher.

Run in a container where /root/reference exists and `make -C oracle ref` has built
oracle/_ref/libhcref.so (the reference's BlockMeshAlgo::Generate and block property
table, compiled from its own sources). The fixtures are data only: the flattened
registry the reference's property table resolves to, and seeded input chunks with
the exact meshes the reference produced for them.

    python tests/golden/make_golden.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import cpu_oracle  # noqa: E402


def main():
    ref = cpu_oracle.Reference()
    ref.export_registry(os.path.join(HERE, "registry.bin"))
    print("wrote registry.bin")


if __name__ == "__main__":
    main()
