"""Write tests/golden/cornell_box.npz from the reference's resources/cornell_box.obj.

The OBJ is a data file the reference ships for its tests (resources/; CC BY
3.0, (c) 2009 Morgan McGuire, https://casual-effects.com/data/).  This script
reads its `o`, `v` and `f` records and stores the mesh as arrays, so that GPU
tests (which run where the reference is absent) can load it:

  vertices  float32 [V, 3]   positions in file order
  triangles uint32  [T, 3]   faces fan-triangulated (v0, vk, vk+1), file order
  object    uint32  [T]      index of the `o` record each triangle belongs to
  names     str     [O]      the `o` record names

Run from the repository root:  python tests/golden/make_cornell_box.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/resources/cornell_box.obj")
OUT = Path(__file__).resolve().parent / "cornell_box.npz"


def parse(text: str):
    verts, tris, obj, names = [], [], [], []
    for line in text.splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "o":
            names.append(parts[1] if len(parts) > 1 else f"object{len(names)}")
        elif parts[0] == "v":
            verts.append([float(x) for x in parts[1:4]])
        elif parts[0] == "f":
            # vertex index of each corner (1-based; negative = relative to the end)
            idx = []
            for c in parts[1:]:
                k = int(c.split("/")[0])
                idx.append(k - 1 if k > 0 else len(verts) + k)
            for j in range(1, len(idx) - 1):
                tris.append([idx[0], idx[j], idx[j + 1]])
                obj.append(max(len(names) - 1, 0))
    return (np.asarray(verts, np.float32), np.asarray(tris, np.uint32), np.asarray(obj, np.uint32),
            np.asarray(names))


def main() -> int:
    v, t, o, names = parse(SRC.read_text())
    np.savez(OUT, vertices=v, triangles=t, object=o, names=names)
    print(f"{OUT}: {v.shape[0]} vertices, {t.shape[0]} triangles, {len(names)} objects")
    return 0


if __name__ == "__main__":
    sys.exit(main())
