"""GPU test of the bvh_analyzer export (rrxWriteAnalyzerDump).

The files the library writes for a GPU-built BLAS must equal, byte for byte,
the oracle's writer applied to the same (downloaded) buffer, and the
reference's own bvh_analyzer (built from its sources into oracle/_ref, when
present) must read them and find the BVH valid with the oracle's SAH.
"""
import math
import os
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

import oracle
from radeonrays_sdk_amd import rr, scenes
from tests.gpu_util import say, to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["soup", "terrain"])
def test_analyzer_dump_matches_oracle_writer(ctx, dev, kind):
    import torch

    say(f"rrxWriteAnalyzerDump {kind}")
    if kind == "soup":
        w = scenes.soup_camera(n_tris=4096, width=32, height=32)
        v, i, rays = w.vertices, w.indices, w.rays
    else:
        v, i = scenes.heightfield_mesh(65)
        rays = scenes.camera_rays(32, 32, (0.5, 0.35, -0.2))
    accel = ctx.build_geometry(to_dev(v, dev), to_dev(i, dev), restructure=True)
    drays = to_dev(rays, dev)
    torch.cuda.synchronize()
    buf = accel.buffer.cpu().numpy()
    with tempfile.TemporaryDirectory() as a, tempfile.TemporaryDirectory() as b:
        a, b = Path(a), Path(b)
        ctx.write_analyzer_dump(accel, a / "bvh.bin", a / "cfg.txt", rays=drays, width=32, height=32,
                                rays_path=a / "rays.bin")
        oracle.write_analyzer_dump(buf, rays, 32, 32, b / "bvh.bin", b / "rays.bin", b / "cfg.txt")
        for f in ("bvh.bin", "rays.bin"):
            assert (a / f).read_bytes() == (b / f).read_bytes(), f
        assert (a / "cfg.txt").read_text().replace(str(a), "D") == (b / "cfg.txt").read_text().replace(str(b), "D")
        lines = (a / "cfg.txt").read_text().splitlines()
        assert lines[1] == "vkbvh2" and int(lines[2]) == i.shape[0] - 1 and int(lines[3]) == i.shape[0]
        if oracle.REF_ANALYZER.exists():
            out = subprocess.run([str(oracle.REF_ANALYZER), str(a / "cfg.txt")], cwd=a, capture_output=True, text=True,
                                 timeout=60, env={**os.environ, "OMP_NUM_THREADS": "2"})
            assert out.returncode == 0, out.stderr
            stats = dict(line.split(": ", 1) for line in out.stdout.splitlines() if ": " in line)
            assert stats["is_valid"].strip() == "1"
            assert math.isclose(float(stats["sah"]), oracle.sah(buf), rel_tol=1e-5)


def test_analyzer_dump_rejects_scenes(ctx, dev):
    v, i = scenes.triangle_soup(100, scenes.rng_for(9))
    blas = ctx.build_geometry(to_dev(v, dev), to_dev(i, dev))
    tlas = ctx.build_scene([blas], np.eye(3, 4, dtype=np.float32)[None])
    with tempfile.TemporaryDirectory() as a:
        with pytest.raises(rr.RRError) as e:
            ctx.write_analyzer_dump(tlas, Path(a) / "b", Path(a) / "c")
        assert e.value.code == rr.RR_ERROR_NOT_IMPLEMENTED
