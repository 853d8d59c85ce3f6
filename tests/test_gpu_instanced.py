"""GPU parity on the two-level instanced scene (64 soups x 4,096 instances,
2048x2048 primary pass, 8,388,608 secondary rays) and a reduced version.

Primary hits, and secondary hits from them, must equal the oracle's
two-level traversal of the same (downloaded) TLAS and BLASes: every ray of
the reduced scene, and an evenly spaced subsample of the full one.
"""
import numpy as np
import pytest

import oracle
from radeonrays_sdk_amd import api, scenes
from tests.gpu_util import RR_INVALID, hits_equal, nodes_equal, say, tlas_tables_equal, to_dev

pytestmark = pytest.mark.gpu


def _build(ctx, dev, w):
    import torch

    blases = [ctx.build_geometry(to_dev(v, dev), to_dev(i, dev), restructure=True) for v, i in w.meshes]
    tlas = ctx.build_scene([blases[m] for m in w.mesh_of_instance], w.transforms)
    torch.cuda.synchronize()
    host = {b.address: b.buffer.cpu().numpy() for b in blases}
    buf = tlas.buffer.cpu().numpy()
    # The oracle's TLAS over the same BLASes: nodes, transforms and BLAS
    # table equal to the GPU's bit for bit; every check below traces it.
    ref = oracle.build_tlas(w.transforms, [blases[m].address for m in w.mesh_of_instance],
                            [host[blases[m].address] for m in w.mesh_of_instance])
    ok, msg = nodes_equal(buf, ref)
    assert ok, msg
    ok, msg = tlas_tables_equal(buf, ref)
    assert ok, msg
    return tlas, ref, host


@pytest.mark.parametrize("size", ["reduced", "full"])
def test_instanced_primary_and_secondary(ctx, dev, size):
    import torch

    if size == "reduced":
        w = scenes.instanced(n_meshes=8, n_inst=64, lo=1000, hi=20000, width=256, height=256, trans=4.0)
        step = 1
    else:
        w = scenes.instanced()
        step = 127
    say(f"instanced {size}: {len(w.meshes)} meshes, {len(w.mesh_of_instance)} instances, "
        f"{sum(i.shape[0] for _, i in w.meshes)} tris")
    tlas, buf, host = _build(ctx, dev, w)
    assert oracle.is_valid(buf)
    primary = scenes.camera_rays(w.width, w.height, w.eye)
    hits = ctx.intersect(tlas, to_dev(primary, dev), "closest", "full_t")
    torch.cuda.synchronize()
    h = api.decode_hits(hits, "full_t")
    sub = np.arange(0, primary.shape[0], step)
    ref = oracle.trace(buf, primary[sub], blas_map=host)
    ok, msg = hits_equal({k: a[sub] for k, a in h.items()}, ref)
    assert ok, "primary: " + msg
    hit_frac = (h["inst_id"] != RR_INVALID).mean()
    assert 0.05 < hit_frac < 0.95, hit_frac

    sec = scenes.secondary_rays(w, primary, h["prim_id"], h["inst_id"], h["t"])
    assert sec.shape[0] == 2 * primary.shape[0]
    say(f"secondary: {sec.shape[0]} rays, {int((sec[:, 7] < 0).sum())} no-op")
    sh = ctx.intersect(tlas, to_dev(sec, dev), "closest", "full")
    torch.cuda.synchronize()
    g = api.decode_hits(sh, "full")
    sub = np.arange(0, sec.shape[0], step)
    ref = oracle.trace(buf, sec[sub], blas_map=host)
    ok, msg = hits_equal({k: a[sub] for k, a in g.items()}, ref, ("prim_id", "inst_id", "u", "v"))
    assert ok, "secondary: " + msg
    # No-op rays (primary miss) never hit.
    assert (g["inst_id"][sec[:, 7] < 0] == RR_INVALID).all()
