"""A compiled C++ client of the drop-in (tests/cpp_client/basic_flow.cpp,
built with the host compiler against include/ only and linked with
libradeonrays.so) replays the reference's basic_test.h flows through the C
ABI and the reference's submission model (record, rrSumbitCommandStream,
rrWaitEvent): CreateContext, BuildSingleTriangle, BuildObj (the Cornell box,
scaled by 20 so the reference's fan origin lies inside it), UpdateObj (every
vertex y -= 40, refit) and BuildObj2Level (a scene of instances).  Every
result it writes is compared with the oracle bit for bit.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from radeonrays_sdk_amd import scenes
from tests.gpu_util import hits_equal, say
from tests.test_gpu_reference_cases import cornell, fan_rays

pytestmark = pytest.mark.gpu

CLIENT_DIR = Path(__file__).resolve().parent / "cpp_client"


def _client() -> Path:
    exe = CLIENT_DIR / "basic_flow"
    if not exe.exists():
        subprocess.run(["make", "-s", "-C", str(CLIENT_DIR)], check=True)
    return exe


def _read_results(path: Path) -> list:
    raw = path.read_bytes()
    assert raw[:4] == b"RRCO"
    pos, out = 4, []
    while pos < len(raw):
        rec, count = np.frombuffer(raw, np.uint32, 2, pos)
        pos += 8
        out.append(np.frombuffer(raw, np.uint8, int(rec) * int(count), pos).reshape(int(count), int(rec)))
        pos += int(rec) * int(count)
    return out


def _full_hits(recs: np.ndarray) -> dict:
    w = recs.view(np.uint32).reshape(-1, 4)
    f = recs.view(np.float32).reshape(-1, 4)
    return {"u": f[:, 0], "v": f[:, 1], "inst_id": w[:, 2], "prim_id": w[:, 3]}


def test_cpp_client_basic_flows(tmp_path):
    v, i, _, _ = cornell(20.0)
    rng = scenes.rng_for(70)
    tri_rays = scenes.make_rays(rng.uniform(-0.5, 0.5, (2000, 3)).astype(np.float32) * np.float32([1, 1, 0]) +
                                np.float32([0, 0, -1]), np.tile(np.float32([0, 0, 1]), (2000, 1)))
    rays = np.concatenate([fan_rays(512), tri_rays,
                           scenes.make_rays(rng.uniform(-25, 25, (20000, 3)).astype(np.float32),
                                            scenes.unit_sphere(20000, rng))])
    xf = np.concatenate([np.eye(3, 4, dtype=np.float32)[None], scenes.random_transforms(rng, 7, trans=30.0)])
    say(f"C++ client: {i.shape[0]} tris, {rays.shape[0]} rays, {xf.shape[0]} instances")
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(b"RRCI")
        f.write(np.array([v.shape[0], i.shape[0], rays.shape[0], xf.shape[0]], np.uint32).tobytes())
        f.write(v.astype(np.float32).tobytes())
        f.write(i.astype(np.uint32).tobytes())
        f.write(rays.astype(np.float32).tobytes())
        f.write(xf.astype(np.float32).tobytes())
    out = tmp_path / "out.bin"
    proc = subprocess.run([str(_client()), str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    res = _read_results(out)
    assert len(res) == 6
    fields = ("prim_id", "inst_id", "u", "v")

    # BuildSingleTriangle.
    tv = np.array([[0, -0.7, 0.7], [-0.7, 0.7, 1.0], [0.7, 0.7, 1.0]], np.float32)
    tri = oracle.build_blas(tv, np.array([[0, 1, 2]], np.int32))
    ref = oracle.trace(tri, rays)
    ok, msg = hits_equal(_full_hits(res[0]), ref, fields)
    assert ok, "single triangle: " + msg
    assert (ref["prim_id"] == 0).sum() > 500

    # BuildObj: closest and any hit (the fast build equals the oracle's node
    # for node, so any-hit visits leaves in the same order).
    blas = oracle.build_blas(v, i)
    ref = oracle.trace(blas, rays)
    ok, msg = hits_equal(_full_hits(res[1]), ref, fields)
    assert ok, "BuildObj closest: " + msg
    ok, msg = hits_equal(_full_hits(res[2]), oracle.trace(blas, rays, any_hit=True), fields)
    assert ok, "BuildObj any: " + msg
    assert (ref["prim_id"] != 0xFFFFFFFF).sum() > 200_000

    # UpdateObj: y -= 40 and refit.
    moved = v.copy()
    moved[:, 1] -= np.float32(40.0)
    ok, msg = hits_equal(_full_hits(res[3]), oracle.trace(oracle.refit(blas, moved, i), rays), fields)
    assert ok, "UpdateObj: " + msg

    # BuildObj2Level: closest FULL_HIT and INSTANCE_ID.
    tlas = oracle.build_tlas(xf, [1] * xf.shape[0], [blas] * xf.shape[0])
    ref = oracle.trace(tlas, rays, blas_map={1: blas})
    ok, msg = hits_equal(_full_hits(res[4]), ref, fields)
    assert ok, "BuildObj2Level: " + msg
    assert np.array_equal(res[5].view(np.uint32).reshape(-1), ref["inst_id"])
    assert len(np.unique(ref["inst_id"])) > 3
