"""Seeded synthetic workloads for the build-and-intersect path.

Each generator returns numpy arrays (float32 vertices [V, 3], int32 triangle
indices [T, 3], float32 rays [R, 8] laid out as RRRay: origin, min_t,
direction, max_t).  Seeds are 0x52520000 + workload index; randomness comes
from numpy's PCG64 generator, normals in float64 rounded to float32.

Workloads (the survey's configs 0-4):
  soup_camera      65,536-triangle soup, 1024x1024 pinhole camera rays
  heightfield      724x724-vertex fBm heightfield (1,045,458 triangles),
                   4,194,304 rays from a slab above it, uniform directions
  instanced        64 soups x 4,096 instances, secondary rays
  batched          8,192 small soups (batched build/refit)
  serving          32 resident scenes and a Poisson request stream
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

SEED_BASE = 0x52520000


def rng_for(workload: int, stream: int = 0) -> np.random.Generator:
    return np.random.default_rng([SEED_BASE + workload, stream])


# --- geometry ----------------------------------------------------------------
def triangle_soup(n_tris: int, rng: np.random.Generator, sigma: float = 0.01):
    """Centroids uniform in [0,1]^3, each vertex = centroid + N(0, sigma^2) per axis."""
    c = rng.uniform(0.0, 1.0, size=(n_tris, 1, 3))
    v = (c + rng.normal(0.0, sigma, size=(n_tris, 3, 3))).astype(np.float32).reshape(-1, 3)
    idx = np.arange(3 * n_tris, dtype=np.int32).reshape(n_tris, 3)
    return v, idx


def _hash2(ix, iz, salt):
    """Integer lattice hash -> uniform float in [0, 1)."""
    h = (ix.astype(np.uint32) * np.uint32(0x8DA6B343)) ^ (iz.astype(np.uint32) * np.uint32(0xD8163841)) ^ np.uint32(salt)
    h ^= h >> np.uint32(16)
    h *= np.uint32(0x7FEB352D)
    h ^= h >> np.uint32(15)
    h *= np.uint32(0x846CA68B)
    h ^= h >> np.uint32(16)
    return (h >> np.uint32(8)).astype(np.float64) / float(1 << 24)


def fbm_heights(res: int, seed: int, octaves: int = 6, base_freq: float = 4.0, lo: float = 0.0, hi: float = 0.2):
    """fBm value noise on a res x res grid over [0,1]^2: lattice hash values,
    smoothstep interpolation, lacunarity 2, gain 0.5, rescaled to [lo, hi]."""
    x = np.linspace(0.0, 1.0, res)
    X, Z = np.meshgrid(x, x, indexing="xy")
    total = np.zeros_like(X)
    amp, freq = 1.0, base_freq
    with np.errstate(over="ignore"):
        for o in range(octaves):
            fx, fz = X * freq, Z * freq
            ix, iz = np.floor(fx).astype(np.int64), np.floor(fz).astype(np.int64)
            tx, tz = fx - ix, fz - iz
            sx, sz = tx * tx * (3 - 2 * tx), tz * tz * (3 - 2 * tz)
            salt = (seed * 0x9E3779B1 + o * 0x85EBCA77) & 0xFFFFFFFF
            v00, v10 = _hash2(ix, iz, salt), _hash2(ix + 1, iz, salt)
            v01, v11 = _hash2(ix, iz + 1, salt), _hash2(ix + 1, iz + 1, salt)
            total += amp * ((v00 * (1 - sx) + v10 * sx) * (1 - sz) + (v01 * (1 - sx) + v11 * sx) * sz)
            amp *= 0.5
            freq *= 2.0
    t = (total - total.min()) / max(total.max() - total.min(), 1e-12)
    return lo + (hi - lo) * t


def heightfield_mesh(res: int = 724, seed: int = SEED_BASE + 1):
    """res x res vertices on [0,1]^2 in x,z with fBm heights y; 2 triangles per cell."""
    h = fbm_heights(res, seed)
    x = np.linspace(0.0, 1.0, res)
    X, Z = np.meshgrid(x, x, indexing="xy")
    v = np.stack([X, h, Z], axis=-1).astype(np.float32).reshape(-1, 3)
    i = np.arange(res - 1)
    I, J = np.meshgrid(i, i, indexing="xy")  # I along x, J along z
    a = (J * res + I).ravel()
    b = a + 1
    c = a + res
    d = c + 1
    tris = np.empty((2 * a.size, 3), dtype=np.int32)
    tris[0::2] = np.stack([a, b, c], axis=1)
    tris[1::2] = np.stack([b, d, c], axis=1)
    return v, tris


# --- rays --------------------------------------------------------------------
def make_rays(origins, dirs, tmin=0.0, tmax=1e30):
    n = origins.shape[0]
    r = np.empty((n, 8), dtype=np.float32)
    r[:, 0:3] = origins
    r[:, 3] = tmin
    r[:, 4:7] = dirs
    r[:, 7] = tmax
    return r


def unit_sphere(n, rng):
    d = rng.normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


def camera_rays(width, height, eye, fov_deg=45.0, tmin=0.0, tmax=1e30):
    """Pinhole camera at `eye` looking along +z with +y up, vertical FOV fov_deg, pixel centres."""
    t = math.tan(math.radians(fov_deg) * 0.5)
    aspect = width / height
    px = (np.arange(width) + 0.5) / width * 2 - 1
    py = 1 - (np.arange(height) + 0.5) / height * 2
    PX, PY = np.meshgrid(px * t * aspect, py * t, indexing="xy")
    d = np.stack([PX, PY, np.ones_like(PX)], axis=-1).reshape(-1, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.broadcast_to(np.asarray(eye, dtype=np.float64), d.shape)
    return make_rays(o.astype(np.float32), d.astype(np.float32), tmin, tmax)


# --- workloads ---------------------------------------------------------------
@dataclass
class MeshWorkload:
    name: str
    vertices: np.ndarray
    indices: np.ndarray
    rays: np.ndarray


def soup_camera(n_tris: int = 65536, width: int = 1024, height: int = 1024) -> MeshWorkload:
    """Config 0: triangle soup, primary rays from (0.5, 0.5, -1.5) along +z, 45 deg FOV, t in [0, 1e30]."""
    rng = rng_for(0)
    v, i = triangle_soup(n_tris, rng)
    rays = camera_rays(width, height, (0.5, 0.5, -1.5))
    return MeshWorkload("soup_camera", v, i, rays)


def heightfield(res: int = 724, n_rays: int = 4 * 1024 * 1024) -> MeshWorkload:
    """Config 1: fBm heightfield; origins U([0,1]x[0.25,0.5]x[0,1]), directions uniform on S^2."""
    v, i = heightfield_mesh(res)
    rng = rng_for(1)
    o = rng.uniform([0.0, 0.25, 0.0], [1.0, 0.5, 1.0], size=(n_rays, 3)).astype(np.float32)
    rays = make_rays(o, unit_sphere(n_rays, rng))
    return MeshWorkload("heightfield", v, i, rays)


def perturbed_heights(vertices: np.ndarray, frame: int, sigma: float = 1e-3) -> np.ndarray:
    """Refit frame `frame`: every vertex height = base + N(0, sigma^2), fresh per frame."""
    rng = rng_for(1, 1000 + frame)
    out = vertices.copy()
    out[:, 1] = (vertices[:, 1].astype(np.float64) + rng.normal(0.0, sigma, size=vertices.shape[0])).astype(np.float32)
    return out


def log_uniform_ints(rng, n, lo, hi):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size=n)).astype(np.int64)


def random_transforms(rng, n, scale=(0.5, 2.0), trans=50.0):
    """Uniform rotation (Shoemake quaternions), uniform scale, translation U[-trans, trans]^3 -> [n, 3, 4]."""
    u1, u2, u3 = rng.uniform(size=(3, n))
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q.T
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], axis=1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], axis=1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], axis=1)], axis=1)
    s = rng.uniform(scale[0], scale[1], size=(n, 1, 1))
    t = rng.uniform(-trans, trans, size=(n, 3, 1))
    return np.concatenate([R * s, t], axis=2).astype(np.float32)


# --- serving stream ------------------------------------------------------------
@dataclass
class ServingScene:
    kind: str  # "soup" (triangle soup in [0,1]^3) or "terrain" (fBm heightfield)
    vertices: np.ndarray
    indices: np.ndarray


@dataclass
class ServingRequest:
    scene: int
    pool: int    # ray pool index (SERVING_POOLS)
    offset: int  # first ray of the pool
    count: int
    query: str   # "closest" or "any"


# Ray pools the requests draw contiguous slices from: primary-camera and
# uniform-sphere rays for each scene kind.
SERVING_POOLS = ("soup_camera", "soup_sphere", "terrain_camera", "terrain_sphere")


def serving_scenes(n_scenes: int = 32, lo: int = 10_000, hi: int = 2_000_000) -> list:
    """Resident scenes of the serving stream: triangle counts log-uniform in
    [lo, hi]; even-numbered scenes are soups (as the 65,536-triangle soup),
    odd-numbered ones fBm heightfields with about that many triangles."""
    rng = rng_for(4)
    sizes = log_uniform_ints(rng, n_scenes, lo, hi)
    out = []
    for k, n in enumerate(sizes):
        if k % 2 == 0:
            v, i = triangle_soup(int(n), rng_for(4, 1 + k))
            out.append(ServingScene("soup", v, i))
        else:
            res = max(2, int(round(math.sqrt(n / 2.0))) + 1)
            v, i = heightfield_mesh(res, seed=SEED_BASE + 4 + 1000 * k)
            out.append(ServingScene("terrain", v, i))
    return out


def serving_ray_pools(pool_rays: int = 1 << 20) -> list:
    """The four ray pools, pool_rays rays each (a square camera image when
    pool_rays is a square)."""
    side = int(math.isqrt(pool_rays))
    rng = rng_for(4, 100)
    soup_cam = camera_rays(side, pool_rays // side, (0.5, 0.5, -1.5))
    o = rng.uniform(0.0, 1.0, size=(pool_rays, 3)).astype(np.float32)
    soup_sph = make_rays(o, unit_sphere(pool_rays, rng))
    terr_cam = camera_rays(side, pool_rays // side, (0.5, 0.35, -0.2))
    o = rng.uniform([0.0, 0.25, 0.0], [1.0, 0.5, 1.0], size=(pool_rays, 3)).astype(np.float32)
    terr_sph = make_rays(o, unit_sphere(pool_rays, rng))
    return [soup_cam, soup_sph, terr_cam, terr_sph]


def serving_requests(n_requests: int, scenes: list, pool_rays: int, lo: int = 1_000, hi: int = 1_000_000,
                     p_closest: float = 0.7, stream: int = 0) -> list:
    """Requests: scene uniform, ray batch log-uniform in [lo, hi] rays drawn
    as a contiguous slice of a camera or sphere pool of the scene's kind
    (even odds), closest-hit with probability p_closest, else any-hit."""
    rng = rng_for(4, 200 + stream)
    counts = np.minimum(log_uniform_ints(rng, n_requests, lo, hi), pool_rays)
    out = []
    for k in range(n_requests):
        s = int(rng.integers(0, len(scenes)))
        pool = (0 if scenes[s].kind == "soup" else 2) + int(rng.integers(0, 2))
        c = int(counts[k])
        off = int(rng.integers(0, pool_rays - c + 1))
        q = "closest" if rng.uniform() < p_closest else "any"
        out.append(ServingRequest(s, pool, off, c, q))
    return out


def poisson_arrivals(n: int, qps: float, stream: int = 0) -> np.ndarray:
    """Arrival offsets in ns of a Poisson process of rate qps (first at 0)."""
    rng = rng_for(4, 300 + stream)
    gaps = rng.exponential(1e9 / qps, size=n)
    gaps[0] = 0.0
    return np.cumsum(gaps).astype(np.int64)


# --- instanced scene ---------------------------------------------------------------
@dataclass
class InstancedWorkload:
    meshes: list            # [(vertices, indices)] per distinct mesh
    mesh_of_instance: np.ndarray  # [I] int
    transforms: np.ndarray  # [I, 3, 4] object -> world, row-major
    eye: tuple
    width: int
    height: int


def instanced(n_meshes: int = 64, n_inst: int = 4096, lo: int = 10_000, hi: int = 500_000, width: int = 2048,
              height: int = 2048, trans: float = 50.0) -> InstancedWorkload:
    """Two-level scene: n_meshes soups (as the 65,536-triangle soup) with
    triangle counts log-uniform in [lo, hi]; n_inst instances, each of a
    uniformly chosen mesh, with a uniform rotation, scale U[0.5, 2] and
    translation U[-trans, trans]^3 (50 by default).  The primary camera sits
    at (0, 0, -3 trans) looking along +z with a 45 degree vertical field of
    view, so the translation cube fills the view."""
    rng = rng_for(2)
    counts = log_uniform_ints(rng, n_meshes, lo, hi)
    meshes = [triangle_soup(int(c), rng_for(2, 1 + k)) for k, c in enumerate(counts)]
    pick = rng.integers(0, n_meshes, size=n_inst)
    xf = random_transforms(rng, n_inst, trans=trans)
    return InstancedWorkload(meshes, pick, xf, (0.0, 0.0, -3.0 * trans), width, height)


def _onb(n):
    """Orthonormal tangents of unit normals n [R, 3] (Duff et al., JCGT 6(1) 2017)."""
    s = np.where(n[:, 2] >= 0, 1.0, -1.0)
    a = -1.0 / (s + n[:, 2])
    b = n[:, 0] * n[:, 1] * a
    t = np.stack([1 + s * n[:, 0] ** 2 * a, s * b, -s * n[:, 0]], axis=1)
    u = np.stack([b, s + n[:, 1] ** 2 * a, -n[:, 1]], axis=1)
    return t, u


def secondary_rays(w: InstancedWorkload, primary: np.ndarray, prim_id: np.ndarray, inst_id: np.ndarray,
                   t_hit: np.ndarray, per_pixel: int = 2, seed: int = 7) -> np.ndarray:
    """per_pixel cosine-weighted (Malley) rays from every primary hit point,
    about the hit triangle's world normal turned towards the camera, with the
    origin pushed off the surface by 1e-4 (1 + max|p|); pixels whose primary
    ray missed emit no-op rays (max_t = -1).  Returns [per_pixel * R, 8],
    the rays of one pixel adjacent."""
    hit = inst_id != 0xFFFFFFFF
    R = primary.shape[0]
    o = primary[:, 0:3].astype(np.float64)
    d = primary[:, 4:7].astype(np.float64)
    n = np.zeros((R, 3))
    idx = np.nonzero(hit)[0]
    if idx.size:
        inst = inst_id[idx].astype(np.int64)
        mesh = w.mesh_of_instance[inst]
        prim = prim_id[idx].astype(np.int64)
        v0 = np.empty((idx.size, 3)); v1 = np.empty((idx.size, 3)); v2 = np.empty((idx.size, 3))
        for m in np.unique(mesh):
            sel = mesh == m
            V, I = w.meshes[m]
            tri = I[prim[sel]]
            v0[sel], v1[sel], v2[sel] = V[tri[:, 0]], V[tri[:, 1]], V[tri[:, 2]]
        n_obj = np.cross(v1 - v0, v2 - v0)
        A = w.transforms[:, :, :3].astype(np.float64)
        inv_t = np.transpose(np.linalg.inv(A), (0, 2, 1))  # normals transform by the inverse transpose
        nw = np.einsum("kij,kj->ki", inv_t[inst], n_obj)
        nw /= np.maximum(np.linalg.norm(nw, axis=1, keepdims=True), 1e-300)
        flip = np.einsum("ki,ki->k", nw, d[idx]) > 0
        nw[flip] = -nw[flip]
        n[idx] = nw
    p = o + t_hit.astype(np.float64)[:, None] * d
    rng = rng_for(2, 100 + seed)
    out = np.zeros((R, per_pixel, 8), np.float32)
    tb, ub = _onb(np.where(hit[:, None], n, np.array([0.0, 0.0, 1.0])))
    off = (1e-4 * (1.0 + np.abs(p).max(axis=1)))[:, None] * n
    for k in range(per_pixel):
        u1, u2 = rng.uniform(size=R), rng.uniform(size=R)
        r, phi = np.sqrt(u1), 2 * np.pi * u2
        dl = (r * np.cos(phi))[:, None] * tb + (r * np.sin(phi))[:, None] * ub + np.sqrt(1 - u1)[:, None] * n
        out[:, k, 0:3] = np.where(hit[:, None], p + off, 0.0)
        out[:, k, 3] = 0.0
        out[:, k, 4:7] = np.where(hit[:, None], dl, np.array([0.0, 0.0, 1.0]))
        out[:, k, 7] = np.where(hit, 1e30, -1.0)
    return out.reshape(-1, 8)


# --- batched build / refit -------------------------------------------------------
@dataclass
class BatchedWorkload:
    vertices: np.ndarray    # all meshes' vertices, concatenated [sum 3 n_k, 3]
    tri_counts: np.ndarray  # [M]
    vertex_offsets: np.ndarray  # [M] first vertex of each mesh
    rays: np.ndarray        # [M * rays_per_mesh, 8], mesh k's rays at k * rays_per_mesh
    rays_per_mesh: int


def batched(n_meshes: int = 8192, lo: int = 64, hi: int = 16384, rays_per_mesh: int = 256) -> BatchedWorkload:
    """Independent soups (as the 65,536-triangle soup) with triangle counts
    log-uniform in [lo, hi]; per mesh, rays_per_mesh rays with origins
    uniform in the mesh's box and directions uniform on the sphere."""
    rng = rng_for(3)
    counts = log_uniform_ints(rng, n_meshes, lo, hi)
    total = int(counts.sum())
    c = rng.uniform(0.0, 1.0, size=(total, 1, 3))
    v = (c + rng.normal(0.0, 0.01, size=(total, 3, 3))).astype(np.float32).reshape(-1, 3)
    offs = np.concatenate([[0], np.cumsum(3 * counts)[:-1]]).astype(np.int64)
    lo_box = np.minimum.reduceat(v, offs, axis=0).astype(np.float64)
    hi_box = np.maximum.reduceat(v, offs, axis=0).astype(np.float64)
    u = rng.uniform(size=(n_meshes, rays_per_mesh, 3))
    o = (lo_box[:, None, :] + u * (hi_box - lo_box)[:, None, :]).astype(np.float32).reshape(-1, 3)
    rays = make_rays(o, unit_sphere(n_meshes * rays_per_mesh, rng))
    return BatchedWorkload(v, counts.astype(np.int64), offs, rays, rays_per_mesh)
