"""Procedural workloads of the sorted-wavefront renderer, generated from seeds.

The reference ships no scenes (its tree has no .yaml or .obj), so the workloads are defined
here and written out in the reference's scene format (two-document YAML + OBJ, the schema its
decoders read: Pathtracer.cpp:32-275) when a file is wanted.

A scene description is a dict:
  shaders:     [{type, colour, reflectance, roughness}]
  meshes:      [{pos4 (V,4) float32, idx (T,3) uint32, uv (V,2) or None, shader}]
  lights:      [{translation, rotation, scale, intensity}]  (QuadLight, QuadLight.h convert<>)
  environment: (r, g, b)
plus a render dict (image, settings, camera, sampler, filter) that maps onto RenderDesc.
"""
from __future__ import annotations

import math
import os

import numpy as np

SHADER_LAMBERT, SHADER_NULL, SHADER_GLOSSY = 0, 1, 2


# ---- geometry ---------------------------------------------------------------------------
def quad_mesh(corners):
    """Two triangles (c0, c1, c2), (c0, c2, c3): front face = (c1-c0) x (c2-c0)."""
    c = np.asarray(corners, dtype=np.float32)
    pos = np.zeros((4, 4), np.float32)
    pos[:, :3] = c
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    return pos, idx, uv


def icosphere(level: int, radius: float, centre):
    """Subdivided icosahedron: 20 * 4^level triangles, outward winding."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = np.array(verts) * radius + np.asarray(centre, np.float64)
    pos = np.zeros((len(v), 4), np.float32)
    pos[:, :3] = v
    idx = np.array(faces, np.uint32)
    # make every face wind outward (normal along the radius)
    p0, p1, p2 = v[idx[:, 0]], v[idx[:, 1]], v[idx[:, 2]]
    n = np.cross(p1 - p0, p2 - p0)
    flip = np.einsum("ij,ij->i", n, (p0 + p1 + p2) / 3.0 - np.asarray(centre)) < 0
    idx[flip] = idx[flip][:, [0, 2, 1]]
    d = (v - np.asarray(centre)) / radius
    uv = np.stack([0.5 + np.arctan2(d[:, 2], d[:, 0]) / (2 * math.pi), np.arccos(np.clip(d[:, 1], -1, 1)) / math.pi],
                  axis=1).astype(np.float32)
    return pos, idx, uv


def value_noise(x, z, seed: int, octaves: int = 6, amplitude: float = 0.1, base_freq: float = 4.0):
    """Seeded lattice value noise with smoothstep interpolation, summed over octaves."""
    rng = np.random.default_rng(seed)
    h = np.zeros_like(x, dtype=np.float64)
    amp, freq, norm = 1.0, base_freq, 0.0
    for _ in range(octaves):
        res = int(freq) + 2
        lattice = rng.random((res + 1, res + 1))
        fx = (x - x.min()) / max(x.max() - x.min(), 1e-12) * freq
        fz = (z - z.min()) / max(z.max() - z.min(), 1e-12) * freq
        ix, iz = np.floor(fx).astype(int), np.floor(fz).astype(int)
        tx, tz = fx - ix, fz - iz
        sx, sz = tx * tx * (3 - 2 * tx), tz * tz * (3 - 2 * tz)
        a = lattice[ix, iz] * (1 - sx) + lattice[ix + 1, iz] * sx
        b = lattice[ix, iz + 1] * (1 - sx) + lattice[ix + 1, iz + 1] * sx
        h += amp * (a * (1 - sz) + b * sz)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return (h / norm - 0.5) * 2.0 * amplitude


def camera_for(eye, yaw_deg: float, pitch_deg: float, focal: float = 55.0):
    """Pinhole camera whose nodal point sits at eye, looking along R * x-hat with
    R = Ry(yaw) * Rx(0) * Rz(pitch) (the reference camera order T*Ry*Rx*Rz, PinHoleCamera.cpp:7-11)."""
    cy, sy = math.cos(math.radians(yaw_deg)), math.sin(math.radians(yaw_deg))
    cz, sz = math.cos(math.radians(pitch_deg)), math.sin(math.radians(pitch_deg))
    # R * x-hat for R = Ry * Rz
    xz = np.array([cz, sz, 0.0])
    n = np.array([cy * xz[0] + sy * xz[2], xz[1], -sy * xz[0] + cy * xz[2]])
    t = np.asarray(eye, np.float64) - n * focal
    return {"type": 0, "translation": t.tolist(), "rotation": [0.0, yaw_deg, pitch_deg], "focal_length": focal,
            "focal_distance": 1000.0, "aperture": 5.6}


# ---- workloads ----------------------------------------------------------------------------
def box_room(seed: int = 0x5EED0001, width: int = 512, height: int = 512, base: int = 4, max_depth: int = 5,
             sphere_level: int = 5, bin_exponent: int = 20):
    """Plain method, CPU-runnable: a box room of 5 axis-aligned quad walls (10 triangles), a
    subdivided icosphere (20 * 4^level triangles, radius 0.3, seed-placed) and one quad light.
    Lambert albedos U[0.2, 0.8] per mesh from the seed."""
    rng = np.random.default_rng(seed)
    walls = {
        "floor": [(-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1)],
        "ceiling": [(-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)],
        "back": [(-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1)],
        "left": [(-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1)],
        "right": [(1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1)],
    }
    shaders, meshes = [], []
    for name, corners in walls.items():
        pos, idx, uv = quad_mesh(corners)
        shaders.append({"type": SHADER_LAMBERT, "colour": rng.uniform(0.2, 0.8, 3).tolist(), "reflectance": 1.0,
                        "roughness": 0.0})
        meshes.append({"name": name, "pos4": pos, "idx": idx, "uv": uv, "shader": len(shaders) - 1})
    centre = [rng.uniform(-0.5, 0.5), 0.3 + rng.uniform(0.0, 0.6), rng.uniform(-0.5, 0.3)]
    pos, idx, uv = icosphere(sphere_level, 0.3, centre)
    shaders.append({"type": SHADER_LAMBERT, "colour": rng.uniform(0.2, 0.8, 3).tolist(), "reflectance": 1.0,
                    "roughness": 0.0})
    meshes.append({"name": "sphere", "pos4": pos, "idx": idx, "uv": uv, "shader": len(shaders) - 1})
    lights = [{"translation": [0.0, 1.98, 0.0], "rotation": [0.0, 0.0, -90.0], "scale": [0.6, 0.6],
               "intensity": 12.0}]
    render = {
        "width": width, "height": height, "base": base, "spp": 0, "sampler": 0, "filter": 0,
        "min_depth": 2, "max_depth": max_depth, "threshold": 0.01, "shading_size": 4096,
        "bin_exponent": bin_exponent, "iteration": 0, "seed": seed,
        "camera": camera_for([0.0, 1.0, 3.4], 90.0, 0.0),
    }
    return {"shaders": shaders, "meshes": meshes, "lights": lights, "environment": (0.0, 0.0, 0.0),
            "render": render}


def terrain(seed: int = 0x5EED0002, grid: int = 1024, groups: int = 64, width: int = 1920, height: int = 1080,
            spp: int = 64, max_depth: int = 8, bin_exponent: int = 23):
    """Plain method at scale: a height-field terrain of a grid x grid quad mesh (2 * grid^2
    triangles) with seeded value-noise displacement (amplitude 0.1, 6 octaves), split into
    `groups` material groups mixing diffuse and glossy lobes (roughness U[0.05, 0.6]); lit by a
    small bright quad (the sun) and a constant environment."""
    rng = np.random.default_rng(seed)
    g = int(round(math.sqrt(groups)))
    assert g * g == groups and grid % g == 0
    xs = np.linspace(-1.0, 1.0, grid + 1)
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    Y = value_noise(X, Z, seed + 1)
    shaders, meshes = [], []
    cell = grid // g
    for gi in range(g):
        for gj in range(g):
            glossy = rng.random() < 0.5
            colour = rng.uniform(0.2, 0.8, 3).tolist()
            if glossy:
                shaders.append({"type": SHADER_GLOSSY, "colour": colour, "reflectance": 1.0,
                                "roughness": float(rng.uniform(0.05, 0.6))})
            else:
                shaders.append({"type": SHADER_LAMBERT, "colour": colour, "reflectance": 1.0, "roughness": 0.0})
            i0, j0 = gi * cell, gj * cell
            xi = X[i0:i0 + cell + 1, j0:j0 + cell + 1]
            yi = Y[i0:i0 + cell + 1, j0:j0 + cell + 1]
            zi = Z[i0:i0 + cell + 1, j0:j0 + cell + 1]
            nv = (cell + 1) * (cell + 1)
            pos = np.zeros((nv, 4), np.float32)
            pos[:, 0], pos[:, 1], pos[:, 2] = xi.ravel(), yi.ravel(), zi.ravel()
            a = np.arange(cell)
            I, J = np.meshgrid(a, a, indexing="ij")
            v00 = (I * (cell + 1) + J).ravel()
            v10 = ((I + 1) * (cell + 1) + J).ravel()
            v01 = (I * (cell + 1) + J + 1).ravel()
            v11 = ((I + 1) * (cell + 1) + J + 1).ravel()
            idx = np.empty((2 * cell * cell, 3), np.uint32)
            idx[0::2] = np.stack([v00, v01, v11], 1)
            idx[1::2] = np.stack([v00, v11, v10], 1)
            uv = np.stack([(xi.ravel() + 1) / 2, (zi.ravel() + 1) / 2], 1).astype(np.float32)
            meshes.append({"name": f"terrain_{gi}_{gj}", "pos4": pos, "idx": idx, "uv": uv,
                           "shader": len(shaders) - 1})
    # a sun-sized quad high above, facing down and tilted
    lights = [{"translation": [2.0, 4.0, 1.5], "rotation": [0.0, 20.0, -70.0], "scale": [0.25, 0.25],
               "intensity": 400.0}]
    render = {
        "width": width, "height": height, "base": int(round(math.sqrt(spp))), "spp": spp, "sampler": 0,
        "filter": 0, "min_depth": 2, "max_depth": max_depth, "threshold": 0.01, "shading_size": 4096,
        "bin_exponent": bin_exponent, "iteration": 0, "seed": seed,
        "camera": camera_for([0.0, 0.9, 1.9], 90.0, -28.0),
    }
    if render["base"] ** 2 != spp:
        render["sampler"] = 1
    return {"shaders": shaders, "meshes": meshes, "lights": lights, "environment": (0.25, 0.3, 0.4),
            "render": render}


def textured_room(seed: int = 0x5EED0006, width: int = 256, height: int = 256, base: int = 2, max_depth: int = 5,
                  sphere_level: int = 4):
    """The box room with image and layered textures (StandardTexture, LayeredTexture,
    ConstantTexture): a seeded checker image on the floor with texture coordinates scaled to
    [-0.3, 2.7] (so the lookup wraps), a layered back wall (a gradient over a constant colour,
    mixed by a seeded noise mask), a left wall whose layer nests another layered texture in its
    mask, and a noise image on the sphere (whose texture seam wraps in s). Images are seeded
    float RGB in [0, 1]."""
    rng = np.random.default_rng(seed)
    room = box_room(seed=seed, width=width, height=height, base=base, max_depth=max_depth,
                    sphere_level=sphere_level)
    shaders = [dict(sh) for sh in room["shaders"]]
    meshes = [dict(m) for m in room["meshes"]]
    checker = np.zeros((16, 16, 3), np.float32)
    cells = rng.uniform(0.15, 0.9, (2, 3)).astype(np.float32)
    yy, xx = np.mgrid[0:16, 0:16]
    checker[:] = np.where(((xx // 4 + yy // 4) % 2 == 0)[..., None], cells[0], cells[1])
    gy, gx = np.mgrid[0:24, 0:40] / np.array([23.0, 39.0])[:, None, None]
    gradient = np.stack([gx, gy, 1.0 - gx], axis=-1).astype(np.float32) * 0.8 + 0.1
    mask = rng.uniform(0.0, 1.0, (8, 8, 3)).astype(np.float32)
    noise = rng.uniform(0.1, 0.9, (32, 64, 3)).astype(np.float32)
    images = [checker, gradient, mask, noise]
    textures = [
        {"type": 1, "image": 0},                                   # 0 checker
        {"type": 1, "image": 1},                                   # 1 gradient
        {"type": 0, "colour": rng.uniform(0.2, 0.8, 3).tolist()},  # 2 constant
        {"type": 1, "image": 2},                                   # 3 mask
        {"type": 2, "upper": 1, "lower": 2, "mask": 3},            # 4 gradient over constant
        {"type": 0, "colour": [0.8, 0.2, 0.2]},                    # 5 constant
        {"type": 2, "upper": 5, "lower": 0, "mask": 4},            # 6 nested layer in the mask
        {"type": 1, "image": 3},                                   # 7 noise
    ]
    names = [m["name"] for m in meshes]
    floor = meshes[names.index("floor")]
    floor["uv"] = floor["uv"] * 3.0 - 0.3
    shaders[floor["shader"]]["texture"] = 0
    shaders[meshes[names.index("back")]["shader"]]["texture"] = 4
    shaders[meshes[names.index("left")]["shader"]]["texture"] = 6
    shaders[meshes[names.index("sphere")]["shader"]]["texture"] = 7
    return {"shaders": shaders, "meshes": meshes, "lights": room["lights"], "environment": (0.0, 0.0, 0.0),
            "images": images, "textures": textures, "render": room["render"]}


def sphere_cluster(seed: int, count: int, level: int, width: int = 256, height: int = 256, base: int = 2,
                   max_depth: int = 6, bin_exponent: int = 23):
    """An icosphere cluster in the box room: `count` spheres of 20 * 4^level triangles with
    seeded centres, radii and Lambert or glossy materials, under the room's ceiling light."""
    rng = np.random.default_rng(seed)
    room = box_room(seed=seed, width=width, height=height, base=base, max_depth=max_depth, sphere_level=0,
                    bin_exponent=bin_exponent)
    shaders = room["shaders"][:5]
    meshes = [m for m in room["meshes"] if m["name"] != "sphere"]
    unit_pos, idx, uv = icosphere(level, 1.0, (0.0, 0.0, 0.0))
    for i in range(count):
        r = float(rng.uniform(0.06, 0.25))
        c = np.array([rng.uniform(-0.8, 0.8), rng.uniform(r, 1.5), rng.uniform(-0.8, 0.6)])
        pos = unit_pos.copy()
        pos[:, :3] = unit_pos[:, :3] * r + c
        colour = rng.uniform(0.2, 0.8, 3).tolist()
        if rng.random() < 0.5:
            shaders.append({"type": SHADER_GLOSSY, "colour": colour, "reflectance": 1.0,
                            "roughness": float(rng.uniform(0.05, 0.6))})
        else:
            shaders.append({"type": SHADER_LAMBERT, "colour": colour, "reflectance": 1.0, "roughness": 0.0})
        meshes.append({"name": f"sphere_{i}", "pos4": pos, "idx": idx, "uv": uv, "shader": len(shaders) - 1})
    return {"shaders": shaders, "meshes": meshes, "lights": room["lights"], "environment": (0.0, 0.0, 0.0),
            "render": room["render"]}


def batch_workload(seed: int = 0x5EED0004, jobs: int = 256, max_depth: int = 6):
    """Batched workload: `jobs` render jobs over 8 resident scenes of 10k-2M triangles (icosphere
    clusters and terrain variants). Each job: a seeded camera on a uniformly chosen scene, pixel
    count log-uniform in [128^2, 1024^2] with aspect U[0.75, 1.33], spp uniform in {1..64}
    (stratified when spp is a square, independent otherwise), max depth 6.
    Returns (scene descriptions, [(scene index, render dict)])."""
    descs = [
        sphere_cluster(seed + 1, 2, 4),      # 10,250 triangles
        sphere_cluster(seed + 2, 8, 4),      # 40,970
        sphere_cluster(seed + 3, 6, 5),      # 122,890
        sphere_cluster(seed + 4, 16, 5),     # 327,690
        terrain(seed + 5, grid=256, groups=16),   # 131,074
        terrain(seed + 6, grid=512, groups=64),   # 524,290
        terrain(seed + 7, grid=768, groups=64),   # 1,179,650
        terrain(seed + 8, grid=1024, groups=64),  # 2,097,154
    ]
    rng = np.random.default_rng(seed)
    out = []
    for j in range(jobs):
        s = int(rng.integers(0, len(descs)))
        side = math.exp(rng.uniform(math.log(128.0), math.log(1024.0)))
        aspect = rng.uniform(0.75, 4.0 / 3.0)
        w = max(1, int(round(side * math.sqrt(aspect))))
        h = max(1, int(round(side / math.sqrt(aspect))))
        spp = int(rng.integers(1, 65))
        b = int(round(math.sqrt(spp)))
        r = dict(descs[s]["render"])
        if s < 4:
            cam = camera_for([rng.uniform(-0.6, 0.6), rng.uniform(0.6, 1.4), rng.uniform(3.0, 3.4)],
                             rng.uniform(80.0, 100.0), rng.uniform(-10.0, 5.0))
        else:
            cam = camera_for([rng.uniform(-0.5, 0.5), rng.uniform(0.6, 1.2), rng.uniform(1.5, 2.0)],
                             rng.uniform(75.0, 105.0), rng.uniform(-35.0, -20.0))
        r.update({"width": w, "height": h, "base": b, "spp": spp, "sampler": 0 if b * b == spp else 1,
                  "max_depth": max_depth, "min_depth": 2, "seed": seed * 1000 + j, "iteration": 0,
                  "camera": cam})
        out.append((s, r))
    return descs, out


def serve_workload(seed: int = 0x5EED0005, requests: int = 1050, max_depth: int = 6):
    """Serving workload: 4 resident scenes of 50k-2M triangles and a list of progressive
    requests, each a seeded camera on a uniformly chosen scene with a pixel count log-uniform in
    [256^2, 1920x1080] (aspect U[1, 16/9]) and a total spp log-uniform in [4, 256], rendered as
    iterations of 16 spp (base 4; fewer when the total is below 16, base 2 or 3).
    Returns (scene descriptions, [(scene index, render dict, iterations)])."""
    descs = [
        sphere_cluster(seed + 1, 3, 5),           # 61,452 triangles
        terrain(seed + 2, grid=256, groups=16),   # 131,074
        terrain(seed + 3, grid=512, groups=64),   # 524,290
        terrain(seed + 4, grid=1024, groups=64),  # 2,097,154
    ]
    rng = np.random.default_rng(seed)
    out = []
    lo, hi = math.log(256.0 * 256.0), math.log(1920.0 * 1080.0)
    for j in range(requests):
        s = int(rng.integers(0, len(descs)))
        px = math.exp(rng.uniform(lo, hi))
        aspect = rng.uniform(1.0, 16.0 / 9.0)
        w = max(1, int(round(math.sqrt(px * aspect))))
        h = max(1, int(round(math.sqrt(px / aspect))))
        spp = math.exp(rng.uniform(math.log(4.0), math.log(256.0)))
        base = 4 if spp >= 16 else max(2, int(round(math.sqrt(spp))))
        iterations = max(1, int(round(spp / (base * base))))
        r = dict(descs[s]["render"])
        if s == 0:
            cam = camera_for([rng.uniform(-0.6, 0.6), rng.uniform(0.6, 1.4), rng.uniform(3.0, 3.4)],
                             rng.uniform(80.0, 100.0), rng.uniform(-10.0, 5.0))
        else:
            cam = camera_for([rng.uniform(-0.5, 0.5), rng.uniform(0.6, 1.2), rng.uniform(1.5, 2.0)],
                             rng.uniform(75.0, 105.0), rng.uniform(-35.0, -20.0))
        r.update({"width": w, "height": h, "base": base, "spp": 0, "sampler": 0, "max_depth": max_depth,
                  "min_depth": 2, "seed": seed * 1000 + j, "iteration": 0, "camera": cam})
        out.append((s, r, iterations))
    return descs, out


def random_soup(seed: int, ntris: int, extent: float = 1.0, size: float = 0.1):
    """Random triangles in a cube (intersection tests)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-extent, extent, (ntris, 1, 3))
    v = c + rng.uniform(-size, size, (ntris, 3, 3))
    pos = np.zeros((ntris * 3, 4), np.float32)
    pos[:, :3] = v.reshape(-1, 3)
    idx = np.arange(ntris * 3, dtype=np.uint32).reshape(-1, 3)
    return {"shaders": [{"type": SHADER_LAMBERT, "colour": [0.5, 0.5, 0.5], "reflectance": 1.0, "roughness": 0.0}],
            "meshes": [{"name": "soup", "pos4": pos, "idx": idx, "uv": None, "shader": 0}],
            "lights": [], "environment": (0.0, 0.0, 0.0)}


def sort_bench_rays(n: int, seed: int = 0x5EED0003):
    """Sort/traversal bench rays: origins uniform in the unit cube, directions uniform on the
    sphere (z = 1 - 2u, phi = 2 pi v); n x 8 floats with tmin 0, tmax 1e5."""
    rng = np.random.default_rng(seed)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = rng.random((n, 3), dtype=np.float32)
    z = 1.0 - 2.0 * rng.random(n, dtype=np.float32)
    phi = (2.0 * np.pi) * rng.random(n, dtype=np.float32)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    rays[:, 4] = r * np.cos(phi)
    rays[:, 5] = r * np.sin(phi)
    rays[:, 6] = z
    rays[:, 7] = 1e5
    return rays


# ---- scene files in the reference schema -------------------------------------------------
def write_obj(path: str, pos4, idx, uv=None):
    pos4 = np.asarray(pos4)
    idx = np.asarray(idx)
    with open(path, "w") as f:
        f.write("# swr procedural mesh\n")
        np.savetxt(f, pos4[:, :3], fmt="v %.9g %.9g %.9g")
        if uv is not None:
            np.savetxt(f, np.asarray(uv), fmt="vt %.9g %.9g")
            np.savetxt(f, np.repeat(idx + 1, 2, axis=1), fmt="f %d/%d %d/%d %d/%d")
        else:
            np.savetxt(f, idx + 1, fmt="f %d %d %d")


def write_pfm(path: str, rgb):
    """An (H, W, 3) float image as a little-endian PFM (scanlines stored bottom-up)."""
    rgb = np.asarray(rgb, dtype="<f4")
    h, w = rgb.shape[:2]
    with open(path, "wb") as f:
        f.write(b"PF\n%d %d\n-1.0\n" % (w, h))
        f.write(np.ascontiguousarray(rgb[::-1]).tobytes())


def write_scene(desc, directory: str, name: str = "scene") -> str:
    """Writes <directory>/<name>.yaml plus one OBJ per mesh in the reference schema."""
    os.makedirs(directory, exist_ok=True)
    r = desc["render"]
    cam = r["camera"]
    lines = [
        "image: {width: %d, height: %d, sample base: %d}" % (r["width"], r["height"], r["base"]),
        "settings: {min depth: %d, max depth: %d, threshold: %r, bucket size: 16, shading size: %d, "
        "bin exponent: %d}" % (r["min_depth"], r["max_depth"], r["threshold"], r["shading_size"], r["bin_exponent"]),
    ]
    fmt3 = lambda v: "[%s]" % ", ".join("%.9g" % float(x) for x in v)
    if cam["type"] == 0:
        lines.append("camera: {type: PinHole, translation: %s, rotation: %s, focal length: %.9g}"
                     % (fmt3(cam["translation"]), fmt3(cam["rotation"]), cam["focal_length"]))
    else:
        lines.append("camera: {type: ThinLens, translation: %s, rotation: %s, focal length: %.9g, "
                     "focal distance: %.9g, aperture: %.9g}"
                     % (fmt3(cam["translation"]), fmt3(cam["rotation"]), cam["focal_length"],
                        cam["focal_distance"], cam["aperture"]))
    lines.append("filter: {type: %s}" % ("Tent" if r["filter"] == 0 else "Box"))
    lines.append("sampler: {type: %s}" % ["Stratified", "Independent", "Grid"][r["sampler"]])
    lines.append("---")
    image_files = []
    for k, im in enumerate(desc.get("images", [])):
        fname = "%s_image%d.pfm" % (name, k)
        write_pfm(os.path.join(directory, fname), im)
        image_files.append(fname)
    textures = desc.get("textures", [])

    def texture_yaml(node):
        t = textures[node]
        if t["type"] == 0:
            return "{type: Constant, colour: %s}" % fmt3(t["colour"])
        if t["type"] == 1:
            return "{type: Standard, string: %s}" % image_files[t["image"]]
        return "{type: Layered, upper: %s, lower: %s, mask: %s}" % (
            texture_yaml(t["upper"]), texture_yaml(t["lower"]), texture_yaml(t["mask"]))

    for sh in desc["shaders"]:
        tex = (texture_yaml(sh["texture"]) if sh.get("texture", -1) >= 0
               else "{type: Constant, colour: %s}" % fmt3(sh["colour"]))
        if sh["type"] == SHADER_GLOSSY:
            lines.append("shader: {type: Glossy, texture: %s, reflectance: %.9g, roughness: %.9g}"
                         % (tex, sh["reflectance"], sh["roughness"]))
        else:
            lines.append("shader: {type: Lambert, texture: %s, reflectance: %.9g}" % (tex, sh["reflectance"]))
    for k, m in enumerate(desc["meshes"]):
        fname = "%s_%s.obj" % (name, m.get("name", k))
        write_obj(os.path.join(directory, fname), m["pos4"], m["idx"], m.get("uv"))
        lines.append("object: {type: Polygon, filename: %s, translation: [0, 0, 0], rotation: [0, 0, 0], "
                     "scale: [1, 1, 1], shader: %d}" % (fname, m["shader"]))
    for l in desc["lights"]:
        lines.append("light: {type: Quad, translation: %s, rotation: %s, scale: [%.9g, %.9g], intensity: %.9g}"
                     % (fmt3(l["translation"]), fmt3(l["rotation"]), l["scale"][0], l["scale"][1], l["intensity"]))
    path = os.path.join(directory, name + ".yaml")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path
