"""Seeded synthetic scenes and ray sets for the traversal workloads.

Every array is produced in float64 with numpy's PCG64 generator and rounded
once to float32, so the same seed gives the same bits on every machine.

Workloads (named by content):
  sphere_ground          UV sphere r=1 (256x128 segments, 65,024 tris) over a 10x10 ground quad at y=-1
  pinhole_rays           1024x1024 pixel-centre primary rays, 45 deg horizontal FOV, eye (0,1,4) -> origin
  uniform_rays           rays with origins uniform in a box, directions uniform on the sphere
  icosphere_field        4096 icospheres (subdiv 3, 1280 tris each), radii log-uniform [0.01,0.5],
                         centres uniform in [0,100]^3, flattened into one mesh (5,242,880 tris)
  bounce_rays            cosine-weighted secondary rays leaving random surface points, path depth U{1..8}
  moving_spheres         2000 UV spheres of 250 tris, 2 motion keys, translation |d| ~ U[0,2]
"""
import math

import numpy as np

from ._lib import RAY_DTYPE


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


# --------------------------------------------------------------------- meshes

def uv_sphere(radius=1.0, nseg=256, nring=128, center=(0.0, 0.0, 0.0)):
    """UV sphere: two poles, nring-1 latitude rings of nseg vertices.

    Triangles: a fan of nseg at each pole plus 2*nseg*(nring-2) in between,
    i.e. 2*nseg*(nring-1) (256x128 -> 65,024).  Returns (verts f32 (V,3), idx u32 (T,3)).
    """
    cx, cy, cz = center
    th = np.pi * np.arange(1, nring) / nring               # ring polar angles
    ph = 2.0 * np.pi * np.arange(nseg) / nseg
    st, ct = np.sin(th)[:, None], np.cos(th)[:, None]
    ring = np.stack([radius * st * np.cos(ph)[None, :] + cx,
                     radius * ct * np.ones_like(ph)[None, :] + cy,
                     radius * st * np.sin(ph)[None, :] + cz], axis=-1).reshape(-1, 3)
    north = np.array([[cx, cy + radius, cz]])
    south = np.array([[cx, cy - radius, cz]])
    verts = np.concatenate([north, ring, south]).astype(np.float32)
    N, S = 0, 1 + (nring - 1) * nseg
    j = np.arange(nseg)
    jn = (j + 1) % nseg
    faces = [np.stack([np.full(nseg, N), 1 + jn, 1 + j], axis=1)]
    for r in range(nring - 2):
        a = 1 + r * nseg
        b = 1 + (r + 1) * nseg
        faces.append(np.stack([a + j, a + jn, b + jn], axis=1))
        faces.append(np.stack([a + j, b + jn, b + j], axis=1))
    last = 1 + (nring - 2) * nseg
    faces.append(np.stack([np.full(nseg, S), last + j, last + jn], axis=1))
    return verts, np.concatenate(faces).astype(np.uint32)


def ground_quad(size=10.0, y=-1.0):
    h = size / 2.0
    verts = np.array([[-h, y, -h], [h, y, -h], [h, y, h], [-h, y, h]], dtype=np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2]], dtype=np.uint32)
    return verts, idx


def icosphere(subdiv=3):
    """Unit icosphere by midpoint subdivision: 20*4^subdiv faces (3 -> 1280), float64 verts."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
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
    return np.array(verts), np.array(faces, dtype=np.int64)


def icosphere_field(count=4096, subdiv=3, rmin=0.01, rmax=0.5, extent=100.0, seed=7):
    """Flattened field of icospheres.  Returns (verts f32, idx u32, sphere_of_face, centres f64, radii f64)."""
    g = rng(seed)
    base_v, base_f = icosphere(subdiv)
    radii = np.exp(g.uniform(np.log(rmin), np.log(rmax), count))
    centres = g.uniform(0.0, extent, (count, 3))
    nv, nf = len(base_v), len(base_f)
    verts = (base_v[None, :, :] * radii[:, None, None] + centres[:, None, :]).reshape(-1, 3).astype(np.float32)
    idx = (base_f[None, :, :] + (np.arange(count) * nv)[:, None, None]).reshape(-1, 3).astype(np.uint32)
    sphere_of_face = np.repeat(np.arange(count), nf)
    return verts, idx, sphere_of_face, centres, radii


# ----------------------------------------------------------------------- rays

def make_rays(P, D, tmax=np.inf, time=0.0):
    n = len(P)
    r = np.zeros(n, RAY_DTYPE)
    r["P"] = np.asarray(P, np.float64).astype(np.float32)
    r["D"] = np.asarray(D, np.float64).astype(np.float32)
    r["tmax"] = np.float32(tmax) if np.isscalar(tmax) else np.asarray(tmax, np.float32)
    r["time"] = np.float32(time) if np.isscalar(time) else np.asarray(time, np.float32)
    return r


def lookat_basis(eye, target, up):
    eye, target, up = (np.asarray(x, np.float64) for x in (eye, target, up))
    w = eye - target
    w /= np.linalg.norm(w)
    u = np.cross(up, w)
    u /= np.linalg.norm(u)
    v = np.cross(w, u)
    return u, v, w


def pinhole_rays(width=1024, height=1024, fov_deg=45.0, eye=(0.0, 1.0, 4.0), target=(0.0, 0.0, 0.0),
                 up=(0.0, 1.0, 0.0)):
    """One ray per pixel centre; fov is horizontal; screen y grows downwards (core/render.go:114-115)."""
    u, v, w = lookat_basis(eye, target, up)
    x = (np.arange(width) + 0.5) / width * 2.0 - 1.0
    y = -((np.arange(height) + 0.5) / height * 2.0 - 1.0)
    sx, sy = np.meshgrid(x, y)
    tn = math.tan(math.radians(fov_deg) / 2.0)
    aspect = width / height
    d = (sx.reshape(-1, 1) * tn) * u + (sy.reshape(-1, 1) * tn / aspect) * v - w
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    P = np.broadcast_to(np.asarray(eye, np.float64), d.shape)
    return make_rays(P, d)


def uniform_sphere_dirs(g, n):
    """Uniform directions on S^2 (as math/sample/sample.go:91-100, z = 1 - 2u)."""
    u1, u2 = g.random(n), g.random(n)
    z = 1.0 - 2.0 * u1
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2.0 * np.pi * u2
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def uniform_rays(n=1_000_000, lo=(-5.0, -1.0, -5.0), hi=(5.0, 1.0, 5.0), seed=1):
    g = rng(seed)
    P = g.uniform(np.asarray(lo), np.asarray(hi), (n, 3))
    return make_rays(P, uniform_sphere_dirs(g, n))


def bounce_rays(verts, idx, n, seed=11, offset_rel=1e-4):
    """Secondary rays: random triangle (uniform over faces), uniform barycentric point, origin pushed off
    the surface along the geometric normal by offset_rel * (edge scale), cosine-weighted direction about
    the outward normal.  Returns (rays, depth) with depth ~ U{1..8} per ray."""
    g = rng(seed)
    f = g.integers(0, len(idx), n)
    tri = idx[f].astype(np.int64)
    v0 = verts[tri[:, 0]].astype(np.float64)
    v1 = verts[tri[:, 1]].astype(np.float64)
    v2 = verts[tri[:, 2]].astype(np.float64)
    r1, r2 = g.random(n), g.random(n)
    s = np.sqrt(r1)
    b0, b1 = 1.0 - s, s * (1.0 - r2)
    p = b0[:, None] * v0 + b1[:, None] * v1 + (1.0 - b0 - b1)[:, None] * v2
    nrm = np.cross(v1 - v0, v2 - v0)
    ln = np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm /= ln
    # the icosphere faces wind counter-clockwise seen from outside, so nrm points outwards
    scale = np.sqrt(ln[:, 0])
    p = p + nrm * (offset_rel * scale)[:, None]
    # cosine-weighted hemisphere about nrm (sample.CosineHemisphere, math/sample/sample.go:18-27)
    u0, u1 = g.random(n), g.random(n)
    rr = np.sqrt(1.0 - u0)
    lx, ly, lz = rr * np.cos(2 * np.pi * u1), rr * np.sin(2 * np.pi * u1), np.sqrt(u0)
    a = np.where(np.abs(nrm[:, :1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(a, nrm)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(nrm, t)
    d = lx[:, None] * t + ly[:, None] * b + lz[:, None] * nrm
    depth = g.integers(1, 9, n).astype(np.int32)
    return make_rays(p, d), depth


def bounce_rays_device(verts, idx, n, seed=17, offset_rel=1e-4, device="cuda"):
    """bounce_rays with the same construction, computed on the GPU with torch's Philox generator
    (float64, rounded once to float32).  Returns (rays as a uint8 device tensor of n*32 bytes in the
    RAY_DTYPE layout, depth int32 device tensor).  Used for the 16M-ray workload, where the host
    construction takes minutes."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    V = torch.as_tensor(np.asarray(verts), device=device)
    F = torch.as_tensor(np.asarray(idx).astype(np.int64), device=device)
    f = torch.randint(0, F.shape[0], (n,), generator=g, device=device)
    tri = F[f]
    v0, v1, v2 = (V[tri[:, k]].double() for k in range(3))
    r1 = torch.rand(n, generator=g, device=device, dtype=torch.float64)
    r2 = torch.rand(n, generator=g, device=device, dtype=torch.float64)
    s = torch.sqrt(r1)
    b0, b1 = 1.0 - s, s * (1.0 - r2)
    p = b0[:, None] * v0 + b1[:, None] * v1 + (1.0 - b0 - b1)[:, None] * v2
    nrm = torch.linalg.cross(v1 - v0, v2 - v0)
    ln = torch.linalg.norm(nrm, dim=1, keepdim=True)
    nrm = nrm / ln
    p = p + nrm * (offset_rel * torch.sqrt(ln))
    u0 = torch.rand(n, generator=g, device=device, dtype=torch.float64)
    u1 = torch.rand(n, generator=g, device=device, dtype=torch.float64)
    rr = torch.sqrt(1.0 - u0)
    lx, ly, lz = rr * torch.cos(2 * np.pi * u1), rr * torch.sin(2 * np.pi * u1), torch.sqrt(u0)
    ey = torch.tensor([[0.0, 1.0, 0.0]], device=device, dtype=torch.float64)
    ex = torch.tensor([[1.0, 0.0, 0.0]], device=device, dtype=torch.float64)
    a = torch.where(nrm[:, :1].abs() > 0.9, ey, ex)
    t = torch.linalg.cross(a, nrm)
    t = t / torch.linalg.norm(t, dim=1, keepdim=True)
    b = torch.linalg.cross(nrm, t)
    d = lx[:, None] * t + ly[:, None] * b + lz[:, None] * nrm
    depth = torch.randint(1, 9, (n,), generator=g, device=device, dtype=torch.int32)
    rec = torch.zeros((n, 8), dtype=torch.float32, device=device)
    rec[:, 0:3] = p.float()
    rec[:, 3] = float("inf")
    rec[:, 4:7] = d.float()
    return rec.view(torch.uint8).reshape(-1), depth


def moving_spheres(count=2000, nseg=25, nring=6, extent=20.0, rmin=0.3, rmax=1.0, max_move=2.0, seed=21):
    """Motion-blur scene: `count` UV spheres (2*nseg*(nring-1) tris each; 25x6 -> 250), two keys:
    key 1 = key 0 translated by a uniform random direction times U[0, max_move].
    Returns (verts f32 (2, V, 3), idx u32)."""
    g = rng(seed)
    sv, sf = uv_sphere(1.0, nseg, nring)
    sv = sv.astype(np.float64)
    radii = g.uniform(rmin, rmax, count)
    centres = g.uniform(-extent, extent, (count, 3))
    move = uniform_sphere_dirs(g, count) * g.uniform(0.0, max_move, count)[:, None]
    k0 = (sv[None] * radii[:, None, None] + centres[:, None, :]).reshape(-1, 3)
    k1 = (sv[None] * radii[:, None, None] + (centres + move)[:, None, :]).reshape(-1, 3)
    idx = (sf[None].astype(np.int64) + (np.arange(count) * len(sv))[:, None, None]).reshape(-1, 3).astype(np.uint32)
    return np.stack([k0, k1]).astype(np.float32), idx


# ------------------------------------------------------------ shaded scenes

# ShaderStd parameters (builtin/shader/std.go:25-47); has_* flags mark parameters that are set.
from ._lib import MATERIAL_DTYPE  # noqa: E402

FRESNEL_DIELECTRIC = 1


def material(diffuse=None, diffuse_strength=None, diffuse_roughness=None, spec=None, spec_strength=None,
             spec_roughness=None, ior=None, emission=None, emission_strength=None):
    m = np.zeros((), MATERIAL_DTYPE)
    for key, val, flag in (("diffuse", diffuse, "has_diffuse_colour"), ("spec1", spec, "has_spec1_colour"),
                           ("emission", emission, "has_emission_colour")):
        if val is not None:
            m[key] = val
            m[flag] = 1
    for key, val in (("diffuse_strength", diffuse_strength), ("diffuse_roughness", diffuse_roughness),
                     ("spec1_strength", spec_strength), ("spec1_roughness", spec_roughness), ("ior", ior),
                     ("emission_strength", emission_strength)):
        if val is not None:
            m[key] = val
            m["has_" + key] = 1
    m["fresnel_model"] = FRESNEL_DIELECTRIC
    return m


def value_noise(p, seed, freq=4.0):
    """Seeded 3-D value noise in [0,1): random values on an integer lattice (hashed from the lattice
    coordinates and the seed), trilinear interpolation with a smoothstep fade."""
    q = np.asarray(p, np.float64) * freq
    i0 = np.floor(q).astype(np.int64)
    f = q - i0
    f = f * f * (3.0 - 2.0 * f)

    def lattice(ix, iy, iz):
        h = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761)
        h = (h ^ (h >> 13)) * 1274126177
        h = h ^ (h >> 16)
        return (h & 0xFFFFFF).astype(np.float64) / float(1 << 24)

    out = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * (f[:, 2] if dz else 1 - f[:, 2])
                out = out + w * lattice(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz)
    return out


def quad(p0, p1, p2, p3):
    return np.array([p0, p1, p2, p3], np.float64), np.array([[0, 1, 2], [0, 2, 3]], np.int64)


def box_spheres_scene(seed=3, nseg=64, nring=32, nspheres=8, amplitude=0.05):
    """Closed box (floor, ceiling, back, left, right walls; the ceiling as four quads around a 0.5 x 0.5
    opening) with `nspheres` UV spheres
    (nseg x nring segments) displaced radially by seeded value noise of the given amplitude, and a
    0.5 x 0.5 light in the ceiling opening made of two triangle lights (radiance 10).  Sphere materials: Lambertian
    albedo U[0.2,0.8] per channel plus a GGX lobe of roughness U[0.05,0.5] (dielectric, IOR 1.5).

    Returns a dict: meshes (verts, idx, shaderidx, shaders, normals), materials, lights, camera.
    Geom insertion order: the flattened box+spheres mesh, then one mesh per light triangle."""
    g = rng(seed)
    mats = [material(diffuse=(0.7, 0.7, 0.7), diffuse_strength=1.0, diffuse_roughness=0.0)]
    verts, faces, sidx = [], [], []
    nv = 0
    h = 0.25   # half size of the light and of the ceiling opening it sits in
    walls = [quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)),        # floor
             # ceiling with a 0.5 x 0.5 opening for the light (four quads around it)
             quad((-1, 2, -1), (-h, 2, -1), (-h, 2, 1), (-1, 2, 1)),
             quad((h, 2, -1), (1, 2, -1), (1, 2, 1), (h, 2, 1)),
             quad((-h, 2, -1), (h, 2, -1), (h, 2, -h), (-h, 2, -h)),
             quad((-h, 2, h), (h, 2, h), (h, 2, 1), (-h, 2, 1)),
             quad((-1, 0, -1), (-1, 2, -1), (1, 2, -1), (1, 0, -1)),      # back wall
             quad((-1, 0, -1), (-1, 0, 1), (-1, 2, 1), (-1, 2, -1)),      # left wall
             quad((1, 0, -1), (1, 2, -1), (1, 2, 1), (1, 0, 1))]          # right wall
    for v, f in walls:
        # the reference shades one side only (core/shader.go:265 skips samples below N): wind inwards
        n = np.cross(v[f[0, 1]] - v[f[0, 0]], v[f[0, 2]] - v[f[0, 0]])
        if np.dot(n, np.array([0.0, 1.0, 0.0]) - v[0]) < 0:
            f = f[:, ::-1]
        verts.append(v)
        faces.append(f + nv)
        sidx.append(np.zeros(len(f), np.int64))
        nv += len(v)
    sv, sf = uv_sphere(1.0, nseg, nring)
    sv = sv.astype(np.float64)
    for s in range(nspheres):
        r = 0.22
        c = np.array([-0.6 + 0.4 * (s % 4), r + 0.05 + 0.9 * (s // 4), -0.3 + 0.3 * (s % 2)])
        disp = r + amplitude * (2.0 * value_noise(sv, seed * 100 + s) - 1.0)
        v = sv * disp[:, None] + c
        mats.append(material(diffuse=tuple(g.uniform(0.2, 0.8, 3)), diffuse_strength=0.6, diffuse_roughness=0.0,
                             spec=(1.0, 1.0, 1.0), spec_strength=0.4, spec_roughness=float(g.uniform(0.05, 0.5)),
                             ior=1.5))
        verts.append(v)
        faces.append(sf.astype(np.int64) + nv)
        sidx.append(np.full(len(sf), 1 + s, np.int64))
        nv += len(v)
    light_mat = len(mats)
    mats.append(material(emission=(1.0, 1.0, 1.0), emission_strength=10.0, diffuse=(0.0, 0.0, 0.0),
                         diffuse_strength=1.0))
    main = {"verts": np.concatenate(verts).astype(np.float32), "idx": np.concatenate(faces).astype(np.uint32),
            "shaderidx": np.concatenate(sidx).astype(np.uint8), "shaders": np.arange(1 + nspheres, dtype=np.int32),
            "normals": None}
    # Light quad recessed above the ceiling opening, facing down (cross(P1-P0, P2-P0) = -y).  Every
    # shading point then lies at least 1 mm below the light's plane: the reference's spherical-triangle
    # sampler takes sqrt of a negative value (NaN pdf) for points within ~1e-4 of that plane
    # (builtin/light/triangle.go:376-411).  The second triangle sits 1 mm higher because two halves of
    # one square share a bounding box, and the scene's leafMax-1 BuildAccel never terminates on
    # coincident centroids (qbvh/build.go:35-43).
    y, y2 = 2.001, 2.002
    L = [((-h, y, -h), (h, y, -h), (h, y, h)), ((-h, y2, -h), (h, y2, h), (-h, y2, h))]
    meshes = [main]
    lights = []
    for k, (p0, p1, p2) in enumerate(L):
        P = np.array([p0, p1, p2], np.float32)
        n = np.cross(P[1] - P[0], P[2] - P[0])
        n = (n / np.linalg.norm(n)).astype(np.float32)
        meshes.append({"verts": P, "idx": np.array([[0, 1, 2]], np.uint32), "shaderidx": None,
                       "shaders": np.array([light_mat], np.int32), "normals": np.tile(n, (3, 1))})
        lights.append({"P0": P[0], "P1": P[1], "P2": P[2], "material": light_mat, "samples": 1, "geom": 1 + k})
    # up = -y: the reference's LookAt basis takes v = u x W (camera.go:164), which points down for up = +y
    camera = {"from": (0.0, 1.0, 3.4), "to": (0.0, 1.0, 0.0), "up": (0.0, -1.0, 0.0), "fov": 45.0, "focal": 12.0}
    return {"meshes": meshes, "materials": np.array(mats, MATERIAL_DTYPE), "lights": lights, "camera": camera}


def pixel_scrambles(width, height, seed=99):
    """6 uint64 per pixel: lensU, lensV, time, lambda, scramble[0], scramble[1] (core/render.go:166-176)."""
    return rng(seed).integers(0, np.iinfo(np.uint64).max, (height, width, 6), dtype=np.uint64, endpoint=True)
