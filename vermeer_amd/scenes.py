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


def van_der_corput(n):
    """Base-2 radical inverse of 0..n-1 (32 bits), as float32."""
    k = np.arange(n, dtype=np.uint64)
    rev = np.zeros(n, np.float64)
    for b in range(32):
        rev += ((k >> np.uint64(b)) & np.uint64(1)).astype(np.float64) * 2.0 ** -(b + 1)
    return rev.astype(np.float32)


def moving_spheres_camera_rays(width=1024, spp=4):
    """Camera rays of the motion-blur workload: width^2 pinhole rays (60 degree FOV from (0, 0, 45)
    towards the origin) repeated spp times, ray k at time VanDerCorput(k)."""
    base = pinhole_rays(width, width, fov_deg=60.0, eye=(0.0, 0.0, 45.0), target=(0.0, 0.0, 0.0))
    rays = np.concatenate([base] * spp)
    rays["time"] = van_der_corput(len(rays))
    return rays


def point_light_shadow_rays(rays, prim, t, light=(10.0, 30.0, 20.0)):
    """One shadow ray per camera ray towards a point light, at the camera ray's time: from the hit
    point P = O + t D, origin P + 1e-4 (L - P), direction L - P, tmax 1 - 1e-4; a ray whose camera ray
    missed gets tmax -1 (it can hit nothing).  float32 throughout."""
    f = np.float32
    O = rays["P"].astype(f)
    D = rays["D"].astype(f)
    hit = np.asarray(prim) >= 0
    tt = np.where(hit, np.asarray(t, f), f(0))
    P = O + D * tt[:, None]
    L = np.asarray(light, f)
    out = np.zeros(len(rays), RAY_DTYPE)
    out["P"] = P + (L - P) * f(1e-4)
    out["D"] = L - P
    out["tmax"] = np.where(hit, f(1.0 - 1e-4), f(-1.0))
    out["time"] = rays["time"]
    return out


# ------------------------------------------------------------ shaded scenes

# ShaderStd parameters (builtin/shader/std.go:25-47); has_* flags mark parameters that are set.
from ._lib import MATERIAL_DTYPE  # noqa: E402

FRESNEL_DIELECTRIC = 1
FRESNEL_METAL = 2


def material(diffuse=None, diffuse_strength=None, diffuse_roughness=None, spec=None, spec_strength=None,
             spec_roughness=None, ior=None, emission=None, emission_strength=None, fresnel="dielectric",
             fresnel_refl=None, fresnel_edge=None):
    """ShaderStd parameters (builtin/shader/std.go:25-47); None leaves a parameter unset.  fresnel:
    "dielectric" (uses ior) or "metal" (fresnel/conductor.go; fresnel_refl / fresnel_edge colours)."""
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
    m["fresnel_model"] = FRESNEL_METAL if fresnel == "metal" else FRESNEL_DIELECTRIC
    for key, val in (("spec1_fresnel_refl", fresnel_refl), ("spec1_fresnel_edge", fresnel_edge)):
        if val is not None:
            m[key] = val
            m["has_" + key] = 1
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


def _normalize32(a):
    """math/vec3_amd64.s:11-42 with the IEEE 1/sqrt estimate, as the renderer's device code computes it."""
    a = np.asarray(a, np.float32)
    l2 = (a[1] * a[1] + a[0] * a[0]) + a[2] * a[2]
    g = np.float32(1.0) / np.sqrt(l2)
    gg = (g * g) * (l2 * np.float32(0.5))
    r = (np.float32(1.5) - gg) * g
    return a * r


def _cross32(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], np.float32)


def disk_light_mesh(P, up, look_at, radius, segments=20):
    """The triangle fan a DiskLight inserts (Disk.createMesh, builtin/light/disk.go:224-262) in float32:
    frame N = normalize(LookAt - P), T = normalize(N x Up), B = N x T (disk.go:84-86); triangle i is
    (P, P + R cos(a) B + R sin(a) T, the same at a + 2 pi / segments), a accumulated in float32.
    Returns (verts (3*segments, 3), idx (segments, 3), normals (3*segments, 3))."""
    P = np.asarray(P, np.float32)
    N = _normalize32(np.asarray(look_at, np.float32) - P)
    T = _normalize32(_cross32(N, up))
    B = _cross32(N, T)
    pi32 = np.float32(3.14159265358)
    R = np.float32(radius)
    dv = (np.float32(2) * pi32) / np.float32(segments)
    ang = np.float32(0)
    verts = []
    for _ in range(segments):
        c0, s0 = np.float32(math.cos(float(ang))), np.float32(math.sin(float(ang)))
        a1 = np.float32(ang + dv)
        c1, s1 = np.float32(math.cos(float(a1))), np.float32(math.sin(float(a1)))
        verts.append(P)
        verts.append(P + (B * (R * c0) + T * (R * s0)))
        verts.append(P + (B * (R * c1) + T * (R * s1)))
        ang = np.float32(ang + dv)
    v = np.array(verts, np.float32)
    idx = np.arange(3 * segments, dtype=np.uint32).reshape(-1, 3)
    return v, idx, np.tile(N, (3 * segments, 1)).astype(np.float32)


def feature_scene(seed=5, nseg=48, nring=24, mirrors=True):
    """Box room exercising the renderer paths beyond box_spheres_scene: a conductor ("Metal") GGX sphere
    with Spec1FresnelEdge set, a dielectric GGX sphere, two facing perfect-mirror spheres (recursion to
    level 3 when mirrors=True), a DiskLight in the ceiling (4 samples, MIS) and a SphereLight (an analytic
    sphere geom, 2 samples).  Geom insertion order: the box+spheres mesh, the disk's fan, the light sphere."""
    g = rng(seed)
    mats = [material(diffuse=(0.7, 0.7, 0.7), diffuse_strength=1.0, diffuse_roughness=0.3)]
    verts, faces, sidx = [], [], []
    nv = 0
    walls = [quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)),
             quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)),
             quad((-1, 0, -1), (-1, 2, -1), (1, 2, -1), (1, 0, -1)),
             quad((-1, 0, -1), (-1, 0, 1), (-1, 2, 1), (-1, 2, -1)),
             quad((1, 0, -1), (1, 2, -1), (1, 2, 1), (1, 0, 1))]
    for v, f in walls:
        n = np.cross(v[f[0, 1]] - v[f[0, 0]], v[f[0, 2]] - v[f[0, 0]])
        if np.dot(n, np.array([0.0, 1.0, 0.0]) - v[0]) < 0:
            f = f[:, ::-1]
        verts.append(v)
        faces.append(f + nv)
        sidx.append(np.zeros(len(f), np.int64))
        nv += len(v)
    sphere_mats = [
        material(diffuse=(0.3, 0.3, 0.3), diffuse_strength=0.2, spec=(1.0, 0.9, 0.8), spec_strength=0.8,
                 spec_roughness=0.3, fresnel="metal", fresnel_refl=(0.95, 0.64, 0.54), fresnel_edge=(0.9, 0.8, 0.6)),
        material(diffuse=tuple(g.uniform(0.2, 0.8, 3)), diffuse_strength=0.6, diffuse_roughness=0.0,
                 spec=(1.0, 1.0, 1.0), spec_strength=0.4, spec_roughness=0.15, ior=1.5),
        material(spec=(1.0, 1.0, 1.0), spec_strength=1.0, spec_roughness=0.0 if mirrors else 0.05, ior=1.5),
        material(diffuse=(0.2, 0.3, 0.6), diffuse_strength=0.3, spec=(0.9, 0.9, 0.9), spec_strength=0.7,
                 spec_roughness=0.0 if mirrors else 0.05, fresnel="metal", fresnel_refl=(0.9, 0.9, 0.9)),
    ]
    centres = [(-0.5, 0.3, -0.2), (0.5, 0.3, -0.4), (-0.15, 0.75, 0.3), (0.45, 0.75, 0.35)]
    sv, sf = uv_sphere(1.0, nseg, nring)
    for k, (c, m) in enumerate(zip(centres, sphere_mats)):
        r = 0.28
        v = sv.astype(np.float64) * r + np.asarray(c)
        mats.append(m)
        verts.append(v)
        faces.append(sf.astype(np.int64) + nv)
        sidx.append(np.full(len(sf), 1 + k, np.int64))
        nv += len(v)
    light_mat = len(mats)
    mats.append(material(emission=(1.0, 0.95, 0.9), emission_strength=8.0, diffuse=(0.0, 0.0, 0.0),
                         diffuse_strength=1.0))
    sphere_light_mat = len(mats)
    mats.append(material(emission=(0.9, 0.95, 1.0), emission_strength=20.0, diffuse=(0.0, 0.0, 0.0),
                         diffuse_strength=1.0))
    main = {"verts": np.concatenate(verts).astype(np.float32), "idx": np.concatenate(faces).astype(np.uint32),
            "shaderidx": np.concatenate(sidx).astype(np.uint8),
            "shaders": np.arange(1 + len(sphere_mats), dtype=np.int32), "normals": None}
    dP, dUp, dLook, dR = (0.0, 1.99, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.3
    dv, di, dn = disk_light_mesh(dP, dUp, dLook, dR, 20)
    disk = {"verts": dv, "idx": di, "shaderidx": None, "shaders": np.array([light_mat], np.int32), "normals": dn}
    sph = {"sphere": True, "centre": (-0.6, 1.45, 0.55), "radius": 0.08, "material": sphere_light_mat}
    lights = [{"type": "disk", "P": dP, "Up": dUp, "LookAt": dLook, "radius": dR, "material": light_mat,
               "samples": 2, "geom": 1},
              {"type": "sphere", "P": sph["centre"], "radius": sph["radius"], "material": sphere_light_mat,
               "samples": 1, "geom": 2}]
    camera = {"from": (0.0, 1.0, 3.4), "to": (0.0, 0.8, 0.0), "up": (0.0, -1.0, 0.0), "fov": 45.0, "focal": 12.0}
    return {"meshes": [main, disk, sph], "materials": np.array(mats, MATERIAL_DTYPE), "lights": lights,
            "camera": camera}


def trs_matrix(t=(0.0, 0.0, 0.0), yaw=0.0, pitch=0.0, scale=(1.0, 1.0, 1.0)):
    """Column-major 4x4 (16 float32): translate(t) * rot_y(yaw) * rot_x(pitch) * scale."""
    cy, sy, cp, sp = np.cos(yaw), np.sin(yaw), np.cos(pitch), np.sin(pitch)
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rx = np.array([[1, 0, 0], [0, cp, -sp], [0, sp, cp]])
    m = np.eye(4)
    m[:3, :3] = ry @ rx @ np.diag(scale)
    m[:3, 3] = t
    return m.astype(np.float32).T.reshape(16)


def instance_scene(seed=11, count=12, nseg=24, nring=12):
    """Transforms in the scene graph: the feature scene's box room and lights (its spheres dropped), with
    three hidden prototype meshes (a displaced UV sphere, an icosphere, a box) held by `count`
    GeomInstances of random rotation, non-uniform scale and position (every third one with two keys: it
    moves and turns over the shutter); a Proc holding two hidden meshes (a small sphere and a slab) under
    a rotation; and a visible mesh with its own two-key Transform (a plank sliding inside its box, so the
    reference's untransformed bounds still cover it).  Materials: Oren-Nayar, dielectric GGX, conductor
    GGX and a perfect mirror.  Geom insertion order: box, disk fan, the three prototypes, the proc's two
    meshes, the plank; then the instances, then the proc."""
    g = rng(seed)
    base = feature_scene(seed=seed, nseg=8, nring=4, mirrors=True)
    mats = [base["materials"][0], base["materials"][5]]   # walls, disk light
    box_v, box_f = [], []
    nv = 0
    for v, f in [quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)),
                 quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)),
                 quad((-1, 0, -1), (-1, 2, -1), (1, 2, -1), (1, 0, -1)),
                 quad((-1, 0, -1), (-1, 0, 1), (-1, 2, 1), (-1, 2, -1)),
                 quad((1, 0, -1), (1, 2, -1), (1, 2, 1), (1, 0, 1))]:
        n = np.cross(v[f[0, 1]] - v[f[0, 0]], v[f[0, 2]] - v[f[0, 0]])
        if np.dot(n, np.array([0.0, 1.0, 0.0]) - v[0]) < 0:
            f = f[:, ::-1]
        box_v.append(v)
        box_f.append(f + nv)
        nv += len(v)
    box = {"verts": np.concatenate(box_v).astype(np.float32), "idx": np.concatenate(box_f).astype(np.uint32),
           "shaderidx": None, "shaders": np.array([0], np.int32), "normals": None}
    disk = dict(base["meshes"][1], shaders=np.array([1], np.int32))
    lights = [dict(base["lights"][0], material=1, geom=1)]
    m_on = len(mats); mats.append(material(diffuse=(0.75, 0.45, 0.3), diffuse_strength=1.0, diffuse_roughness=0.4))
    m_gx = len(mats); mats.append(material(diffuse=(0.2, 0.5, 0.3), diffuse_strength=0.5, spec=(1, 1, 1),
                                           spec_strength=0.5, spec_roughness=0.25, ior=1.5))
    m_mt = len(mats); mats.append(material(diffuse=(0.3, 0.3, 0.3), diffuse_strength=0.2, spec=(1.0, 0.9, 0.8),
                                           spec_strength=0.8, spec_roughness=0.3, fresnel="metal",
                                           fresnel_refl=(0.95, 0.64, 0.54)))
    m_mr = len(mats); mats.append(material(spec=(1, 1, 1), spec_strength=1.0, spec_roughness=0.0, ior=1.5))
    sv, sf = uv_sphere(1.0, nseg, nring)
    sv = sv * (1.0 + 0.08 * (2.0 * value_noise(sv, seed) - 1.0))[:, None]
    blob = {"verts": sv.astype(np.float32), "idx": sf.astype(np.uint32), "shaderidx": None,
            "shaders": np.array([m_on], np.int32), "normals": None, "hidden": True}
    iv, i_f = icosphere(2)
    ico = {"verts": iv.astype(np.float32), "idx": i_f.astype(np.uint32), "shaderidx": None,
           "shaders": np.array([m_mt], np.int32), "normals": None, "hidden": True}
    cube_v = np.array([(x, y, z) for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float32)
    cube_f = np.array([[0, 1, 3], [0, 3, 2], [4, 6, 7], [4, 7, 5], [0, 4, 5], [0, 5, 1],
                       [2, 3, 7], [2, 7, 6], [0, 2, 6], [0, 6, 4], [1, 5, 7], [1, 7, 3]], np.uint32)
    cube = {"verts": cube_v, "idx": cube_f, "shaderidx": None, "shaders": np.array([m_gx], np.int32),
            "normals": None, "hidden": True}
    small = {"verts": (sv * 0.12 + np.array([0.0, 0.15, 0.0])).astype(np.float32), "idx": sf.astype(np.uint32),
             "shaderidx": None, "shaders": np.array([m_mr], np.int32), "normals": None, "hidden": True}
    slab_v = cube_v * np.array([0.25, 0.02, 0.12], np.float32)
    slab = {"verts": slab_v, "idx": cube_f, "shaderidx": None, "shaders": np.array([m_gx], np.int32),
            "normals": None, "hidden": True}
    plank_v = cube_v * np.array([0.3, 0.03, 0.1], np.float32) + np.array([0.0, 0.05, 0.6], np.float32)
    plank = {"verts": plank_v, "idx": cube_f, "shaderidx": None, "shaders": np.array([m_on], np.int32),
             "normals": None,
             "transform": np.stack([trs_matrix((-0.05, 0.0, 0.0), yaw=0.05), trs_matrix((0.05, 0.0, 0.0), yaw=-0.05)])}
    meshes = [box, disk, blob, ico, cube, small, slab, plank]
    instances = []
    for k in range(count):
        proto = 2 + k % 3
        sc = g.uniform(0.08, 0.2, 3) * (np.array([1.0, 1.0, 1.0]) if proto != 4 else np.array([1.0, 0.6, 1.4]))
        pos = np.array([g.uniform(-0.75, 0.75), g.uniform(0.15, 1.6), g.uniform(-0.75, 0.5)])
        yaw, pitch = g.uniform(0, 2 * np.pi), g.uniform(-0.6, 0.6)
        mats_k = [trs_matrix(pos, yaw, pitch, sc)]
        if k % 3 == 0:
            mats_k.append(trs_matrix(pos + g.uniform(-0.08, 0.08, 3), yaw + 0.4, pitch, sc))
        r = float(np.max(sc)) * 1.75 + 0.1
        instances.append({"geom": proto, "matrices": np.stack(mats_k), "bmin": pos - r, "bmax": pos + r})
    procs = [{"geoms": [5, 6], "matrices": trs_matrix((0.55, 0.35, 0.25), yaw=0.7, pitch=0.2)[None, :],
              "bmin": (0.2, 0.0, -0.1), "bmax": (0.9, 0.7, 0.6)}]
    camera = dict(base["camera"])
    return {"meshes": meshes, "materials": np.array(mats, MATERIAL_DTYPE), "lights": lights, "camera": camera,
            "instances": instances, "procs": procs}


def _rot_y(theta):
    c, s_ = np.cos(theta), np.sin(theta)
    return np.array([[c, 0.0, s_], [0.0, 1.0, 0.0], [-s_, 0.0, c]])


def moving_scene(seed=7, nseg=32, nring=16, keys=3, mirrors=True):
    """Motion-blur room for the renderer: the feature scene's box (static) and lights, with three moving
    meshes.  Mesh "spheres": three UV spheres (diffuse, conductor GGX, perfect mirror) that translate and
    spin about y over `keys` keys, with per-corner normals for every key.  Mesh "blob": one sphere with a
    single-key normal array (ignored: N = Ng) and the reference's slot indexing, two shaders by
    hemisphere.  Geom insertion order: box, spheres, blob, disk fan."""
    g = rng(seed)
    base = feature_scene(seed=seed, nseg=8, nring=4, mirrors=mirrors)
    mats = [base["materials"][0]]
    box_v, box_f = [], []
    nv = 0
    for v, f in [quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)),
                 quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)),
                 quad((-1, 0, -1), (-1, 2, -1), (1, 2, -1), (1, 0, -1)),
                 quad((-1, 0, -1), (-1, 0, 1), (-1, 2, 1), (-1, 2, -1)),
                 quad((1, 0, -1), (1, 2, -1), (1, 2, 1), (1, 0, 1))]:
        n = np.cross(v[f[0, 1]] - v[f[0, 0]], v[f[0, 2]] - v[f[0, 0]])
        if np.dot(n, np.array([0.0, 1.0, 0.0]) - v[0]) < 0:
            f = f[:, ::-1]
        box_v.append(v)
        box_f.append(f + nv)
        nv += len(v)
    box = {"verts": np.concatenate(box_v).astype(np.float32), "idx": np.concatenate(box_f).astype(np.uint32),
           "shaderidx": None, "shaders": np.array([0], np.int32), "normals": None}
    sphere_mats = [
        material(diffuse=(0.8, 0.3, 0.2), diffuse_strength=1.0, diffuse_roughness=0.2),
        material(diffuse=(0.3, 0.3, 0.3), diffuse_strength=0.2, spec=(1.0, 0.9, 0.8), spec_strength=0.8,
                 spec_roughness=0.25, fresnel="metal", fresnel_refl=(0.95, 0.64, 0.54)),
        material(spec=(1.0, 1.0, 1.0), spec_strength=1.0, spec_roughness=0.0 if mirrors else 0.05, ior=1.5),
    ]
    sv, sf = uv_sphere(1.0, nseg, nring)
    sv = sv.astype(np.float64)
    corners = sf.reshape(-1).astype(np.int64)
    kv, kn, faces, sidx = [[] for _ in range(keys)], [[] for _ in range(keys)], [], []
    nv = 0
    for k, m in enumerate(sphere_mats):
        c0 = np.array([-0.5 + 0.5 * k, 0.35 + 0.1 * k, -0.2 + 0.1 * k])
        move = g.uniform(-0.25, 0.25, 3)
        spin = g.uniform(0.3, 1.2)
        for key in range(keys):
            t = key / (keys - 1)
            R = _rot_y(spin * t)
            r = 0.25 * (1.0 + 0.2 * t)
            kv[key].append((sv @ R.T) * r + c0 + move * t)
            kn[key].append((sv @ R.T)[corners])
        mats.append(m)
        faces.append(sf.astype(np.int64) + nv)
        sidx.append(np.full(len(sf), k, np.int64))
        nv += len(sv)
    spheres = {"keys": keys, "verts": np.stack([np.concatenate(v) for v in kv]).astype(np.float32),
               "idx": np.concatenate(faces).astype(np.uint32), "shaderidx": np.concatenate(sidx).astype(np.uint8),
               "shaders": np.arange(1, 1 + len(sphere_mats), dtype=np.int32),
               "normals": np.stack([np.concatenate(n) for n in kn]).astype(np.float32)}
    blob_mats = [material(diffuse=(0.2, 0.7, 0.3), diffuse_strength=1.0),
                 material(diffuse=(0.2, 0.3, 0.8), diffuse_strength=0.5, spec=(1.0, 1.0, 1.0), spec_strength=0.5,
                          spec_roughness=0.4)]
    bv = [sv * 0.18 + np.array([0.55, 0.25 + 0.3 * key / (keys - 1), 0.45]) for key in range(keys)]
    blob = {"keys": keys, "verts": np.stack(bv).astype(np.float32), "idx": sf.astype(np.uint32),
            "shaderidx": (sv[sf[:, 0], 1] > 0).astype(np.uint8),
            "shaders": np.array([len(mats), len(mats) + 1], np.int32),
            "normals": sv[corners][None].astype(np.float32), "reference_indexing": True}
    mats += blob_mats
    light_mat = len(mats)
    mats.append(material(emission=(1.0, 0.95, 0.9), emission_strength=8.0, diffuse=(0.0, 0.0, 0.0),
                         diffuse_strength=1.0))
    dP, dUp, dLook, dR = (0.0, 1.99, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.35
    dv, di, dn = disk_light_mesh(dP, dUp, dLook, dR, 20)
    disk = {"verts": dv, "idx": di, "shaderidx": None, "shaders": np.array([light_mat], np.int32), "normals": dn}
    lights = [{"type": "disk", "P": dP, "Up": dUp, "LookAt": dLook, "radius": dR, "material": light_mat,
               "samples": 2, "geom": 3}]
    return {"meshes": [box, spheres, blob, disk], "materials": np.array(mats, MATERIAL_DTYPE), "lights": lights,
            "camera": base["camera"]}


def moving_spheres_render_scene(count=2000, seed=21):
    """The 2000 moving spheres (500k triangles, two keys) as a render: a diffuse shaded motion mesh in
    front of a diffuse backdrop (z = -25), lit by a DiskLight (radius 5, one sample: one BSDF and one
    light shadow ray per hit).  Camera at (0, 0, 45) looking at the origin, fov 60."""
    verts, idx = moving_spheres(count=count, seed=seed)
    mats = [material(diffuse=(0.7, 0.6, 0.5), diffuse_strength=1.0, diffuse_roughness=0.2),
            material(diffuse=(0.5, 0.5, 0.55), diffuse_strength=1.0)]
    spheres = {"keys": 2, "verts": verts, "idx": idx, "shaderidx": None, "shaders": np.array([0], np.int32),
               "normals": None}
    bv, bf = quad((-60, -60, -25), (60, -60, -25), (60, 60, -25), (-60, 60, -25))
    backdrop = {"verts": bv.astype(np.float32), "idx": bf.astype(np.uint32), "shaderidx": None,
                "shaders": np.array([1], np.int32), "normals": None}
    light_mat = len(mats)
    mats.append(material(emission=(1.0, 0.95, 0.9), emission_strength=30.0, diffuse=(0.0, 0.0, 0.0),
                         diffuse_strength=1.0))
    dP, dUp, dLook, dR = (0.0, 30.0, 30.0), (1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 5.0
    dv, di, dn = disk_light_mesh(dP, dUp, dLook, dR, 20)
    disk = {"verts": dv, "idx": di, "shaderidx": None, "shaders": np.array([light_mat], np.int32), "normals": dn}
    lights = [{"type": "disk", "P": dP, "Up": dUp, "LookAt": dLook, "radius": dR, "material": light_mat,
               "samples": 1, "geom": 2}]
    camera = {"from": (0.0, 0.0, 45.0), "to": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0), "fov": 60.0, "focal": 12.0}
    return {"meshes": [spheres, backdrop, disk], "materials": np.array(mats, MATERIAL_DTYPE), "lights": lights,
            "camera": camera}


def pixel_scrambles(width, height, seed=99):
    """6 uint64 per pixel: lensU, lensV, time, lambda, scramble[0], scramble[1] (core/render.go:166-176)."""
    return rng(seed).integers(0, np.iinfo(np.uint64).max, (height, width, 6), dtype=np.uint64, endpoint=True)


# --------------------------------------------------------------------- path mode: the occluded-light room

PT_MATERIAL_DTYPE = np.dtype([("albedo", "<f4", 3), ("emission", "<f4", 3), ("one_sided", "<i4"), ("pad", "<i4")])


def pt_material(albedo=(0.0, 0.0, 0.0), emission=(0.0, 0.0, 0.0), one_sided=False):
    m = np.zeros((), PT_MATERIAL_DTYPE)
    m["albedo"] = albedo
    m["emission"] = emission
    m["one_sided"] = int(one_sided)
    return m


def _grid_wall(origin, eu, ev, nu, nv, inward, amp, seed, window, hole=None):
    """nu x nv quads spanning origin + [0,1] eu + [0,1] ev, displaced along `inward` by value noise of
    amplitude amp times window(point) (0 at the wall border); quads whose centre lies in `hole`
    (a predicate on the centre) are left out.  Triangles are wound to face `inward`."""
    o, eu, ev, inward = (np.asarray(a, np.float64) for a in (origin, eu, ev, inward))
    i, j = np.meshgrid(np.arange(nu + 1), np.arange(nv + 1), indexing="ij")
    P = o + (i[..., None] / nu) * eu + (j[..., None] / nv) * ev
    P = P.reshape(-1, 3)
    d = amp * (2.0 * value_noise(P, seed, freq=2.0) - 1.0) * window(P)
    P = P + d[:, None] * inward
    ci, cj = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    ci, cj = ci.reshape(-1), cj.reshape(-1)
    if hole is not None:
        ctr = o + ((ci[:, None] + 0.5) / nu) * eu + ((cj[:, None] + 0.5) / nv) * ev
        keep = ~hole(ctr)
        ci, cj = ci[keep], cj[keep]
    a = ci * (nv + 1) + cj
    b = (ci + 1) * (nv + 1) + cj
    c = (ci + 1) * (nv + 1) + cj + 1
    dd = ci * (nv + 1) + cj + 1
    f = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, dd], 1)])
    if np.dot(np.cross(eu, ev), inward) < 0:
        f = f[:, ::-1]
    return P, f


def _unit_box_tris():
    corners = np.array([[x, y, z] for z in (-1, 1) for y in (-1, 1) for x in (-1, 1)], np.float64)
    quads = [(0, 2, 6, 4), (1, 3, 7, 5), (0, 1, 5, 4), (2, 3, 7, 6), (0, 1, 3, 2), (4, 5, 7, 6)]
    tris = []
    for q in quads:
        for t in ((q[0], q[1], q[2]), (q[0], q[2], q[3])):
            p = corners[list(t)]
            n = np.cross(p[1] - p[0], p[2] - p[0])
            if np.dot(n, p.mean(0)) < 0:
                t = (t[0], t[2], t[1])
            tris.append(t)
    return corners, np.array(tris, np.int64)


def _random_rotations(g, n):
    q = g.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
                     np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
                     np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)


ROOM_SLIT = (-1.2, 1.2, 2.8, 3.2)          # z0, z1, y0, y1 of the slit in the x = +4 wall
ROOM_ALCOVE = (4.0, 4.8, 2.4, 3.6, -1.6, 1.6)


def guided_room_scene(seed=11, res=20, nboxes=66667, emission=2000.0, amplitude=0.03):
    """The occluded-light room: a sealed 8 x 4 x 8 room whose six walls are res-per-metre grids of
    noise-displaced quads (res 20: 204,000 triangles), a slit of 2.4 x 0.4 in the x = +4 wall opening
    into an alcove behind it, and `nboxes` randomly rotated boxes of 12 triangles (66,667: 800,004
    triangles) cluttering the lower 1.5 m.  The only light is a 2.0 x 0.05 rectangle inside the alcove
    facing away from the slit (one-sided), so light reaches the room only after bouncing off the
    alcove's walls.  Materials are Lambertian.

    Returns a dict: meshes (main mesh, light mesh), materials (PT_MATERIAL_DTYPE), camera."""
    g = rng(seed)
    z0, z1, y0, y1 = ROOM_SLIT
    ax0, ax1, ay0, ay1, az0, az1 = ROOM_ALCOVE
    margin = 0.3

    def border_window(lo, hi):
        lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)

        def w(P):
            d = np.full(len(P), np.inf)
            for k in range(3):
                if hi[k] > lo[k]:
                    d = np.minimum(d, np.minimum(P[:, k] - lo[k], hi[k] - P[:, k]))
            t = np.clip(d / margin, 0.0, 1.0)
            return t * t * (3 - 2 * t)
        return w

    def slit_wall_window(P):
        w = border_window((4, 0, -4), (4, 4, 4))(P)
        dy = np.maximum(np.maximum(ay0 - P[:, 1], P[:, 1] - ay1), 0.0)
        dz = np.maximum(np.maximum(az0 - P[:, 2], P[:, 2] - az1), 0.0)
        t = np.clip(np.hypot(dy, dz) / margin, 0.0, 1.0)
        return w * t * t * (3 - 2 * t)

    def slit_hole(C):
        return (C[:, 2] > z0) & (C[:, 2] < z1) & (C[:, 1] > y0) & (C[:, 1] < y1)

    n8, n4 = 8 * res, 4 * res
    walls = [  # origin, eu, ev, nu, nv, inward, shader, window, hole
        ((-4, 0, -4), (8, 0, 0), (0, 0, 8), n8, n8, (0, 1, 0), 1, border_window((-4, 0, -4), (4, 0, 4)), None),
        ((-4, 4, -4), (8, 0, 0), (0, 0, 8), n8, n8, (0, -1, 0), 2, border_window((-4, 4, -4), (4, 4, 4)), None),
        ((-4, 0, -4), (0, 0, 8), (0, 4, 0), n8, n4, (1, 0, 0), 0, border_window((-4, 0, -4), (-4, 4, 4)), None),
        ((4, 0, -4), (0, 0, 8), (0, 4, 0), n8, n4, (-1, 0, 0), 0, slit_wall_window, slit_hole),
        ((-4, 0, -4), (8, 0, 0), (0, 4, 0), n8, n4, (0, 0, 1), 0, border_window((-4, 0, -4), (4, 4, -4)), None),
        ((-4, 0, 4), (8, 0, 0), (0, 4, 0), n8, n4, (0, 0, -1), 0, border_window((-4, 0, 4), (4, 4, 4)), None),
    ]
    verts, faces, sidx = [], [], []
    nv = 0
    for k, (o, eu, ev, nu, nvv, inward, sh, win, hole) in enumerate(walls):
        P, f = _grid_wall(o, eu, ev, nu, nvv, inward, amplitude, seed * 10 + k, win, hole)
        verts.append(P)
        faces.append(f + nv)
        sidx.append(np.full(len(f), sh, np.int64))
        nv += len(P)
    alcove = [quad((ax1, ay0, az0), (ax1, ay0, az1), (ax1, ay1, az1), (ax1, ay1, az0)),
              quad((ax0, ay1, az0), (ax1, ay1, az0), (ax1, ay1, az1), (ax0, ay1, az1)),
              quad((ax0, ay0, az0), (ax1, ay0, az0), (ax1, ay0, az1), (ax0, ay0, az1)),
              quad((ax0, ay0, az0), (ax1, ay0, az0), (ax1, ay1, az0), (ax0, ay1, az0)),
              quad((ax0, ay0, az1), (ax1, ay0, az1), (ax1, ay1, az1), (ax0, ay1, az1))]
    centre = np.array([0.5 * (ax0 + ax1), 0.5 * (ay0 + ay1), 0.0])
    for P, f in alcove:
        n = np.cross(P[f[0, 1]] - P[f[0, 0]], P[f[0, 2]] - P[f[0, 0]])
        if np.dot(n, centre - P[0]) < 0:
            f = f[:, ::-1]
        verts.append(P)
        faces.append(f + nv)
        sidx.append(np.full(len(f), 3, np.int64))
        nv += len(P)
    cam_from = np.array([2.6, 1.7, -2.6])
    if nboxes:
        corners, btris = _unit_box_tris()
        C = np.empty((0, 3))
        while len(C) < nboxes:
            c = np.stack([g.uniform(-3.7, 3.7, nboxes), g.uniform(0.05, 1.5, nboxes), g.uniform(-3.7, 3.7, nboxes)], 1)
            c = c[np.linalg.norm(c - cam_from, axis=1) > 0.8]
            C = np.concatenate([C, c])[:nboxes]
        H = 0.5 * np.exp(g.uniform(math.log(0.03), math.log(0.2), (nboxes, 3)))
        R = _random_rotations(g, nboxes)
        bv = np.einsum("nij,nkj->nki", R, corners[None] * H[:, None, :]) + C[:, None, :]
        verts.append(bv.reshape(-1, 3))
        faces.append((btris[None] + (nv + 8 * np.arange(nboxes))[:, None, None]).reshape(-1, 3))
        pal = g.integers(0, 32, nboxes)
        sidx.append(np.repeat(4 + pal, 12))
        nv += 8 * nboxes
    mats = [pt_material((0.7, 0.7, 0.7)), pt_material((0.5, 0.5, 0.5)), pt_material((0.8, 0.8, 0.8)),
            pt_material((0.8, 0.8, 0.8))]
    mats += [pt_material(tuple(g.uniform(0.2, 0.8, 3))) for _ in range(32)]
    light_mat = len(mats)
    mats.append(pt_material((0.0, 0.0, 0.0), (emission, emission, emission), one_sided=True))
    main = {"verts": np.concatenate(verts).astype(np.float32), "idx": np.concatenate(faces).astype(np.uint32),
            "shaderidx": np.concatenate(sidx).astype(np.uint8), "shaders": np.arange(light_mat, dtype=np.int32),
            "normals": None}
    lx, ly0, ly1 = 4.35, 2.975, 3.025
    LP = np.array([(lx, ly0, 1.0), (lx, ly0, -1.0), (lx, ly1, -1.0), (lx, ly1, 1.0)], np.float32)
    light = {"verts": LP, "idx": np.array([[0, 1, 2], [0, 2, 3]], np.uint32), "shaderidx": None,
             "shaders": np.array([light_mat], np.int32), "normals": None}
    # looking away from the slit (it is behind the camera): every visible surface is lit only by light
    # that left the alcove through the slit
    camera = {"from": tuple(cam_from), "to": (-4.0, 1.3, 1.2), "up": (0.0, -1.0, 0.0), "fov": 75.0, "focal": 1.0}
    return {"meshes": [main, light], "materials": np.array(mats, PT_MATERIAL_DTYPE), "camera": camera}


def furnace_box_scene(albedo=0.5, emission=1.0, subdiv=2):
    """Closed cube [-1,1]^3 whose inward-facing walls (subdiv x subdiv quads each) all have Lambertian
    albedo `albedo` and emit `emission` (one-sided, inwards).  Radiance seen from inside with paths of
    at most D segments: emission * (1 - albedo^D) / (1 - albedo)."""
    verts, faces = [], []
    nv = 0
    for axis in range(3):
        for s in (-1.0, 1.0):
            u, v = [k for k in range(3) if k != axis]
            o = np.zeros(3); o[axis] = s; o[u] = -1; o[v] = -1
            eu = np.zeros(3); eu[u] = 2
            ev = np.zeros(3); ev[v] = 2
            inward = np.zeros(3); inward[axis] = -s
            P, f = _grid_wall(o, eu, ev, subdiv, subdiv, inward, 0.0, 0, lambda p: np.zeros(len(p)))
            verts.append(P)
            faces.append(f + nv)
            nv += len(P)
    m = pt_material((albedo,) * 3, (emission,) * 3, one_sided=True)
    main = {"verts": np.concatenate(verts).astype(np.float32), "idx": np.concatenate(faces).astype(np.uint32),
            "shaderidx": None, "shaders": np.array([0], np.int32), "normals": None}
    camera = {"from": (0.1, -0.2, 0.3), "to": (0.5, 0.3, -1.0), "up": (0.0, -1.0, 0.0), "fov": 90.0, "focal": 1.0}
    return {"meshes": [main], "materials": np.array([m], PT_MATERIAL_DTYPE), "camera": camera}
