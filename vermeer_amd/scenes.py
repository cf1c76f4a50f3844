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
