"""Synthetic scenes and ray sets for the kd-tree build + traversal path.

Every generator is seeded with PCG32 (O'Neill 2014, ``pcg32_srandom_r`` /
``pcg32_random_r``), one stream per (purpose, element), so each array is a
pure function of (seed, sizes) and can be regenerated anywhere.  Geometry is
produced in float64 and rounded once to float32; -0.0 is folded to +0.0.

Triangle sets are returned as three ``float32[N, 4]`` arrays (p0, p1, p2 with
w = 1), the structure-of-arrays layout the reference's GeometryList keeps on
the device (Utils/CUDA/GeometryList.h:58-68).  Rays are ``float32[R, 4]``
origin (w = 1) and direction (w = 0) arrays, normalised directions, as
IRayTracer::CreateRays writes them (Utils/CUDA/IRayTracer.cu:55-70).
"""
from __future__ import annotations

import math

import numpy as np

_MULT = np.uint64(6364136223846793005)
_M32 = np.uint64(0xFFFFFFFF)

# stream purposes
P_CAMERA, P_NOISE, P_CLUSTER, P_TRI, P_PHOTON, P_BOUNCE, P_MESH, P_PERTURB, P_RAYS = range(1, 10)


class Pcg32:
    """Vectorised PCG32: one independent stream per element of ``index``."""

    def __init__(self, seed: int, purpose: int, index):
        idx = np.asarray(index, dtype=np.uint64)
        stream = (np.uint64(purpose) << np.uint64(40)) | idx
        self.inc = (stream << np.uint64(1)) | np.uint64(1)
        state = np.zeros_like(self.inc)
        state = state * _MULT + self.inc
        state = state + np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.state = state * _MULT + self.inc

    def next_u32(self) -> np.ndarray:
        old = self.state
        self.state = old * _MULT + self.inc
        xorshifted = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)) & _M32
        rot = old >> np.uint64(59)
        left = (xorshifted << ((np.uint64(32) - rot) & np.uint64(31))) & _M32
        return ((xorshifted >> rot) | left).astype(np.uint32)

    def uniform(self) -> np.ndarray:
        """float64 in [0, 1) with 24 random bits (exactly representable in float32)."""
        return (self.next_u32() >> np.uint32(8)).astype(np.float64) * (1.0 / 16777216.0)

    def normal(self) -> np.ndarray:
        """Standard normal via Box-Muller (one draw pair per call, cosine branch)."""
        u1 = 1.0 - self.uniform()
        u2 = self.uniform()
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def _f32(a) -> np.ndarray:
    a = np.asarray(a, dtype=np.float32)
    return a + np.float32(0.0)   # fold -0.0 to +0.0


def _pack(v: np.ndarray, w: float) -> np.ndarray:
    out = np.empty((v.shape[0], 4), dtype=np.float32)
    out[:, :3] = _f32(v)
    out[:, 3] = w
    return out


def tris_from(v0, v1, v2):
    """(N,3) vertex arrays -> (p0, p1, p2) float32[N,4] with w = 1."""
    return _pack(np.asarray(v0), 1.0), _pack(np.asarray(v1), 1.0), _pack(np.asarray(v2), 1.0)


# --------------------------------------------------------------------------
# meshes
# --------------------------------------------------------------------------
def quad(a, b, c, d):
    """two triangles (a,b,c), (a,c,d)"""
    return [(a, b, c), (a, c, d)]


def _subdivide_to_sphere(verts, faces, subdiv: int, radius: float, center):
    """midpoint subdivision, every new vertex pushed onto the unit sphere"""
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    for _ in range(subdiv):
        cache = {}
        new_faces = []

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = verts[i] + verts[j]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new_faces += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new_faces
    V = np.array(verts) * radius + np.asarray(center, dtype=np.float64)
    F = np.array(faces, dtype=np.int64)
    return V, F


def icosphere(subdiv: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)):
    """Icosahedron subdivided ``subdiv`` times: 20 * 4**subdiv triangles."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
             (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    return _subdivide_to_sphere(verts, faces, subdiv, radius, center)


def tetra_sphere(subdiv: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)):
    """Tetrahedron subdivided ``subdiv`` times onto the sphere: 4 * 4**subdiv
    triangles, every vertex shared by at most 6 triangles (a UV sphere's
    poles are fans of n_phi triangles through one point, which no kd-tree
    with fewer than n_phi triangles per leaf can separate)."""
    verts = [(1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)]
    faces = [(0, 1, 2), (0, 3, 1), (0, 2, 3), (1, 3, 2)]
    return _subdivide_to_sphere(verts, faces, subdiv, radius, center)


def uv_sphere(n_theta: int, n_phi: int, radius=1.0, center=(0.0, 0.0, 0.0)):
    """UV sphere with 2*n_phi*(n_theta-1) triangles (poles as fans)."""
    th = np.linspace(0.0, math.pi, n_theta + 1)
    ph = np.linspace(0.0, 2.0 * math.pi, n_phi + 1)[:-1]
    T, P = np.meshgrid(th, ph, indexing="ij")
    V = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], axis=-1).reshape(-1, 3)
    idx = np.arange((n_theta + 1) * n_phi).reshape(n_theta + 1, n_phi)
    faces = []
    for i in range(n_theta):
        a = idx[i]
        b = idx[i + 1]
        a1 = np.roll(a, -1)
        b1 = np.roll(b, -1)
        if i != n_theta - 1:
            faces.append(np.stack([a, b, b1], axis=1))
        if i != 0:
            faces.append(np.stack([a, b1, a1], axis=1))
    F = np.concatenate(faces)
    return V * radius + np.asarray(center, dtype=np.float64), F


# --------------------------------------------------------------------------
# cameras
# --------------------------------------------------------------------------
def pinhole_rays(width, height, eye, target, up=(0.0, 1.0, 0.0), fov_deg=45.0, seed=1,
                 jitter=True):
    """One ray per pixel, row-major (id = y * width + x), seeded U[0,1)^2
    jitter inside the pixel.  Screen position (x + jx) / W * 2 - 1 matches the
    reference's unprojection (IRayTracer.cu:62-65) apart from the jitter."""
    eye = np.asarray(eye, dtype=np.float64)
    fwd = np.asarray(target, dtype=np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(up, dtype=np.float64))
    right /= np.linalg.norm(right)
    upv = np.cross(right, fwd)
    n = width * height
    ids = np.arange(n, dtype=np.uint64)
    if jitter:
        rng = Pcg32(seed, P_CAMERA, ids)
        jx = rng.uniform()
        jy = rng.uniform()
    else:
        jx = np.zeros(n)
        jy = np.zeros(n)
    px = (ids % np.uint64(width)).astype(np.float64)
    py = (ids // np.uint64(width)).astype(np.float64)
    sx = (px + jx) / width * 2.0 - 1.0
    sy = (py + jy) / height * 2.0 - 1.0
    th = math.tan(math.radians(fov_deg) * 0.5)
    aspect = width / height
    d = fwd[None, :] + (sx * th * aspect)[:, None] * right[None, :] + (sy * th)[:, None] * upv[None, :]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    org = np.repeat(eye[None, :], n, axis=0)
    return _pack(org, 1.0), _pack(d, 0.0)


def camera_inverse(eye, target, up=(0.0, 1.0, 0.0), fov_deg=45.0, width=256, height=256,
                   near=0.1, far=100.0) -> np.ndarray:
    """The 16 floats IRayTracer's Unproject reads (IRayTracer.cu:45-53) for a
    pinhole camera at ``eye`` looking at ``target``: the inverse of the
    OpenGL projection * view matrix, computed in FP64, stored so that
    m[4*j + i] is row i, column j (the screen point (sx, sy, 1, 1) maps to
    c_k = m[k] sx + m[4+k] sy + m[8+k] + m[12+k]); screen y = -1 is the
    bottom row.  Rounded once to float32."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, np.asarray(up, np.float64))
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = s, u, -f
    view[:3, 3] = -view[:3, :3] @ eye
    t = 1.0 / math.tan(math.radians(fov_deg) * 0.5)
    proj = np.zeros((4, 4))
    proj[0, 0] = t / (width / height)
    proj[1, 1] = t
    proj[2, 2] = (far + near) / (near - far)
    proj[2, 3] = 2.0 * far * near / (near - far)
    proj[3, 2] = -1.0
    inv = np.linalg.inv(proj @ view)
    return np.ascontiguousarray(inv.T.reshape(16), dtype=np.float32)


# the configurations' cameras: (eye, target, fov in degrees)
CAMERA_C1 = ((0.0, 0.0, 3.5), (0.0, 0.0, 0.0), 45.0)
CAMERA_C2 = ((0.5, 0.8, 0.0), (0.5, 0.0, 0.55), 32.0)
CAMERA_C5 = ((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 45.0)


def camera(config_camera, width, height):
    """(inverse view-projection [16], camera position [3]) of a configuration's camera"""
    eye, target, fov = config_camera
    return camera_inverse(eye, target, fov_deg=fov, width=width, height=height), np.asarray(eye, np.float32)


def cosine_dirs(normal: np.ndarray, u1: np.ndarray, u2: np.ndarray) -> np.ndarray:
    """Cosine-weighted directions about unit normals (Malley's method)."""
    r = np.sqrt(u1)
    phi = 2.0 * math.pi * u2
    x = r * np.cos(phi)
    y = r * np.sin(phi)
    z = np.sqrt(np.maximum(0.0, 1.0 - u1))
    n = normal
    a = np.where(np.abs(n[:, 0:1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(a, n)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(n, t)
    d = x[:, None] * t + y[:, None] * b + z[:, None] * n
    return d / np.linalg.norm(d, axis=1, keepdims=True)


# --------------------------------------------------------------------------
# C1: Cornell room (12 axis-aligned quads) + icosphere (5,120 tris)
# --------------------------------------------------------------------------
def cornell_icosphere(subdiv: int = 4):
    """24 room triangles (5 walls, ceiling light quad, inner box of 6 faces)
    followed by an icosphere of radius 0.3 (20*4**subdiv triangles)."""
    q = []
    L, H = -1.0, 1.0
    q += quad((L, L, L), (H, L, L), (H, L, H), (L, L, H))      # floor
    q += quad((L, H, L), (L, H, H), (H, H, H), (H, H, L))      # ceiling
    q += quad((L, L, L), (L, H, L), (H, H, L), (H, L, L))      # back
    q += quad((L, L, L), (L, L, H), (L, H, H), (L, H, L))      # left
    q += quad((H, L, L), (H, H, L), (H, H, H), (H, L, H))      # right
    ly = 0.98
    q += quad((-0.25, ly, -0.25), (0.25, ly, -0.25), (0.25, ly, 0.25), (-0.25, ly, 0.25))  # light
    x0, x1, y0, y1, z0, z1 = -0.7, -0.15, -0.999, -0.1, -0.65, -0.1
    q += quad((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1))   # box bottom
    q += quad((x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (x1, y1, z0))   # box top
    q += quad((x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0))   # box back
    q += quad((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1))   # box front
    q += quad((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0))   # box left
    q += quad((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1))   # box right
    room = np.array(q, dtype=np.float64)                                # (24, 3, 3)
    V, F = icosphere(subdiv, 0.3, (0.4, -0.69, 0.25))
    sph = V[F]                                                          # (5120, 3, 3)
    allt = np.concatenate([room, sph], axis=0)
    return tris_from(allt[:, 0], allt[:, 1], allt[:, 2])


def config1(width=256, height=256, seed=1, subdiv=4):
    """C1: scene + 256x256 jittered pinhole rays from (0, 0, 3.5) looking at -z."""
    p0, p1, p2 = cornell_icosphere(subdiv)
    org, dirs = pinhole_rays(width, height, (0.0, 0.0, 3.5), (0.0, 0.0, 0.0), fov_deg=45.0, seed=seed)
    return (p0, p1, p2), (org, dirs)


# --------------------------------------------------------------------------
# C2: heightfield terrain (value noise)
# --------------------------------------------------------------------------
def _lattice_values(seed, octave, n):
    ids = np.arange(n * n, dtype=np.uint64) + (np.uint64(octave) << np.uint64(32))
    return (Pcg32(seed, P_NOISE, ids).uniform() * 2.0 - 1.0).reshape(n, n)


def value_noise_heights(res: int, seed: int = 2, octaves: int = 6, base_freq: float = 4.0,
                        lacunarity: float = 2.0, gain: float = 0.5, amplitude: float = 0.1):
    """heights[res, res] on the unit square (x = i/(res-1), z = j/(res-1))."""
    x = np.linspace(0.0, 1.0, res)
    h = np.zeros((res, res))
    freq, amp = base_freq, amplitude
    for o in range(octaves):
        n = int(math.ceil(freq)) + 2
        lat = _lattice_values(seed, o, n)
        gx = x * freq
        i0 = np.floor(gx).astype(np.int64)
        f = gx - i0
        s = f * f * (3.0 - 2.0 * f)
        a = lat[i0][:, i0]
        b = lat[i0 + 1][:, i0]
        c = lat[i0][:, i0 + 1]
        d = lat[i0 + 1][:, i0 + 1]
        sx = s[:, None]
        sz = s[None, :]
        h += amp * ((a * (1 - sx) + b * sx) * (1 - sz) + (c * (1 - sx) + d * sx) * sz)
        freq *= lacunarity
        amp *= gain
    return h


def heightfield(res: int = 1024, seed: int = 2):
    """(res-1)^2 * 2 triangles of a res x res vertex grid."""
    h = value_noise_heights(res, seed)
    x = np.linspace(0.0, 1.0, res)
    X, Z = np.meshgrid(x, x, indexing="ij")
    V = np.stack([X, h, Z], axis=-1)
    a = V[:-1, :-1].reshape(-1, 3)
    b = V[1:, :-1].reshape(-1, 3)
    c = V[1:, 1:].reshape(-1, 3)
    d = V[:-1, 1:].reshape(-1, 3)
    v0 = np.concatenate([a, a])
    v1 = np.concatenate([b, c])
    v2 = np.concatenate([c, d])
    # interleave the two triangles of each cell
    order = np.arange(v0.shape[0]).reshape(2, -1).T.reshape(-1)
    return tris_from(v0[order], v1[order], v2[order])


def config2_primary(width=1920, height=1080, res=1024, seed=2):
    """C2 scene and its primary rays (the bounce rays depend on the hits)."""
    tris = heightfield(res, seed)
    org, dirs = pinhole_rays(width, height, (0.5, 0.8, 0.0), (0.5, 0.0, 0.55),
                             fov_deg=32.0, seed=seed)
    return tris, (org, dirs)


# --------------------------------------------------------------------------
# C3: clustered triangle soup + area light + photons
# --------------------------------------------------------------------------
def cluster_soup(n_tris: int = 1_000_000, n_clusters: int = 64, sigma: float = 0.05,
                 seed: int = 3, with_light: bool = True):
    """n_tris random triangles around Gaussian cluster centres, then the two
    light triangles (ids n_tris, n_tris+1): a 0.2 x 0.2 square at y = 1.05."""
    cc = Pcg32(seed, P_CLUSTER, np.arange(n_clusters, dtype=np.uint64))
    centers = np.stack([0.1 + 0.8 * cc.uniform() for _ in range(3)], axis=1)
    r = Pcg32(seed, P_TRI, np.arange(n_tris, dtype=np.uint64))
    cid = np.minimum((r.uniform() * n_clusters).astype(np.int64), n_clusters - 1)
    cen = centers[cid] + sigma * np.stack([r.normal(), r.normal(), r.normal()], axis=1)
    edge = np.exp(math.log(1e-3) + (math.log(2e-2) - math.log(1e-3)) * r.uniform())
    # random unit normal, random in-plane rotation, equilateral triangle
    zc = 2.0 * r.uniform() - 1.0
    ph = 2.0 * math.pi * r.uniform()
    rs = np.sqrt(np.maximum(0.0, 1.0 - zc * zc))
    nrm = np.stack([rs * np.cos(ph), rs * np.sin(ph), zc], axis=1)
    a = np.where(np.abs(nrm[:, 0:1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(a, nrm)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(nrm, t)
    th0 = 2.0 * math.pi * r.uniform()
    rad = edge / math.sqrt(3.0)
    vs = []
    for k in range(3):
        ang = th0 + 2.0 * math.pi * k / 3.0
        vs.append(cen + rad[:, None] * (np.cos(ang)[:, None] * t + np.sin(ang)[:, None] * b))
    v0, v1, v2 = vs
    if with_light:
        s, y = 0.1, 1.05
        lq = np.array(quad((0.5 - s, y, 0.5 - s), (0.5 + s, y, 0.5 - s), (0.5 + s, y, 0.5 + s),
                           (0.5 - s, y, 0.5 + s)), dtype=np.float64)
        v0 = np.concatenate([v0, lq[:, 0]])
        v1 = np.concatenate([v1, lq[:, 1]])
        v2 = np.concatenate([v2, lq[:, 2]])
    return tris_from(v0, v1, v2)


def photon_emission(n_photons: int = 4_000_000, seed: int = 3):
    """Photon origins uniform on the light square (1e-4 below it), directions
    cosine-weighted about -y."""
    r = Pcg32(seed, P_PHOTON, np.arange(n_photons, dtype=np.uint64))
    ox = 0.4 + 0.2 * r.uniform()
    oz = 0.4 + 0.2 * r.uniform()
    org = np.stack([ox, np.full(n_photons, 1.05 - 1e-4), oz], axis=1)
    nrm = np.repeat(np.array([[0.0, -1.0, 0.0]]), n_photons, axis=0)
    d = cosine_dirs(nrm, r.uniform(), r.uniform())
    return _pack(org, 1.0), _pack(d, 0.0)


def bounce_rays(tris, org, dirs, hits, seed, bounce, offset=1e-4):
    """Diffuse continuation rays for every hit ray (ids >= 0): origin at the
    hit point pushed ``offset`` along the geometric normal towards the
    incoming side, direction cosine-weighted about that normal.  Returns
    (org, dir, src_index) for the rays that continue."""
    p0, p1, p2 = tris
    ids = hits["id"] if isinstance(hits, np.ndarray) and hits.dtype.names else hits[:, 0]
    alive = np.nonzero(ids >= 0)[0]
    tri = ids[alive].astype(np.int64)
    a = p0[tri, :3].astype(np.float64)
    e1 = p1[tri, :3].astype(np.float64) - a
    e2 = p2[tri, :3].astype(np.float64) - a
    n = np.cross(e1, e2)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    d = dirs[alive, :3].astype(np.float64)
    flip = np.sum(n * d, axis=1) > 0.0
    n[flip] = -n[flip]
    t = (hits["t"] if hits.dtype.names else hits[:, 1])[alive].astype(np.float64)
    pos = org[alive, :3].astype(np.float64) + t[:, None] * d + offset * n
    r = Pcg32(seed + 1000 * bounce, P_BOUNCE, alive.astype(np.uint64))
    nd = cosine_dirs(n, r.uniform(), r.uniform())
    return _pack(pos, 1.0), _pack(nd, 0.0), alive


# --------------------------------------------------------------------------
# small random scenes for parity tests
# --------------------------------------------------------------------------
def random_soup(n: int, seed: int, extent: float = 1.0, size: float = 0.1):
    r = Pcg32(seed, P_TRI, np.arange(n, dtype=np.uint64))
    c = np.stack([r.uniform(), r.uniform(), r.uniform()], axis=1) * extent
    vs = [c + size * (np.stack([r.uniform(), r.uniform(), r.uniform()], axis=1) - 0.5)
          for _ in range(3)]
    return tris_from(*vs)


def random_rays(n: int, seed: int, lo=-0.5, hi=1.5, target_lo=0.0, target_hi=1.0):
    r = Pcg32(seed, P_RAYS, np.arange(n, dtype=np.uint64))
    o = lo + (hi - lo) * np.stack([r.uniform(), r.uniform(), r.uniform()], axis=1)
    t = target_lo + (target_hi - target_lo) * np.stack([r.uniform(), r.uniform(), r.uniform()], axis=1)
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return _pack(o, 1.0), _pack(d, 0.0)


# --------------------------------------------------------------------------
# C5: perturbed sphere for the differentiable intersection
# --------------------------------------------------------------------------
def perturbed_sphere(subdiv: int = 7, sigma: float = 0.01, seed: int = 5):
    """tetrahedron sphere (radius 1, 4 * 4**subdiv triangles; 7 gives 65,536)
    as (vertices[V,3] float32 unperturbed, perturbed by N(0, sigma^2), faces[F,3] int32)."""
    V, F = tetra_sphere(subdiv)
    r = Pcg32(seed, P_PERTURB, np.arange(V.shape[0], dtype=np.uint64))
    noise = np.stack([r.normal(), r.normal(), r.normal()], axis=1) * sigma
    return _f32(V), _f32(V + noise), F.astype(np.int32)


def config5(width=512, height=512, seed=5, subdiv=7):
    """C5: sphere vertices (unperturbed, perturbed), faces and 512x512 rays
    from (0, 0, 3) towards the origin."""
    v_ref, v_pert, faces = perturbed_sphere(subdiv, seed=seed)
    org, dirs = pinhole_rays(width, height, (0.0, 0.0, 3.0), (0.0, 0.0, 0.0), fov_deg=45.0, seed=seed)
    return v_ref, v_pert, faces, (org, dirs)


# --------------------------------------------------------------------------
# C4: batched rebuild of many animated meshes
# --------------------------------------------------------------------------
def cube_sphere(k: int):
    """Unit sphere from a cube whose six faces are k x k quad grids pushed
    onto the sphere: 12 k^2 triangles, no vertex in more than 6 of them (no
    pole fans).  Returns (unit vertices float64 [V, 3], faces int32 [F, 3])."""
    k = max(1, int(k))
    g = np.arange(k + 1)
    u, v = np.meshgrid(g, g, indexing="ij")
    u, v = u.reshape(-1), v.reshape(-1)
    pts, quads = [], []
    base = 0
    q = (u < k) & (v < k)
    a = np.nonzero(q)[0]
    qi = np.stack([a, a + (k + 1), a + (k + 1) + 1, a + 1], axis=1)  # (u,v) (u+1,v) (u+1,v+1) (u,v+1)
    for axis in range(3):
        for side in (0, k):
            c = np.zeros((u.shape[0], 3), dtype=np.int64)
            c[:, axis] = side
            c[:, (axis + 1) % 3] = u
            c[:, (axis + 2) % 3] = v
            pts.append(c)
            quads.append(qi + base)
            base += u.shape[0]
    pts = np.concatenate(pts)
    quads = np.concatenate(quads)
    uniq, inv = np.unique(pts, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    quads = inv[quads]
    F = np.concatenate([quads[:, [0, 1, 2]], quads[:, [0, 2, 3]]])
    V = 2.0 * uniq.astype(np.float64) / k - 1.0
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    # outward winding
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flip = np.sum(n * (V[F[:, 0]] + V[F[:, 1]] + V[F[:, 2]]), axis=1) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V, F.astype(np.int32)


# radial displacement r = 1 + A sin(8 theta + 0.3 f) cos(6 phi + 0.2 f)
WAVE = dict(amplitude=0.05, k_theta=8.0, w_theta=0.3, k_phi=6.0, w_phi=0.2)


def displace_sphere64(unit: np.ndarray, frame: float) -> np.ndarray:
    """FP64 restatement of the per-frame sphere deformer: unit vertices
    (float32 [V, 4] or [V, 3]) scaled by the radial wave at frame f."""
    u = np.asarray(unit)[:, :3].astype(np.float64)
    theta = np.arccos(np.clip(u[:, 2], -1.0, 1.0))
    phi = np.arctan2(u[:, 1], u[:, 0])
    w = WAVE
    r = 1.0 + w["amplitude"] * np.sin(w["k_theta"] * theta + w["w_theta"] * frame) * \
        np.cos(w["k_phi"] * phi + w["w_phi"] * frame)
    return u * r[:, None]


def c4_sizes(num_meshes: int = 256, seed: int = 4, lo: float = 1e3, hi: float = 2e5):
    """triangle counts n_k = round(exp U(ln lo, ln hi)) and the cube-sphere
    grid k with 12 k^2 closest to n_k"""
    r = Pcg32(seed, P_MESH, np.arange(num_meshes, dtype=np.uint64))
    n = np.rint(np.exp(math.log(lo) + r.uniform() * (math.log(hi) - math.log(lo))))
    k = np.maximum(1, np.rint(np.sqrt(n / 12.0))).astype(np.int64)
    return n.astype(np.int64), k


def config4(num_meshes: int = 256, rays_per_mesh: int = 65536, seed: int = 4, lo: float = 1e3,
            hi: float = 2e5):
    """C4: ``num_meshes`` cube spheres of log-uniform size with their unit
    vertices concatenated, and ``rays_per_mesh`` rays per mesh from radius 2
    towards a uniform point of the unit ball.  Returns a dict with
      unit      float32 [V, 4] unit vertices (w = 1), meshes concatenated
      faces     int32 [F, 3] global vertex indices, meshes concatenated
      vert_off  int64 [M+1] vertex offsets, tri_off int32 [M+1] triangle offsets
      org, dirs float32 [M * rays_per_mesh, 4]; ray_mesh int32 mesh of each ray"""
    _, ks = c4_sizes(num_meshes, seed, lo, hi)
    cache = {}
    units, faces, voff, toff = [], [], [0], [0]
    for k in ks:
        if k not in cache:
            cache[k] = cube_sphere(int(k))
        V, F = cache[k]
        units.append(V)
        faces.append(F + voff[-1])
        voff.append(voff[-1] + V.shape[0])
        toff.append(toff[-1] + F.shape[0])
    unit = _pack(np.concatenate(units), 1.0)
    n = num_meshes * rays_per_mesh
    r = Pcg32(seed, P_RAYS, np.arange(n, dtype=np.uint64))
    z = 1.0 - 2.0 * r.uniform()
    ph = 2.0 * math.pi * r.uniform()
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    o = 2.0 * np.stack([s * np.cos(ph), s * np.sin(ph), z], axis=1)
    rad = np.cbrt(r.uniform())
    z2 = 1.0 - 2.0 * r.uniform()
    ph2 = 2.0 * math.pi * r.uniform()
    s2 = np.sqrt(np.maximum(0.0, 1.0 - z2 * z2))
    t = rad[:, None] * np.stack([s2 * np.cos(ph2), s2 * np.sin(ph2), z2], axis=1)
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ray_mesh = np.repeat(np.arange(num_meshes, dtype=np.int32), rays_per_mesh)
    return dict(unit=unit, faces=np.concatenate(faces).astype(np.int32),
                vert_off=np.asarray(voff, dtype=np.int64), tri_off=np.asarray(toff, dtype=np.int32),
                org=_pack(o, 1.0), dirs=_pack(d, 0.0), ray_mesh=ray_mesh, k=ks)


# --------------------------------------------------------------------------
# oscillating grid (OscCUDAMeshNode)
# --------------------------------------------------------------------------
def osc_grid(width: float, depth: float, height: float, detail: int):
    """OscCUDAMeshNode::Init (OscCUDAMeshNode.cu:26-68): (detail+1)^2 float4
    vertices, vertex (i, j) at i + j*(detail+1) = (i w/detail - w/2, height,
    j d/detail - d/2, 1) in FP32, and two triangles per grid quad."""
    d = detail + 1
    ux = np.float32(width) / np.float32(detail)
    uy = np.float32(depth) / np.float32(detail)
    j, i = np.meshgrid(np.arange(d), np.arange(d), indexing="ij")   # index = i + j*d
    i, j = i.reshape(-1), j.reshape(-1)
    V = np.empty((d * d, 4), np.float32)
    V[:, 0] = i.astype(np.float32) * ux - np.float32(width) * np.float32(0.5)
    V[:, 1] = np.float32(height)
    V[:, 2] = j.astype(np.float32) * uy - np.float32(depth) * np.float32(0.5)
    V[:, 3] = 1.0
    m, n = np.meshgrid(np.arange(detail), np.arange(detail), indexing="ij")
    m, n = m.reshape(-1), n.reshape(-1)
    q = np.stack([m + n * d, m + (n + 1) * d, m + 1 + n * d,
                  m + 1 + n * d, m + (n + 1) * d, m + 1 + (n + 1) * d], axis=1)
    return V, q.reshape(-1).astype(np.uint32)


def osc_surface64(V: np.ndarray, detail: int, height: float, time: float, amplitude: float = 0.2,
                  frequency: float = 0.3):
    """FP64 restatement of UpdateSurface (OscCUDAMeshNode.cu:72-93):
    returns (vertices with the new y, normals)"""
    d = detail + 1
    idx = np.arange(d * d)
    x, y = (idx % d).astype(np.float64), (idx // d).astype(np.float64)
    ax, ay = frequency * x + time, frequency * y + time
    out = np.asarray(V, np.float64).copy()
    out[:, 1] = height + amplitude * np.cos(ax) + amplitude * np.sin(ay)
    n = np.zeros((d * d, 4))
    n[:, 0] = -amplitude * np.sin(ax) * frequency
    n[:, 2] = amplitude * np.cos(ay) * frequency
    n[:, 1] = 1.0 - n[:, 0] ** 2 - n[:, 2] ** 2
    return out, n
