"""Deterministic synthetic scenes for the traversal workloads.

Everything here is input synthesis (meshes, cameras, instance transforms) in
numpy, seeded with PCG32 (O'Neill 2014, ``pcg32_srandom`` / ``pcg32_random``)
so that the GPU path and the CPU checker see byte-identical inputs.

Workloads (sizes as used by the benchmark and the tests):

* ``soup``      -- 12,582 random triangles: v0 uniform in [-5,5]^3, v1 and v2
  are v0 plus a uniform offset in [0,1]^3; 640x640 pinhole camera at
  (0,0,-18) looking down +z with a 90 degree vertical field of view.
* ``icosphere`` -- frequency-n geodesic icosphere (20 n^2 triangles) of
  radius R, each vertex displaced radially, r = R (1 + amp * fbm(p)), by
  4-octave seeded value noise.  n = 256 gives 1,310,720 triangles, n = 100
  gives 200,000, n = 64 gives 81,920.

The camera model is the reference's pinhole (eye-space viewport corners at
the near plane plus an eye-to-world matrix, bvhtracer/src/camera/camera.rs:
219-252, 809-834, 994-1010).
"""
import math

import numpy as np

PCG_MULT = np.uint64(6364136223846793005)
F32_MAX = np.float32(np.finfo(np.float32).max)


# ---------------------------------------------------------------- PCG32 ---

class Pcg32:
    """Vectorised PCG32 (XSH-RR 64/32): one independent stream per lane.

    ``Pcg32(seed, stream_ids)`` seeds lane k like ``pcg32_srandom(seed,
    stream_ids[k])``; ``next_u32()`` advances every lane once.
    """

    def __init__(self, seed, stream_ids):
        sid = np.asarray(stream_ids, dtype=np.uint64)
        with np.errstate(over="ignore"):
            self.inc = (sid << np.uint64(1)) | np.uint64(1)
            self.state = np.zeros_like(sid)
            self._step()
            self.state = self.state + np.uint64(seed)
            self._step()

    def _step(self):
        with np.errstate(over="ignore"):
            self.state = self.state * PCG_MULT + self.inc

    def next_u32(self):
        old = self.state
        self._step()
        with np.errstate(over="ignore"):
            xorshifted = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)) & np.uint64(0xFFFFFFFF)
            rot = (old >> np.uint64(59)).astype(np.uint64)
            res = (xorshifted >> rot) | (xorshifted << ((np.uint64(32) - rot) & np.uint64(31)))
        return (res & np.uint64(0xFFFFFFFF)).astype(np.uint32)

    def next_f32(self):
        """Uniform in [0, 1): (x >> 8) * 2^-24, exact in binary32."""
        return (self.next_u32() >> np.uint32(8)).astype(np.float32) * np.float32(2.0 ** -24)


def pcg32_scalar_sequence(seed, stream, count):
    """Reference scalar PCG32 (used by tests to pin the vectorised one)."""
    mask = (1 << 64) - 1
    mult = 6364136223846793005
    inc = ((stream << 1) | 1) & mask
    state = 0
    state = (state * mult + inc) & mask
    state = (state + seed) & mask
    state = (state * mult + inc) & mask
    out = []
    for _ in range(count):
        old = state
        state = (old * mult + inc) & mask
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        out.append(((xs >> rot) | (xs << ((32 - rot) & 31))) & 0xFFFFFFFF)
    return out


# ------------------------------------------------------------ triangle soup

def triangle_soup(n=12582, seed=0x5EED0001):
    """n triangles as an (n, 9) float32 array (v0, v1, v2 row-major)."""
    rng = Pcg32(seed, np.arange(n, dtype=np.uint64))
    u = np.stack([rng.next_f32() for _ in range(9)], axis=1)
    ten, five = np.float32(10.0), np.float32(5.0)
    v0 = ten * u[:, 0:3] - five
    v1 = v0 + u[:, 3:6]
    v2 = v0 + u[:, 6:9]
    return np.ascontiguousarray(np.concatenate([v0, v1, v2], axis=1).astype(np.float32))


# ------------------------------------------------------------- icosphere

def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([
        [-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
        [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
        [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1],
    ], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([
        [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
        [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
        [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
        [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
    ], dtype=np.int64)
    return v, f


def _lattice_hash(ix, iy, iz, seed):
    """Integer lattice hash -> float64 in [-1, 1] (seeded value noise)."""
    with np.errstate(over="ignore"):
        h = (ix.astype(np.int64).astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
             ^ iy.astype(np.int64).astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)
             ^ iz.astype(np.int64).astype(np.uint64) * np.uint64(0x165667B19E3779F9)
             ^ np.uint64(seed) * np.uint64(0x27D4EB2F165667C5))
        h ^= h >> np.uint64(29)
        h *= np.uint64(0xBF58476D1CE4E5B9)
        h ^= h >> np.uint64(32)
    return (h >> np.uint64(11)).astype(np.float64) * (2.0 / float(1 << 53)) - 1.0


def value_noise(p, seed):
    """3-D value noise in [-1, 1], trilinear with smoothstep fade."""
    fl = np.floor(p)
    i = fl.astype(np.int64)
    f = p - fl
    w = f * f * (3.0 - 2.0 * f)
    acc = 0.0
    for dx in (0, 1):
        wx = w[:, 0] if dx else 1.0 - w[:, 0]
        for dy in (0, 1):
            wy = w[:, 1] if dy else 1.0 - w[:, 1]
            for dz in (0, 1):
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                acc = acc + wx * wy * wz * _lattice_hash(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
    return acc


def fbm(p, seed, octaves=4, base_freq=4.0, lacunarity=2.0, gain=0.5):
    total = np.zeros(p.shape[0])
    norm = 0.0
    amp, freq = 1.0, base_freq
    for o in range(octaves):
        total += amp * value_noise(p * freq, seed + o)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def icosphere(freq, radius=32.0, amp=0.15, seed=0x1C05, octaves=4):
    """Frequency-``freq`` geodesic icosphere as an (20 freq^2, 9) float32 array.

    Grid points are exact integer barycentric combinations of icosahedron
    vertices, summed in ascending vertex-id order so that points on shared
    edges are computed identically from both faces (watertight).
    """
    V, F = _icosahedron()
    n = int(freq)
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    keep = (ii + jj) <= n
    I, J = ii[keep], jj[keep]
    K = n - I - J
    idx = -np.ones((n + 1, n + 1), dtype=np.int64)
    idx[I, J] = np.arange(I.size)
    up_i, up_j = np.nonzero((ii + jj) < n)
    dn_i, dn_j = np.nonzero((ii + jj) < n - 1)
    tris = []
    for face in F:
        order = np.argsort(face)
        ids = face[order]
        wts = np.stack([K, I, J], axis=1).astype(np.float64)[:, order]
        P = (wts[:, 0:1] * V[ids[0]] + wts[:, 1:2] * V[ids[1]]) + wts[:, 2:3] * V[ids[2]]
        P = P / np.sqrt((P * P).sum(axis=1, keepdims=True))
        r = radius * (1.0 + amp * fbm(P, seed, octaves))
        Q = (P * r[:, None]).astype(np.float32)
        a = Q[idx[up_i, up_j]]
        b = Q[idx[up_i + 1, up_j]]
        c = Q[idx[up_i, up_j + 1]]
        up = np.concatenate([a, b, c], axis=1)
        a = Q[idx[dn_i + 1, dn_j]]
        b = Q[idx[dn_i + 1, dn_j + 1]]
        c = Q[idx[dn_i, dn_j + 1]]
        dn = np.concatenate([a, b, c], axis=1)
        tris.append(up)
        tris.append(dn)
    out = np.ascontiguousarray(np.concatenate(tris, axis=0).astype(np.float32))
    assert out.shape[0] == 20 * n * n
    return out


def icosphere_frequency(tri_count):
    """Frequency n with 20 n^2 closest to ``tri_count`` (n >= 1)."""
    return max(1, int(round(math.sqrt(tri_count / 20.0))))


# ---------------------------------------------------------------- camera

def camera(eye, forward, right, up, fovy_deg=90.0, aspect=1.0, near=2.0):
    """Pinhole camera record: eye-space viewport corners at the near plane
    (Frustum::from(SymmetricFovSpec), camera.rs:219-252) and the eye-to-world
    matrix [right | up | -forward | eye] (the inverse of the view matrix of
    CameraAttitude::from_spec, camera.rs:809-834), column-major float32."""
    top = near * math.tan(math.radians(fovy_deg) / 2.0)
    left = -aspect * top
    rgt = aspect * top
    ext_x, ext_y = rgt - left, 2.0 * top
    tl = np.array([left, top, -near], np.float64)
    tr = np.array([left + ext_x, top, -near], np.float64)
    bl = np.array([left, top - ext_y, -near], np.float64)
    m = np.zeros((4, 4), np.float64)  # m[col][row]
    m[0, :3] = right
    m[1, :3] = up
    m[2, :3] = -np.asarray(forward, np.float64)
    m[3, :3] = eye
    m[3, 3] = 1.0
    rec = np.zeros(1, dtype=CAMERA_DTYPE)
    rec["tl"][0] = tl.astype(np.float32)
    rec["tr"][0] = tr.astype(np.float32)
    rec["bl"][0] = bl.astype(np.float32)
    rec["view_inv"][0] = m.reshape(16).astype(np.float32)
    return rec


CAMERA_DTYPE = np.dtype([("tl", "<f4", 3), ("tr", "<f4", 3), ("bl", "<f4", 3), ("view_inv", "<f4", 16)])


def soup_camera():
    return camera(eye=(0.0, 0.0, -18.0), forward=(0.0, 0.0, 1.0), right=(1.0, 0.0, 0.0),
                  up=(0.0, 1.0, 0.0), fovy_deg=90.0, aspect=1.0, near=2.0)


def sphere_camera(radius=32.0, width=1920, height=1080, distance=1.6, fovy_deg=60.0):
    return camera(eye=(0.0, 0.0, -distance * radius), forward=(0.0, 0.0, 1.0),
                  right=(1.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fovy_deg=fovy_deg,
                  aspect=width / height, near=2.0)


# ------------------------------------------------------------ transforms

def random_affines(count, seed, scale_range=(0.5, 1.5), trans_range=20.0):
    """Seeded affine transforms: uniform random rotation (Shoemake's unit
    quaternion), isotropic scale and translation.  Returns (M, Minv) as
    (count, 16) float32 column-major arrays, composed like Transform3::new
    (scale applied to the rotation columns, translation in column 3;
    bvhtracer/src/transform.rs:24-43); Minv is the float64 inverse rounded."""
    rng = Pcg32(seed, np.arange(count, dtype=np.uint64))
    u1, u2, u3 = [rng.next_f32().astype(np.float64) for _ in range(3)]
    s = rng.next_f32().astype(np.float64)
    tx, ty, tz = [rng.next_f32().astype(np.float64) for _ in range(3)]
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    qx = a * np.sin(2 * np.pi * u2)
    qy = a * np.cos(2 * np.pi * u2)
    qz = b * np.sin(2 * np.pi * u3)
    qw = b * np.cos(2 * np.pi * u3)
    scale = scale_range[0] + (scale_range[1] - scale_range[0]) * s
    M = np.zeros((count, 4, 4))  # [k, col, row]
    M[:, 0, 0] = 1 - 2 * (qy * qy + qz * qz)
    M[:, 0, 1] = 2 * (qx * qy + qw * qz)
    M[:, 0, 2] = 2 * (qx * qz - qw * qy)
    M[:, 1, 0] = 2 * (qx * qy - qw * qz)
    M[:, 1, 1] = 1 - 2 * (qx * qx + qz * qz)
    M[:, 1, 2] = 2 * (qy * qz + qw * qx)
    M[:, 2, 0] = 2 * (qx * qz + qw * qy)
    M[:, 2, 1] = 2 * (qy * qz - qw * qx)
    M[:, 2, 2] = 1 - 2 * (qx * qx + qy * qy)
    M[:, :3, :3] *= scale[:, None, None]
    M[:, 3, 0] = (2 * tx - 1) * trans_range
    M[:, 3, 1] = (2 * ty - 1) * trans_range
    M[:, 3, 2] = (2 * tz - 1) * trans_range
    M[:, 3, 3] = 1.0
    M32 = M.astype(np.float32)
    # inverse of the float32 matrix, computed in float64 (row-major algebra on M^T)
    Minv = np.linalg.inv(np.transpose(M32.astype(np.float64), (0, 2, 1)))
    Minv32 = np.transpose(Minv, (0, 2, 1)).astype(np.float32)
    return M32.reshape(count, 16), Minv32.reshape(count, 16)


def deform_wave(tris_orig, frame, amp=0.2, kx=4.0, w=0.1):
    """Per-frame deformation y' = y + amp*sin(kx*x + frame*w) of the original
    vertices (float32 arithmetic), as the per-frame refit workload uses."""
    t = np.asarray(tris_orig, np.float32).reshape(-1, 3, 3)
    x = t[:, :, 0]
    y = t[:, :, 1]
    yn = y + np.float32(amp) * np.sin(np.float32(kx) * x + np.float32(frame * w)).astype(np.float32)
    out = t.copy()
    out[:, :, 1] = yn.astype(np.float32)
    return out.reshape(-1, 9)


# ---------------------------------------------------------- batched scenes

RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmax", "<f4"), ("d", "<f4", 3), ("pad", "<f4")])


def _log_uniform(u, lo, hi):
    return np.exp(np.log(lo) + u.astype(np.float64) * (np.log(hi) - np.log(lo)))


def batch_workload(n_scenes=64, seed=0xBA7C4, tri_range=(1e3, 1e6), ray_range=(4e3, 2e6), max_instances=8,
                   radius=32.0):
    """Uneven multi-scene workload: scene s is a displaced icosphere with a
    log-uniform triangle count in ``tri_range`` (frequency round(sqrt(T/20))),
    1..``max_instances`` seeded affine instances, and a log-uniform number of
    rays in ``ray_range``: half pinhole camera rays looking at the scene from
    outside its box, half rays with origins uniform in the scene box and
    directions uniform on the sphere.  Returns a list of dicts with keys
    mesh, M, Minv, rays."""
    rng = Pcg32(seed, np.arange(n_scenes, dtype=np.uint64))
    u_tri, u_inst, u_ray = rng.next_f32(), rng.next_f32(), rng.next_f32()
    tris = _log_uniform(u_tri, *tri_range)
    n_inst = 1 + np.minimum((u_inst * max_instances).astype(np.int64), max_instances - 1)
    n_rays = np.round(_log_uniform(u_ray, *ray_range)).astype(np.int64)
    out = []
    for s in range(n_scenes):
        freq = icosphere_frequency(tris[s])
        mesh = icosphere(freq, radius=radius, seed=0x1C05 + s)
        M, Minv = random_affines(int(n_inst[s]), seed=seed + 1000 + s, trans_range=2.5 * radius)
        # scene box from the transformed mesh box corners
        lo, hi = mesh.reshape(-1, 3).min(0).astype(np.float64), mesh.reshape(-1, 3).max(0).astype(np.float64)
        corners = np.array([[(hi if c & 1 else lo)[0], (hi if c & 2 else lo)[1], (hi if c & 4 else lo)[2], 1.0]
                            for c in range(8)])
        w = np.concatenate([corners @ M[k].reshape(4, 4).astype(np.float64) for k in range(n_inst[s])])[:, :3]
        blo, bhi = w.min(0), w.max(0)
        center, ext = 0.5 * (blo + bhi), float((bhi - blo).max())
        n = int(n_rays[s])
        n_cam = n // 2
        rays = np.zeros(n, RAY_DTYPE)
        # camera rays: square grid of ceil(sqrt(n_cam)) per side, first n_cam taken
        side = int(np.ceil(np.sqrt(max(n_cam, 1))))
        k = np.arange(n_cam)
        px = (k % side + 0.5) / side * 2.0 - 1.0
        py = (k // side + 0.5) / side * 2.0 - 1.0
        eye = center + np.array([0.0, 0.0, -1.6 * ext])
        d = np.stack([px * 0.6, py * 0.6, np.ones(n_cam)], 1)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rays["o"][:n_cam] = eye.astype(np.float32)
        rays["d"][:n_cam] = d.astype(np.float32)
        # random rays inside the box
        r2 = Pcg32(seed + 7 * s + 3, np.arange(n - n_cam, dtype=np.uint64))
        uo = np.stack([r2.next_f32() for _ in range(3)], 1).astype(np.float64)
        z = 2.0 * r2.next_f32().astype(np.float64) - 1.0
        phi = 2.0 * np.pi * r2.next_f32().astype(np.float64)
        rr = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        rays["o"][n_cam:] = (blo + uo * (bhi - blo)).astype(np.float32)
        rays["d"][n_cam:] = np.stack([rr * np.cos(phi), rr * np.sin(phi), z], 1).astype(np.float32)
        rays["tmax"] = F32_MAX
        out.append({"mesh": mesh, "M": M, "Minv": Minv, "rays": rays})
    return out


# ------------------------------------------------ differentiable rendering

def indexed(tris):
    """(N, 9) triangle soup -> (V float32 (nV, 3), F uint32 (N, 3)), merging
    vertices with identical coordinates."""
    v = np.ascontiguousarray(np.asarray(tris, np.float32).reshape(-1, 3))
    uniq, inv = np.unique(v.view(np.dtype((np.void, 12))).ravel(), return_index=False, return_inverse=True)
    V = np.ascontiguousarray(uniq.view(np.float32).reshape(-1, 3))
    return V, np.ascontiguousarray(inv.reshape(-1, 3).astype(np.uint32))


def radial_perturb(V, sigma=0.05, seed=0xD1FF):
    """Scale each vertex radially by (1 + sigma * g), g ~ N(0, 1) from PCG32
    streams (Box-Muller)."""
    rng = Pcg32(seed, np.arange(V.shape[0], dtype=np.uint64))
    u1 = np.maximum(rng.next_f32().astype(np.float64), 1e-12)
    u2 = rng.next_f32().astype(np.float64)
    g = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return (V.astype(np.float64) * (1.0 + sigma * g)[:, None]).astype(np.float32)


def vertex_normals(V, F):
    """Area-weighted vertex normals (float64 accumulate, float32 result)."""
    v = V.astype(np.float64)
    n = np.cross(v[F[:, 1]] - v[F[:, 0]], v[F[:, 2]] - v[F[:, 0]])
    acc = np.zeros_like(v)
    for k in range(3):
        np.add.at(acc, F[:, k], n)
    return (acc / np.linalg.norm(acc, axis=1, keepdims=True)).astype(np.float32)


def orbit_cameras(count=8, distance=96.0, fovy_deg=40.0, seed=0xCA3):
    """``count`` seeded pinhole cameras on a sphere of radius ``distance``
    looking at the origin (up vector +y, or +z near the poles)."""
    rng = Pcg32(seed, np.arange(count, dtype=np.uint64))
    z = 2.0 * rng.next_f32().astype(np.float64) - 1.0
    phi = 2.0 * np.pi * rng.next_f32().astype(np.float64)
    r = np.sqrt(1.0 - z * z)
    cams = []
    for k in range(count):
        eye = distance * np.array([r[k] * np.cos(phi[k]), z[k], r[k] * np.sin(phi[k])])
        fwd = -eye / np.linalg.norm(eye)
        up0 = np.array([0.0, 1.0, 0.0]) if abs(fwd[1]) < 0.95 else np.array([0.0, 0.0, 1.0])
        right = np.cross(fwd, up0)
        right /= np.linalg.norm(right)
        up = np.cross(right, fwd)
        cams.append(camera(eye, fwd, right, up, fovy_deg=fovy_deg, aspect=1.0, near=2.0))
    return np.concatenate(cams)


DIFF_LIGHT = (0.48, 0.6, -0.64)  # unit model-space light direction


def diff_workload(freq=64, radius=32.0, sigma=0.05, n_cameras=8, width=512, seed=0xD1FF):
    """Differentiable depth/shading workload: the indexed icosphere of
    ``freq`` (20 freq^2 faces; 64 -> 81,920) as the starting mesh, a target
    mesh with seeded radial vertex perturbation ``sigma``, one object with a
    seeded affine transform (rotation, scale in [0.9, 1.1], translation in
    [-2, 2]^3), ``n_cameras`` seeded orbit cameras at width^2 pixels and the
    model-space light direction.  Returns a dict."""
    V, F = indexed(icosphere(freq, radius=radius, amp=0.0))
    Vt = radial_perturb(V, sigma, seed)
    M, Minv = random_affines(1, seed ^ 0x7F, scale_range=(0.9, 1.1), trans_range=2.0)
    cams = orbit_cameras(n_cameras, distance=3.0 * radius, fovy_deg=45.0, seed=seed ^ 0xCA3)
    return {"V": V, "F": F, "V_target": Vt, "M": M[0], "Minv": Minv[0], "cameras": cams, "width": width,
            "light": np.array(DIFF_LIGHT, np.float32)}

# SGD for the differentiable workload (loss = sum over pixels): step size
# and per-vertex step clip (silhouette pixels carry gradients up to ~40x the
# median vertex gradient).  Sweep over 100 steps: lr 1e-3 with a 0.05 clip
# lowers the loss fastest among lr 1e-3..1e-1, clip 0.02..0.1.
DIFF_LR = 1e-3
DIFF_MAX_STEP = 0.05


def animate(tris, r):
    """The refit benchmark's deformation (benches/bench_bvh_refit_rebuild.rs:
    animate): every vertex rotated about z by s = (a (y - 0.2)) 0.2 with
    a = sin(r) 0.5, in binary32 (numpy's float32 sin/cos; input synthesis,
    the same buffer goes to the GPU and to the oracle)."""
    t = np.asarray(tris, np.float32).reshape(-1, 3, 3)
    a = np.float32(np.sin(np.float32(r))) * np.float32(0.5)
    x, y = t[..., 0], t[..., 1]
    s = (a * (y - np.float32(0.2))) * np.float32(0.2)
    c, sn = np.cos(s), np.sin(s)
    out = t.copy()
    out[..., 0] = x * c - y * sn
    out[..., 1] = x * sn + y * c
    return np.ascontiguousarray(out.reshape(-1, 9))
