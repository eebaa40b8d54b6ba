"""Small analytic scenes for the closed-form tests: triangle soups built in
numpy and wrapped in the binding's Scene (materials, lights, camera).

Everything here is geometry set-up; no estimator arithmetic lives here.
"""
import math

import numpy as np

import pppm_amd as P


def icosphere(freq, radius=1.0, center=(0.0, 0.0, 0.0)):
    """Geodesic sphere: an icosahedron with vertices at the poles (0, +-1, 0),
    each face split into freq^2 triangles, vertices pushed onto the sphere.
    Triangles are wound with their geometric normal pointing outward."""
    top, bot = np.array([0.0, 1.0, 0.0]), np.array([0.0, -1.0, 0.0])
    lat = math.atan(0.5)
    up = [np.array([math.cos(lat) * math.cos(k * 2 * math.pi / 5), math.sin(lat),
                    math.cos(lat) * math.sin(k * 2 * math.pi / 5)]) for k in range(5)]
    dn = [np.array([math.cos(lat) * math.cos((k + 0.5) * 2 * math.pi / 5), -math.sin(lat),
                    math.cos(lat) * math.sin((k + 0.5) * 2 * math.pi / 5)]) for k in range(5)]
    faces = []
    for k in range(5):
        k1 = (k + 1) % 5
        faces.append((top, up[k1], up[k]))
        faces.append((up[k], up[k1], dn[k]))
        faces.append((dn[k], up[k1], dn[k1]))
        faces.append((bot, dn[k], dn[k1]))
    tris = []
    f = freq
    for a, b, c in faces:
        def pt(i, j):
            p = a + (b - a) * (i / f) + (c - a) * (j / f)
            return p / np.linalg.norm(p)
        for i in range(f):
            for j in range(f - i):
                tris.append((pt(i, j), pt(i + 1, j), pt(i, j + 1)))
                if i + j + 1 < f:
                    tris.append((pt(i + 1, j), pt(i + 1, j + 1), pt(i, j + 1)))
    t = np.array(tris, np.float64) * radius + np.asarray(center, np.float64)
    # orient outward
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    out = np.einsum("ij,ij->i", n, t.mean(axis=1) - np.asarray(center))
    flip = out < 0
    t[flip] = t[flip][:, [0, 2, 1]]
    return t.reshape(-1, 9).astype(np.float32)


def quad(corner, eu, ev):
    """Two triangles of the parallelogram corner + s eu + t ev, normal along eu x ev."""
    c, u, v = (np.asarray(x, np.float64) for x in (corner, eu, ev))
    return np.array([np.concatenate([c, c + u, c + u + v]), np.concatenate([c, c + u + v, c + v])], np.float32)


def tri_area(tri):
    t = np.asarray(tri, np.float64).reshape(-1, 3, 3)
    return 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)


def bsdf(kind, rho=(0.5, 0.5, 0.5), emission=(0.0, 0.0, 0.0), ior=1.5, sigma_a=(0, 0, 0), sigmap_s=(0, 0, 0),
         eta=1.0):
    b = P.Bsdf()
    b.type = kind
    b.reflectance[:] = rho
    b.emission[:] = emission
    b.ior = ior
    b.sigma_a[:] = sigma_a
    b.sigmap_s[:] = sigmap_s
    b.eta = eta
    return b


def light(corner, eu, ev, emission):
    L = P.Light()
    L.corner[:] = corner
    L.edge_u[:] = eu
    L.edge_v[:] = ev
    L.emission[:] = emission
    return L


def camera(center, direction, up, fov_deg, width, height):
    c = P.Camera()
    c.center[:] = center
    c.direction[:] = direction
    c.up[:] = up
    c.fov_y_deg = fov_deg
    c.width = width
    c.height = height
    return c


def scene(parts, bsdfs, lights, cam):
    """parts: list of (triangles [n][9], bsdf index)."""
    tri = np.concatenate([p for p, _ in parts]).astype(np.float32)
    ids = np.concatenate([np.full(len(p), m, np.int32) for p, m in parts])
    pts = tri.reshape(-1, 3)
    return P.Scene(tri, ids, list(bsdfs), list(lights), cam, tuple(pts.min(0).astype(float)),
                   tuple(pts.max(0).astype(float)))


def emitter_quad(corner, eu, ev, le, index):
    """A parallelogram light and its emitter triangles (emitting along eu x ev)."""
    return (quad(corner, eu, ev), index), light(corner, eu, ev, le)


def sphere_furnace(rho=(0.3, 0.5, 0.7), freq=12, radius=2.0, le=(4.0, 4.0, 4.0), light_side=0.1, res=32):
    """A closed geodesic sphere of uniform albedo lit by a small downward-facing
    light just under its top pole; the camera at the centre looks sideways.
    Every point of a sphere's inside receives the same irradiance from a
    cosine emitter on its surface and from any uniform radiosity, so the
    exitant radiance is uniform: L = rho Phi / (pi A (1 - rho))."""
    sph = icosphere(freq, radius)
    y0 = radius - 0.01
    h = light_side / 2
    lq, L = emitter_quad((-h, y0, -h), (2 * h, 0, 0), (0, 0, 2 * h), le, 1)
    b = [bsdf(P.LAMBERTIAN, rho), bsdf(P.EMITTER, (0, 0, 0), emission=le)]
    cam = camera((0, 0, 0), (1, 0, 0), (0, 1, 0), 50.0, res, res)
    return scene([(sph, 0), lq], b, [L], cam), sph
