"""Geometries: synthetic elevation models (src/Geometries/*.jl).

make_synthetic / AABB / extents / dilate / DataElevation / rotate mirror the
reference, extract_dem the DEM preparation of the glacier driver; the seeded
eight-bump glacier generator builds the benchmark geometry (a parabolic ice
surface over a bed of Gaussian bumps).
"""
import math

import numpy as np


class AABB:
    """Axis-aligned bounding box (elevation_data.jl:1-15)."""

    def __init__(self, xmin, xmax, ymin, ymax, zmin, zmax):
        self.xmin, self.xmax, self.ymin, self.ymax, self.zmin, self.zmax = map(float, (xmin, xmax, ymin, ymax, zmin, zmax))

    @classmethod
    def from_coords(cls, xs, ys, zs):
        return cls(np.min(xs), np.max(xs), np.min(ys), np.max(ys), np.nanmin(zs), np.nanmax(zs))


def extents(box):
    """geometry_helpers.jl:6"""
    return box.xmax - box.xmin, box.ymax - box.ymin, box.zmax - box.zmin


def center(box):
    """geometry_helpers.jl:9-12"""
    return 0.5 * (box.xmin + box.xmax), 0.5 * (box.ymin + box.ymax), 0.5 * (box.zmin + box.zmax)


def dilate(box, fractions):
    """geometry_helpers.jl:15-18"""
    fr = fractions if np.ndim(fractions) else (fractions,) * 3
    dx, dy, dz = (e * f for e, f in zip(extents(box), fr))
    return AABB(box.xmin - dx, box.xmax + dx, box.ymin - dy, box.ymax + dy, box.zmin - dz, box.zmax + dz)


def union(a, b):
    """geometry_helpers.jl:24-28"""
    return AABB(min(a.xmin, b.xmin), max(a.xmax, b.xmax), min(a.ymin, b.ymin), max(a.ymax, b.ymax),
                min(a.zmin, b.zmin), max(a.zmax, b.zmax))


class SyntheticElevation:
    """SyntheticElevation(lx, ly, zmin, zmax, z_bed, z_surf) (elevation_data.jl:35-46)."""

    def __init__(self, lx, ly, zmin, zmax, z_bed, z_surf):
        self.z_bed = z_bed
        self.z_surf = z_surf
        self.domain = AABB(-lx / 2, lx / 2, -ly / 2, ly / 2, zmin, zmax)
        self.rotated_domain = self.domain
        self.rotation = np.eye(3)


def filtered(X):
    """The values of X that are not NaN (geometry_helpers.jl:21)."""
    X = np.asarray(X, dtype=np.float64).ravel()
    return X[~np.isnan(X)]


def rotate(X, Y, Z, R):
    """(R (x, y, z)) for every point of the arrays X, Y, Z (geometry_helpers.jl:31-36)."""
    R = np.asarray(R, dtype=np.float64).reshape(3, 3)
    X, Y, Z = (np.asarray(a, dtype=np.float64) for a in (X, Y, Z))
    return (R[0, 0] * X + R[0, 1] * Y + R[0, 2] * Z, R[1, 0] * X + R[1, 1] * Y + R[1, 2] * Z,
            R[2, 0] * X + R[2, 1] * Y + R[2, 2] * Z)


def rotate_minmax(X, Y, Z, R):
    """rotate of the two points (min x, min y, min z) and (max x, max y, max z),
    z without NaNs (geometry_helpers.jl:39).  Note that this rotates two
    corners of the box, not all eight, as the reference does."""
    zf = filtered(Z)
    return rotate(np.array([np.min(X), np.max(X)]), np.array([np.min(Y), np.max(Y)]),
                  np.array([np.min(zf), np.max(zf)]), R)


class DataElevation:
    """DataElevation(x, y, offsets, z_bed, z_surf, R) (elevation_data.jl:19-33).

    x, y: vertex coordinates of the DEM; z_bed, z_surf: elevations (NaN
    where there is no data); R: rotation applied to the points before the
    distance to the DEM is taken (compute_levelset_from_dem's R).  domain is
    the box of x, y and the NaN-free elevations; rotated_domain the union of
    the boxes of the rotated bed and surface corner points."""

    def __init__(self, x, y, offsets, z_bed, z_surf, R):
        self.x = np.asarray(x, dtype=np.float64)
        self.y = np.asarray(y, dtype=np.float64)
        self.offsets = offsets
        self.z_bed = np.asarray(z_bed, dtype=np.float64)
        self.z_surf = np.asarray(z_surf, dtype=np.float64)
        self.rotation = np.asarray(R, dtype=np.float64).reshape(3, 3)
        zb, zs = filtered(self.z_bed), filtered(self.z_surf)
        self.domain = AABB(np.min(self.x), np.max(self.x), np.min(self.y), np.max(self.y),
                           min(np.min(zb), np.min(zs)), max(np.max(zb), np.max(zs)))
        bed = AABB.from_coords(*rotate_minmax(self.x, self.y, self.z_bed, self.rotation))
        surf = AABB.from_coords(*rotate_minmax(self.x, self.y, self.z_surf, self.rotation))
        self.rotated_domain = union(bed, surf)


def extract_dem(dem, dilation=0.05):
    """The DEM preparation of scripts_future_API/test_vavilov.jl:65-95 from a
    DataElevation (the reference loads it from a JLD2 file first): dilate the
    domain by 5 %, put z = 0 at its bottom and scale z by its height, set
    the horizontal extents to 10 and the height to 1.  Returns a dict with
    bed, ice (elevations at the DEM vertices, Julia index order), the 2D
    vertex grid of the DEM (dem_grid) and lx, ly, lz."""
    from .grids import UniformGrid
    dm = dilate(dem.domain, dilation)
    lx, ly, lz = extents(dm)
    nx, ny = (n - 1 for n in dem.z_surf.shape)
    ice = (dem.z_surf - dm.zmin) / lz
    bed = (dem.z_bed - dm.zmin) / lz
    lx /= lx / 10
    ly /= ly / 10
    lz /= lz
    grid_2d = UniformGrid(origin=(-lx / 2, -ly / 2), extent=(lx, ly), dims=(nx, ny))
    return {"bed": bed, "ice": ice, "dem_grid": grid_2d, "lx": lx, "ly": ly, "lz": lz}


def _linrange(a, b, n):
    return np.linspace(a, b, n)


def generate_turtle(nx, ny, lx, ly, zmin, zmax):
    """synthetic_geometries.jl:1-18 (y spans from -lx/2 as in the reference)."""
    amp, om, tanb, el, gl = 0.1, 10 * math.pi, math.tan(-math.pi / 6), 0.35, 0.9
    x = _linrange(-lx / 2, lx / 2, nx + 1)
    y = _linrange(-lx / 2, ly / 2, ny + 1)
    X, Y = np.meshgrid(x / lx, y / ly, indexing="ij")
    bed = (zmax - zmin) * (amp * np.sin(om * X) * np.sin(om * Y) + tanb * X + el + (1.5 * Y) ** 2)
    surf = gl * (1.0 - ((1.7 * X + 0.22) ** 2 + (0.5 * Y) ** 2))
    return bed, surf


def generate_dome(nx, ny, lx, ly):
    """synthetic_geometries.jl:20-33"""
    h0 = 0.8
    x = _linrange(-lx / 2, lx / 2, nx + 1)
    y = _linrange(-lx / 2, ly / 2, ny + 1)
    X, Y = np.meshgrid(x, y, indexing="ij")
    bed = np.zeros_like(X)
    surf = h0 - (X ** 2 + Y ** 2)
    surf[surf < bed] = 0.0
    return bed, surf


def generate_valley(nx, ny, lx, ly, ox, oy, dz, rgl, ogx, ogy, ogz):
    """synthetic_geometries.jl:35-52 (y built from ox as in the reference)."""
    x = _linrange(ox, ox + lx, nx + 1)
    y = _linrange(ox, oy + ly, ny + 1)
    X, Y = np.meshgrid(x, y, indexing="ij")
    bed = 0.025 * rgl * np.sin(2 * math.pi * 10 * (X - ox) / lx) * np.cos(2 * math.pi * 10 * (Y - oy) / lx)
    bed = bed - bed.min() + dz
    surf = np.sqrt(np.maximum(rgl ** 2 - (X - ogx) ** 2 - (Y - ogy) ** 2, 0.0)) + ogz
    return bed, surf


def make_synthetic(nx, ny, lx, ly, zmin, zmax, type):
    """make_synthetic(nx, ny, lx, ly, zmin, zmax, type) (synthetic_geometries.jl:72-89)."""
    if type in ("dome", ":dome"):
        bed, surf = generate_dome(nx, ny, lx, ly)
    elif type in ("turtle", ":turtle"):
        bed, surf = generate_turtle(nx, ny, lx, ly, zmin, zmax)
    elif type in ("valley", ":valley"):
        lx, ly = 5.0, 2.5
        bed, surf = generate_valley(nx, ny, lx, ly, -lx / 2, -ly / 2, 0.1, 4.0, 0.0, 0.0, -3.2)
    else:
        raise ValueError("Synthetic dem type not defined")
    return SyntheticElevation(lx, ly, zmin, zmax, bed, surf)


def splitmix64(seed):
    """Deterministic 64-bit generator (SplitMix64) yielding floats in [0, 1)."""
    state = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        z ^= z >> 31
        yield (z >> 11) * (1.0 / 9007199254740992.0)


class GaussianGlacier:
    """Seeded synthetic glacier: a bed of 8 Gaussian bumps under a parabolic ice surface.

    Domain [-lx/2, lx/2] x [-ly/2, ly/2] x [0, H].  Bed:
    z_b = z0 + sum_k a_k exp(-((x-x_k)^2 + (y-y_k)^2) / (2 w_k^2)) with
    a_k ~ U(0.05, 0.2) H, w_k ~ U(0.1, 0.3) L (L = min(lx, ly)), centres
    uniform in the xy domain; surface z_s = h0 - c (x^2 + y^2).
    """

    def __init__(self, seed=1, lx=2.0, ly=2.0, H=1.0, z0=0.05, h0=0.85, c=0.6, nbumps=8):
        rng = splitmix64(seed)
        L = min(lx, ly)
        self.lx, self.ly, self.H, self.z0, self.h0, self.c = lx, ly, H, z0, h0, c
        self.bumps = []
        for _ in range(nbumps):
            a = (0.05 + 0.15 * next(rng)) * H
            w = (0.1 + 0.2 * next(rng)) * L
            xk = -lx / 2 + lx * next(rng)
            yk = -ly / 2 + ly * next(rng)
            self.bumps.append((a, w, xk, yk))

    def bed(self, x, y):
        z = np.full(np.broadcast(x, y).shape, self.z0)
        for a, w, xk, yk in self.bumps:
            z = z + a * np.exp(-((x - xk) ** 2 + (y - yk) ** 2) / (2.0 * w * w))
        return z

    def surface(self, x, y):
        return self.h0 - self.c * (x * x + y * y)

    def domain(self):
        return AABB(-self.lx / 2, self.lx / 2, -self.ly / 2, self.ly / 2, 0.0, self.H)
