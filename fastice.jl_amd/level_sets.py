"""Level sets and volume fractions on the HIP library.

Mirrors FastIce.LevelSets (src/LevelSets/LevelSets.jl:3-4):
    compute_levelset_from_dem(psi, dem, dem_grid, grid, R=None)  # compute_levelset_from_dem!
    invert_levelset(psi)                                         # invert_levelset!
    omega_from_psi(omega, psi)                                   # ω_from_ψ!
psi and omega are the device fields of an immersed-boundary model
(model.level_sets.na / ns, model.field_masks.na / ns).
"""
import numpy as np

from . import _abi as A
from .grids import UniformGrid


def dem_parent(dem, dem_grid):
    """DEM values on the parent box of the 2D vertex grid, (dnx+3, dny+3), Julia order.

    A callable is evaluated at every parent vertex (FunctionField semantics);
    an interior array of (dnx+1, dny+1) vertex values gets zero halos, as a
    freshly allocated Field filled by set! has (scripts_future_API/test_vavilov.jl:207-210).
    """
    nx, ny = dem_grid.dims
    if callable(dem):
        X, Y = dem_grid.parent_coords(0b11)
        a = np.asarray(dem(X, Y), dtype=np.float64)
        return np.asfortranarray(np.broadcast_to(a, (nx + 3, ny + 3)))
    a = np.asarray(dem, dtype=np.float64)
    if a.shape == (nx + 3, ny + 3):
        return np.asfortranarray(a)
    if a.shape == (nx + 1, ny + 1):
        p = np.zeros((nx + 3, ny + 3), order="F")
        p[1:nx + 2, 1:ny + 2] = a
        return p
    raise ValueError("DEM of shape %s does not fit a %dx%d-cell grid" % (a.shape, nx, ny))


def compute_levelset_from_dem(psi, dem, dem_grid, grid=None, R=None):
    """Signed distance (+ above, - below) of every vertex to the DEM surface
    (compute_levelsets.jl:16-25, signed_distances.jl:78-92)."""
    if not isinstance(dem_grid, UniformGrid) or dem_grid.ndims != 2:
        raise TypeError("dem_grid must be a 2D UniformGrid")
    m = psi.model
    p = dem_parent(dem, dem_grid)
    o = np.array(dem_grid.origin, dtype=np.float64)
    e = np.array(dem_grid.extent_, dtype=np.float64)
    r = None if R is None else np.ascontiguousarray(np.asarray(R, dtype=np.float64).reshape(3, 3))
    A.check(A.lib().fi_levelset_from_dem(m._h, psi.phase, A.dptr(p), dem_grid.dims[0], dem_grid.dims[1],
                                         A.dptr(o), A.dptr(e), None if r is None else A.dptr(r)))


def compute_levelset_from_polygon(psi, poly):
    """2D level set of a closed polygon (n x 2 points): signed distance of
    every vertex, negative inside (sd_poly, signed_distances.jl:3-19)."""
    xy = np.ascontiguousarray(np.asarray(poly, dtype=np.float64).reshape(-1, 2))
    A.check(A.lib().fi_levelset_from_polygon(psi.model._h, psi.phase, A.dptr(xy), xy.shape[0]))


def invert_levelset(psi):
    """psi <- -psi (compute_levelsets.jl:27-35)."""
    A.check(A.lib().fi_invert_levelset(psi.model._h, psi.phase))


def omega_from_psi(omega, psi):
    """Volume fractions of {psi < 0} at the eight staggered locations (volume_fractions_kernels.jl:101-110)."""
    if omega.phase != psi.phase:
        raise ValueError("mask and level set belong to different phases")
    m = psi.model
    A.check(A.lib().fi_omega_from_psi(m._h, psi.phase))
