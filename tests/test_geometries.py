"""DEM geometry helpers (src/Geometries/elevation_data.jl, geometry_helpers.jl)
and the glacier driver's DEM preparation (scripts_future_API/test_vavilov.jl:65-95),
against closed forms: no rotation, a quarter turn about z, NaN filtering,
the normalisation to a 10 x 10 x 1 box."""
import math

import numpy as np


def _dem(pkg, R):
    x = np.linspace(-3.0, 5.0, 9)
    y = np.linspace(1.0, 4.0, 7)
    X, Y = np.meshgrid(x, y, indexing="ij")
    bed = 0.1 * X + 0.2 * Y - 1.0
    surf = bed + 2.0 + 0.05 * X * Y
    surf[0, 0] = np.nan  # no data
    return x, y, bed, surf, pkg.DataElevation(x, y, None, bed, surf, R)


def test_data_elevation_domains(pkg):
    x, y, bed, surf, dem = _dem(pkg, np.eye(3))
    zmin = min(np.nanmin(bed), np.nanmin(surf))
    zmax = max(np.nanmax(bed), np.nanmax(surf))
    d = dem.domain
    assert (d.xmin, d.xmax, d.ymin, d.ymax) == (-3.0, 5.0, 1.0, 4.0)
    assert (d.zmin, d.zmax) == (zmin, zmax) and not math.isnan(d.zmax)
    r = dem.rotated_domain
    assert (r.xmin, r.xmax, r.ymin, r.ymax, r.zmin, r.zmax) == (d.xmin, d.xmax, d.ymin, d.ymax, d.zmin, d.zmax)


def test_data_elevation_quarter_turn(pkg):
    # (x, y, z) -> (-y, x, z): the two rotated corners span x in [-ymax, -ymin], y in [xmin, xmax]
    R = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    x, y, bed, surf, dem = _dem(pkg, R)
    r = dem.rotated_domain
    assert (r.xmin, r.xmax, r.ymin, r.ymax) == (-4.0, -1.0, -3.0, 5.0)
    assert r.zmin == dem.domain.zmin and r.zmax == dem.domain.zmax
    xr, yr, zr = pkg.rotate([1.0, 2.0], [3.0, 4.0], [5.0, 6.0], R)
    assert list(xr) == [-3.0, -4.0] and list(yr) == [1.0, 2.0] and list(zr) == [5.0, 6.0]
    assert list(pkg.filtered([1.0, np.nan, 2.0])) == [1.0, 2.0]


def test_extract_dem_normalisation(pkg):
    x, y, bed, surf, dem = _dem(pkg, np.eye(3))
    out = pkg.extract_dem(dem)
    dm = pkg.dilate(dem.domain, 0.05)
    lz = dm.zmax - dm.zmin
    assert (out["lx"], out["ly"], out["lz"]) == (10.0, 10.0, 1.0)
    assert np.array_equal(out["ice"], (surf - dm.zmin) / lz, equal_nan=True)
    assert np.array_equal(out["bed"], (bed - dm.zmin) / lz)
    assert 0.0 < np.nanmin(out["bed"]) and np.nanmax(out["ice"]) < 1.0
    g = out["dem_grid"]
    assert g.dims == (8, 6) and g.origin == (-5.0, -5.0) and g.extent_ == (10.0, 10.0)
