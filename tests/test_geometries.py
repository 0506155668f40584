"""DEM geometry helpers (src/Geometries/elevation_data.jl, geometry_helpers.jl)
and the glacier driver's DEM preparation (scripts_future_API/test_vavilov.jl:65-95),
against closed forms: no rotation, a quarter turn about z, NaN filtering,
the normalisation to a 10 x 10 x 1 box."""
import math

import numpy as np
import pytest


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


def test_make_synthetic_shapes_and_values(pkg):
    """make_synthetic (synthetic_geometries.jl:72-89): shapes (nx+1, ny+1) at
    the DEM vertices, values at points where the closed forms reduce to
    constants, and the reference's y axis that starts at -lx/2 (kept)."""
    G = pkg.geometries
    nx, ny = 40, 20
    # dome: bed 0; surface 0.8 - r^2 clipped at the bed; centre vertex (0, 0)
    d = G.make_synthetic(nx, ny, 2.0, 2.0, 0.0, 1.0, "dome")
    assert d.z_bed.shape == (nx + 1, ny + 1) and d.z_surf.shape == (nx + 1, ny + 1)
    assert np.all(d.z_bed == 0.0) and np.all(d.z_surf >= 0.0)
    assert d.z_surf[nx // 2, ny // 2] == 0.8
    assert d.z_surf[0, 0] == 0.0 and d.z_surf[-1, -1] == 0.0
    # lx != ly: the first y vertex sits at -lx/2, not -ly/2
    d2 = G.make_synthetic(nx, ny, 1.6, 1.0, 0.0, 1.0, "dome")
    assert abs(d2.z_surf[nx // 2, 0] - (0.8 - 0.8 ** 2)) < 1e-15
    # turtle at the centre vertex: bed (zmax - zmin) el, surface gl (1 - 0.22^2)
    t = G.make_synthetic(nx, ny, 5.0, 5.0, 0.5, 2.5, "turtle")
    assert abs(t.z_bed[nx // 2, ny // 2] - 2.0 * 0.35) < 1e-15
    assert abs(t.z_surf[nx // 2, ny // 2] - 0.9 * (1.0 - 0.22 ** 2)) < 1e-15
    # valley: its own extents; bed shifted to a minimum of 0.1, surface the
    # cap of a sphere of radius 4 centred 3.2 below the origin (0.8 on top)
    v = G.make_synthetic(nx, ny, 1.0, 1.0, 0.0, 1.0, "valley")
    dm = v.domain
    assert (dm.xmax - dm.xmin, dm.ymax - dm.ymin) == (5.0, 2.5)
    assert v.z_bed.min() == 0.1
    assert np.all(v.z_surf >= -3.2) and v.z_surf.max() <= 0.8 + 1e-15
    with pytest.raises(ValueError):
        G.make_synthetic(nx, ny, 1.0, 1.0, 0.0, 1.0, "cone")
