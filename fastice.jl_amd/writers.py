"""HDF5 datasets and XDMF3 metadata of model fields (src/Writers.jl).

write_h5(grid, path, fields) and write_xdmf(grid, path, fields, h5_names,
timesteps) follow Writers.jl:17-24 and :54-64 on one device: every field
(name -> DeviceField, or an array in Julia index order) is brought to the
host once, as its interior, and stored as dataset "/name" of the grid's
Center size.  HDF5 stores C order, so a Julia-order (nx, ny, nz) interior
becomes dims (nz, ny, nx) with x fastest, which is what HDF5.jl writes for
the reference (and what the XDMF Dimensions attribute lists).  The XDMF
document is the reference's (Writers.jl:80-137) byte for byte in 3D; 2D
grids get 2DCoRectMesh / ORIGIN_DXDY where the reference writes the 3D
names regardless of the dimension.
"""
import numpy as np

from . import _hdf5
from .grids import Center


def _interior(field):
    """Julia-order interior of a DeviceField or an array."""
    if hasattr(field, "interior"):
        return field.interior()
    return np.asarray(field, dtype=np.float64)


def write_dset(io, fields, grid_size):
    """write_dset (Writers.jl:47-53): one dataset of the Center size per field."""
    for name, field in fields.items():
        a = _interior(field)
        if tuple(a.shape) != tuple(grid_size):
            raise ValueError("field %s has interior size %s, the grid's Center size is %s"
                             % (name, tuple(a.shape), tuple(grid_size)))
        # Julia order (x fastest) -> C order with the dims reversed
        io.write("/" + name, np.asfortranarray(a).T)


def write_h5(grid, path, fields):
    """write_h5(arch, grid, path, fields) (Writers.jl:17-24)."""
    with _hdf5.File(path, "w") as io:
        write_dset(io, fields, grid.size(Center()))


def read_h5(path, name):
    """h5read(path, name) in Julia index order (the inverse of write_h5)."""
    with _hdf5.File(path, "r") as io:
        return np.asfortranarray(io.read("/" + name).T)


def velocity_magnitude(model):
    """Vm = sqrt(lerp(Vx)^2 + lerp(Vy)^2 + lerp(Vz)^2) at cell centres, in
    Julia index order: the output field of scripts_future_API/test_vavilov.jl:183,238-240
    (lerp from faces to centres: (V[i] + V[i+1]) / 2 in the face's dimension)."""
    nd = model.grid.ndims
    total = None
    for d, name in enumerate(("VX", "VY", "VZ")[:nd]):
        v = model.field(name).interior()
        lo = [slice(None)] * nd
        hi = [slice(None)] * nd
        lo[d] = slice(0, -1)
        hi[d] = slice(1, None)
        c = 0.5 * (v[tuple(lo)] + v[tuple(hi)])
        total = c * c if total is None else total + c * c
    return np.sqrt(total)


def glacier_output_fields(model):
    """The fields the glacier driver saves (scripts_future_API/test_vavilov.jl:206):
    pressure, velocity magnitude and the Center volume fractions."""
    return {"Pr": model.field("P"), "Vm": velocity_magnitude(model), "wt_ns_c": model.field_masks.ns.ccc,
            "wt_na_c": model.field_masks.na.ccc}


def julia_float(x):
    """Float64 text as Julia prints it: the shortest round-trip digits, in
    positional form for 1e-4 <= |x| < 1e6 (and 0), else d.ddde-n / d.ddden.
    The golden XDMF of test_writers.jl pins the positional form; the
    exponent form is parity unpinned."""
    x = float(x)
    if x != x:
        return "NaN"
    if x in (float("inf"), float("-inf")):
        return "Inf" if x > 0 else "-Inf"
    r = repr(x)
    ax = abs(x)
    if x == 0.0 or 1e-4 <= ax < 1e6:
        if "e" in r or "E" in r:
            r = np.format_float_positional(x, unique=True, trim="0")
        return r if ("." in r) else r + ".0"
    m, e = np.format_float_scientific(x, unique=True, trim="0").split("e")
    if "." not in m:
        m += ".0"
    elif m.endswith("."):
        m += "0"
    return "%se%d" % (m, int(e))


def generate_xdmf(grid_size, grid_spacing, grid_origin, names, h5_names, timesteps):
    """The XDMF3 document (Writers.jl:80-137) as text."""
    nd = len(grid_size)
    topo = "3DCoRectMesh" if nd == 3 else "2DCoRectMesh"
    geom = "ORIGIN_DXDYDZ" if nd == 3 else "ORIGIN_DXDY"
    rev = lambda t: list(reversed(list(t)))
    ts = list(timesteps) if np.ndim(timesteps) else [timesteps]
    h5s = list(h5_names)
    if len(h5s) < len(ts):
        raise ValueError("one HDF5 file name per time step")
    lines = ['<?xml version="1.0" encoding="utf-8"?>', '<Xdmf Version="3.0">', "  <Domain>",
             '    <Grid GridType="Collection" CollectionType="Temporal">']
    for it, tt in enumerate(ts):
        h5_path = h5s[it]
        if not h5_path.endswith(".h5"):
            raise ValueError("HDF5 file names must end in .h5 (got %s)" % h5_path)
        lines += ['      <Grid GridType="Uniform">',
                  '        <Topology TopologyType="%s" Dimensions="%s"/>' % (topo, " ".join(str(n + 1) for n in rev(grid_size))),
                  '        <Time Value="%s"/>' % julia_float(tt),
                  '        <Geometry GeometryType="%s">' % geom,
                  '          <DataItem Format="XML" NumberType="Float" Dimensions="%d">%s</DataItem>'
                  % (nd, " ".join(julia_float(v) for v in rev(grid_origin))),
                  '          <DataItem Format="XML" NumberType="Float" Dimensions="%d">%s</DataItem>'
                  % (nd, " ".join(julia_float(v) for v in rev(grid_spacing))),
                  "        </Geometry>"]
        for name in names:
            lines += ['        <Attribute Name="%s" Center="Cell">' % name,
                      '          <DataItem Format="HDF" NumberType="Float" Precision="8" Dimensions="%s">%s:/%s</DataItem>'
                      % (" ".join(str(n) for n in rev(grid_size)), h5_path, name),
                      "        </Attribute>"]
        lines.append("      </Grid>")
    lines += ["    </Grid>", "  </Domain>", "</Xdmf>"]
    return "\n".join(lines) + "\n"


def write_xdmf(grid, path, fields, h5_names, timesteps=0.0):
    """write_xdmf(arch, grid, path, fields, h5_names, timesteps=0.0) (Writers.jl:54-64)."""
    text = generate_xdmf(grid.size(Center()), grid.spacing(), grid.origin_of(Center()), list(fields), h5_names,
                         timesteps)
    with open(path, "w") as f:
        f.write(text)
