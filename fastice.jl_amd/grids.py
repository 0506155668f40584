"""Uniform staggered grids (the parts of Chmy's UniformGrid the PT path uses).

Coordinates follow the reference's grid tests (test/test_chmy_grids.jl:113-160):
x_v(i) = x0 + (i-1)*dx and x_c(i) = x0 + (i-1/2)*dx for the 1-based index i,
extended to halo indices by the same formulas.
"""
import numpy as np


class Location:
    __slots__ = ()


class _Center(Location):
    def __repr__(self):
        return "Center()"


class _Vertex(Location):
    def __repr__(self):
        return "Vertex()"


_C = _Center()
_V = _Vertex()


def Center():
    return _C


def Vertex():
    return _V


def flip(loc):
    """flip(Center()) == Vertex() and vice versa (test_chmy_grids.jl:9-10)."""
    return _V if loc is _C else _C


def loc_bits(locs):
    """Tuple of Locations -> bit mask (bit d set = Vertex in dimension d)."""
    return sum((1 << d) for d, l in enumerate(locs) if l is _V)


def bits_loc(bits, ndims):
    return tuple(_V if (bits >> d) & 1 else _C for d in range(ndims))


class UniformGrid:
    """UniformGrid(; origin, extent, dims) on one device."""

    def __init__(self, origin, extent, dims, arch=None):
        self.origin = tuple(float(o) for o in origin)
        self.extent_ = tuple(float(e) for e in extent)
        self.dims = tuple(int(n) for n in dims)
        if not (len(self.origin) == len(self.extent_) == len(self.dims)):
            raise ValueError("origin, extent and dims must have the same length")
        if any(n < 1 for n in self.dims) or any(e <= 0 for e in self.extent_):
            raise ValueError("dims must be >= 1 and extents > 0")
        self.arch = arch
        self.spacing_ = tuple(e / n for e, n in zip(self.extent_, self.dims))

    @property
    def ndims(self):
        return len(self.dims)

    def _locs(self, loc):
        if isinstance(loc, Location):
            return (loc,) * self.ndims
        return tuple(loc)

    def size(self, loc=None):
        loc = self._locs(_C if loc is None else loc)
        return tuple(n + (1 if l is _V else 0) for n, l in zip(self.dims, loc))

    def spacing(self, loc=None, dim=None):
        return self.spacing_ if dim is None else self.spacing_[dim]

    def inv_spacing(self, loc=None, dim=None):
        inv = tuple(1.0 / d for d in self.spacing_)
        return inv if dim is None else inv[dim]

    def coord(self, loc, dim, i):
        """Coordinate of 1-based index i at location loc in dimension dim (0-based)."""
        l = self._locs(loc)[dim]
        if l is _V:
            return self.origin[dim] + (i - 1) * self.spacing_[dim]
        return self.origin[dim] + (i - 0.5) * self.spacing_[dim]

    def coords(self, loc, dim, lo=1, hi=None):
        l = self._locs(loc)[dim]
        if hi is None:
            hi = self.dims[dim] + (1 if l is _V else 0)
        i = np.arange(lo, hi + 1, dtype=np.float64)
        if l is _V:
            return self.origin[dim] + (i - 1.0) * self.spacing_[dim]
        return self.origin[dim] + (i - 0.5) * self.spacing_[dim]

    def bounds(self, loc, dim):
        l = self._locs(loc)[dim]
        n = self.dims[dim] + (1 if l is _V else 0)
        return self.coord(loc, dim, 1), self.coord(loc, dim, n)

    def extent(self, loc=None, dim=None):
        loc = _V if loc is None else loc
        ext = tuple(self.bounds(loc, d)[1] - self.bounds(loc, d)[0] for d in range(self.ndims))
        return ext if dim is None else ext[dim]

    def origin_of(self, loc, dim=None):
        o = tuple(self.bounds(loc, d)[0] for d in range(self.ndims))
        return o if dim is None else o[dim]

    def parent_shape(self):
        """Uniform parent box of every field: (nx+3, ny+3[, nz+3])."""
        return tuple(n + 3 for n in self.dims)

    def parent_coords(self, bits):
        """Coordinate arrays (meshgrid, Julia index order) of the whole parent box at a location."""
        axes = []
        for d in range(self.ndims):
            i = np.arange(0, self.dims[d] + 3, dtype=np.float64)
            if (bits >> d) & 1:
                axes.append(self.origin[d] + (i - 1.0) * self.spacing_[d])
            else:
                axes.append(self.origin[d] + (i - 0.5) * self.spacing_[d])
        return np.meshgrid(*axes, indexing="ij")

    def __repr__(self):
        return "UniformGrid(origin=%s, extent=%s, dims=%s)" % (self.origin, self.extent_, self.dims)
