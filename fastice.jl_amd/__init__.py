"""fastice.jl_amd -- MI355X-native pseudo-transient Stokes path of FastIce.jl.

Host-side mirror of the reference API over libfasticehip.so (C ABI in
include/fastice_hip.h).  The directory name contains a dot, so load it with
`load_package()` from the repository root (it registers the module as
`fastice_jl_amd`):

    import importlib.util, sys
    spec = importlib.util.spec_from_file_location(
        "fastice_jl_amd", "fastice.jl_amd/__init__.py",
        submodule_search_locations=["fastice.jl_amd"])
"""
from . import _abi
from ._abi import FastIceError, NonFiniteError
from .batch import Batch, model_bytes
from .full_stokes import (BoundaryCondition, IsothermalFullStokesModel, IsothermalImmersedBoundaryFullStokesModel,
                          SBC, Slip, TBC, Traction, VBC, Velocity, advance_iteration, build_desc, compute_residuals,
                          norm_inf)
from .geometries import (AABB, DataElevation, GaussianGlacier, SyntheticElevation, center, dilate, extents, extract_dem,
                         filtered, make_synthetic, rotate, rotate_minmax, splitmix64, union)
from .grids import Center, UniformGrid, Vertex, flip
from .level_sets import compute_levelset_from_dem, compute_levelset_from_polygon, invert_levelset, omega_from_psi
from .physics import LinearViscousRheology, PowerLawRheology, Relaxation, ViscoplasticRegularisation
from .problems import (batch_glaciers, glacier_2d, glacier_2d_spec, glacier_3d, glacier_3d_spec, manufactured_2d,
                       manufactured_3d, pt_params)
from .serving import Request, Server, percentiles, poisson_requests
from . import _hdf5, heat, writers
from .heat import HeatDiffusionModel
from .writers import read_h5, write_h5, write_xdmf


def mem_info():
    """(free, total) device memory in bytes."""
    import ctypes as C
    f, t = C.c_size_t(0), C.c_size_t(0)
    _abi.check(_abi.lib().fi_mem_info(C.byref(f), C.byref(t)))
    return f.value, t.value
