"""ctypes wrapper of the CPU FP64 oracle (oracle/fastice_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker of the HIP library.  The product
package never imports this module.

OracleModel keeps every field as a numpy array in the parent layout of the
HIP library ((nx+3, ny+3[, nz+3]), Julia index order, x fastest), takes the
same fi_model_desc, and exposes set/get/advance/compute_residuals/solve and
the level-set calls.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "fastice_oracle.c")
LIB = os.path.join(HERE, "liboracle.so")

CFLAGS = ["-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-std=c11", "-Wall"]


def build(verbose=False):
    """Compile liboracle.so with gcc (FP contraction off)."""
    tmp = LIB + ".tmp"
    cmd = ["gcc"] + CFLAGS + ["-o", tmp, SRC, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


_DP = C.POINTER(C.c_double)


class or_model(C.Structure):
    _fields_ = [
        ("ndims", C.c_int),
        ("n", C.c_int * 3),
        ("origin", C.c_double * 3),
        ("extent", C.c_double * 3),
        ("dx", C.c_double * 3),
        ("idx", C.c_double * 3),
        ("sy", C.c_long),
        ("sz", C.c_long),
        ("bc_kind", (C.c_int * 2) * 3),
        ("bc_val", ((C.c_double * 3) * 2) * 3),
        ("dPr", C.c_double),
        ("dtau", C.c_double * 6),
        ("dV", C.c_double * 3),
        ("rh_kind", C.c_int),
        ("rh_eta", C.c_double),
        ("rh_A", C.c_double),
        ("rh_n", C.c_double),
        ("rh_has_reg", C.c_int),
        ("rh_inv_eta_reg", C.c_double),
        ("rh_relax", C.c_int),
        ("rh_rate", C.c_double),
        ("grav_field", C.c_int),
        ("grav", C.c_double * 3),
    ]


_FNAMES = ["P", "txx", "tyy", "tzz", "txy", "txz", "tyz", "Vx", "Vy", "Vz", "eta", "eta_next",
           "rP", "rVx", "rVy", "rVz", "gx", "gy", "gz", "rh_field"]


class or_fields(C.Structure):
    _fields_ = [(n, _DP) for n in _FNAMES] + [("na", _DP * 8), ("ns", _DP * 8)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        M, Fp = C.POINTER(or_model), C.POINTER(or_fields)
        sigs = {
            "or_init_grid": (None, [M]),
            "or_advance": (None, [M, Fp, C.c_int, C.c_int]),
            "or_update_sigma": (None, [M, Fp, C.c_int]),
            "or_update_V": (None, [M, Fp, C.c_int]),
            "or_bc_stress": (None, [M, Fp]),
            "or_bc_velocity": (None, [M, Fp]),
            "or_bc_neumann_center": (None, [M, _DP]),
            "or_bc_neumann_vertex": (None, [M, _DP]),
            "or_compute_residuals": (None, [M, Fp, C.c_int]),
            "or_compute_tau": (None, [M, Fp, C.c_int]),
            "or_ib_flags": (None, [M, Fp, C.POINTER(C.c_uint8)]),
            "or_residual_norms": (None, [M, Fp, _DP]),
            "or_solve": (C.c_int, [M, Fp, C.c_int, C.c_int, C.c_double, C.c_int, _DP]),
            "or_rheology": (C.c_double, [M, Fp, C.c_long, C.c_double]),
            "or_omega_from_psi": (None, [M, _DP, _DP * 8]),
            "or_psi2omega3": (C.c_double, [_DP, _DP]),
            "or_psi2omega2": (C.c_double, [_DP, _DP]),
            "or_levelset_from_dem": (None, [M, _DP, _DP, C.POINTER(C.c_int), _DP, _DP, _DP]),
            "or_levelset_from_polygon": (None, [M, _DP, _DP, C.c_int]),
            "or_sd_poly": (C.c_double, [C.c_double, C.c_double, _DP, C.c_int]),
            "or_invert_levelset": (None, [M, _DP]),
            "or_heat_advance": (None, [C.POINTER(or_heat), C.POINTER(or_heat_fields), C.c_double, C.c_int]),
            "or_heat_bc": (None, [C.POINTER(or_heat), C.POINTER(or_heat_fields)]),
            "or_heat_residual_norms": (None, [C.POINTER(or_heat), C.POINTER(or_heat_fields), C.c_double, _DP]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_DP) if a is not None else None


# fi field ids -> oracle names (matches fastice.jl_amd/_abi.FIELD_NAMES order)
FIELD_NAMES = ["P", "TXX", "TYY", "TZZ", "TXY", "TXZ", "TYZ", "VX", "VY", "VZ", "ETA", "ETA_NEXT",
               "RP", "RVX", "RVY", "RVZ", "GX", "GY", "GZ", "RHEO", "PSI_NA", "PSI_NS"]
_TO_OR = dict(zip(FIELD_NAMES[:20], _FNAMES))


def psi2omega3(psi8, d):
    p = np.ascontiguousarray(psi8, dtype=np.float64)
    dd = np.ascontiguousarray(d, dtype=np.float64)
    return lib().or_psi2omega3(_p(p), _p(dd))


def sd_poly(p, xy):
    """or_sd_poly: signed distance of point p to the closed polygon xy (n x 2)."""
    pts = np.ascontiguousarray(np.asarray(xy, dtype=np.float64).reshape(-1, 2))
    return lib().or_sd_poly(float(p[0]), float(p[1]), _p(pts), pts.shape[0])


def psi2omega2(psi4, d):
    p = np.ascontiguousarray(psi4, dtype=np.float64)
    dd = np.ascontiguousarray(d, dtype=np.float64)
    return lib().or_psi2omega2(_p(p), _p(dd))


class OracleModel:
    """CPU restatement of one PT-Stokes model, driven from an fi_model_desc."""

    def __init__(self, desc):
        self.m = or_model()
        m = self.m
        nd = desc.ndims
        m.ndims = nd
        for q in range(3):
            m.n[q] = desc.n[q] if q < nd else 0
            m.origin[q] = desc.origin[q]
            m.extent[q] = desc.extent[q] if q < nd else 1.0
            for s in range(2):
                m.bc_kind[q][s] = desc.bc_kind[q][s]
                for c in range(3):
                    m.bc_val[q][s][c] = desc.bc_value[q][s][c]
            m.dV[q] = desc.dtau_V[q]
            m.grav[q] = desc.gravity[q]
        m.dPr = desc.dtau_P
        for q in range(6):
            m.dtau[q] = desc.dtau_tau[q]
        r = desc.rheology
        m.rh_kind = r.kind
        m.rh_eta = r.eta
        m.rh_A = r.A
        m.rh_n = r.n
        m.rh_has_reg = r.has_reg
        m.rh_inv_eta_reg = (1.0 / r.eta_reg) if r.has_reg else 0.0
        m.rh_relax = r.relax
        m.rh_rate = r.rate
        m.grav_field = desc.gravity_is_field
        self.ib = bool(desc.immersed_boundary)
        lib().or_init_grid(C.byref(m))
        self.ndims = nd
        self.shape = tuple(desc.n[q] + 3 for q in range(nd))
        self.arrays = {}
        names = FIELD_NAMES[:16]
        if desc.gravity_is_field:
            names = names + ["GX", "GY"] + (["GZ"] if nd == 3 else [])
        rh_field = r.kind == 1 or (r.kind == 2 and r.A_is_field)
        if rh_field:
            names = names + ["RHEO"]
        for nme in names:
            self.arrays[nme] = np.zeros(self.shape, order="F")
        if self.ib:
            self.arrays["PSI_NA"] = np.zeros(self.shape, order="F")
            self.arrays["PSI_NS"] = np.zeros(self.shape, order="F")
            for ph in ("NA", "NS"):
                for q in range(8 if nd == 3 else 4):
                    self.arrays["OM_%s%d" % (ph, q)] = np.zeros(self.shape, order="F")
        self.eta_cur = 0
        self._rh_field = rh_field

    # -- fields -------------------------------------------------------------
    def _buf(self, name):
        if name == "ETA":
            return self.arrays["ETA_NEXT" if self.eta_cur else "ETA"]
        if name == "ETA_NEXT":
            return self.arrays["ETA" if self.eta_cur else "ETA_NEXT"]
        return self.arrays[name]

    def get(self, name):
        return self._buf(name)

    def set(self, name, value, bits=None, grid=None):
        a = self._buf(name)
        if np.isscalar(value):
            a[...] = value
        elif callable(value):
            a[...] = np.broadcast_to(value(*grid.parent_coords(bits)), self.shape)
        else:
            v = np.asarray(value, dtype=np.float64)
            if v.shape == self.shape:
                a[...] = v
            else:
                sl = tuple(slice(1, 1 + s) for s in v.shape)
                a[sl] = v
        return a

    def _fields(self):
        f = or_fields()
        for fid_name, or_name in _TO_OR.items():
            a = self.arrays.get(fid_name)
            if fid_name == "ETA":
                a = self._buf("ETA")
            elif fid_name == "ETA_NEXT":
                a = self._buf("ETA_NEXT")
            if a is not None:
                setattr(f, or_name, _p(a))
        if not self._rh_field:
            f.rh_field = None
        if self.ib:
            nq = 8 if self.ndims == 3 else 4
            for q in range(nq):
                f.na[q] = _p(self.arrays["OM_NA%d" % q])
                f.ns[q] = _p(self.arrays["OM_NS%d" % q])
        return f

    def _sync_eta(self, f):
        cur = C.cast(f.eta, C.c_void_p).value
        self.eta_cur = 0 if cur == self.arrays["ETA"].ctypes.data else 1

    # -- driver -------------------------------------------------------------
    def advance(self, niter=1):
        f = self._fields()
        lib().or_advance(C.byref(self.m), C.byref(f), int(niter), int(self.ib))
        self._sync_eta(f)

    def apply_bc(self, group):
        f = self._fields()
        if group == "stress":
            lib().or_bc_stress(C.byref(self.m), C.byref(f))
        elif group == "velocity":
            # velocity fields only (eta_next Neumann is part of the iteration's batch)
            save = self._buf("ETA_NEXT").copy()
            lib().or_bc_velocity(C.byref(self.m), C.byref(f))
            self._buf("ETA_NEXT")[...] = save
        elif group in ("eta", "viscosity"):
            lib().or_bc_neumann_center(C.byref(self.m), _p(self._buf("ETA")))
            lib().or_bc_neumann_center(C.byref(self.m), _p(self._buf("ETA_NEXT")))
        else:
            raise ValueError(group)

    def compute_residuals(self):
        f = self._fields()
        lib().or_compute_residuals(C.byref(self.m), C.byref(f), int(self.ib))
        out = np.zeros(4)
        lib().or_residual_norms(C.byref(self.m), C.byref(f), _p(out))
        return out

    def compute_tau(self):
        f = self._fields()
        lib().or_compute_tau(C.byref(self.m), C.byref(f), int(self.ib))

    def update_sigma(self):
        """update_σ! alone (no boundary conditions)."""
        f = self._fields()
        lib().or_update_sigma(C.byref(self.m), C.byref(f), int(self.ib))

    def update_V(self):
        """update_V! alone (no boundary conditions, no viscosity swap)."""
        f = self._fields()
        lib().or_update_V(C.byref(self.m), C.byref(f), int(self.ib))

    def residuals_only(self):
        """compute_residuals! (kernel and residual boundary conditions), no norms."""
        f = self._fields()
        lib().or_compute_residuals(C.byref(self.m), C.byref(f), int(self.ib))

    def ib_flags(self):
        """Validity bits (or_ib_flags) as a uint8 array of the parent shape."""
        out = np.zeros(self.shape, dtype=np.uint8, order="F")
        f = self._fields()
        lib().or_ib_flags(C.byref(self.m), C.byref(f), out.ctypes.data_as(C.POINTER(C.c_uint8)))
        return out

    def solve(self, max_iter, ncheck, tol):
        f = self._fields()
        out = np.zeros(4)
        it = lib().or_solve(C.byref(self.m), C.byref(f), int(max_iter), int(ncheck), float(tol), int(self.ib), _p(out))
        self._sync_eta(f)
        return it, out

    def rheology_at(self, c, tII):
        f = self._fields()
        return lib().or_rheology(C.byref(self.m), C.byref(f), int(c), float(tII))

    # -- level sets -----------------------------------------------------------
    def levelset_from_dem(self, phase, dem_parent, dem_n, dem_origin, dem_extent, R=None):
        psi = self.arrays["PSI_NA" if phase == 0 else "PSI_NS"]
        dem = np.asfortranarray(dem_parent, dtype=np.float64)
        n = (C.c_int * 2)(*dem_n)
        o = np.asarray(dem_origin, dtype=np.float64)
        e = np.asarray(dem_extent, dtype=np.float64)
        r = None if R is None else np.ascontiguousarray(np.asarray(R, dtype=np.float64).reshape(9))
        lib().or_levelset_from_dem(C.byref(self.m), _p(psi), _p(dem), n, _p(o), _p(e), _p(r))

    def levelset_from_polygon(self, phase, xy):
        psi = self.arrays["PSI_NA" if phase == 0 else "PSI_NS"]
        pts = np.ascontiguousarray(np.asarray(xy, dtype=np.float64).reshape(-1, 2))
        lib().or_levelset_from_polygon(C.byref(self.m), _p(psi), _p(pts), pts.shape[0])

    def invert_levelset(self, phase):
        lib().or_invert_levelset(C.byref(self.m), _p(self.arrays["PSI_NA" if phase == 0 else "PSI_NS"]))

    def omega_from_psi(self, phase):
        ph = "NA" if phase == 0 else "NS"
        psi = self.arrays["PSI_" + ph]
        om = (_DP * 8)()
        nq = 8 if self.ndims == 3 else 4
        for q in range(nq):
            om[q] = _p(self.arrays["OM_%s%d" % (ph, q)])
        lib().or_omega_from_psi(C.byref(self.m), _p(psi), om)


class or_heat(C.Structure):
    _fields_ = [("ndims", C.c_int), ("n", C.c_int * 3), ("dx", C.c_double * 3), ("sy", C.c_long), ("sz", C.c_long),
                ("bc_kind", (C.c_int * 2) * 3), ("bc_val", (C.c_double * 2) * 3), ("lam", C.c_double),
                ("dtq", C.c_double), ("dtT", C.c_double)]


class or_heat_fields(C.Structure):
    _fields_ = [("T", _DP), ("To", _DP), ("q", _DP * 3)]


class OracleHeat:
    """CPU restatement of the PT heat-diffusion model (or_heat_*): fields T,
    To, QX, QY[, QZ] as parent-layout arrays in Julia index order."""

    def __init__(self, ndims, n, extent, bc_kind, bc_val, lam, dtq, dtT):
        h = self.h = or_heat()
        h.ndims = ndims
        for d in range(3):
            h.n[d] = n[d] if d < ndims else 0
            h.dx[d] = extent[d] / n[d] if d < ndims else 1.0
            for s in range(2):
                h.bc_kind[d][s] = bc_kind[d][s] if d < ndims else 0
                h.bc_val[d][s] = bc_val[d][s] if d < ndims else 0.0
        h.sy = n[0] + 3
        h.sz = (n[0] + 3) * (n[1] + 3) if ndims == 3 else 0
        h.lam, h.dtq, h.dtT = lam, dtq, dtT
        self.shape = tuple(n[d] + 3 for d in range(ndims))
        names = ["T", "TO", "QX", "QY"] + (["QZ"] if ndims == 3 else [])
        self.arrays = {k: np.zeros(self.shape, order="F") for k in names}
        self.f = or_heat_fields()
        self.f.T = self.arrays["T"].ctypes.data_as(_DP)
        self.f.To = self.arrays["TO"].ctypes.data_as(_DP)
        for d, k in enumerate(("QX", "QY", "QZ")[:ndims]):
            self.f.q[d] = self.arrays[k].ctypes.data_as(_DP)

    def set(self, name, a):
        self.arrays[name][...] = np.asarray(a, dtype=np.float64)

    def get(self, name):
        return self.arrays[name].copy(order="F")

    def apply_bc(self):
        lib().or_heat_bc(C.byref(self.h), C.byref(self.f))

    def advance(self, dt, niter):
        lib().or_heat_advance(C.byref(self.h), C.byref(self.f), C.c_double(dt), int(niter))

    def residual_norms(self, dt):
        out = (C.c_double * 2)()
        lib().or_heat_residual_norms(C.byref(self.h), C.byref(self.f), C.c_double(dt), out)
        return np.array(out[:])
