"""GPU parity of compute_τ! and of the immersed-boundary validity bits, and the
two IB special cases of tests/test_oracle_ib_pins.py run on the HIP path.

  * fi_compute_tau (k_tau3 / k_tau2) against the oracle's compute_τ!
    (FullStokes kernels_3d.jl:44-64, kernels_2d.jl:33-44; IB
    kernels_3d.jl:78-116), plain and IB, 3D and 2D, bit for bit;
  * fi_ib_validity (k_flags) against the oracle's or_ib_flags and the
    generic null_spaces.jl transcription, on random masks around 1e-6;
  * all-one masks: the IB model (fused sweep included) equals the plain
    model bit for bit;
  * the hydrostatic column on a cut bed stays at rest bit for bit.
"""
import math

import numpy as np
import pytest

import ib_transcription as T
from helpers import make_pair, state_names
from test_oracle_ib_pins import _random_masks, slab_masks, slab_state

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _kw(pkg, dims, extent, gravity=None, rheology=None):
    nd = len(dims)
    grid = pkg.UniformGrid(origin=(0.0,) * nd, extent=extent, dims=dims)
    t = pkg.TBC(*([0.0] * nd))
    bcs = {k: (t, t) for k in ("x", "y", "z")[:nd]}
    if gravity is None:
        gravity = dict(zip(("x", "y", "z")[:nd], [0.0] * (nd - 1) + [1.0]))
    return dict(grid=grid, boundary_conditions=bcs, gravity=gravity,
                rheology=rheology or pkg.LinearViscousRheology(1.0),
                solver_params=pkg.pt_params(grid, r=0.9, re_mech=5 * math.pi))


def _upload_masks(pkg, model, ref, na, ns):
    A = pkg._abi
    for q, a in enumerate(na):
        pkg.full_stokes.DeviceField(model, A.FI_F_OM_NA0 + q).set(a)
        ref.set("OM_NA%d" % q, a)
    for q, a in enumerate(ns):
        pkg.full_stokes.DeviceField(model, A.FI_F_OM_NS0 + q).set(a)
        ref.set("OM_NS%d" % q, a)


@pytest.mark.parametrize("ib", [False, True])
@pytest.mark.parametrize("dims", [(20, 17, 12), (70, 36, 9), (40, 23), (2, 3, 1), (1, 3, 2), (4, 3)])
def test_compute_tau_bitwise(pkg, orc, dims, ib):
    _say("case: compute_tau, dims %s, immersed boundary %s" % (dims, ib))
    nd = len(dims)
    kw = _kw(pkg, dims, (1.0, 0.8, 0.6)[:nd])
    rng = np.random.default_rng(31)
    shape = tuple(n + 3 for n in dims)
    eta = np.asfortranarray(rng.uniform(0.5, 2.0, size=shape))
    model, ref = make_pair(pkg, orc, kw, ib=ib, eta=eta, seed=17, amp=1.0)
    if ib:
        nq = 8 if nd == 3 else 4
        _upload_masks(pkg, model, ref, _random_masks(rng, shape, nq), _random_masks(rng, shape, nq))
    model.compute_tau()
    ref.compute_tau()
    names = state_names(nd)
    for name in names:
        a, b = model.field(name).parent(), ref.get(name)
        assert np.array_equal(a, b), name
    # teeth: compute_τ! rewrote the stress (random before)
    assert np.count_nonzero(ref.get("TXY")[tuple(slice(1, n + 2) for n in dims)]) > 0
    model.close()


@pytest.mark.parametrize("dims", [(5, 4, 6), (70, 36, 9), (9, 7), (1, 2, 1), (2, 1)])
def test_validity_bits_bitwise(pkg, orc, dims):
    _say("case: validity bits on random masks, dims %s" % (dims,))
    nd = len(dims)
    kw = _kw(pkg, dims, tuple(1.0 for _ in dims))
    model, ref = make_pair(pkg, orc, kw, ib=True)
    rng = np.random.default_rng(41)
    shape = tuple(n + 3 for n in dims)
    nq = 8 if nd == 3 else 4
    na, ns = _random_masks(rng, shape, nq), _random_masks(rng, shape, nq)
    _upload_masks(pkg, model, ref, na, ns)
    got = model.validity()
    assert np.array_equal(got, ref.ib_flags())
    g = T.Grid(dims, ref.m.dx[:nd])
    assert np.array_equal(got, T.flags(g, dict(zip(T.MASK3 if nd == 3 else T.MASK2, na)),
                                       dict(zip(T.MASK3 if nd == 3 else T.MASK2, ns))))
    model.close()


@pytest.mark.parametrize("dims", [(70, 36, 24), (40, 23)])
def test_all_fluid_ib_equals_plain(pkg, orc, dims):
    _say("case: all-one masks, IB model vs plain model, dims %s" % (dims,))
    nd = len(dims)
    rheo = pkg.Relaxation("eta", pkg.ViscoplasticRegularisation(pkg.PowerLawRheology(0.7, 3), 20.0), 0.3)
    kw = _kw(pkg, dims, (1.0, 0.8, 0.6)[:nd], gravity=dict(zip(("x", "y", "z")[:nd], (0.3, -0.2, 1.0)[:nd])),
             rheology=rheo)
    ib = pkg.IsothermalImmersedBoundaryFullStokesModel(**kw)
    pl = pkg.IsothermalFullStokesModel(**kw)
    A = pkg._abi
    for q in range(8 if nd == 3 else 4):
        pkg.full_stokes.DeviceField(ib, A.FI_F_OM_NA0 + q).set(1.0)
        pkg.full_stokes.DeviceField(ib, A.FI_F_OM_NS0 + q).set(1.0)
    rng = np.random.default_rng(23)
    shape = tuple(n + 3 for n in dims)
    eta = np.asfortranarray(rng.uniform(0.5, 2.0, size=shape))
    for name in state_names(nd):
        a = np.asfortranarray(1e-2 * rng.standard_normal(shape))
        ib.field(name).set(a)
        pl.field(name).set(a)
    for m in (ib, pl):
        m.viscosity.eta.set(eta)
        m.viscosity.eta_next.set(eta)
        m.apply_bc("eta")
    for n_it in (1, 2, 37):  # first, second and later iterations (the IB sweep's skipping states)
        ib.advance(n_it)
        pl.advance(n_it)
        for name in state_names(nd) + ["ETA", "ETA_NEXT"]:
            assert np.array_equal(ib.field(name).parent(), pl.field(name).parent()), (n_it, name)
    assert np.array_equal(ib.compute_residuals(), pl.compute_residuals())
    ib.compute_tau()
    pl.compute_tau()
    for name in state_names(nd)[1:]:
        assert np.array_equal(ib.field(name).parent(), pl.field(name).parent()), name
    ib.close()
    pl.close()


@pytest.mark.parametrize("dims", [(70, 36, 32), (40, 32)])
def test_hydrostatic_slab_at_rest(pkg, orc, dims):
    _say("case: hydrostatic column on a cut bed, dims %s" % (dims,))
    nd = len(dims)
    extent = tuple(n / 32.0 for n in dims)               # spacing 1/32: exact balance
    kw = _kw(pkg, dims, extent)
    model = pkg.IsothermalImmersedBoundaryFullStokesModel(**kw)
    A = pkg._abi
    dx = [e / n for e, n in zip(extent, dims)]
    na, ns = slab_masks(dims, dx, 0.3, kb=5)
    for q, a in enumerate(na):
        pkg.full_stokes.DeviceField(model, A.FI_F_OM_NA0 + q).set(a)
    for q, a in enumerate(ns):
        pkg.full_stokes.DeviceField(model, A.FI_F_OM_NS0 + q).set(a)
    flags = model.validity()
    model.field("P").set(slab_state(dims, dx, 1.0, flags))
    for name in state_names(nd)[1:]:
        model.field(name).set(0.0)
    model.viscosity.eta.set(1.0)
    model.viscosity.eta_next.set(1.0)
    model.apply_bc("stress")
    before = {k: model.field(k).parent() for k in state_names(nd)}
    assert np.all(model.compute_residuals() == 0.0)
    model.advance(20)
    for k, a in before.items():
        assert np.array_equal(model.field(k).parent(), a), k
    model.close()
