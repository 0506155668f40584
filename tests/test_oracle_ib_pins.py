"""Pins of the oracle's immersed-boundary path (no GPU).

The reference ships no vectors for the masked kernels and its Julia code
cannot run here, so the oracle's IB branches (oracle/fastice_oracle.c:
null3/valid3/null2/valid2, sigma3/velocity3/residuals3/tau3 and their 2D
forms) are pinned three ways:

  * against an independent numpy transcription of null_spaces.jl:1-43 and
    IB kernels_3d.jl:3-143 (tests/ib_transcription.py) that derives every
    neighbour from the reference's generic flip/left/right rule, on random
    masks with values on and around the 1e-6 threshold, halos included:
    a wrong neighbour in the oracle's per-location case table, a swapped
    na/ns or a transposed operand fails it;
  * with every volume fraction 1 (halos included) the masked kernels must
    reduce bit for bit to the plain FullStokes kernels (kernels_3d.jl:3-75),
    whose oracle is pinned by closed forms in test_oracle_pins.py;
  * a hydrostatic column above a bed cut by the solid mask is an exact
    discrete rest state: pressure rho g (z_top - z), no stress, no flow.
    Grid spacings are powers of two so that the balance is exact in
    floating point: the state must stay unchanged bit for bit, with zero
    residuals.  Weighting the pressure by the solid mask instead of the ice
    mask, or the wrong sign of gravity, breaks it.
"""
import math

import numpy as np
import pytest

import ib_transcription as T

STATE3 = ["P", "TXX", "TYY", "TZZ", "TXY", "TXZ", "TYZ", "VX", "VY", "VZ"]
STATE2 = ["P", "TXX", "TYY", "TXY", "VX", "VY"]


def _desc(pkg, kw, ib=True):
    d, _, _ = pkg.build_desc(kw["grid"], kw["boundary_conditions"], kw["gravity"], kw["rheology"],
                             kw["solver_params"], immersed_boundary=ib)
    return d


def _kw(pkg, dims, extent, rheology=None, gravity=None):
    nd = len(dims)
    grid = pkg.UniformGrid(origin=(0.0,) * nd, extent=extent, dims=dims)
    t = pkg.TBC(*([0.0] * nd))
    bcs = {k: (t, t) for k in ("x", "y", "z")[:nd]}
    if gravity is None:
        gravity = dict(zip(("x", "y", "z")[:nd], (0.3, -0.2, 1.0)[:nd] if nd == 3 else (0.3, 1.0)))
    return dict(grid=grid, boundary_conditions=bcs, gravity=gravity,
                rheology=rheology or pkg.LinearViscousRheology(1.0),
                solver_params=pkg.pt_params(grid, r=0.9, re_mech=5 * math.pi))


def _random_masks(rng, shape, nq, near_threshold=True):
    """Volume fractions on and around the 1e-6 threshold, exact 0 and 1,
    and generic values; halos random too (the null-space rule reads them)."""
    vals = np.array([0.0, 1e-7, np.nextafter(1e-6, 0.0), 1e-6, np.nextafter(1e-6, 1.0), 0.25, 0.5, 1.0])
    out = []
    for _ in range(nq):
        if near_threshold:
            pick = rng.integers(0, len(vals), size=shape)
            a = vals[pick]
            gen = rng.random(shape) < 0.4
            a = np.where(gen, rng.uniform(0.05, 1.0, size=shape), a)
        else:
            a = rng.uniform(0.05, 1.0, size=shape)
        out.append(np.asfortranarray(a))
    return out


def _set_masks(ref, na, ns):
    for q, a in enumerate(na):
        ref.set("OM_NA%d" % q, a)
    for q, a in enumerate(ns):
        ref.set("OM_NS%d" % q, a)


def _mask_dict(nd, arrays):
    names = T.MASK3 if nd == 3 else T.MASK2
    return dict(zip(names, arrays))


@pytest.mark.parametrize("dims", [(5, 4, 6), (3, 5, 2), (6, 5)])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_validity_bits_generic_rule(pkg, orc, dims, seed):
    """or_ib_flags == the generic null_spaces.jl rule on random masks."""
    nd = len(dims)
    kw = _kw(pkg, dims, tuple(1.0 for _ in dims))
    ref = orc.OracleModel(_desc(pkg, kw))
    rng = np.random.default_rng(seed)
    shape = tuple(n + 3 for n in dims)
    nq = 8 if nd == 3 else 4
    na = _random_masks(rng, shape, nq)
    ns = _random_masks(rng, shape, nq)
    _set_masks(ref, na, ns)
    got = ref.ib_flags()
    g = T.Grid(dims, ref.m.dx[:nd])
    want = T.flags(g, _mask_dict(nd, na), _mask_dict(nd, ns))
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]
    # the test has teeth: both valid and invalid locations of every kind occur
    bits = range(7) if nd == 3 else (0, 1, 4, 5)
    inner = g.view(want, g.zero_off())
    for b in bits:
        on = (inner >> b) & 1
        assert 0 < on.sum() < on.size, b


def _random_state(rng, shape, names, amp=1.0):
    return {k: np.asfortranarray(amp * rng.standard_normal(shape)) for k in names}


@pytest.mark.parametrize("dims", [(6, 5, 4), (7, 6)])
@pytest.mark.parametrize("seed", [4, 5])
def test_ib_kernels_generic_transcription(pkg, orc, dims, seed):
    """The oracle's IB update_σ!, update_V!, compute_τ! and
    compute_residuals! equal the generic transcription bit for bit."""
    nd = len(dims)
    rheo = pkg.ViscoplasticRegularisation(pkg.PowerLawRheology(0.7, 3), 20.0)
    kw = _kw(pkg, dims, (1.0, 0.8, 0.6)[:nd], rheology=rheo)
    desc = _desc(pkg, kw)
    ref = orc.OracleModel(desc)
    rng = np.random.default_rng(seed)
    shape = tuple(n + 3 for n in dims)
    nq = 8 if nd == 3 else 4
    na = _random_masks(rng, shape, nq)
    ns = _random_masks(rng, shape, nq)
    _set_masks(ref, na, ns)
    names = STATE3 if nd == 3 else STATE2
    st = _random_state(rng, shape, names)
    st["ETA"] = np.asfortranarray(rng.uniform(0.5, 2.0, size=shape))
    st["ETA_NEXT"] = np.asfortranarray(rng.uniform(0.5, 2.0, size=shape))
    for k, a in st.items():
        ref.set(k, a)
    g = T.Grid(dims, ref.m.dx[:nd])
    NA, NS = _mask_dict(nd, na), _mask_dict(nd, ns)
    dt = dict(zip(["TXX", "TYY", "TZZ", "TXY", "TXZ", "TYZ"] if nd == 3 else ["TXX", "TYY", "TXY"],
                  list(desc.dtau_tau)[:6 if nd == 3 else 3]))
    grav = list(desc.gravity)[:nd]
    inv_reg = 1.0 / desc.rheology.eta_reg
    A, n = desc.rheology.A, desc.rheology.n

    def rheology(tII):
        assert n == 3.0
        return 1.0 / (1.0 / (0.5 / (A * (tII * tII))) + inv_reg)

    stress = names[:7] if nd == 3 else names[:4]
    s = {k: v.copy(order="F") for k, v in st.items()}
    # update_σ!
    ref.update_sigma()
    T.update_sigma(g, s, NA, NS, dt, desc.dtau_P)
    for k in stress:
        assert np.array_equal(ref.get(k), s[k]), ("sigma", k)
    # update_V! reads the new stress
    ref.update_V()
    T.update_V(g, s, NA, NS, list(desc.dtau_V)[:nd], grav, rheology)
    for k in names[-nd:] + ["ETA_NEXT"]:
        assert np.array_equal(ref.get(k), s[k]), ("velocity", k)
    # compute_residuals! (traction everywhere: no residual boundary rows)
    for k in ("RP", "RVX", "RVY") + (("RVZ",) if nd == 3 else ()):
        s[k] = np.zeros(shape, order="F")
    ref.residuals_only()
    T.compute_residuals(g, s, NA, NS, grav)
    for k in ("RP", "RVX", "RVY") + (("RVZ",) if nd == 3 else ()):
        assert np.array_equal(ref.get(k), s[k]), ("residual", k)
    # compute_τ!
    ref.compute_tau()
    T.compute_tau(g, s, NA, NS)
    for k in stress[1:]:
        assert np.array_equal(ref.get(k), s[k]), ("tau", k)
    # teeth: the masked kernels touched both valid and invalid locations
    assert np.count_nonzero(g.view(s["P"], g.zero_off())) not in (0, g.view(s["P"], g.zero_off()).size)


def test_transcription_catches_a_swapped_neighbour(pkg, orc):
    """A planted mistake (the x-neighbours of a Center location read one cell
    to the left) makes the generic rule disagree with the oracle."""
    dims = (5, 4, 6)
    kw = _kw(pkg, dims, (1.0, 1.0, 1.0))
    ref = orc.OracleModel(_desc(pkg, kw))
    rng = np.random.default_rng(9)
    shape = tuple(n + 3 for n in dims)
    na, ns = _random_masks(rng, shape, 8), _random_masks(rng, shape, 8)
    _set_masks(ref, na, ns)
    g = T.Grid(dims, ref.m.dx[:3])
    save = T.left_off, T.right_off
    try:
        T.left_off = lambda loc, d: (-1 if d == 0 else 0) if loc[d] == "c" else -1
        T.right_off = lambda loc, d: (0 if d == 0 else 1) if loc[d] == "c" else 0
        bad = T.flags(g, _mask_dict(3, na), _mask_dict(3, ns))
    finally:
        T.left_off, T.right_off = save
    assert not np.array_equal(ref.ib_flags(), bad)


@pytest.mark.parametrize("dims", [(6, 5, 4), (7, 5)])
def test_all_fluid_masks_reduce_to_plain_kernels(pkg, orc, dims):
    """With every fraction 1 (halos included) the IB model is the FullStokes
    model bit for bit: iterations, residuals and compute_τ!."""
    nd = len(dims)
    rheo = pkg.Relaxation("eta", pkg.ViscoplasticRegularisation(pkg.PowerLawRheology(0.7, 3), 20.0), 0.3)
    kw = _kw(pkg, dims, (1.0, 0.8, 0.6)[:nd], rheology=rheo)
    ib = orc.OracleModel(_desc(pkg, kw, ib=True))
    pl = orc.OracleModel(_desc(pkg, kw, ib=False))
    shape = tuple(n + 3 for n in dims)
    nq = 8 if nd == 3 else 4
    ones = [np.ones(shape, order="F") for _ in range(nq)]
    _set_masks(ib, ones, ones)
    rng = np.random.default_rng(21)
    names = STATE3 if nd == 3 else STATE2
    st = _random_state(rng, shape, names, amp=1e-2)
    eta = np.asfortranarray(rng.uniform(0.5, 2.0, size=shape))
    for m in (ib, pl):
        for k, a in st.items():
            m.set(k, a)
        m.set("ETA", eta)
        m.set("ETA_NEXT", eta)
        m.apply_bc("eta")
    assert np.all(ib.ib_flags()[tuple(slice(1, n + 2) for n in dims)] == (0x7F if nd == 3 else 0x33))
    ib.advance(25)
    pl.advance(25)
    for k in names + ["ETA", "ETA_NEXT"]:
        assert np.array_equal(ib.get(k), pl.get(k)), k
    assert np.array_equal(ib.compute_residuals(), pl.compute_residuals())
    for k in ("RP", "RVX", "RVY") + (("RVZ",) if nd == 3 else ()):
        assert np.array_equal(ib.get(k), pl.get(k)), k
    ib.compute_tau()
    pl.compute_tau()
    for k in names[1:]:
        assert np.array_equal(ib.get(k), pl.get(k)), k
    # teeth: the state moved
    assert not np.array_equal(ib.get("VX"), st["VX"])


def slab_masks(dims, dx, zb_frac, kb):
    """Solid mask ns of a flat bed at z_b = (kb - 1 + zb_frac) dz (Chmy cell kb,
    1-based) in the last dimension, the exact fraction of every staggered
    dual box above it; ice mask na = 1.  Halos are 0."""
    nd = len(dims)
    shape = tuple(n + 3 for n in dims)
    dz = dx[-1]
    zb = (kb - 1 + zb_frac) * dz
    names = T.MASK3 if nd == 3 else T.MASK2
    na, ns = [], []
    idx = np.arange(shape[-1])
    for loc in names:
        interior = tuple(slice(1, n + 1 + (1 if c == "v" else 0)) for n, c in zip(dims, loc))
        a = np.zeros(shape, order="F")
        a[interior] = 1.0
        na.append(a)
        if loc[-1] == "c":
            lo, hi = (idx - 1) * dz, idx * dz          # cell idx spans [(idx-1) dz, idx dz]
        else:
            lo, hi = (idx - 1.5) * dz, (idx - 0.5) * dz  # dual box of vertex idx
        frac = np.clip((hi - zb) / dz, 0.0, 1.0)
        b = np.zeros(shape, order="F")
        b[interior] = np.broadcast_to(frac.reshape((1,) * (nd - 1) + (-1,)), shape)[interior]
        ns.append(b)
    return na, ns


def slab_state(dims, dx, rho_g, flags):
    """P = rho g (z_top - z) at valid pressure points (0 elsewhere), no stress, no flow."""
    nd = len(dims)
    shape = tuple(n + 3 for n in dims)
    dz = dx[-1]
    ztop = dims[-1] * dz
    zc = (np.arange(shape[-1]) - 0.5) * dz
    P = np.broadcast_to((rho_g * (ztop - zc)).reshape((1,) * (nd - 1) + (-1,)), shape).copy(order="F")
    P[(flags & 1) == 0] = 0.0
    return P


@pytest.mark.parametrize("dims", [(6, 5, 8), (7, 8)])
@pytest.mark.parametrize("zb_frac", [0.25, 0.5])
def test_hydrostatic_slab_is_an_exact_rest_state(pkg, orc, dims, zb_frac):
    nd = len(dims)
    extent = (0.75, 0.625, 1.0)[:nd - 1] + (1.0,)       # dz = 1/8: exact arithmetic
    grav = dict(zip(("x", "y", "z")[:nd], [0.0] * (nd - 1) + [1.0]))
    kw = _kw(pkg, dims, extent, gravity=grav)
    ref = orc.OracleModel(_desc(pkg, kw))
    dx = list(ref.m.dx[:nd])
    na, ns = slab_masks(dims, dx, zb_frac, kb=3)
    _set_masks(ref, na, ns)
    flags = ref.ib_flags()
    P = slab_state(dims, dx, 1.0, flags)
    ref.set("P", P)
    ref.set("ETA", 1.0)
    ref.set("ETA_NEXT", 1.0)
    ref.apply_bc("stress")
    before = {k: ref.get(k).copy() for k in (STATE3 if nd == 3 else STATE2)}
    r = ref.compute_residuals()
    assert np.all(r == 0.0), r
    ref.advance(5)
    for k, a in before.items():
        assert np.array_equal(ref.get(k), a), k
    # the column is really resolved: some pressure points lie on the cut cell
    # and some velocity faces above the bed are valid
    kz = tuple([slice(1, n + 1) for n in dims[:-1]] + [2])  # the cell below the cut cell
    assert np.all(flags[kz] & 1 == 0)
    assert np.any(flags & (1 << (6 if nd == 3 else 5)))


def test_slab_catches_a_swapped_mask(pkg, orc):
    """Weighting the pressure by the solid mask instead of the ice mask (the
    swap a shared misreading would make) breaks the rest state."""
    dims = (6, 5, 8)
    kw = _kw(pkg, dims, (0.75, 0.625, 1.0), gravity=dict(x=0.0, y=0.0, z=1.0))
    ref = orc.OracleModel(_desc(pkg, kw))
    dx = list(ref.m.dx[:3])
    na, ns = slab_masks(dims, dx, 0.3, kb=3)
    _set_masks(ref, ns, ns)      # planted: ice mask replaced by the solid mask
    flags = ref.ib_flags()
    ref.set("P", slab_state(dims, dx, 1.0, flags))
    ref.set("ETA", 1.0)
    ref.set("ETA_NEXT", 1.0)
    assert max(ref.compute_residuals()) > 0.1
