"""Independent numpy transcription of the immersed-boundary kernels, written
from the reference text with a generic location algebra, for pinning the
oracle's hand-indexed C restatement.

Nothing here is derived from the oracle: every neighbour index comes from
the rule the reference states once for all locations, not from a per-location
case table.

  null_spaces.jl:1-11   flipped(loc, D) flips dimension D of a location;
                        left/right(w, loc, D, I) read w at the flipped
                        location, at the index il/ir gives: a Center point's
                        Vertex neighbours are I and I+1, a Vertex point's
                        Center neighbours are I-1 and I.
  null_spaces.jl:13-37  isnullspace: some neighbour below 1e-6 across any
                        dimension (generic), across the two Vertex dimensions
                        only for the 3D edge locations (vvc, vcv, cvv).
  null_spaces.jl:41-43  isempty: w(loc) < 1e-6; isvalid = !(null || empty).
  IB kernels_3d.jl:3-143  update_σ!, update_V!, compute_τ!,
                        compute_residuals!, with the masked difference
                        d(f, w) = (w_f f)[right] - (w_f f)[left], times 1/Δ,
                        w_f the mask at f's location, and lerp nested with x
                        innermost.  2D restates the same kernels in two
                        dimensions, gravity unweighted as in 3D.

Arrays are parent-layout numpy arrays (n+3 per dimension, Chmy index =
array index); the kernels cover Chmy indices [1, n+1] in every dimension.
Each expression keeps the left-to-right evaluation order of the cited
Julia line, so results are comparable bit for bit.
"""
import numpy as np

TOL = 1e-6

# field locations (x, y[, z]) as 'c'/'v' strings
LOC3 = dict(P="ccc", TXX="ccc", TYY="ccc", TZZ="ccc", TXY="vvc", TXZ="vcv", TYZ="cvv", VX="vcc", VY="cvc",
            VZ="ccv", ETA="ccc")
LOC2 = dict(P="cc", TXX="cc", TYY="cc", TXY="vv", VX="vc", VY="cv", ETA="cc")
MASK3 = ["ccc", "vvv", "vcc", "cvc", "ccv", "vvc", "vcv", "cvv"]
MASK2 = ["cc", "vv", "vc", "cv"]


class Grid:
    def __init__(self, dims, spacing):
        self.dims = tuple(dims)
        self.nd = len(dims)
        self.inv = tuple(1.0 / s for s in spacing)

    def view(self, a, off):
        """a at I + off for every I of the work range."""
        return a[tuple(slice(1 + o, n + 2 + o) for o, n in zip(off, self.dims))]

    def zero_off(self):
        return (0,) * self.nd


def flip(c):
    return "v" if c == "c" else "c"


def flipped(loc, d):
    return loc[:d] + flip(loc[d]) + loc[d + 1:]


def left_off(loc, d):
    # il(flip(loc[d]), loc[d]): Center point -> Vertex neighbour I; Vertex -> Center I-1
    return 0 if loc[d] == "c" else -1


def right_off(loc, d):
    return 1 if loc[d] == "c" else 0


def _shift(off, d, by):
    o = list(off)
    o[d] += by
    return tuple(o)


def side(g, w, loc, d, which):
    off = _shift(g.zero_off(), d, left_off(loc, d) if which == "l" else right_off(loc, d))
    return g.view(w[flipped(loc, d)], off)


def isnullspace(g, w, loc):
    if g.nd == 3 and loc.count("v") == 2:
        dims = [d for d in range(3) if loc[d] == "v"]  # null_spaces.jl:24-37
    else:
        dims = range(g.nd)  # null_spaces.jl:13-22
    out = np.zeros(g.view(w[loc], g.zero_off()).shape, dtype=bool)
    for d in dims:
        out |= (side(g, w, loc, d, "l") < TOL) | (side(g, w, loc, d, "r") < TOL)
    return out


def isempty(g, w, loc):
    return g.view(w[loc], g.zero_off()) < TOL


def isvalid(g, w, loc):
    return ~(isnullspace(g, w, loc) | isempty(g, w, loc))


def flags(g, na, ns):
    """Validity bits in the layout of fi_ib_validity / or_ib_flags."""
    out = np.zeros(tuple(n + 3 for n in g.dims), dtype=np.uint8, order="F")
    if g.nd == 3:
        sig = ["ccc", "vvc", "vcv", "cvv"]
        vel = ["vcc", "cvc", "ccv"]
    else:
        sig = ["cc", "vv"]
        vel = ["vc", "cv"]
    acc = np.zeros(g.view(out, g.zero_off()).shape, dtype=np.uint8)
    for q, loc in enumerate(sig):
        acc |= ((isvalid(g, na, loc) & ~isempty(g, ns, loc)).astype(np.uint8) << q)
    for q, loc in enumerate(vel):
        acc |= ((isvalid(g, ns, loc) & ~isempty(g, na, loc)).astype(np.uint8) << (4 + q))
    g.view(out, g.zero_off())[...] = acc
    return out


def dpart(g, f, floc, d, w=None):
    """Masked (or plain) derivative of f (at floc) along d at the flipped location."""
    off = g.zero_off()
    if floc[d] == "v":
        lo, hi = off, _shift(off, d, 1)
    else:
        lo, hi = _shift(off, d, -1), off
    if w is None:
        return (g.view(f, hi) - g.view(f, lo)) * g.inv[d]
    wf = w[floc]
    return (g.view(wf, hi) * g.view(f, hi) - g.view(wf, lo) * g.view(f, lo)) * g.inv[d]


def lerp(g, f, floc, tloc):
    dims = [d for d in range(g.nd) if floc[d] != tloc[d]]

    def rec(ds, off):
        if not ds:
            return g.view(f, off)
        d = ds[-1]
        lo, hi = (-1, 0) if floc[d] == "c" else (0, 1)
        return 0.5 * (rec(ds[:-1], _shift(off, d, lo)) + rec(ds[:-1], _shift(off, d, hi)))

    return rec(dims, g.zero_off())


def at(g, w, loc):
    return g.view(w[loc], g.zero_off())


def _put(g, s, name, val, valid):
    g.view(s[name], g.zero_off())[...] = np.where(valid, val, 0.0)


def update_sigma(g, s, na, ns, dt, dPr):
    """IB update_σ! (IB kernels_3d.jl:3-44), in place on the state dict s."""
    L = LOC3 if g.nd == 3 else LOC2
    eta = s["ETA"]
    V = [s["VX"], s["VY"]] + ([s["VZ"]] if g.nd == 3 else [])
    VL = [L["VX"], L["VY"]] + ([L["VZ"]] if g.nd == 3 else [])
    old = {k: g.view(s[k], g.zero_off()).copy() for k in L if k.startswith("T") or k == "P"}
    with np.errstate(all="ignore"):
        cl = L["P"]
        ok = isvalid(g, na, cl) & ~isempty(g, ns, cl)
        e = [dpart(g, V[d], VL[d], d, ns) for d in range(g.nd)]
        div = e[0] + e[1] + e[2] if g.nd == 3 else e[0] + e[1]
        _put(g, s, "P", old["P"] - div * lerp(g, eta, L["ETA"], cl) * dPr, ok)
        diag = ["TXX", "TYY", "TZZ"][:g.nd]
        for d, k in enumerate(diag):
            t = old[k]
            _put(g, s, k, t - (t - 2.0 * lerp(g, eta, L["ETA"], L[k]) * (e[d] - div / 3.0)) * dt[k], ok)
        pairs = [("TXY", 0, 1), ("TXZ", 0, 2), ("TYZ", 1, 2)] if g.nd == 3 else [("TXY", 0, 1)]
        for k, a, b in pairs:
            loc = L[k]
            ok = isvalid(g, na, loc) & ~isempty(g, ns, loc)
            # 0.5 * (d_a V_b + d_b V_a)
            exy = 0.5 * (dpart(g, V[b], VL[b], a, ns) + dpart(g, V[a], VL[a], b, ns))
            t = old[k]
            _put(g, s, k, t - (t - 2.0 * lerp(g, eta, L["ETA"], loc) * exy) * dt[k], ok)


def _momentum(g, s, na, d, grav):
    """-d_d(P) + d_d(tau_dd) + the other two shear derivatives - rho g (IB kernels_3d.jl:51,57,63)."""
    L = LOC3 if g.nd == 3 else LOC2
    if g.nd == 3:
        shear = {0: [("TXY", 1), ("TXZ", 2)], 1: [("TXY", 0), ("TYZ", 2)], 2: [("TXZ", 0), ("TYZ", 1)]}[d]
    else:
        shear = {0: [("TXY", 1)], 1: [("TXY", 0)]}[d]
    diag = ["TXX", "TYY", "TZZ"][d]
    acc = -dpart(g, s["P"], L["P"], d, na) + dpart(g, s[diag], L[diag], d, na)
    for k, dd in shear:
        acc = acc + dpart(g, s[k], L[k], dd, na)
    gv = grav[d]
    gv = gv if np.isscalar(gv) else g.view(gv, g.zero_off())
    return acc - gv


def update_V(g, s, na, ns, dV, grav, rheology):
    """IB update_V! (IB kernels_3d.jl:46-74): velocity then eta_next."""
    L = LOC3 if g.nd == 3 else LOC2
    names = ["VX", "VY", "VZ"][:g.nd]
    with np.errstate(all="ignore"):
        for d, k in enumerate(names):
            loc = L[k]
            ok = isvalid(g, ns, loc) & ~isempty(g, na, loc)
            v = g.view(s[k], g.zero_off()).copy()
            new = v + _momentum(g, s, na, d, grav) / lerp(g, s["ETA"], L["ETA"], loc) / at(g, ns, loc) * dV[d]
            _put(g, s, k, new, ok)
        c = L["ETA"]
        if g.nd == 3:
            t2 = 0.5 * (at_f(g, s["TXX"]) ** 2 + at_f(g, s["TYY"]) ** 2 + at_f(g, s["TZZ"]) ** 2) + \
                lerp(g, s["TXY"], L["TXY"], c) ** 2 + lerp(g, s["TXZ"], L["TXZ"], c) ** 2 + \
                lerp(g, s["TYZ"], L["TYZ"], c) ** 2
        else:
            t2 = 0.5 * (at_f(g, s["TXX"]) ** 2 + at_f(g, s["TYY"]) ** 2) + lerp(g, s["TXY"], L["TXY"], c) ** 2
        g.view(s["ETA_NEXT"], g.zero_off())[...] = rheology(np.sqrt(t2))


def at_f(g, f):
    return g.view(f, g.zero_off())


def compute_tau(g, s, na, ns):
    """IB compute_τ! (IB kernels_3d.jl:78-116; the undefined `Pr` of :81 read as
    the Center location of τ.xx)."""
    L = LOC3 if g.nd == 3 else LOC2
    eta = s["ETA"]
    V = [s["VX"], s["VY"]] + ([s["VZ"]] if g.nd == 3 else [])
    VL = [L["VX"], L["VY"]] + ([L["VZ"]] if g.nd == 3 else [])
    with np.errstate(all="ignore"):
        cl = L["TXX"]
        ok = isvalid(g, na, cl) & ~isempty(g, ns, cl)
        e = [dpart(g, V[d], VL[d], d, ns) for d in range(g.nd)]
        div = e[0] + e[1] + e[2] if g.nd == 3 else e[0] + e[1]
        for d, k in enumerate(["TXX", "TYY", "TZZ"][:g.nd]):
            _put(g, s, k, 2.0 * lerp(g, eta, L["ETA"], L[k]) * (e[d] - div / 3.0), ok)
        pairs = [("TXY", 0, 1), ("TXZ", 0, 2), ("TYZ", 1, 2)] if g.nd == 3 else [("TXY", 0, 1)]
        for k, a, b in pairs:
            loc = L[k]
            ok = isvalid(g, na, loc) & ~isempty(g, ns, loc)
            exy = 0.5 * (dpart(g, V[b], VL[b], a, ns) + dpart(g, V[a], VL[a], b, ns))
            _put(g, s, k, 2.0 * lerp(g, eta, L["ETA"], loc) * exy, ok)


def compute_residuals(g, s, na, ns, grav):
    """IB compute_residuals! (IB kernels_3d.jl:118-143)."""
    L = LOC3 if g.nd == 3 else LOC2
    V = [s["VX"], s["VY"]] + ([s["VZ"]] if g.nd == 3 else [])
    VL = [L["VX"], L["VY"]] + ([L["VZ"]] if g.nd == 3 else [])
    with np.errstate(all="ignore"):
        cl = L["P"]
        ok = isvalid(g, na, cl) & ~isempty(g, ns, cl)
        e = [dpart(g, V[d], VL[d], d, ns) for d in range(g.nd)]
        _put(g, s, "RP", e[0] + e[1] + e[2] if g.nd == 3 else e[0] + e[1], ok)
        for d, k in enumerate(["VX", "VY", "VZ"][:g.nd]):
            loc = L[k]
            ok = isvalid(g, ns, loc) & ~isempty(g, na, loc)
            _put(g, s, "R" + k, _momentum(g, s, na, d, grav), ok)
