"""The committed Julia binding (julia/FastIceHIP/src/FastIceHIP.jl) against the
C header (include/fastice_hip.h), without Julia: every `ccall` names a
declared entry point with the same number of arguments and compatible
argument and return types, the Julia records mirror fi_rheology and
fi_model_desc field by field, the enum values agree, and the reference's
entry points are exported with the reference's argument lists
(compute_levelsets.jl:21-35, volume_fractions_kernels.jl:101-110,
Isothermal.jl:60-125)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "julia", "FastIceHIP", "src", "FastIceHIP.jl")
HDR = os.path.join(ROOT, "include", "fastice_hip.h")


def _header():
    text = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    return text


def c_canon(t):
    """Canonical form of a C parameter or return type."""
    t = t.strip()
    arr = re.search(r"\[\d*\]\s*$", t)
    t = re.sub(r"\b(const)\b", "", t)
    t = re.sub(r"\s+[A-Za-z_]\w*\s*(\[\d*\])?$", lambda m: "*" if m.group(1) else "", t) if re.search(r"[\w\*]\s+[A-Za-z_]\w*(\s*\[\d*\])?$", t) else t
    t = t.replace(" ", "")
    if arr and not t.endswith("*"):
        t += "*"
    base = {"int": "int", "int32_t": "int", "double": "double", "size_t": "size_t", "void": "void",
            "fi_model_t": "handle", "fi_batch_t": "handle", "fi_heat_t": "handle", "char*": "cstr"}
    if t in base:
        return base[t]
    if t.endswith("*"):
        inner = t[:-1]
        return {"double": "ptr:double", "int": "ptr:int", "uint8_t": "ptr:uint8", "size_t": "ptr:size_t",
                "longlong": "ptr:longlong", "fi_model_t": "ptr:handle", "fi_batch_t": "ptr:handle",
                "fi_model_desc": "ptr:desc", "char": "cstr"}.get(inner, "ptr:" + inner)
    return t


def header_functions():
    text = _header()
    out = {}
    for m in re.finditer(r"^\s*((?:const\s+)?[A-Za-z_][\w\s\*]*?)\b(fi_\w+)\s*\(([^;]*?)\)\s*;", text, flags=re.M | re.S):
        ret, name, args = m.group(1), m.group(2), m.group(3)
        params = [] if args.strip() in ("", "void") else [a for a in args.split(",")]
        out[name] = (c_canon(ret + " x").replace("ptr:char", "cstr"), [c_canon(p) for p in params])
    return out


JL_TYPES = {"Cint": "int", "Cdouble": "double", "Csize_t": "size_t", "Ptr{Cvoid}": "handle", "Cstring": "cstr",
            "Ptr{Cdouble}": "ptr:double", "Ref{FiModelDesc}": "ptr:desc", "Ref{Ptr{Cvoid}}": "ptr:handle",
            "Ptr{UInt8}": "ptr:uint8", "Ptr{Cint}": "ptr:int"}


def julia_ccalls():
    text = open(JL).read()
    calls = []
    for m in re.finditer(r"ccall\(\(:(fi_\w+),\s*lib\),\s*([\w{}]+),\s*\(([^()]*)\)", text):
        name, ret, args = m.group(1), m.group(2), m.group(3)
        types = [a.strip() for a in args.split(",") if a.strip()]
        calls.append((name, ret, types))
    return calls


def test_every_ccall_matches_the_header():
    hdr = header_functions()
    calls = julia_ccalls()
    assert len(calls) >= 12
    for name, ret, types in calls:
        assert name in hdr, name
        cret, cparams = hdr[name]
        assert JL_TYPES[ret] == cret, (name, ret, cret)
        assert len(types) == len(cparams), (name, types, cparams)
        for jt, ct in zip(types, cparams):
            assert JL_TYPES[jt] == ct, (name, jt, ct)


def _c_struct(name):
    text = _header()
    m = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, flags=re.S)
    fields = []
    for line in m.group(1).split(";"):
        line = line.strip()
        if not line:
            continue
        f = re.match(r"(\w+)\s+(\w+)((?:\[\d+\])*)$", line)
        typ, fname, dims = f.group(1), f.group(2), [int(x) for x in re.findall(r"\[(\d+)\]", f.group(3))]
        n = 1
        for d in dims:
            n *= d
        fields.append((fname, {"int32_t": "i32", "double": "f64", "fi_rheology": "rheology"}[typ], n))
    return fields


def _jl_struct(name):
    text = open(JL).read()
    m = re.search(r"struct %s\n(.*?)\nend" % name, text, flags=re.S)
    fields = []
    for line in m.group(1).splitlines():
        f = re.match(r"\s*(\w+)::(.+)$", line)
        if not f:
            continue
        fname, typ = f.group(1), f.group(2).strip()
        n = 1
        while True:
            t = re.match(r"NTuple\{(\d+),(.+)\}$", typ)
            if not t:
                break
            n *= int(t.group(1))
            typ = t.group(2)
        fields.append((fname, {"Int32": "i32", "Float64": "f64", "FiRheology": "rheology"}[typ], n))
    return fields


def test_records_mirror_the_c_structs():
    assert _jl_struct("FiRheology") == _c_struct("fi_rheology")
    assert _jl_struct("FiModelDesc") == _c_struct("fi_model_desc")


def test_enum_values_agree():
    hdr = _header()
    jl = open(JL).read()
    for name in ("FI_BC_TRACTION", "FI_BC_VELOCITY", "FI_BC_SLIP", "FI_RHEO_LINEAR_CONST", "FI_RHEO_LINEAR_FIELD",
                 "FI_RHEO_POWERLAW", "FI_RELAX_NONE", "FI_RELAX_LIN", "FI_RELAX_LOG", "FI_RELAX_LOG10",
                 "FI_LAYOUT_PARENT", "FI_LAYOUT_INTERIOR", "FI_BCGROUP_STRESS", "FI_BCGROUP_VELOCITY",
                 "FI_BCGROUP_ETA", "FI_PHASE_NA", "FI_PHASE_NS"):
        c = int(re.search(r"\b%s\s*=\s*(\d+)" % name, hdr).group(1))
        # Julia: a tuple assignment "const A, B = Int32(0), Int32(1)"
        line = next(l for l in jl.splitlines() if re.search(r"\b%s\b" % name, l) and l.startswith("const"))
        lhs, rhs = line[len("const"):].split("=")
        names = [s.strip() for s in lhs.split(",")]
        vals = [int(re.search(r"Int32\((\d+)\)", v).group(1)) for v in rhs.split(",")]
        assert dict(zip(names, vals))[name] == c, name
    # field ids: the header's enum order
    ids = dict(re.findall(r"(\w+)=(\d+)", re.search(r"const FIELD_IDS = \((.*?)\)\n", jl, flags=re.S).group(1)))
    order = ["P", "TXX", "TYY", "TZZ", "TXY", "TXZ", "TYZ", "VX", "VY", "VZ", "ETA", "ETA_NEXT", "RP", "RVX", "RVY",
             "RVZ", "GX", "GY", "GZ", "RHEO", "PSI_NA", "PSI_NS", "OM_NA0"]
    for q, n in enumerate(order):
        assert int(ids[n]) == q, n
    assert int(ids["OM_NS0"]) == 30


def test_reference_signatures_are_kept():
    jl = open(JL).read()
    want = {
        "compute_levelset_from_dem!": ["arch", "launch", "Ψ", "dem", "dem_grid2D", "grid", "R"],
        "invert_levelset!": ["arch", "launch", "Ψ", "grid"],
        "ω_from_ψ!": ["arch", "launch", "ω", "ψ", "grid"],
        "advance_iteration!": ["m", "t", "Δt"],
    }
    for fn, args in want.items():
        m = re.search(r"^(?:function\s+)?%s\(([^)]*)\)" % re.escape(fn), jl, flags=re.M)
        assert m, fn
        got = [re.split(r"[:=]", a.strip())[0].strip() for a in m.group(1).split(",")]
        assert got == args, (fn, got)
    for kw in ("arch", "grid", "boundary_conditions", "gravity", "rheology", "solver_params", "outer_width"):
        assert re.search(r"IsothermalFullStokesModel\(;[^)]*\b%s\b" % kw, jl), kw
    assert re.search(r"IsothermalImmersedBoundaryFullStokesModel\(;[^)]*\blevel_sets\b", jl)
    assert "function lower_desc(" in jl
