"""HDF5 + XDMF3 writers (src/Writers.jl) against the reference's writer test.

test/test_writers.jl:36-78 writes two Center fields of a (4, 5, 6) grid on
origin (-0.4, -0.5, 0.0), extent (1.0, 1.1, 1.2), reads the datasets back
with h5read and compares the XDMF3 document with a literal; that literal is
the fixture golden/xdmf3_test_writers.xml (made by golden/make_xdmf_fixture.py).
HDF5.jl stores a Julia (nx, ny, nz) array as C-order dims (nz, ny, nx); the
layout test checks that index map element by element with distinct values.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def w(pkg):
    try:
        pkg._hdf5.lib()
    except RuntimeError as e:
        pytest.skip(str(e))
    return pkg.writers


@pytest.fixture
def grid(pkg):
    return pkg.UniformGrid(origin=(-0.4, -0.5, 0.0), extent=(1.0, 1.1, 1.2), dims=(4, 5, 6))


def test_xdmf3_matches_reference_literal(pkg, grid, tmp_path):
    fields = {"Fa": np.ones((4, 5, 6)), "Fb": np.full((4, 5, 6), 2.0)}
    path = tmp_path / "test.xdmf3"
    pkg.writers.write_xdmf(grid, str(path), fields, ("test.h5",))
    want = open(os.path.join(HERE, "golden", "xdmf3_test_writers.xml")).read()
    assert path.read_text() == want


def test_xdmf3_time_series_and_errors(pkg, grid):
    text = pkg.writers.generate_xdmf((4, 5, 6), (0.25, 0.22, 0.2), (0.0, 0.0, 0.0), ["P"], ["a.h5", "b.h5"],
                                     [0.0, 1.5])
    assert text.count('<Grid GridType="Uniform">') == 2
    assert '<Time Value="1.5"/>' in text and "b.h5:/P" in text
    with pytest.raises(ValueError):
        pkg.writers.generate_xdmf((4, 5, 6), (1, 1, 1), (0, 0, 0), ["P"], ["a.hdf"], 0.0)


def test_julia_float_text(pkg):
    jf = pkg.writers.julia_float
    assert jf(0.0) == "0.0"
    assert jf(-0.275) == "-0.275"
    assert jf(0.09999999999999999) == "0.09999999999999999"
    assert jf(0.0001) == "0.0001"
    assert jf(123456.0) == "123456.0"
    assert jf(1e-5) == "1.0e-5"
    assert jf(1e6) == "1.0e6"
    assert jf(-2.5e-7) == "-2.5e-7"
    for x in (0.1, 1.0 / 3.0, 7e-3, 12345.678, 3.0e8, 6.02e23):
        assert float(jf(x)) == x  # round trip


def test_h5_round_trip_and_layout(pkg, w, grid, tmp_path):
    rng = np.random.default_rng(3)
    fa = rng.standard_normal((4, 5, 6))
    fb = np.arange(4 * 5 * 6, dtype=np.float64).reshape((4, 5, 6), order="F")  # value = Julia linear index
    path = str(tmp_path / "test.h5")
    w.write_h5(grid, path, {"Fa": fa, "Fb": fb})
    assert np.array_equal(w.read_h5(path, "Fa"), fa)
    assert np.array_equal(w.read_h5(path, "Fb"), fb)
    with pkg._hdf5.File(path, "r") as io:
        raw = io.read("/Fb")
    assert raw.shape == (6, 5, 4)
    for i, j, k in ((0, 0, 0), (3, 0, 0), (0, 4, 0), (0, 0, 5), (2, 3, 4)):
        assert raw[k, j, i] == i + 4 * j + 20 * k


def test_h5_rejects_other_sizes(w, grid, tmp_path):
    with pytest.raises(ValueError):
        w.write_h5(grid, str(tmp_path / "x.h5"), {"V": np.zeros((5, 5, 6))})


@pytest.mark.gpu
def test_h5_device_fields(pkg, tmp_path):
    print("case: write_h5 / write_xdmf of device fields, 3D manufactured 20x18x16", flush=True)
    try:
        pkg._hdf5.lib()
    except RuntimeError as e:
        pytest.skip(str(e))
    kw, cad = pkg.manufactured_3d((20, 18, 16))
    m = pkg.IsothermalFullStokesModel(**kw)
    m.viscosity.eta.set(1.0)
    m.viscosity.eta_next.set(1.0)
    m.advance(25)
    fields = {"P": m.field("P"), "TXX": m.field("TXX"), "ETA": m.field("ETA")}
    path = str(tmp_path / "state.h5")
    pkg.writers.write_h5(m.grid, path, fields)
    pkg.writers.write_xdmf(m.grid, str(tmp_path / "state.xdmf3"), fields, ("state.h5",))
    for name, f in fields.items():
        assert np.array_equal(pkg.writers.read_h5(path, name), f.interior()), name
    m.close()
