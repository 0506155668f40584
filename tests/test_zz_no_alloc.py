"""fi_field_download_noalloc (the entry point that allocates nothing) against
fi_field_download with FI_LAYOUT_INTERIOR: identical bytes on two calls in a
row, the second with a smaller field, and no change in free device memory."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(pkg, model, name):
    A = pkg._abi
    f = model.field(name)
    out = np.full(f.interior_shape(), np.nan, order="F")
    A.check(A.lib().fi_field_download_noalloc(model._h, A.F[name], A.dptr(out), out.size))
    return out


def _free(pkg):
    free, total = C.c_size_t(0), C.c_size_t(0)
    pkg._abi.check(pkg._abi.lib().fi_mem_info(C.byref(free), C.byref(total)))
    return free.value


def test_noalloc_download_equals_download(pkg):
    print("case: no-alloc interior download, 3D 170x150x110 then 2D 40x23", flush=True)
    rng = np.random.default_rng(3)
    # first call: an x-face field of 171*150*110 doubles (22.6 MB), more than
    # the 16 MiB staging area, so it moves in two chunks
    kw3, _ = pkg.manufactured_3d((170, 150, 110))
    big = pkg.IsothermalFullStokesModel(**kw3)
    vx = np.asfortranarray(rng.standard_normal(big.field("VX").interior_shape()))
    big.field("VX").set(vx)
    big.advance(3)
    want3 = big.field("VX").interior()
    assert not np.array_equal(want3, vx)
    kw2, _ = pkg.manufactured_2d((40, 23))
    small = pkg.IsothermalFullStokesModel(**kw2)
    small.field("VY").set(np.asfortranarray(rng.standard_normal(small.field("VY").interior_shape())))
    small.advance(2)
    want2 = small.field("VY").interior()
    free0 = _free(pkg)
    got3 = _noalloc(pkg, big, "VX")
    got2 = _noalloc(pkg, small, "VY")
    free1 = _free(pkg)
    assert got3.tobytes() == want3.tobytes()
    assert got2.tobytes() == want2.tobytes()
    assert free1 == free0
    # a wrong count is rejected with a clear error
    A = pkg._abi
    bad = np.zeros(5)
    rc = A.lib().fi_field_download_noalloc(small._h, A.F["VY"], A.dptr(bad), bad.size)
    assert rc == A.FI_ERR_ARG and b"interior transfer" in A.lib().fi_last_error()
    big.close()
    small.close()
