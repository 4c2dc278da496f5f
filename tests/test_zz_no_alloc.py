"""mspms_sfgridgen_noalloc (the grid generator with the solid atoms staged
in a fixed device array and the tables written into the caller's device
buffer, no allocation) against mspms_sfgridgen on the reference's CH3 /
(11,11) nanotube input: bit for bit, on two calls in a row, the second with
a smaller input (fewer solid atoms, coarser grid), so stale staging from
the first call would show.  More than 1000 solid atoms is refused."""
import ctypes as C
import os

import numpy as np
import pytest

from mspms2_amd import abi
from tests.util import GOLDEN

pytestmark = pytest.mark.gpu

CASE = os.path.join(GOLDEN, "sfgridgen_ch3_1A")


def _noalloc(g, torch):
    from mspms2_amd import sfgridgen as S
    from mspms2_amd.engine import lib
    L = lib()
    L.mspms_sfgridgen_noalloc.argtypes = [C.c_int, C.POINTER(abi.SfGridGenSpec), C.c_void_p,
                                          C.c_void_p]
    nx, ny, nz = S.dims(g)
    out = torch.full((13, nx, ny, nz), float("nan"), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    s, keep = S._spec(g)
    rc = L.mspms_sfgridgen_noalloc(0, C.byref(s), out.data_ptr(), None)
    return rc, out.cpu().numpy()


def test_noalloc_equals_allocating_generator_twice():
    import torch
    from mspms2_amd import sfgridgen as S
    g = S.read_input(os.path.join(CASE, "grid.in"))
    g.grid_itvl = 0.5
    print(f"call 1: {g.s_natom} solid atoms, {S.dims(g)} nodes", flush=True)
    rc, a = _noalloc(g, torch)
    assert rc == 0
    assert np.array_equal(a, S.generate(g))
    # second call: every other solid atom, a coarser grid
    g.s_xyz = np.ascontiguousarray(g.s_xyz[:, ::2])
    g.s_names = g.s_names[::2]
    g.grid_itvl = 1.0
    print(f"call 2: {g.s_natom} solid atoms, {S.dims(g)} nodes", flush=True)
    rc, b = _noalloc(g, torch)
    assert rc == 0
    assert np.array_equal(b, S.generate(g))


def test_noalloc_refuses_more_solid_atoms_than_its_buffer():
    import torch
    from mspms2_amd import sfgridgen as S
    from mspms2_amd.engine import lib
    g = S.read_input(os.path.join(CASE, "grid.in"))
    g.s_xyz = np.ascontiguousarray(np.tile(g.s_xyz, (1, 1001 // g.s_natom + 1))[:, :1001])
    g.s_names = ["C"] * 1001
    assert g.s_natom == 1001
    rc, _ = _noalloc(g, torch)
    assert rc == abi.MSPMS_EINVAL
    assert b"1000" in lib().mspms_last_error()
