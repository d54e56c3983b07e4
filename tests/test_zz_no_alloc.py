"""pfg_download_fields_noalloc, called through the C ABI: bit for bit the
fields of pfg_download_fields, on calls in a row whose inputs shrink (the last
a small lattice with solid cells, so nothing the earlier calls left in the
fixed device array may show), over one or several windows of z-planes, with
a subset of the outputs, and the size limit of the array."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCRATCH_FLOATS = 8 * 1024 * 1024


def lattice(pfg, nx, ny, nz, store=0, flags=False, seed=3):
    p = pfg.default_params(nx=nx, ny=ny, nz=nz, tau=0.8, store=store, use_flags=int(flags))
    p.wall[2] = pfg.WALL_BOUNCE
    p.wall_u_hi[2][0] = 0.04
    p.body_force[0] = 2e-6
    p.body_force[1] = -1e-6
    return p


def noalloc(pfg, s, want=(True, True, True, True)):
    """The four outputs (NaN where the call writes nothing; None where not
    asked for) and the status of the call."""
    out = [np.full(s.N, np.nan, np.float32) if w else None for w in want]
    ptrs = [None if a is None else a.ctypes.data_as(C.POINTER(C.c_float)) for a in out]
    st = pfg.lib().pfg_download_fields_noalloc(C.c_void_p(s._h.value), *ptrs)
    return st, out


def stepped(pfg, p, steps, flags=None):
    s = pfg.Solver(p)
    if flags is not None:
        s.upload_flags(flags)
    s.set_fluid_only(True)
    s.step(steps)
    return s


def check_equal(pfg, s, want=(True, True, True, True)):
    ref = s.fields()
    st, got = noalloc(pfg, s, want)
    assert st == 0
    for k, (g, r) in enumerate(zip(got, ref)):
        if g is not None:
            assert np.array_equal(g, r), ("field", k, int(np.sum(g != r)))
    return ref


def test_noalloc_fields_equal_download_fields(gpu):
    pfg = gpu
    # 1: a 1024x512 plane, four fields: windows of 4, 4 and 2 planes
    p = lattice(pfg, 1024, 512, 10)
    assert SCRATCH_FLOATS // (1024 * 512 * 4) == 4
    with stepped(pfg, p, 6) as s:
        check_equal(pfg, s)
    # 2: smaller, 16-bit storage, one window of 65 planes and one of 55
    p = lattice(pfg, 200, 160, 120, store=1)
    with stepped(pfg, p, 12) as s:
        check_equal(pfg, s)
        # two outputs: twice the planes per window
        check_equal(pfg, s, (True, False, False, True))
    # 3: rows of x-blocks per plane not a multiple of the four per workgroup,
    # two windows (635 and 65 planes)
    p = lattice(pfg, 100, 33, 700)
    with stepped(pfg, p, 4) as s:
        check_equal(pfg, s)
    # 4: smaller again, a tenth of the cells solid (they read 0 on both paths)
    p = lattice(pfg, 70, 12, 10, flags=True)
    rng = np.random.default_rng(5)
    flags = (rng.random(70 * 12 * 10) < 0.1).astype(np.uint8)
    with stepped(pfg, p, 20, flags) as s:
        ref = check_equal(pfg, s)
        assert (ref[0][flags != 0] == 0).all() and (ref[0][flags == 0] > 0.5).all()


def test_noalloc_size_limit_and_batches(gpu):
    pfg = gpu
    # a 2048x1100 plane: three outputs fit the array (6.76M floats), four do not
    p = lattice(pfg, 2048, 1100, 3)
    assert 2048 * 1100 * 3 <= SCRATCH_FLOATS < 2048 * 1100 * 4
    with stepped(pfg, p, 2) as s:
        st, out = noalloc(pfg, s)
        assert st == 1  # PFG_ERR_ARG
        check_equal(pfg, s, (True, True, False, True))
    ps = [lattice(pfg, 40, 24, 20), lattice(pfg, 30, 20, 10)]
    with pfg.Batch(ps) as b:
        st = pfg.lib().pfg_download_fields_noalloc(
            C.c_void_p(b._h.value), *[np.empty(1, np.float32).ctypes.data_as(
                C.POINTER(C.c_float)) for _ in range(4)])
        assert st == 5  # PFG_ERR_STATE: a batch reads its lattices one at a time
