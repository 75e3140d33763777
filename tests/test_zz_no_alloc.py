"""pcbgpu_solve_grid_noalloc (every device buffer from the library's fixed scratch array) against
pcbgpu_solve_grid (buffers it allocates): two calls in a row, the second on a smaller board, each equal bit for bit
in rectangle, best state, sweeps, tries and field, and equal to the oracle; a search too large for the scratch is
refused with an error and the next call still solves."""
import ctypes as C

import numpy as np
import pytest

from pcbroute_amd.abi import INF, Out, Problem

from cases import random_case

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20  # PCBGPU_NOALLOC_SCRATCH_BYTES in include/pcbgpu.h


def _say(*a):
    print(*a, flush=True)


def _noalloc(backend, grid, prob):
    """The no-alloc solve through the C ABI: (rc, out, field as (nl, rh, rw) or None)."""
    cap = prob.window[2] * prob.window[3] * prob.nl
    field = np.full(cap, -7, np.int32)
    out = Out()
    rc = backend.lib.pcbgpu_solve_grid_noalloc(grid.h, C.byref(prob.query()),
                                               field.ctypes.data_as(C.POINTER(C.c_int32)), cap, C.byref(out))
    if rc < 0:
        return rc, out, None
    n = out.nl * out.rw * out.rh
    assert not out.d, "the no-alloc solve must leave out->d alone"
    assert np.all(field[n:] == -7), "values written past the solved rectangle"
    return rc, out, field[:n].reshape(out.nl, out.rh, out.rw)


def _check(backend, oracle, prob):
    g = backend.grid(prob)
    try:
        rc, out, d = _noalloc(backend, g, prob)
        assert rc >= 0, backend.error()
        ref = backend.solve_grid(g, prob)
    finally:
        g.close()
    assert (out.rx0, out.ry0, out.rw, out.rh) == ref.rect
    # pass counts may differ by a few between two solves (the field is relaxed in place); results may not
    assert out.best == ref.best and out.sweeps == rc > 0 and ref.sweeps > 0 and out.tries == ref.tries
    np.testing.assert_array_equal(d, ref.d)
    orc = oracle.solve(prob)
    assert ref.rect == orc["rect"] and ref.best == orc["best"]
    np.testing.assert_array_equal(d, orc["d"])


def test_noalloc_equals_solve_grid_twice_then_smaller(backend, oracle):
    big = random_case(1500, nx=300, ny=220, nl=4, sym=True)
    small = random_case(1501, nx=110, ny=80, nl=2)
    _say("no-alloc: first call", big.window, big.nl)
    _check(backend, oracle, big)
    _say("no-alloc: second call, smaller", small.window, small.nl)
    _check(backend, oracle, small)


def test_noalloc_refuses_a_search_beyond_its_scratch(backend, oracle):
    nx, ny, nl = 1600, 1200, 4  # 7.7M states: the field and cost words (8 B each) and via costs exceed the scratch
    assert 8 * nx * ny * nl + 4 * nx * ny > SCRATCH_BYTES
    cst = np.ones((nl, ny, nx), np.int32)
    st = np.zeros((nl, ny, nx), np.uint16)
    vcst = np.ones((ny, nx), np.int32)
    vst = np.zeros((ny, nx), np.uint16)
    far = Problem(cst, st, vcst, vst, starts=[(0, 0)], targets=[((nl - 1) * ny + ny - 1) * nx + nx - 1], via_cost=50)
    g = backend.grid(far)
    try:
        _say("no-alloc: too large", far.window, nl)
        rc, out, d = _noalloc(backend, g, far)
        assert rc == -1 and "scratch" in backend.error()
    finally:
        g.close()
    small = random_case(1502, nx=100, ny=70, nl=3)
    _say("no-alloc: after the refusal", small.window, small.nl)
    _check(backend, oracle, small)
