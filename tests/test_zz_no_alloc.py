"""GPU: bem_far_field_noalloc (far field through a fixed device scratch
array, no device allocation) against bem_far_field, through the C ABI.

Both run the same kernels with the same launch shapes, so the amplitudes must
be identical to the bit; two calls in a row, the second on a smaller input
(fewer directions, a smaller surface), check that nothing left in the shared
array by the first call leaks into the second.
"""
import ctypes
import sys

import numpy as np
import pytest

import bemhip as bh

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _far(lib, sol, fn, r, dirs):
    nd = 1 if dirs is None else dirs.shape[0]
    out = np.empty(nd, dtype=np.complex128)
    dp = None if dirs is None else np.ascontiguousarray(dirs, dtype=np.float64).ctypes.data_as(
        ctypes.POINTER(ctypes.c_double))
    rc = getattr(lib, fn)(sol._h, r, nd, dp, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return rc, out


def test_far_field_noalloc_matches_far_field():
    say("[gpu] bem_far_field_noalloc vs bem_far_field")
    ctx = bh.Context(0)
    try:
        lib = ctx.lib
        rng = np.random.default_rng(2)
        big = bh.mesh(bh.Spheroid(0.02, 0.01), level=4, ctx=ctx)
        small = bh.mesh(bh.Sphere(0.01), level=2, ctx=ctx)
        cases = [(big, bh.FluidFilled(1.0357, 1.0279), 500.0, 1024), (small, bh.Rigid(), 150.0, 5)]
        for surface, bc, k, nd in cases:
            op = bh.Operator(surface, bc, k)
            sol = op.solve([(0.3, 0.2), (1.1, 0.0)])
            d = rng.standard_normal((nd, 3))
            d /= np.linalg.norm(d, axis=1, keepdims=True)
            for r in (0, 1):
                rc0, ref = _far(lib, sol, "bem_far_field", r, d)
                rc1, got = _far(lib, sol, "bem_far_field_noalloc", r, d)
                assert rc0 == 0 and rc1 == 0
                say(f"[gpu]   N={surface.n} r={r} {nd} directions: max |diff| {np.abs(got - ref).max():.1e}")
                assert np.array_equal(got, ref)
                # backscatter (null directions)
                rc0, ref1 = _far(lib, sol, "bem_far_field", r, None)
                rc1, got1 = _far(lib, sol, "bem_far_field_noalloc", r, None)
                assert rc0 == 0 and rc1 == 0 and np.array_equal(got1, ref1)
        # beyond the fixed maximum: a clear error, no launch
        d = np.tile([1.0, 0.0, 0.0], (1025, 1))
        rc, _ = _far(lib, sol, "bem_far_field_noalloc", 0, d)
        assert rc == bh.BEM_ERR_ARGUMENT
        assert b"1024" in lib.bem_last_error(ctx._h)
    finally:
        ctx.close()
