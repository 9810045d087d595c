"""nbx_rdf_noalloc (the allocation-free rdf entry point) against nbx_rdf, through the C ABI.

The two entry points share the histogram kernel; the bar is bit equality of rs and g(r).
Two calls in a row (the second with fewer frames, fewer atoms and fewer bins) check that
the library-owned histogram is cleared between calls.  Sizes above the fixed histogram
fail with NBX_ERR_CAPACITY and leave the outputs unwritten.
"""
import ctypes as C

import numpy as np
import pytest

from nbodysimulator import _lib
from nbodysimulator import generators as gen

pytestmark = pytest.mark.gpu


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def cubic_lj(s):
    n = s.u.shape[0]
    ctx = _lib.Context(n, _lib.SYS_POTENTIAL, s.masses)
    ctx.set_boundary(_lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
    ctx.add_lennard_jones(1.0, 1.0, 2.5)
    return ctx


def frames_of(s, count, seed):
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(
        np.stack([s.u + 0.05 * rng.standard_normal(s.u.shape) for _ in range(count)]))


def last_error(ctx):
    return _lib.load().nbx_last_error(ctx._h).decode()


def raw_rdf(ctx, fn, frames, maxbin):
    rs = np.full(maxbin, np.nan)
    gr = np.full(maxbin, np.nan)
    rc = fn(ctx._h, dp(frames), frames.shape[0], maxbin, dp(rs), dp(gr))
    return rc, rs, gr


def test_noalloc_rdf_matches_rdf_on_two_calls():
    lib = _lib.load()
    big = gen.lj_liquid(6, rho=0.8442, T=0.72, seed=11)     # N = 864
    small = gen.lj_liquid(4, rho=0.8442, T=0.72, seed=12)   # N = 256
    cases = [(big, frames_of(big, 3, 1), 1000), (small, frames_of(small, 2, 2), 300)]
    for s, frames, maxbin in cases:
        ctx = cubic_lj(s)
        rc_a, rs_a, gr_a = raw_rdf(ctx, lib.nbx_rdf_noalloc, frames, maxbin)
        assert rc_a == _lib.NBX_OK, last_error(ctx)
        rc_b, rs_b, gr_b = raw_rdf(ctx, lib.nbx_rdf, frames, maxbin)
        assert rc_b == _lib.NBX_OK, last_error(ctx)
        np.testing.assert_array_equal(rs_a, rs_b)
        np.testing.assert_array_equal(gr_a, gr_b)
        assert gr_a.sum() > 0


def test_noalloc_rdf_repeated_calls_do_not_accumulate():
    s = gen.lj_liquid(3, rho=0.8442, T=0.72, seed=13)        # N = 108
    ctx = cubic_lj(s)
    frames = frames_of(s, 2, 3)
    first = ctx.rdf_noalloc(frames, 500)
    second = ctx.rdf_noalloc(frames[:1], 200)
    again = ctx.rdf_noalloc(frames, 500)
    ref_second = ctx.rdf(frames[:1], 200)
    np.testing.assert_array_equal(second[1], ref_second[1])
    np.testing.assert_array_equal(first[1], again[1])


def test_noalloc_rdf_rejects_bins_beyond_its_histogram():
    lib = _lib.load()
    s = gen.lj_liquid(3, rho=0.8442, T=0.72, seed=14)
    ctx = cubic_lj(s)
    frames = frames_of(s, 1, 4)
    rc, rs, gr = raw_rdf(ctx, lib.nbx_rdf_noalloc, frames, _lib.RDF_NOALLOC_MAXBIN + 1)
    assert rc == _lib.NBX_ERR_CAPACITY
    assert "exceeds" in last_error(ctx)
    assert np.isnan(gr).all() and np.isnan(rs).all()
    rc, _, gr = raw_rdf(ctx, lib.nbx_rdf_noalloc, frames, _lib.RDF_NOALLOC_MAXBIN)
    assert rc == _lib.NBX_OK
    rs_ref, gr_ref = ctx.rdf(frames, _lib.RDF_NOALLOC_MAXBIN)
    np.testing.assert_array_equal(gr, gr_ref)
