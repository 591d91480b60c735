"""frico_fields_noalloc (device buffers in and out, no allocation) against frico_fields.

The caller-side device buffers are made with the HIP runtime the library itself links
(libamdhip64.so.7), through ctypes.  Two calls in a row, the second with fewer columns and points,
are compared with frico_fields on the same currents and points: the two entry points share their
kernels, so the fields must agree bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FREQ = 300e6
H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        from frico_amd import _lib
        _lib.load()  # the library's own dependency: the same runtime instance
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def upload(self, a):
        if not (a.flags.c_contiguous or a.flags.f_contiguous):
            a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def empty(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _solved_context(nang):
    import frico_amd
    from frico_amd import meshes
    ctx = frico_amd.Context(0)
    sm = meshes.plate(meshes.wavelength(FREQ), 6)
    ctx.load_mesh(sm, 5)
    ctx.fill(FREQ, "quad", "c128")
    th = np.radians(np.linspace(0.0, 80.0, nang))
    ctx.rhs_planewave(FREQ, th, np.zeros(nang), 1.0, 0.0)
    ctx.factor("c128")
    ctx.solve()
    return ctx


def _noalloc(ctx, hip, X, ncols, pts, cols, h_sign=1.0):
    N = ctx.N
    npts = pts.shape[0]
    dX = hip.upload(np.asfortranarray(X))
    dP = hip.upload(np.ascontiguousarray(pts, np.float64))
    dC = None if cols is None else hip.upload(np.ascontiguousarray(cols, np.int32))
    dE = hip.empty(npts * 3 * 16)
    dH = hip.empty(npts * 3 * 16)
    print(f"frico_fields_noalloc: N={N} T={ctx.T} ncols={ncols} npts={npts}", flush=True)
    rc = ctx._L.frico_fields_noalloc(ctx._h, FREQ, dX, X.shape[0], ncols, dP, npts, dC, h_sign, dE, dH)
    assert rc == 0, ctx._L.frico_last_error().decode()
    ctx.synchronize()
    return hip.download(dE, (npts, 3), np.complex128), hip.download(dH, (npts, 3), np.complex128)


def test_fields_noalloc_matches_fields_on_two_calls():
    hip = _Hip()
    ctx = _solved_context(6)
    try:
        X = ctx.solution()
        rng = np.random.default_rng(7)
        lam = 299792458.0 / FREQ
        d = rng.normal(size=(40, 3))
        pts = 3 * lam * d / np.linalg.norm(d, axis=1)[:, None]
        cols = np.arange(40) % 6
        # first call: six columns, a column per point
        E1, H1 = _noalloc(ctx, hip, X, 6, pts, cols)
        Er, Hr = ctx.fields(FREQ, pts, cols=cols)
        assert np.all(np.isfinite(E1)) and np.all(np.isfinite(H1))
        assert np.array_equal(E1, Er) and np.array_equal(H1, Hr)
        # second call, smaller: two columns of the same currents, nine points, column 0 for all
        E2, H2 = _noalloc(ctx, hip, X[:, :2], 2, pts[:9], None)
        Er2, Hr2 = ctx.fields(FREQ, pts[:9])
        assert np.array_equal(E2, Er2) and np.array_equal(H2, Hr2)
        # the physically signed H (exp(-jkR)) through both entry points
        E3, H3 = _noalloc(ctx, hip, X[:, :2], 2, pts[:9], np.array([1, 0] * 4 + [1]), h_sign=-1.0)
        Er3, Hr3 = ctx.fields(FREQ, pts[:9], cols=np.array([1, 0] * 4 + [1]), h_sign=-1.0)
        assert np.array_equal(E3, Er3) and np.array_equal(H3, Hr3)
    finally:
        hip.free()
        ctx.close()


def test_fields_noalloc_flags_bad_columns_and_sizes():
    from frico_amd import _lib
    hip = _Hip()
    ctx = _solved_context(3)
    try:
        X = ctx.solution()
        pts = np.array([[0.0, 0.0, 5.0], [5.0, 0.0, 0.0], [0.0, 5.0, 1.0]])
        # a column outside [0, ncols) gives NaN fields for that point only
        E, H = _noalloc(ctx, hip, X, 3, pts, np.array([0, 3, 2]))
        Er, Hr = ctx.fields(FREQ, pts[[0, 2]], cols=np.array([0, 2]))
        assert np.all(np.isnan(E[1])) and np.all(np.isnan(H[1]))
        assert np.array_equal(E[[0, 2]], Er) and np.array_equal(H[[0, 2]], Hr)
        # scratch limit: 64 * T * ncols bytes; rejected on the host before any launch
        T = ctx.T
        too_many = (64 * 1024 * 1024) // (64 * T) + 1
        dX = hip.upload(np.asfortranarray(X))
        dP = hip.upload(pts)
        dE = hip.empty(3 * 3 * 16)
        rc = ctx._L.frico_fields_noalloc(ctx._h, FREQ, dX, ctx.N, too_many, dP, 3, None, 1.0, dE, None)
        assert rc == _lib.FRICO_ERR_LIMIT, rc
        assert b"scratch" in ctx._L.frico_last_error()
        # leading dimension below N and null outputs are argument errors
        assert ctx._L.frico_fields_noalloc(ctx._h, FREQ, dX, ctx.N - 1, 1, dP, 3, None, 1.0, dE, None) == -4
        assert ctx._L.frico_fields_noalloc(ctx._h, FREQ, dX, ctx.N, 1, dP, 3, None, 1.0, None, None) == -10
    finally:
        hip.free()
        ctx.close()
