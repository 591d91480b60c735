"""frico_amd -- MI355X (gfx950) EFIE Method-of-Moments hot path behind FRICO's entry points.

Python front-end of libfrico_hip.so.  The names follow the reference: `Context.fill` is
compute_matrix / compute_matrix_approx, `factor` + `solve` are solve_general's zgetrf/zgetrs,
`fields` is eval_fields, and `antenna_sweep` / `bistatic_sweep` are the two do_sweep entry
points (src/emw_solver.h:85-88, src/lapack_wrappers.h:28-35, src/emw_postpro_common.h:33,
src/emw_ana_antenna.h:52, src/emw_ana_bistatic_rcs.h:45).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib, meshes
from ._lib import (C64, C128, FILL_APPROX, FILL_QUAD, SOLVE_C64_IR, SOLVE_C64_NP_IR, SOLVE_C64_SYM_IR,
                   SOLVE_C128)

__all__ = ["Context", "Batch", "AntennaResponse", "FricoError", "SingularMatrixError", "NotConvergedError", "SolveInfo",
           "meshes", "FILL_QUAD", "FILL_APPROX", "C128", "C64", "SOLVE_C128", "SOLVE_C64_IR",
           "SOLVE_C64_SYM_IR", "SOLVE_C64_NP_IR", "SamplingPlane", "parse_sampling_plane", "sampling_plane_mesh",
           "read_msh", "sweep_frequencies", "format_double", "write_port_sweep", "write_radiation_diagram", "write_bistatic_rcs",
           "write_antenna_text_outputs"]


class FricoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class SingularMatrixError(FricoError):
    """zgetrf info > 0: exactly zero pivot at column `code` (1-based)."""


class NotConvergedError(FricoError):
    """Iterative refinement stopped above the tolerance (the iterate is still available)."""


@dataclass
class SolveInfo:
    iterations: int
    converged: bool
    max_rel_residual: float


@dataclass
class SamplingPlane:
    """A -P plane (src/sampling_planes.h:33-38): `axis` is the reference's `normal`, the axis
    about which the rectangle width x height (in z = 0) is turned by 90 degrees; h = element size."""
    axis: tuple
    width: float
    height: float
    h: float


def parse_sampling_plane(spec: str) -> SamplingPlane:
    """parse_sampling_plane (src/sampling_planes.cpp:84-106): "xy|yz|xz:w:h:size" or
    "nx:ny:nz:w:h:size".  Raises FricoError for an invalid specification."""
    L = _lib.load()
    ax = np.zeros(3)
    w, hh, sz = C.c_double(), C.c_double(), C.c_double()
    rc = L.frico_sampling_plane_parse(spec.encode(), _p(ax), C.byref(w), C.byref(hh), C.byref(sz))
    if rc:
        raise FricoError(rc, L.frico_last_error().decode())
    return SamplingPlane(tuple(float(a) for a in ax), w.value, hh.value, sz.value)


def sampling_plane_mesh(sp: SamplingPlane):
    """The plane's mesh (make_sampling_plane_mesh, src/sampling_planes.cpp:108-156) as a structured
    mesh of the same rectangle: returns (vertices nv x 3, triangles nt x 3)."""
    L = _lib.load()
    ax = np.ascontiguousarray(sp.axis, np.float64)
    nv, nt = C.c_int64(), C.c_int64()
    rc = L.frico_sampling_plane_mesh(_p(ax), sp.width, sp.height, sp.h, None, 0, None, 0, C.byref(nv), C.byref(nt))
    if rc:
        raise FricoError(rc, L.frico_last_error().decode())
    V = np.zeros((nv.value, 3))
    T = np.zeros((nt.value, 3), np.int64)
    rc = L.frico_sampling_plane_mesh(_p(ax), sp.width, sp.height, sp.h, _p(V), nv.value, _p(T, C.c_int64),
                                     nt.value, C.byref(nv), C.byref(nt))
    if rc:
        raise FricoError(rc, L.frico_last_error().decode())
    return V, T


def read_msh(path, feed_group=None):
    """A gmsh 4.1 ASCII .msh file read as load_from_gmsh reads it through gmsh (see frico_msh_read):
    returns (SurfaceMesh, groups) with groups = {name: (dim, tag, [entity tags])}.  feed_group
    names the physical group of the delta-gap curves, as FRICO's source option does; its entity
    tags become the mesh's feed."""
    L = _lib.load()
    h = C.c_void_p()
    rc = L.frico_msh_read(str(path).encode(), C.byref(h))
    if rc:
        raise FricoError(rc, L.frico_last_error().decode())
    try:
        nv, nt, nb, ng = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        L.frico_msh_sizes(h, C.byref(nv), C.byref(nt), C.byref(nb), C.byref(ng))
        V = np.zeros((nv.value, 3))
        T = np.zeros((nt.value, 3), np.int64)
        tg = np.zeros(nt.value, np.int32)
        B = np.zeros((nb.value, 2), np.int64)
        bt = np.zeros(nb.value, np.int32)
        L.frico_msh_arrays(h, _p(V), _p(T, C.c_int64), _p(tg, C.c_int32), _p(B, C.c_int64), _p(bt, C.c_int32))
        groups = {}
        for i in range(ng.value):
            name = C.create_string_buffer(256)
            dim, tag, ne = C.c_int32(), C.c_int32(), C.c_int32()
            L.frico_msh_group(h, i, name, 256, C.byref(dim), C.byref(tag), None, 0, C.byref(ne))
            ents = np.zeros(ne.value, np.int32)
            L.frico_msh_group(h, i, None, 0, None, None, _p(ents, C.c_int32), ne.value, C.byref(ne))
            groups[name.value.decode()] = (dim.value, tag.value, [int(e) for e in ents])
    finally:
        L.frico_msh_free(h)
    feed = []
    if feed_group is not None:
        if feed_group not in groups:
            raise FricoError(-2, f'Unknown physical group "{feed_group}": cannot enable source')
        feed = groups[feed_group][2]
    sm = meshes.SurfaceMesh(V, T, tg, B, bt, name=str(path), feed=feed)
    return sm, groups


def sweep_frequencies(start: float, end: float, step: float) -> np.ndarray:
    """The frequencies of a sweep as init_sweep builds them (src/emw_solver.cpp:460-471): start, then
    freq += step while freq <= end; the accumulation can drop the end point."""
    L = _lib.load()
    n = C.c_int64()
    rc = L.frico_sweep_frequencies(float(start), float(end), float(step), None, 0, C.byref(n))
    if rc:
        raise FricoError(rc, "bad sweep range")
    f = np.zeros(n.value)
    rc = L.frico_sweep_frequencies(float(start), float(end), float(step), _p(f), f.size, C.byref(n))
    if rc:
        raise FricoError(rc, "bad sweep range")
    return f


def format_double(v: float) -> str:
    """std::format("{}", v) (std::to_chars shortest form), as the reference's text outputs print."""
    L = _lib.load()
    buf = C.create_string_buffer(64)
    rc = L.frico_format_double(float(v), buf, 64)
    if rc:
        raise FricoError(rc, L.frico_last_error().decode())
    return buf.value.decode()


def _write(fn, *args):
    L = _lib.load()
    rc = getattr(L, fn)(*args)
    if rc:
        raise FricoError(rc, L.frico_last_error().decode())


def write_port_sweep(path, Z0, freqs, zin, swr):
    """port_sweep.txt (write_sweep_data, src/emw_ana_antenna.cpp:150-169)."""
    f = np.ascontiguousarray(freqs, np.float64)
    z = np.ascontiguousarray(np.asarray(zin, np.complex128).view(np.float64))
    w = np.ascontiguousarray(swr, np.float64)
    _write("frico_write_port_sweep", str(path).encode(), float(Z0), f.size, _p(f), _p(z), _p(w))


def write_radiation_diagram(path, freq, gain):
    """polar_<n>.txt (write_radiation_diagrams, src/emw_ana_antenna.cpp:171-203); gain 3 x 360."""
    g = np.ascontiguousarray(gain, np.float64).reshape(3, 360)
    _write("frico_write_radiation_diagram", str(path).encode(), float(freq), _p(g))


def write_antenna_h5(path, freqs, port, gain, rdiag_dist=10.0, Z0=50.0, voltage=1.0 + 0.0j):
    """The antenna analysis in HDF5 (write_hdf5, src/emw_ana_antenna.cpp:207-234): port = nsteps x 10
    and gain = nsteps x 1080 as antenna_sweep returns them.  Read back with frico_amd.h5.H5File."""
    f = np.ascontiguousarray(freqs, np.float64).ravel()
    po = np.ascontiguousarray(port, np.float64).reshape(f.size, 10)
    g = np.ascontiguousarray(gain, np.float64).reshape(f.size, 1080)
    v = complex(voltage)
    _write("frico_write_antenna_h5", str(path).encode(), float(rdiag_dist), float(Z0), f.size, _p(f), v.real, v.imag,
           _p(po), _p(g))


def write_antenna_h5_results(path, freqs, results, rdiag_dist=10.0, Z0=50.0, voltage=1.0 + 0.0j):
    """write_antenna_h5 from Context.antenna_sweep's list of per-frequency results."""
    port = np.array([[r["current"].real, r["current"].imag, r["power"].real, r["power"].imag, r["impedance"].real,
                      r["impedance"].imag, r["gamma"].real, r["gamma"].imag, r["swr"], r["residual"]]
                     for r in results])
    gain = np.array([np.asarray(r["gain"]).reshape(1080) for r in results])
    write_antenna_h5(path, freqs, port, gain, rdiag_dist, Z0, voltage)


def write_linear_system_h5(path, Z, V, I=None):
    """The -d dump (src/emw_ana_antenna.cpp:260-268): /frico/Z ([i][j] = Z[i, j]), /frico/V and, if
    given, /frico/I (n x nrhs; a vector is one column)."""
    Zf = np.asfortranarray(Z, np.complex128)
    n = Zf.shape[0]
    Vf = np.asfortranarray(np.asarray(V, np.complex128).reshape(n, -1))
    If = None if I is None else np.asfortranarray(np.asarray(I, np.complex128).reshape(n, -1))
    if If is not None and If.shape != Vf.shape:
        raise ValueError("I must have V's shape")
    _write("frico_write_linear_system_h5", str(path).encode(), n, _p(Zf), n, Vf.shape[1], _p(Vf), n,
           None if If is None else _p(If), n)


def write_bistatic_rcs(path, table):
    """bistatic_rcs_<n>.txt (src/emw_ana_bistatic_rcs.cpp:97-117); table 359 x 4."""
    t = np.ascontiguousarray(table, np.float64).reshape(359, 4)
    _write("frico_write_bistatic_rcs", str(path).encode(), _p(t))


def write_antenna_text_outputs(directory, freqs, results, Z0=50.0):
    """The text-output block of do_sweep (src/emw_ana_antenna.cpp:276-285): polar_<n>.txt for each
    sweep step and port_sweep.txt, from antenna_sweep's results."""
    import os
    for n, r in enumerate(results):
        write_radiation_diagram(os.path.join(directory, f"polar_{n}.txt"), freqs[n], r["gain"])
    write_port_sweep(os.path.join(directory, "port_sweep.txt"), Z0, freqs, [r["impedance"] for r in results],
                     [r["swr"] for r in results])


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def _modes(mode, prec=None):
    m = {"quad": FILL_QUAD, "approx": FILL_APPROX}.get(mode, mode)
    if prec is None:
        return m
    return m, {"c128": C128, "c64": C64}.get(prec, prec)


def _form(form):
    return {"c128": SOLVE_C128, "c64_ir": SOLVE_C64_IR, "c64_sym_ir": SOLVE_C64_SYM_IR,
            "c64_np_ir": SOLVE_C64_NP_IR}.get(form, form)


class Context:
    """One device, one stream; owns the mesh, Z, its factors, the RHS block and the solution."""

    def __init__(self, device: int = 0):
        self._L = _lib.load()
        h = C.c_void_p()
        self._check(self._L.frico_ctx_create(device, C.byref(h)))
        self._h = h
        self.mesh = None

    def close(self):
        if getattr(self, "_h", None):
            self._L.frico_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, ok=(0,)):
        if rc in ok:
            return rc
        msg = self._L.frico_last_error().decode()
        if rc > 0:
            raise SingularMatrixError(rc, msg)
        if rc == _lib.FRICO_ERR_NOCONV:
            raise NotConvergedError(rc, msg)
        raise FricoError(rc, msg)

    # ---- configuration / profiling --------------------------------------------------------
    def set_block(self, nb: int):
        self._check(self._L.frico_ctx_set_block(self._h, nb))

    def set_profiling(self, on: bool = True):
        self._check(self._L.frico_ctx_set_profiling(self._h, int(on)))

    def profile(self) -> dict:
        p = _lib.frico_profile()
        self._check(self._L.frico_ctx_get_profile(self._h, C.byref(p)))
        return {k: getattr(p, k) for k, _ in p._fields_}

    def synchronize(self):
        self._check(self._L.frico_ctx_synchronize(self._h))

    # ---- mesh -------------------------------------------------------------------------------
    def load_mesh(self, sm: meshes.SurfaceMesh, degree: int = 5):
        self._check(self._L.frico_mesh_load(
            self._h, _p(sm.vertices), sm.nv, _p(sm.triangles, C.c_int64), _p(sm.tags, C.c_int32), sm.nt,
            _p(sm.bedges, C.c_int64), _p(sm.btags, C.c_int32), sm.bedges.shape[0], degree))
        self.mesh = sm
        self.degree = degree
        return self

    @property
    def N(self) -> int:
        return int(self._L.frico_num_unknowns(self._h))

    @property
    def T(self) -> int:
        return int(self._L.frico_num_triangles(self._h))

    def rwg_table(self):
        arr = (_lib.frico_rwg * self.N)()
        self._check(self._L.frico_rwg_table(self._h, arr))
        return arr

    def reorient_deltagap(self, ifaces, force=False):
        ifa = np.ascontiguousarray(ifaces, np.int32)
        signs = np.ones(self.N)
        applied = C.c_int(0)
        self._check(self._L.frico_reorient_deltagap(self._h, _p(ifa, C.c_int32), len(ifa), int(force),
                                                    _p(signs), C.byref(applied)))
        return signs if applied.value else None

    # ---- matrix -------------------------------------------------------------------------------
    def fill(self, freq: float, mode="quad", prec="c128"):
        m, p = _modes(mode, prec)
        self._check(self._L.frico_fill(self._h, float(freq), m, p))

    def symmetrize(self):
        self._check(self._L.frico_symmetrize(self._h))

    def matrix(self) -> np.ndarray:
        N = self.N
        Z = np.empty((N, N), np.complex128, order="F")
        self._check(self._L.frico_matrix_get(self._h, _p(Z), N))
        return Z

    def set_matrix(self, Z):
        Z = np.asfortranarray(Z, np.complex128)
        self._check(self._L.frico_matrix_set(self._h, _p(Z), Z.shape[0], Z.shape[0]))

    # ---- right-hand sides ---------------------------------------------------------------------
    def rhs_deltagap(self, freq, ifaces, voltage=1.0, signs=None):
        ifa = np.ascontiguousarray(ifaces, np.int32)
        v = complex(voltage)
        self._check(self._L.frico_rhs_deltagap(self._h, float(freq), _p(ifa, C.c_int32), len(ifa), v.real,
                                               v.imag, None if signs is None else _p(np.ascontiguousarray(signs, np.float64))))

    def rhs_planewave(self, freq, theta, phi, eth=1.0, eph=0.0):
        th = np.atleast_1d(np.asarray(theta, np.float64))
        ph = np.broadcast_to(np.asarray(phi, np.float64), th.shape).copy()
        et = np.ascontiguousarray(np.broadcast_to(np.asarray(eth, np.complex128), th.shape))
        ep = np.ascontiguousarray(np.broadcast_to(np.asarray(eph, np.complex128), th.shape))
        self._check(self._L.frico_rhs_planewave(self._h, float(freq), th.size, _p(th), _p(ph), _p(et), _p(ep)))

    def set_rhs(self, B):
        B = np.asfortranarray(np.atleast_2d(np.asarray(B, np.complex128).T).T)
        self._check(self._L.frico_rhs_set(self._h, _p(B), B.shape[0], B.shape[1]))

    def rhs(self) -> np.ndarray:
        n = int(self._L.frico_num_rhs(self._h))
        B = np.empty((self.N, n), np.complex128, order="F")
        self._check(self._L.frico_rhs_get(self._h, _p(B), self.N))
        return B

    # ---- factor / solve -----------------------------------------------------------------------
    def factor(self, form="c128"):
        return self._check(self._L.frico_factor(self._h, _form(form)))

    @property
    def factored_form(self) -> int:
        """Form of the current factorisation (SOLVE_*), -1 if none: the symmetric form reports
        SOLVE_C64_IR after falling back to the pivoted LU."""
        return int(self._L.frico_factored_form(self._h))

    def solve(self, tol=1e-12, maxit=10) -> SolveInfo:
        info = _lib.frico_solve_info()
        self._check(self._L.frico_solve(self._h, float(tol), int(maxit), C.byref(info)))
        return SolveInfo(info.iterations, bool(info.converged), info.max_rel_residual)

    def dump_linear_system(self, path):
        """The -d dump of this context (frico_dump_linear_system): its complex128 matrix as it stands,
        its right-hand sides and, if solved, the solution."""
        self._check(self._L.frico_dump_linear_system(self._h, str(path).encode()))

    def solution(self) -> np.ndarray:
        n = int(self._L.frico_num_rhs(self._h))
        X = np.empty((self.N, n), np.complex128, order="F")
        self._check(self._L.frico_solution_get(self._h, _p(X), self.N))
        return X

    def set_solution(self, X):
        X = np.asfortranarray(np.atleast_2d(np.asarray(X, np.complex128).T).T)
        self._check(self._L.frico_solution_set(self._h, _p(X), X.shape[0], X.shape[1]))

    def residual(self) -> np.ndarray:
        n = int(self._L.frico_num_rhs(self._h))
        r = np.empty(n)
        self._check(self._L.frico_residual(self._h, _p(r)))
        return r

    # ---- post-processing ----------------------------------------------------------------------
    def triangle_currents(self, col=0):
        T = self.T
        AJ = np.empty((T, 3), np.complex128)
        AdivJ = np.empty(T, np.complex128)
        self._check(self._L.frico_triangle_currents(self._h, int(col), _p(AJ), _p(AdivJ)))
        return AJ, AdivJ

    def fields(self, freq, pts, cols=None, h_sign=1.0):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        n = pts.shape[0]
        E = np.empty((n, 3), np.complex128)
        H = np.empty((n, 3), np.complex128)
        cc = None if cols is None else np.ascontiguousarray(cols, np.int64)
        self._check(self._L.frico_fields(self._h, float(freq), _p(pts), n,
                                         None if cc is None else _p(cc, C.c_int64), float(h_sign), _p(E), _p(H)))
        return E, H

    def fields_on_plane(self, freq, plane, col=0):
        """eval_fields(sim, ctx, smpmsh, E, H) (src/emw_postpro_common.cpp:77-90) on a -P plane
        (a spec string or a SamplingPlane): E and H at every vertex of the plane's mesh from
        solution column `col`, as write_fields evaluates them (:180-189).  Returns (V, T, E, H)."""
        sp = parse_sampling_plane(plane) if isinstance(plane, str) else plane
        V, T = sampling_plane_mesh(sp)
        E, H = self.fields(freq, V, cols=np.full(V.shape[0], col, np.int64))
        return V, T, E, H

    # ---- whole analyses (do_sweep mirrors) ----------------------------------------------------
    def antenna_sweep(self, freqs, ifaces, voltage=1.0, mode="quad", R=10.0, Z0=50.0,
                      force_symmetry=False, force_reorient=False, form="c128", tol=1e-12, maxit=10):
        f = np.atleast_1d(np.asarray(freqs, np.float64))
        ifa = np.ascontiguousarray(ifaces, np.int32)
        port = np.zeros((f.size, 10))
        gain = np.zeros((f.size, 3, 360))
        v = complex(voltage)
        self._check(self._L.frico_antenna_sweep(self._h, _p(f), f.size, _modes(mode), _p(ifa, C.c_int32),
                                                len(ifa), v.real, v.imag, R, Z0, int(force_symmetry),
                                                int(force_reorient), _form(form), tol, maxit, _p(port), _p(gain)))
        out = []
        for i in range(f.size):
            p = port[i]
            out.append({"current": complex(p[0], p[1]), "power": complex(p[2], p[3]),
                        "impedance": complex(p[4], p[5]), "gamma": complex(p[6], p[7]), "swr": p[8],
                        "residual": p[9], "gain": gain[i]})
        return out

    def bistatic_sweep(self, freqs, theta_r, phi_r, eth=1.0, eph=1.0, mode="quad", R=10.0, form="c128",
                       tol=1e-12, maxit=10):
        f = np.atleast_1d(np.asarray(freqs, np.float64))
        tab = np.zeros((f.size, 359, 4))
        self._check(self._L.frico_bistatic_sweep(self._h, _p(f), f.size, _modes(mode), theta_r, phi_r, eth, eph,
                                                 R, _form(form), tol, maxit, _p(tab)))
        return tab

    def monostatic_sweep(self, freqs, theta, phi, eth=1.0, eph=0.0, mode="quad", R=10.0, form="c128",
                         tol=1e-10, maxit=10):
        f = np.atleast_1d(np.asarray(freqs, np.float64))
        th = np.ascontiguousarray(theta, np.float64)
        ph = np.ascontiguousarray(np.broadcast_to(np.asarray(phi, np.float64), th.shape))
        sig = np.zeros((f.size, th.size, 2))
        infos = (_lib.frico_solve_info * f.size)()
        self._check(self._L.frico_monostatic_sweep(self._h, _p(f), f.size, _modes(mode), th.size, _p(th), _p(ph),
                                                   eth, eph, R, _form(form), tol, maxit, _p(sig), infos))
        return sig, [SolveInfo(i.iterations, bool(i.converged), i.max_rel_residual) for i in infos]

    # ---- adjoint (autograd-Function-shaped pair) ---------------------------------------------
    def _adjoint_mode(self, mode):
        self._check(self._L.frico_adjoint_set_fill(self._h, {"quad": FILL_QUAD, "approx": FILL_APPROX}[mode]))

    def adjoint_forward(self, freq, ifaces, voltage=1.0, obs=(0.0, 0.0, 10.0), mode="quad"):
        """Antenna analysis (quadrature or approximate mode) kept for a backward pass: returns
        (Zin, gain at obs)."""
        self._adjoint_mode(mode)
        ifa = np.ascontiguousarray(ifaces, np.int32)
        ob = np.ascontiguousarray(obs, np.float64)
        out = np.zeros(4)
        v = complex(voltage)
        self._check(self._L.frico_adjoint_forward(self._h, float(freq), _p(ifa, C.c_int32), len(ifa), v.real, v.imag,
                                                  _p(ob), _p(out)))
        return complex(out[0], out[1]), float(out[2])

    def adjoint_backward(self, g_zin_re=0.0, g_zin_im=0.0, g_gain=0.0):
        """Vertex-coordinate gradient of g_re Re(Zin) + g_im Im(Zin) + g_gain G, shape (nv, 3)."""
        g = np.zeros(self.mesh.nv * 3)
        self._check(self._L.frico_adjoint_backward(self._h, float(g_zin_re), float(g_zin_im), float(g_gain), _p(g)))
        return g.reshape(-1, 3)

    def rcs_forward(self, freq, theta_i, phi_i, eth=1.0, eph=0.0, theta_s=None, phi_s=None, R=10.0, mode="quad"):
        """RCS (quadrature or approximate mode) kept for a backward pass: (sigma_VV, sigma_HH) at
        (R, theta_s, phi_s) for the plane wave from (theta_i, phi_i); monostatic when theta_s, phi_s
        are omitted."""
        self._adjoint_mode(mode)
        th = complex(eth)
        ph_ = complex(eph)
        a = np.array([th.real, th.imag])
        b = np.array([ph_.real, ph_.imag])
        out = np.zeros(2)
        ts = theta_i if theta_s is None else theta_s
        ps = phi_i if phi_s is None else phi_s
        self._check(self._L.frico_adjoint_rcs_forward(self._h, float(freq), float(theta_i), float(phi_i), _p(a), _p(b),
                                                      float(ts), float(ps), float(R), _p(out)))
        return float(out[0]), float(out[1])

    def rcs_backward(self, g_vv=0.0, g_hh=0.0):
        """Vertex-coordinate gradient of g_vv sigma_VV + g_hh sigma_HH, shape (nv, 3)."""
        g = np.zeros(self.mesh.nv * 3)
        self._check(self._L.frico_adjoint_rcs_backward(self._h, float(g_vv), float(g_hh), _p(g)))
        return g.reshape(-1, 3)

    def rcs_gradients(self, freq, theta_i, phi_i, eth=1.0, eph=0.0, theta_s=None, phi_s=None, R=10.0, mode="quad"):
        """sigma_VV, sigma_HH and their vertex gradients (each (nv, 3))."""
        s = self.rcs_forward(freq, theta_i, phi_i, eth, eph, theta_s, phi_s, R, mode)
        return {"sigma": np.array(s), "dsigma_vv": self.rcs_backward(1.0, 0.0),
                "dsigma_hh": self.rcs_backward(0.0, 1.0)}

    def antenna_gradients(self, freq, ifaces, voltage=1.0, obs=(0.0, 0.0, 10.0), mode="quad"):
        """Zin, G and their gradients: dZin (complex, (nv,3)) and dG (real, (nv,3))."""
        zin, gain = self.adjoint_forward(freq, ifaces, voltage, obs, mode)
        dre = self.adjoint_backward(1.0, 0.0, 0.0)
        dim = self.adjoint_backward(0.0, 1.0, 0.0)
        dg = self.adjoint_backward(0.0, 0.0, 1.0)
        return {"zin": zin, "gain": gain, "dzin": dre + 1j * dim, "dgain": dg}


class AntennaResponse:
    """Forward/backward pair shaped like torch.autograd.Function over the C ABI:
    forward(vertices) -> [Re Zin, Im Zin, G]; backward(grad_out[3]) -> d/d vertices."""

    def __init__(self, ctx: "Context", sm, freq, degree=5, voltage=1.0, obs=(0.0, 0.0, 10.0)):
        self.ctx, self.sm, self.freq, self.degree, self.voltage, self.obs = ctx, sm, freq, degree, voltage, obs

    def forward(self, vertices):
        sm = meshes.SurfaceMesh(np.asarray(vertices, np.float64).reshape(-1, 3), self.sm.triangles, self.sm.tags,
                                self.sm.bedges, self.sm.btags, feed=self.sm.feed)
        self.ctx.load_mesh(sm, self.degree)
        zin, g = self.ctx.adjoint_forward(self.freq, sm.feed, self.voltage, self.obs)
        return np.array([zin.real, zin.imag, g])

    def backward(self, grad_out):
        go = np.asarray(grad_out, np.float64)
        return self.ctx.adjoint_backward(go[0], go[1], go[2])


class Batch:
    """Many meshes, each at many frequencies, swept together on one device (frico_batch_*).

    The reference sweeps one mesh at a time (do_sweep, src/emw_ana_antenna.cpp:236-295); here the
    meshes run largest first on `workers` concurrent contexts, with HIP-graph replay of the
    complex64 LU and refinement sweeps and a per-size-class choice of the LU block width."""

    def __init__(self, device: int = 0, workers: int = 8):
        self._L = _lib.load()
        h = C.c_void_p()
        self._check(self._L.frico_batch_create(device, workers, C.byref(h)))
        self._h = h
        self.meshes = []

    def _check(self, rc, ok=(0,)):
        return Context._check(self, rc, ok)

    def close(self):
        if getattr(self, "_h", None):
            self._L.frico_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add_mesh(self, sm: meshes.SurfaceMesh, degree: int = 5, feed=None) -> int:
        fd = np.ascontiguousarray(sm.feed if feed is None else feed, np.int32)
        rc = self._L.frico_batch_add_mesh(self._h, _p(sm.vertices), sm.nv, _p(sm.triangles, C.c_int64),
                                          _p(sm.tags, C.c_int32), sm.nt, _p(sm.bedges, C.c_int64),
                                          _p(sm.btags, C.c_int32), sm.bedges.shape[0], degree, _p(fd, C.c_int32),
                                          len(fd))
        if rc < 0:
            self._check(rc)
        self.meshes.append(sm)
        return rc

    def antenna(self, freqs, fill="c64", form="c64_ir", tol=1e-10, maxit=10, graphs=True, autotune=True,
                voltage=1.0):
        """Input impedance Zin[m, f] and the solve status of every (mesh, frequency)."""
        f = np.ascontiguousarray(np.atleast_1d(freqs), np.float64)
        nm = len(self.meshes)
        z = np.zeros((nm, f.size, 2))
        infos = (_lib.frico_solve_info * (nm * f.size))()
        flags = (_lib.BATCH_GRAPHS if graphs else 0) | (_lib.BATCH_AUTOTUNE if autotune else 0)
        v = complex(voltage)
        self._check(self._L.frico_batch_antenna(self._h, _p(f), f.size, {"c128": C128, "c64": C64}[fill],
                                                _form(form), tol, maxit, flags, v.real, v.imag, _p(z), infos))
        info = [[SolveInfo(infos[m * f.size + i].iterations, bool(infos[m * f.size + i].converged),
                           infos[m * f.size + i].max_rel_residual) for i in range(f.size)] for m in range(nm)]
        return z[..., 0] + 1j * z[..., 1], info

    def workers_used(self) -> int:
        """Concurrent contexts of the last run (capped by GPU_MAX_HW_QUEUES and co-residency)."""
        return int(self._L.frico_batch_workers_used(self._h))

    def factored_forms(self, nfreq):
        """The factorisation (SOLVE_*) each (mesh, frequency) of the last run was solved with, after
        any fall-back: an array [mesh, frequency]."""
        forms = np.zeros(len(self.meshes) * nfreq, np.int32)
        n = self._L.frico_batch_factored_forms(self._h, _p(forms, C.c_int32))
        if n < 0:
            self._check(n)
        return forms.reshape(len(self.meshes), nfreq)

    def block_sizes(self):
        """(block width per mesh, number of meshes that replayed graphs) of the last run."""
        nb = np.zeros(len(self.meshes), np.int32)
        n = self._L.frico_batch_block_sizes(self._h, _p(nb, C.c_int32))
        if n < 0:
            self._check(n)
        return nb, n
