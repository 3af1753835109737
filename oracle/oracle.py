"""ctypes binding of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the measured or
shipped path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_DIR, "liboracle.so")


def build(force: bool = False) -> None:
    subprocess.run(["make", "-C", _DIR] + (["-B"] if force else []), check=True,
                   stdout=subprocess.DEVNULL)


class AlphaResult(ctypes.Structure):
    _fields_ = [("alpha", ctypes.c_double), ("iters", ctypes.c_int), ("status", ctypes.c_int),
                ("resid", ctypes.c_double)]


class Params27(ctypes.Structure):
    _fields_ = [("beta", ctypes.c_double), ("tol", ctypes.c_double), ("maxit", ctypes.c_int),
                ("alpha_mode", ctypes.c_int), ("nthreads", ctypes.c_int)]


class OStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("iters", "nodes", "n_bound", "n_nonpos",
                                              "n_trivial", "n_noconv")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class OBox(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("origin", ctypes.c_double * 3),
                ("face_code", ctypes.c_int32 * 6)]


class BoxParams27(ctypes.Structure):
    _fields_ = [("beta", ctypes.c_double), ("tol", ctypes.c_double), ("maxit", ctypes.c_int),
                ("alpha_mode", ctypes.c_int), ("nthreads", ctypes.c_int),
                ("n_in", ctypes.c_int), ("n_out", ctypes.c_int),
                ("feq_in", ctypes.POINTER(ctypes.c_double)),
                ("feq_out", ctypes.POINTER(ctypes.c_double)),
                ("feq_fs", ctypes.POINTER(ctypes.c_double))]


class BcStats(ctypes.Structure):
    _fields_ = [("n_ibb_links", ctypes.c_int64), ("n_repair", ctypes.c_int64)]


class Consts39(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_int) for n in ("shifted", "turbulent", "lvl", "startup", "maxit")] +
                [(n, ctypes.c_double) for n in (
                    "gama", "R", "cv", "cp", "lead_g", "t_ref", "p_ref", "rho_ref", "mu_ref",
                    "kappa_ref", "sos", "mach", "molec_weight", "dt", "eps", "init_beta1",
                    "init_beta2", "beta1_scale", "beta2_scale")])


class Stats39(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("nodes", "lm_iters", "lm_noconv", "n_alpha2",
                                              "n_alpha_lo", "n_alpha_hi", "n_nan",
                                              "n_kappa_err", "n_repair")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


FACE_PERIODIC = 2147483647
_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        D, I = ctypes.c_double, ctypes.c_int
        L.o27_feq.argtypes = [D, D, D, D, dp]
        L.o27_moments.argtypes = [dp, dp, dp]
        L.o27_alpha.argtypes = [dp, dp, D, D, I]
        L.o27_alpha.restype = AlphaResult
        L.o27_alpha_series.argtypes = [dp, dp, D, D, I, I]
        L.o27_alpha_series.restype = AlphaResult
        L.o27_entropy_g.argtypes = [dp, dp, D, dp, dp]
        L.o27_collide_node.argtypes = [dp, D, D, D, I, I, dp, dp, dp]
        L.o27_collide_node.restype = AlphaResult
        L.o27_step_periodic.argtypes = [I, I, I, dp, dp, ctypes.POINTER(Params27), dp,
                                        ctypes.POINTER(OStats)]
        L.o27_init_tgv.argtypes = [I, I, I, D, D, ctypes.c_uint64, dp]
        L.o27_macro.argtypes = [I, dp, dp, dp, dp, dp]
        L.o_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_uint32)]
        L.o_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64]
        L.o_uniform.restype = D
        L.o_build_links27.argtypes = [ctypes.POINTER(OBox), I, dp, dp,
                                      ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)]
        L.o_build_links27.restype = ctypes.c_int64
        L.o_build_links39.argtypes = L.o_build_links27.argtypes
        L.o_build_links39.restype = ctypes.c_int64
        L.o_build_links39s.argtypes = L.o_build_links27.argtypes
        L.o_build_links39s.restype = ctypes.c_int64
        L.o27_step_box.argtypes = [ctypes.POINTER(OBox), ctypes.POINTER(ctypes.c_int32), dp, dp,
                                   ctypes.POINTER(BoxParams27), dp, ctypes.POINTER(OStats),
                                   ctypes.POINTER(BcStats)]
        c39 = ctypes.POINTER(Consts39)
        L.o39_consts_make.argtypes = [c39, D, D, D, D, D, I, I, I, I]
        L.o39_sutherland.argtypes = [D, c39, dp, dp]
        L.o39_sutherland.restype = D
        L.o39_therm_cond.argtypes = [D, D, c39, dp]
        L.o39_feq.argtypes = [dp, D, I, dp]
        L.o39_lu_step.argtypes = [dp, dp, dp, I]
        L.o39_lu_step.restype = D
        L.o39_lm_solve.argtypes = [D, dp, D, dp, dp, dp, I, I, I]
        L.o39_alpha_estimate.argtypes = [dp, dp, D, ctypes.c_int64, c39, ctypes.POINTER(I)]
        L.o39_alpha_estimate.restype = D
        L.o39_collide.argtypes = [dp, dp, dp, D, dp, D, D, ctypes.c_int64, c39, dp, dp]
        L.o39_qbar.argtypes = [dp, dp, D, D, I, dp]
        L.o39_little_qbar.argtypes = [dp, dp, D, D, I, dp]
        L.o39_moments.argtypes = [dp, dp, I, dp, dp, dp, dp, D]
        L.o39_damp_periodic.argtypes = [I, I, I, dp, I, I]
        L.o39_damp_periodic.restype = ctypes.c_int64
        L.o39_init_fields.argtypes = [I, I, I, dp, dp, dp, dp, dp, c39, dp, dp, dp,
                                      ctypes.POINTER(ctypes.c_int64)]
        L.o39_step_periodic.argtypes = [I, I, I, dp, dp, dp, dp, c39, ctypes.c_int64, I, I, I,
                                        ctypes.POINTER(Stats39)]
        ip = ctypes.POINTER(ctypes.c_int32)
        L.o39_damp_faces.argtypes = [I, I, I, ip, dp, I, I]
        L.o39_damp_faces.restype = ctypes.c_int64
        L.o39_step_ghost.argtypes = [I, I, I, I, dp, dp, dp, dp, c39, ctypes.c_int64, I, I, I,
                                     ctypes.POINTER(Stats39)]
        L.o39_source.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, I, I, I, I, ip, I,
                                 ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(I)]
        L.o39_source.restype = I
        L.o39_step_faces.argtypes = [I, I, I, ip, dp, dp, ip, dp, dp, dp, dp, dp, c39,
                                     ctypes.c_int64, I, I, I, ctypes.POINTER(Stats39)]
        L.o39_step_walls.argtypes = [I, I, I, ip, ip, dp, dp, ip, dp, dp, dp, dp, dp, c39,
                                     ctypes.c_int64, ctypes.c_int64, I, I,
                                     ctypes.POINTER(Stats39), ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_int64)]
        L.o39_wall_derivs.argtypes = [I, dp, dp, dp, ip, I, dp, dp]
        L.o39_wall_grad.argtypes = [I, D, dp, D, D, dp, dp, D, D, D, c39, dp, dp]
        L.o39_wall_rebuild.argtypes = [I, I, I, ip, ip, dp, dp, dp, ctypes.c_int64, dp, c39,
                                       ctypes.c_int64, ctypes.c_int64, dp, dp,
                                       ctypes.POINTER(I), ctypes.POINTER(I)]
        L.o39_wall_rebuild.restype = I
        L.o39p_set_slips.argtypes = [dp, ctypes.POINTER(I)]
        L.o39p_wall_neq.argtypes = [I, D, dp, dp, dp, D, D, D, dp, dp]
        for nm in ("o_tri_volume", "o_box_test"):
            getattr(L, nm).argtypes = [dp, dp, dp]
        L.o_link_code.argtypes = [dp, dp, D]
        L.o_link_code.restype = ctypes.c_int32
        _lib = L
    return _lib


REF_PATH = os.path.join(_DIR, "_ref", "libuclbs_ref.so")
_ref = None


def ref_available():
    """Whether the compiled reference routines exist, without loading them
    (test collection, also under `-m gpu`, must not map the library)."""
    return os.path.exists(REF_PATH)


def ref_lib():
    """The reference's own FP64 routines compiled from its sources (built by
    oracle/ref/Makefile where the reference tree is present), or None."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = ctypes.CDLL(REF_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        D, I = ctypes.c_double, ctypes.c_int
        R.ref_tri_volume.argtypes = [dp, dp, dp]
        R.ref_box_test.argtypes = [dp, dp, dp]
        R.ref_link_code.argtypes = [dp, dp, I, D]
        R.ref_link_code.restype = ctypes.c_int32
        R.ref_set_constants.argtypes = [D] * 6
        R.ref_sutherland.argtypes = [D]
        R.ref_sutherland.restype = D
        R.ref_therm_cond.argtypes = [D, D]
        R.ref_therm_cond.restype = D
        R.ref_lu_decomp.argtypes = [dp, dp, dp]
        R.ref_lu_decomp.restype = D
        _ref = R
    return _ref


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _table(name, n):
    return np.array((ctypes.c_int * n).in_dll(lib(), name)[:], dtype=np.int64)


def lattice27():
    L = lib()
    w = np.array((ctypes.c_double * 27).in_dll(L, "O27_W")[:])
    return (_table("O27_CX", 27), _table("O27_CY", 27), _table("O27_CZ", 27),
            _table("O27_OPP", 27), w)


def lattice39():
    return (_table("O39_CX", 39), _table("O39_CY", 39), _table("O39_CZ", 39),
            _table("O39_OPP", 39))


def philox(ctr, key):
    out = (ctypes.c_uint32 * 4)()
    lib().o_philox4x32_10((ctypes.c_uint32 * 4)(*ctr), (ctypes.c_uint32 * 2)(*key), out)
    return list(out)


def uniform(node, comp, seed):
    return lib().o_uniform(node, comp, seed)


def feq27(rho, ux, uy, uz):
    out = np.empty(27)
    lib().o27_feq(rho, ux, uy, uz, _dp(out))
    return out


def moments27(f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    r = ctypes.c_double()
    u = np.empty(3)
    lib().o27_moments(_dp(f), ctypes.byref(r), _dp(u))
    return r.value, u


def alpha27(f, feq, alpha0=0.0, tol=1e-10, maxit=30, series_order=0):
    f = np.ascontiguousarray(f, dtype=np.float64)
    feq = np.ascontiguousarray(feq, dtype=np.float64)
    if series_order:
        r = lib().o27_alpha_series(_dp(f), _dp(feq), alpha0, tol, maxit, series_order)
    else:
        r = lib().o27_alpha(_dp(f), _dp(feq), alpha0, tol, maxit)
    return r.alpha, r.iters, r.status, r.resid


def entropy_g27(f, feq, a):
    f = np.ascontiguousarray(f, dtype=np.float64)
    feq = np.ascontiguousarray(feq, dtype=np.float64)
    g, gp = ctypes.c_double(), ctypes.c_double()
    lib().o27_entropy_g(_dp(f), _dp(feq), a, ctypes.byref(g), ctypes.byref(gp))
    return g.value, gp.value


def collide27(f, beta, alpha0=0.0, tol=1e-10, maxit=30, fixed=False):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.empty(27)
    rho = ctypes.c_double()
    u = np.empty(3)
    r = lib().o27_collide_node(_dp(f), beta, alpha0, tol, maxit, int(fixed), _dp(out),
                               ctypes.byref(rho), _dp(u))
    return out, r.alpha, r.iters, r.status


def init_tgv27(shape, u0, noise_amp, seed):
    nx, ny, nz = shape
    f = np.empty((27, nx * ny * nz))
    lib().o27_init_tgv(nx, ny, nz, u0, noise_amp, seed, _dp(f))
    return f


def step_periodic27(shape, f, beta, nsteps=1, tol=1e-10, maxit=30, fixed=False,
                    nthreads=0, want_alpha=True):
    """Advance f (27, n) by nsteps fused steps; returns (f, alpha, stats)."""
    nx, ny, nz = shape
    n = nx * ny * nz
    a = np.ascontiguousarray(f, dtype=np.float64).reshape(27, n).copy()
    b = np.empty_like(a)
    alpha = np.empty(n) if want_alpha else None
    p = Params27(beta, tol, maxit, int(fixed), nthreads)
    st = OStats()
    for _ in range(nsteps):
        lib().o27_step_periodic(nx, ny, nz, _dp(a), _dp(b), ctypes.byref(p),
                                _dp(alpha) if want_alpha else None, ctypes.byref(st))
        a, b = b, a
    return a, alpha, st.as_dict()


def macro27(f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    n = f.shape[1]
    out = [np.empty(n) for _ in range(4)]
    lib().o27_macro(n, _dp(f), *[_dp(o) for o in out])
    return out


def make_box(shape, faces, dx=1.0, origin=(0.0, 0.0, 0.0)):
    b = OBox()
    b.nx, b.ny, b.nz = shape
    b.dx = dx
    b.origin = (ctypes.c_double * 3)(*origin)
    b.face_code = (ctypes.c_int32 * 6)(*faces)
    return b


def build_links27(shape, faces, tri, nrm, dx=1.0, origin=(0.0, 0.0, 0.0)):
    """Reference-form link table state (27, n) int32; returns (state, n_solid, n_tie)."""
    b = make_box(shape, faces, dx, origin)
    n = shape[0] * shape[1] * shape[2]
    st = np.empty((27, n), dtype=np.int32)
    tri = np.ascontiguousarray(tri, dtype=np.float64)
    nrm = np.ascontiguousarray(nrm, dtype=np.float64)
    ties = ctypes.c_int64()
    ns = lib().o_build_links27(ctypes.byref(b), len(tri), _dp(tri) if len(tri) else None,
                               _dp(nrm) if len(nrm) else None,
                               st.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                               ctypes.byref(ties))
    return st, int(ns), int(ties.value)


def build_links39(shape, faces, tri, nrm, dx=1.0, origin=(0.0, 0.0, 0.0)):
    """Reference-form D3Q39 link table (39, n) int32 on the plain lattice;
    returns (state, n_solid, n_tie)."""
    b = make_box(shape, faces, dx, origin)
    n = shape[0] * shape[1] * shape[2]
    st = np.empty((39, n), dtype=np.int32)
    tri = np.ascontiguousarray(tri, dtype=np.float64)
    nrm = np.ascontiguousarray(nrm, dtype=np.float64)
    ties = ctypes.c_int64()
    ns = lib().o_build_links39(ctypes.byref(b), len(tri), _dp(tri) if len(tri) else None,
                               _dp(nrm) if len(nrm) else None,
                               st.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                               ctypes.byref(ties))
    return st, int(ns), int(ties.value)


def build_links39s(shape, faces, tri, nrm, dx=1.0, origin=(0.0, 0.0, 0.0)):
    """Reference-form D3Q39 link table (39, n) int32 on the shifted lattice
    (every link keyed on its upstream node); returns (state, n_solid, n_tie)."""
    b = make_box(shape, faces, dx, origin)
    n = shape[0] * shape[1] * shape[2]
    st = np.empty((39, n), dtype=np.int32)
    tri = np.ascontiguousarray(tri, dtype=np.float64)
    nrm = np.ascontiguousarray(nrm, dtype=np.float64)
    ties = ctypes.c_int64()
    ns = lib().o_build_links39s(ctypes.byref(b), len(tri), _dp(tri) if len(tri) else None,
                                _dp(nrm) if len(nrm) else None,
                                st.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                ctypes.byref(ties))
    return st, int(ns), int(ties.value)


def step_box27(shape, faces, state, f, beta, nsteps=1, inlets=(), outlets=(), freestream=None,
               tol=1e-10, maxit=30, fixed=False, nthreads=0, want_alpha=True):
    """Advance f (27, n) on a box with link table `state`; returns (f, alpha, stats, bcstats).
    inlets / outlets / freestream are (rho, u, v, w) states tabulated with feq27."""
    b = make_box(shape, faces)
    n = shape[0] * shape[1] * shape[2]
    fin = np.array([feq27(*s_) for s_ in inlets]).reshape(-1) if len(inlets) else np.zeros(1)
    fou = np.array([feq27(*s_) for s_ in outlets]).reshape(-1) if len(outlets) else np.zeros(1)
    ffs = feq27(*freestream) if freestream is not None else None
    p = BoxParams27(beta, tol, maxit, int(fixed), nthreads, len(inlets), len(outlets),
                    _dp(fin), _dp(fou), _dp(ffs) if ffs is not None else None)
    st = np.ascontiguousarray(state, dtype=np.int32)
    a = np.ascontiguousarray(f, dtype=np.float64).reshape(27, n).copy()
    bb = np.empty_like(a)
    alpha = np.empty(n) if want_alpha else None
    ost, bst = OStats(), BcStats()
    for _ in range(nsteps):
        lib().o27_step_box(ctypes.byref(b), st.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                           _dp(a), _dp(bb), ctypes.byref(p),
                           _dp(alpha) if want_alpha else None, ctypes.byref(ost),
                           ctypes.byref(bst))
        a, bb = bb, a
    return a, alpha, ost.as_dict(), {"n_ibb_links": int(bst.n_ibb_links),
                                     "n_repair": int(bst.n_repair)}


def uniform27(state, rho, u):
    """feq(rho, u) on fluid nodes, rest weights on solid / null nodes."""
    n = state.shape[1]
    f = np.repeat(feq27(rho, *u)[:, None], n, axis=1)
    w = lattice27()[4]
    solid = state[0] < 0
    f[:, solid] = w[:, None]
    return f


def tri_volume(xc, e, t):
    return lib().o_tri_volume(_dp(np.ascontiguousarray(xc, np.float64)),
                              _dp(np.ascontiguousarray(e, np.float64)),
                              _dp(np.ascontiguousarray(t, np.float64)))


def box_test(p1, p2, bb):
    return lib().o_box_test(_dp(np.ascontiguousarray(p1, np.float64)),
                            _dp(np.ascontiguousarray(p2, np.float64)),
                            _dp(np.ascontiguousarray(bb, np.float64)))


def link_code(e1, xc, link):
    return int(lib().o_link_code(_dp(np.ascontiguousarray(e1, np.float64)),
                                 _dp(np.ascontiguousarray(xc, np.float64)), float(link)))


# ---- D3Q39 -----------------------------------------------------------------
def consts39(gama=1.4, temperature=300.0, pressure=101325.0, mach=0.3, dx=1e-3, startup=25,
             shifted=False, turbulent=True, lvl=0):
    c = Consts39()
    lib().o39_consts_make(ctypes.byref(c), gama, temperature, pressure, mach, dx, startup,
                          int(shifted), int(turbulent), lvl)
    return c


def velocities39(shifted=False):
    cx, cy, cz = lattice39()[:3]
    return cx + (1 if shifted else 0), cy, cz


def feq39(lm, rho, shifted=False):
    out = np.zeros(39)
    lib().o39_feq(_dp(np.ascontiguousarray(lm, np.float64)), float(rho), int(shifted), _dp(out))
    return out


def lu_step39(J, tgt, lm, ref_mode=False):
    """J: 13x13; tgt: rho, u, v, w, P(6), q(3); returns (new lm, max|delta|)."""
    J = np.ascontiguousarray(J, np.float64)
    tgt = np.ascontiguousarray(tgt, np.float64)
    lm = np.array(lm, np.float64)
    mx = lib().o39_lu_step(_dp(J), _dp(tgt), _dp(lm), int(ref_mode))
    return lm, mx


def lm_solve39(rho, u, T, P, lm0, shifted=False, ref_mode=False, maxit=50):
    lm = np.array(lm0, np.float64)
    e = np.zeros(39)
    it = lib().o39_lm_solve(float(rho), _dp(np.ascontiguousarray(u, np.float64)), float(T),
                            _dp(np.ascontiguousarray(P, np.float64)), _dp(lm), _dp(e),
                            int(shifted), int(ref_mode), maxit)
    return lm, e, it


def sutherland39(T, c):
    dm, dt_ = ctypes.c_double(), ctypes.c_double()
    mu = lib().o39_sutherland(float(T), ctypes.byref(c), ctypes.byref(dm), ctypes.byref(dt_))
    return mu, dm.value, dt_.value


def therm_cond39(rho, T, c):
    k = ctypes.c_double()
    rc = lib().o39_therm_cond(float(rho), float(T), ctypes.byref(c), ctypes.byref(k))
    return (k.value if rc == 0 else None)


def alpha_est39(f, diff, T, step, c):
    st = ctypes.c_int()
    a = lib().o39_alpha_estimate(_dp(np.ascontiguousarray(f, np.float64)),
                                 _dp(np.ascontiguousarray(diff, np.float64)), float(T), int(step),
                                 ctypes.byref(c), ctypes.byref(st))
    return a, st.value


def moments39(f, g, c):
    r, T = ctypes.c_double(), ctypes.c_double()
    u = np.zeros(3)
    P = np.zeros(6)
    lib().o39_moments(_dp(np.ascontiguousarray(f, np.float64)), _dp(np.ascontiguousarray(g, np.float64)),
                      c.shifted, ctypes.byref(r), _dp(u), ctypes.byref(T), _dp(P), c.cv)
    return r.value, u, T.value, P


def collide39(f, g, lm, rho, u, T, alpha, step, c):
    fo, go = np.zeros(39), np.zeros(39)
    lib().o39_collide(_dp(np.ascontiguousarray(f, np.float64)), _dp(np.ascontiguousarray(g, np.float64)),
                      _dp(np.ascontiguousarray(lm, np.float64)), float(rho),
                      _dp(np.ascontiguousarray(u, np.float64)), float(T), float(alpha), int(step),
                      ctypes.byref(c), _dp(fo), _dp(go))
    return fo, go


def qbar39(df, u, T, eps, shifted=False):
    """find_q_bar: the quasi-equilibrium correction of the deviation df."""
    out = np.zeros(39)
    lib().o39_qbar(_dp(np.ascontiguousarray(df, np.float64)), _dp(np.ascontiguousarray(u, np.float64)),
                   float(T), float(eps), int(shifted), _dp(out))
    return out


def little_qbar39(dg, u, T, eps, shifted=False):
    """find_little_q_bar: the energy-flux correction of the g deviation dg."""
    out = np.zeros(39)
    lib().o39_little_qbar(_dp(np.ascontiguousarray(dg, np.float64)),
                          _dp(np.ascontiguousarray(u, np.float64)), float(T), float(eps),
                          int(shifted), _dp(out))
    return out


def init39(shape, rho, ux, uy, uz, T, c):
    nx, ny, nz = shape
    n = nx * ny * nz
    f = np.zeros((39, n))
    g = np.zeros((39, n))
    lm = np.zeros((13, n))
    bad = ctypes.c_int64()
    arrs = [np.ascontiguousarray(np.broadcast_to(a, (n,)), np.float64) for a in (rho, ux, uy, uz, T)]
    lib().o39_init_fields(nx, ny, nz, *[_dp(a) for a in arrs], ctypes.byref(c), _dp(f), _dp(g),
                          _dp(lm), ctypes.byref(bad))
    assert bad.value == 0
    return f, g, lm


def step_periodic39(shape, f, g, lm, c, step0=0, nsteps=1, ref_lu=False, jacobi=True, keep_alpha=False):
    nx, ny, nz = shape
    f = np.array(f, np.float64)
    g = np.array(g, np.float64)
    lm = np.array(lm, np.float64)
    alpha = np.zeros(nx * ny * nz) if keep_alpha else None
    st = Stats39()
    lib().o39_step_periodic(nx, ny, nz, _dp(f), _dp(g), _dp(lm),
                            _dp(alpha) if alpha is not None else None, ctypes.byref(c), step0,
                            nsteps, int(ref_lu), int(jacobi), ctypes.byref(st))
    return f, g, lm, alpha, st.as_dict()


FACE_PERIODIC = 2147483647


def _faces39(faces, shifted):
    fc = np.ascontiguousarray(faces, np.int32)
    assert fc.shape == (6,)
    for k, v in enumerate(fc):
        assert v in (FACE_PERIODIC, -1000, -6, -10, -11) or -299 <= v <= -101, v
        assert (v == FACE_PERIODIC) == (fc[(k + 3) % 6] == FACE_PERIODIC)
        assert not (shifted and v == -11 and k % 3 == 0), "no inviscid x face on the shifted lattice"
    return fc


def freestream39(c):
    """fs_fieq / fs_gieq: the equilibrium of rho = 1, u = (Mach, 0, 0), T = 1/gamma
    (freestream_basis_calcs.f90:35-71, basic_fluids_calcs.f95:31-48)."""
    T = 1.0 / c.gama
    u = np.array([c.mach, 0.0, 0.0])
    P = np.array([T + u[0] ** 2, T, T, 0.0, 0.0, 0.0])
    lm0 = np.zeros(13)
    lm0[0] = -1.5 * np.log(2 * np.pi * T) - u[0] ** 2 / (2 * T)
    lm0[1] = u[0] / T
    lm0[4:7] = -1.0 / (2 * T)
    lm, _, it = lm_solve39(1.0, u, T, P, lm0, bool(c.shifted), maxit=200)
    assert it >= 0
    fe = feq39(lm, 1.0, bool(c.shifted))
    return fe, c.lead_g * T * fe


def bc_tables39(c, inlets=(), outlets=()):
    """Inlet / outlet equilibrium table of inlet_outlet_basis_calcs.f90, one row
    (39 f, then 39 g) per inlet, then per outlet; each item (rho, u, v, w, T,
    type).  Types 1, 2, 4, 7: the multiplier solve of the state from the
    reference's starting guesses (chi -2.5, pi_aa -0.3; shifted-lattice
    outlets zeta = u, lambda = u / 25); g of every row of the family is then
    (2 c_v - 3) T f with this state's T (the reference assigns the whole
    array).  Type 3: the freestream pair.  Type 6 (outlets): no row values."""
    fsf, fsg = freestream39(c)
    rows = np.zeros((max(1, len(inlets) + len(outlets)), 78))

    def family(items, base, outlet):
        for j, it in enumerate(items):
            rho, u, v, w, T, t = it
            r = rows[base + j]
            if t in (1, 2, 4, 7):
                lm0 = np.array([-2.5, 0, 0, 0, -0.3, -0.3, -0.3, 0, 0, 0, 0, 0, 0], float)
                if outlet and c.shifted:
                    lm0[1:4] = (u, v, w)
                    lm0[10:13] = (u / 25, v / 25, w / 25)
                P = rho * np.array([T + u * u, T + v * v, T + w * w, u * v, u * w, v * w])
                lm, _, itn = lm_solve39(rho, [u, v, w], T, P, lm0, bool(c.shifted), maxit=200)
                assert itn >= 0
                r[:39] = feq39(lm, rho, bool(c.shifted))
                for k in range(len(items)):
                    rows[base + k, 39:] = rows[base + k, :39] * c.lead_g * T
            elif t == 3:
                r[:39], r[39:] = fsf, fsg
    family(list(inlets), 0, False)
    family(list(outlets), len(inlets), True)
    return rows


def face_rows39(faces, inlets=(), outlets=()):
    """Table row each face reads (inlet_bc.f95:31-53, outlet_bc.f95:33-58):
    -2 freestream (type 3), -1 pure outflow (type 6) or no inlet / outlet,
    an outlet of type 2 the row of the inlet of its number."""
    out = np.full(6, -1, np.int32)
    nin = len(inlets)
    for f, v in enumerate(faces):
        if -199 <= v <= -101:
            j = -v - 100
            out[f] = -2 if inlets[j - 1][5] == 3 else j - 1
        elif -299 <= v <= -201:
            j = -v - 200
            t = outlets[j - 1][5]
            out[f] = -1 if t == 6 else -2 if t == 3 else (j - 1) if t == 2 else nin + j - 1
    return out


def step_faces39(shape, faces, f, g, lm, c, step0=0, nsteps=1, ref_lu=False, jacobi=True,
                 keep_alpha=False, inlets=(), outlets=()):
    """Steps of a box whose faces (x-, y-, z-, x+, y+, z+) carry FACE_PERIODIC,
    -1000 (freestream), -6 (pure outflow), -10 (bounce-back), -11 (inviscid),
    -100-j (inlet j) or -200-j (outlet j) with inlets / outlets given as
    (rho, u, v, w, T, type)."""
    nx, ny, nz = shape
    fc = _faces39(faces, bool(c.shifted))
    fsf, fsg = freestream39(c)
    f = np.array(f, np.float64)
    g = np.array(g, np.float64)
    lm = np.array(lm, np.float64)
    alpha = np.zeros(nx * ny * nz) if keep_alpha else None
    st = Stats39()
    rows = np.ascontiguousarray(face_rows39(faces, inlets, outlets))
    tab = np.ascontiguousarray(bc_tables39(c, inlets, outlets))
    ip = ctypes.POINTER(ctypes.c_int32)
    lib().o39_step_faces(nx, ny, nz, fc.ctypes.data_as(ip), _dp(fsf), _dp(fsg),
                         rows.ctypes.data_as(ip), _dp(tab), _dp(f), _dp(g), _dp(lm),
                         _dp(alpha) if alpha is not None else None, ctypes.byref(c), step0,
                         nsteps, int(ref_lu), int(jacobi), ctypes.byref(st))
    return f, g, lm, alpha, st.as_dict()


_IBB_SLIPS = None


def ibb_slips39():
    """The plain-lattice wall rule's slip terms (tests/golden/ibb39_plain.json,
    written from the reference's cases by tests/golden/make_ibb39_plain.py),
    handed to the C oracle: (flat (n, 8) table, start indices)."""
    global _IBB_SLIPS
    if _IBB_SLIPS is None:
        import json
        path = os.path.join(os.path.dirname(_DIR), "tests", "golden", "ibb39_plain.json")
        d = json.load(open(path))
        rows, start = [], [0]
        for part in ("slips_f", "slips_g"):
            for t in d[part]:
                rows += t
                start.append(len(rows))
        tab = np.ascontiguousarray(np.array(rows, np.float64).reshape(-1, 8))
        st = np.ascontiguousarray(start, np.int32)
        lib().o39p_set_slips(_dp(tab), st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        _IBB_SLIPS = (tab, st)
    return _IBB_SLIPS


def wall_neq39p(i, T, ut, du, dT, b1, b2, cv, slips=True):
    """F_i, G_i of the plain-lattice wall rule (direction i, 1..38): the
    Hermite projection plus the case's slips (slips=False: the projection
    alone).  du[a, b] = d u_b / d x_a."""
    if slips:
        ibb_slips39()
    else:
        lib().o39p_set_slips(None, None)
    F, G = ctypes.c_double(), ctypes.c_double()
    lib().o39p_wall_neq(int(i), float(T), _dp(np.ascontiguousarray(ut, np.float64)),
                        _dp(np.ascontiguousarray(du, np.float64)),
                        _dp(np.ascontiguousarray(dT, np.float64)), float(b1), float(b2),
                        float(cv), ctypes.byref(F), ctypes.byref(G))
    if not slips:
        global _IBB_SLIPS
        _IBB_SLIPS = None
    return F.value, G.value


def step_walls39(shape, faces, state, f, g, lm, c, step0=0, nsteps=1, lvl_step0=None,
                 jacobi=True, keep_alpha=False, inlets=(), outlets=()):
    """Steps of a box with a link table: on the shifted lattice (table of
    build_links39s) wall links rebuilt by the reference's shifted-lattice wall
    rule, on the plain lattice (table of build_links39) by its plain-lattice
    rule.  Returns (f, g, lm, alpha, stats, n_wall_links, n_wall_repairs)."""
    if not c.shifted:
        ibb_slips39()
    nx, ny, nz = shape
    fc = _faces39(faces, bool(c.shifted))
    fsf, fsg = freestream39(c)
    f = np.array(f, np.float64)
    g = np.array(g, np.float64)
    lm = np.array(lm, np.float64)
    st_ = np.ascontiguousarray(state, np.int32)
    alpha = np.zeros(nx * ny * nz) if keep_alpha else None
    st = Stats39()
    nwl, nwr = ctypes.c_int64(), ctypes.c_int64()
    ip = ctypes.POINTER(ctypes.c_int32)
    rows = np.ascontiguousarray(face_rows39(faces, inlets, outlets))
    tab = np.ascontiguousarray(bc_tables39(c, inlets, outlets))
    lib().o39_step_walls(nx, ny, nz, fc.ctypes.data_as(ip), st_.ctypes.data_as(ip), _dp(fsf),
                         _dp(fsg), rows.ctypes.data_as(ip), _dp(tab), _dp(f), _dp(g), _dp(lm),
                         _dp(alpha) if alpha is not None else None, ctypes.byref(c), step0,
                         step0 if lvl_step0 is None else lvl_step0, nsteps, int(jacobi),
                         ctypes.byref(st), ctypes.byref(nwl), ctypes.byref(nwr))
    return f, g, lm, alpha, st.as_dict(), nwl.value, nwr.value


def wall_derivs39(i, ut, su, sT, scode, shifted=True):
    """Gradients of wall link i from the 7-point stencil (su (7, 3), sT (7,),
    node codes scode (7,)): returns (du (3, 3) with du[a, b] = d u_b / d x_a,
    dT (3,))."""
    du = np.zeros((3, 3))
    dT = np.zeros(3)
    lib().o39_wall_derivs(int(i), _dp(np.ascontiguousarray(ut, np.float64)),
                          _dp(np.ascontiguousarray(su, np.float64)),
                          _dp(np.ascontiguousarray(sT, np.float64)),
                          np.ascontiguousarray(scode, np.int32).ctypes.data_as(
                              ctypes.POINTER(ctypes.c_int32)), int(shifted), _dp(du), _dp(dT))
    return du, dT


def wall_grad39(i, rho_t, ut, Tt, feq_i, du, dT, eps, b1, b2, c):
    """Rebuilt (f, g) of wall link i before the magnitude repair."""
    fo, go = np.zeros(1), np.zeros(1)
    lib().o39_wall_grad(int(i), float(rho_t), _dp(np.ascontiguousarray(ut, np.float64)),
                        float(Tt), float(feq_i), _dp(np.ascontiguousarray(du, np.float64)),
                        _dp(np.ascontiguousarray(dT, np.float64)), float(eps), float(b1),
                        float(b2), ctypes.byref(c), _dp(fo), _dp(go))
    return float(fo[0]), float(go[0])


def wall_rebuild39(shape, faces, state, f, g, lm, k, fi, c, step, lvl_step):
    """Wall-link populations of fluid node k rebuilt from the input buffers
    f, g, lm and its pulled populations fi: (fw, gw, n_repair, noconv, kerr);
    entries of non-wall links are left at NaN."""
    nx, ny, nz = shape
    fc = _faces39(faces, True)
    fw = np.full(39, np.nan)
    gw = np.full(39, np.nan)
    nc, ke = ctypes.c_int(), ctypes.c_int()
    ip = ctypes.POINTER(ctypes.c_int32)
    st = np.ascontiguousarray(state, np.int32)
    nrep = lib().o39_wall_rebuild(nx, ny, nz, fc.ctypes.data_as(ip), st.ctypes.data_as(ip),
                                  _dp(np.ascontiguousarray(f, np.float64)),
                                  _dp(np.ascontiguousarray(g, np.float64)),
                                  _dp(np.ascontiguousarray(lm, np.float64)), int(k),
                                  _dp(np.ascontiguousarray(fi, np.float64)), ctypes.byref(c),
                                  int(step), int(lvl_step), _dp(fw), _dp(gw), ctypes.byref(nc),
                                  ctypes.byref(ke))
    return fw, gw, nrep, nc.value, ke.value


def source39(node, k, shape, faces, shifted=False):
    """(None, None) for the freestream value, else (source node (x, y, z), direction)."""
    nx, ny, nz = shape
    fc = _faces39(faces, shifted)
    src, d = ctypes.c_int64(), ctypes.c_int()
    fs = lib().o39_source(node[0], node[1], node[2], k, nx, ny, nz,
                          fc.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), int(shifted),
                          ctypes.byref(src), ctypes.byref(d))
    if fs:
        return None, None
    v = src.value
    return (v % nx, (v // nx) % ny, v // (nx * ny)), d.value


def step_ghost39(shape, ghost, f, g, lm, c, step0=0, nsteps=1, ref_lu=False, jacobi=True,
                 keep_alpha=False):
    """Steps of a refined level stored with a ghost layer of depth `ghost`:
    only the valid nodes inside it are updated."""
    nx, ny, nz = shape
    f = np.array(f, np.float64)
    g = np.array(g, np.float64)
    lm = np.array(lm, np.float64)
    alpha = np.zeros(nx * ny * nz) if keep_alpha else None
    st = Stats39()
    lib().o39_step_ghost(nx, ny, nz, int(ghost), _dp(f), _dp(g), _dp(lm),
                         _dp(alpha) if alpha is not None else None, ctypes.byref(c), step0, nsteps,
                         int(ref_lu), int(jacobi), ctypes.byref(st))
    return f, g, lm, alpha, st.as_dict()


def coarse_to_fine39(cprev_f, cnew_f, cprev_g, cnew_g, clm, cshape, corigin, fshape, forigin,
                     theta, c, nodes=None):
    """Fine f, g, multipliers at the stored fine nodes (all if None) from the
    coarse level (interlevel_interpolation.f95:114-493): f, g, rho, T
    interpolated in space and time, the multipliers in space; a negative or NaN
    f or g becomes the equilibrium of the interpolated coarse state
    (:417-431).  Returns (f, g, lm, n_repaired)."""
    import o_amr
    args = (cshape, corigin, fshape, forigin, theta, nodes)
    fi = o_amr.coarse_to_fine(cprev_f, cnew_f, *args)
    gi = o_amr.coarse_to_fine(cprev_g, cnew_g, *args)
    lmi = o_amr.coarse_to_fine(clm, clm, *args)

    def rho_t(f, g):
        rho, ux, uy, uz, T = macro39(f, g, c)
        return np.stack([rho, T])
    rt = o_amr.coarse_to_fine(rho_t(cprev_f, cprev_g), rho_t(cnew_f, cnew_g), *args)
    bad = ~((fi >= 0) & (gi >= 0))
    nrep = int(bad.sum())
    for k in np.nonzero(bad.any(0))[0]:
        e = feq39(lmi[:, k], rt[0, k], bool(c.shifted))
        m = bad[:, k]
        fi[m, k] = e[m]
        gi[m, k] = c.lead_g * rt[1, k] * e[m]
    return fi, gi, lmi, nrep


def ghost_nodes39(fshape, ghost=4):
    nx, ny, nz = fshape
    k = np.arange(nx * ny * nz)
    x, y, z = k % nx, (k // nx) % ny, k // (nx * ny)
    g = (x < ghost) | (y < ghost) | (z < ghost) | (x >= nx - ghost) | (y >= ny - ghost) | \
        (z >= nz - ghost)
    return np.nonzero(g)[0]


def damp_faces39(shape, faces, fo, shifted=False, jacobi=True):
    nx, ny, nz = shape
    fc = _faces39(faces, shifted)
    fo = np.array(fo, np.float64)
    nrep = lib().o39_damp_faces(nx, ny, nz, fc.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                _dp(fo), int(shifted), int(jacobi))
    return fo, nrep


def macro39(f, g, c):
    cx, cy, cz = velocities39(bool(c.shifted))
    rho = f.sum(0)
    ux = (cx[:, None] * f).sum(0) / rho
    uy = (cy[:, None] * f).sum(0) / rho
    uz = (cz[:, None] * f).sum(0) / rho
    cm = (cx ** 2 + cy ** 2 + cz ** 2)[:, None]
    T = ((cm * f).sum(0) + g.sum(0) - rho * (ux ** 2 + uy ** 2 + uz ** 2)) / (2 * c.cv * rho)
    return rho, ux, uy, uz, T
