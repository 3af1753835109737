"""Python host binding of libuclbs_hip.so (ctypes over the C ABI in include/uclbs_hip.h).

The production host of this library is Fortran (fortran/uclbs_hip_mod.f90 binds the
same symbols through iso_c_binding, as a drop-in for the phase calls of
looping/timestepping/loop.f95:104-139).  This module binds the identical C ABI
for tests and the benchmark.  There is no CPU fallback: if the shared library
is missing or a call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG_DIR, "libuclbs_hip.so")

# exported symbols, in the order include/uclbs_hip.h declares them
EXPORTS = (
    "uclbs_last_error", "uclbs_abi_version", "uclbs_create", "uclbs_destroy",
    "uclbs_define_level", "uclbs_upload_f", "uclbs_download_f", "uclbs_download_macro",
    "uclbs_set_keep_alpha", "uclbs_download_alpha", "uclbs_init_tgv",
    "uclbs_step_level", "uclbs_step_level_async", "uclbs_time_steps", "uclbs_sync",
    "uclbs_get_stats", "uclbs_reset_stats", "uclbs_set_faces", "uclbs_set_bc_states",
    "uclbs_build_links", "uclbs_set_links", "uclbs_download_links", "uclbs_link_counts",
    "uclbs_init_uniform", "uclbs_geom_icosphere", "uclbs_geom_wedge", "uclbs_download_macro39", "uclbs_checkpoint_write",
    "uclbs_checkpoint_read", "uclbs_set_bc_states39", "uclbs_import_fab", "uclbs_export_fab", "uclbs_time_kernels",
    "uclbs_copy_bandwidth", "uclbs_set_gas", "uclbs_init_fields", "uclbs_upload_state",
    "uclbs_download_state", "uclbs_set_timestep", "uclbs_mlp_init", "uclbs_mlp_set_params",
    "uclbs_mlp_get_params", "uclbs_mlp_loss_grad", "uclbs_mlp_forward", "uclbs_mlp_reserve",
    "uclbs_mlp_collect", "uclbs_mlp_download_samples", "uclbs_mlp_download_sample_nodes",
    "uclbs_mlp_train", "uclbs_interlevel",
    "uclbs_init_from_coarse", "uclbs_fine_to_coarse", "uclbs_define_level_boxes",
    "uclbs_advance", "uclbs_autotune", "uclbs_mlp_forward_noalloc", "uclbs_set_placement",
    "uclbs_plot3d_write", "uclbs_output_wait", "uclbs_mic_setup", "uclbs_mic_sample",
    "uclbs_mic_flush", "uclbs_mic_weights", "uclbs_regrid_level", "uclbs_regrid_copied",
    "uclbs_clear_level", "uclbs_tag_ggm", "uclbs_cluster_tags",
)
MLP_PARAMS = 9217
FACE_PERIODIC = 2147483647

LATTICE = {"d3q27": 27, "d3q39": 39}
PRECISION = {"fp32s": 0, "fp64": 1}
ALPHA_MODE = {"newton": 0, "fixed2": 1, "learned": 2}


class UclbsError(RuntimeError):
    pass


class Gas(ctypes.Structure):
    """uclbs_gas: the D3Q39 gas and freestream (namelist gama, temperature,
    pressure, Mach, dx, startup_timesteps, turbulent)."""
    _fields_ = [("gama", ctypes.c_double), ("temperature", ctypes.c_double),
                ("pressure", ctypes.c_double), ("mach", ctypes.c_double), ("dx", ctypes.c_double),
                ("startup_timesteps", ctypes.c_int32), ("turbulent", ctypes.c_int32),
                ("shifted", ctypes.c_int32), ("lm_maxit", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 4)]


class Plot3DHeader(ctypes.Structure):
    """uclbs_plot3d_header: the values of each solution block's header record
    (plot3d_output.f95:256-258), the timestep naming the files, the last level
    written and the Fortran subrecord length (0: gfortran's default)."""
    _fields_ = [(k, ctypes.c_double) for k in ("mach", "reynolds", "time", "gamma", "t_ref",
                                               "stag_temperature", "cv", "cp", "gas_r")] + [
        ("timestep", ctypes.c_int32), ("max_level", ctypes.c_int32),
        ("max_subrecord", ctypes.c_int64)]


class Config(ctypes.Structure):
    _fields_ = [
        ("lattice", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("alpha_mode", ctypes.c_int32),
        ("alpha_maxit", ctypes.c_int32),
        ("alpha_tol", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("reserved", ctypes.c_int32 * 8),
    ]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "node_updates", "newton_iters", "n_trivial", "n_bound", "n_nonpos", "n_noconv",
        "n_fallback", "n_logpath", "n_ibb_links", "n_repair")]

    def as_dict(self) -> dict:
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


_lib: Optional[ctypes.CDLL] = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise UclbsError(f"{path} is missing: run __graft_entry__.build() (make -C <pkg>)")
    lib = ctypes.CDLL(path)
    P, I, I64, D, U64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_uint64
    i3 = ctypes.POINTER(ctypes.c_int32)
    dp = ctypes.POINTER(ctypes.c_double)
    fp = ctypes.POINTER(ctypes.c_float)
    sig = {
        "uclbs_last_error": (ctypes.c_char_p, []),
        "uclbs_abi_version": (I, []),
        "uclbs_create": (I, [ctypes.POINTER(Config), ctypes.POINTER(P)]),
        "uclbs_destroy": (I, [P]),
        "uclbs_define_level": (I, [P, I, i3, i3, i3]),
        "uclbs_upload_f": (I, [P, I, dp, I64]),
        "uclbs_download_f": (I, [P, I, dp, I64]),
        "uclbs_download_macro": (I, [P, I, dp, dp, dp, dp]),
        "uclbs_set_keep_alpha": (I, [P, I]),
        "uclbs_download_alpha": (I, [P, I, dp]),
        "uclbs_init_tgv": (I, [P, I, D, D, U64]),
        "uclbs_step_level": (I, [P, I, I]),
        "uclbs_step_level_async": (I, [P, I, I]),
        "uclbs_time_steps": (I, [P, I, I, dp]),
        "uclbs_sync": (I, [P]),
        "uclbs_get_stats": (I, [P, ctypes.POINTER(Stats)]),
        "uclbs_reset_stats": (I, [P]),
        "uclbs_set_faces": (I, [P, I, i3]),
        "uclbs_set_bc_states": (I, [P, I, dp, I, dp, dp]),
        "uclbs_build_links": (I, [P, I, dp, D, I64, dp, dp, ctypes.POINTER(I64), ctypes.POINTER(I64)]),
        "uclbs_set_links": (I, [P, I, i3, I64]),
        "uclbs_download_links": (I, [P, I, i3, I64]),
        "uclbs_link_counts": (I, [P, I, ctypes.POINTER(I64), ctypes.POINTER(I64)]),
        "uclbs_init_uniform": (I, [P, I, D, D, D, D]),
        "uclbs_geom_icosphere": (I, [I, D, dp, U64, dp, dp, I64, ctypes.POINTER(I64)]),
        "uclbs_geom_wedge": (I, [D, D, D, D, D, D, I, I, U64, dp, dp, I64, ctypes.POINTER(I64)]),
        "uclbs_download_macro39": (I, [P, I, dp, dp, dp, dp, dp, dp, I64]),
        "uclbs_checkpoint_write": (I, [P, ctypes.c_char_p]),
        "uclbs_set_bc_states39": (I, [P, I, dp, i3, I, dp, i3]),
        "uclbs_import_fab": (I, [P, I, I, ctypes.c_void_p, I, I64]),
        "uclbs_export_fab": (I, [P, I, I, ctypes.c_void_p, I, I64]),
        "uclbs_checkpoint_read": (I, [P, ctypes.c_char_p]),
        "uclbs_time_kernels": (I, [P, I, I, dp, dp]),
        "uclbs_copy_bandwidth": (I, [I64, I, dp]),
        "uclbs_set_gas": (I, [P, ctypes.POINTER(Gas)]),
        "uclbs_init_fields": (I, [P, I, dp, dp, dp, dp, dp]),
        "uclbs_upload_state": (I, [P, I, dp, dp, dp, I64]),
        "uclbs_download_state": (I, [P, I, dp, dp, dp, I64]),
        "uclbs_set_timestep": (I, [P, I, I64]),
        "uclbs_mlp_init": (I, [P, U64]),
        "uclbs_mlp_set_params": (I, [P, fp, ctypes.c_float, ctypes.c_float]),
        "uclbs_mlp_get_params": (I, [P, fp, fp, fp]),
        "uclbs_mlp_loss_grad": (I, [P, fp, fp, I, fp, dp]),
        "uclbs_mlp_forward": (I, [P, fp, I, I, fp]),
        "uclbs_mlp_forward_noalloc": (I, [P, fp, I, fp]),
        "uclbs_mlp_reserve": (I, [P, I64]),
        "uclbs_mlp_collect": (I, [P, I, I, ctypes.POINTER(I64)]),
        "uclbs_mlp_download_samples": (I, [P, fp, fp, I64, ctypes.POINTER(I64)]),
        "uclbs_define_level_boxes": (I, [P, I, I, ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_int32)]),
        "uclbs_advance": (I, [P, I, I]),
        "uclbs_autotune": (I, [P, I, ctypes.POINTER(I)]),
        "uclbs_mlp_download_sample_nodes": (I, [P, ctypes.POINTER(ctypes.c_uint32), I64,
                                                ctypes.POINTER(I64)]),
        "uclbs_mlp_train": (I, [P, I64, I, D, U64, dp, dp]),
        "uclbs_interlevel": (I, [P, I, D]),
        "uclbs_init_from_coarse": (I, [P, I]),
        "uclbs_regrid_level": (I, [P, I, I, ctypes.POINTER(ctypes.c_int32),
                                   ctypes.POINTER(ctypes.c_int32)]),
        "uclbs_regrid_copied": (I, [P, ctypes.POINTER(I64)]),
        "uclbs_clear_level": (I, [P, I]),
        "uclbs_tag_ggm": (I, [P, I, I, I, D, dp, ctypes.POINTER(ctypes.c_uint8),
                              ctypes.POINTER(I64)]),
        "uclbs_cluster_tags": (I, [ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_int32),
                                   ctypes.POINTER(ctypes.c_int32), D, I, I, I, I,
                                   ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                   ctypes.POINTER(I)]),
        "uclbs_fine_to_coarse": (I, [P, I]),
        "uclbs_set_placement": (I, [P, dp, dp]),
        "uclbs_plot3d_write": (I, [P, ctypes.c_char_p, ctypes.POINTER(Plot3DHeader)]),
        "uclbs_output_wait": (I, [P]),
        "uclbs_mic_setup": (I, [P, ctypes.c_char_p, I, dp, i3, i3, i3]),
        "uclbs_mic_sample": (I, [P]),
        "uclbs_mic_flush": (I, [P]),
        "uclbs_mic_weights": (I, [P, I, i3, dp, ctypes.POINTER(I64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def icosphere(level: int, radius: float, center, seed: int):
    """Triangle soup (tri (T,9), nrm (T,3)) of a randomly rotated icosphere."""
    lib = load_library()
    nt = ctypes.c_int64()
    c3 = np.ascontiguousarray(center, dtype=np.float64)
    rc = lib.uclbs_geom_icosphere(level, radius, _dptr(c3), seed, None, None, 0, ctypes.byref(nt))
    if rc:
        raise UclbsError(f"icosphere size query failed ({rc})")
    tri = np.empty((nt.value, 9))
    nrm = np.empty((nt.value, 3))
    rc = lib.uclbs_geom_icosphere(level, radius, _dptr(c3), seed, _dptr(tri), _dptr(nrm),
                                  nt.value, ctypes.byref(nt))
    if rc:
        raise UclbsError(f"icosphere failed ({rc})")
    return tri, nrm


def wedge(apex, length: float, half_angle_deg: float, z_span, n_len: int, n_span: int,
          seed: int):
    """Triangle soup (tri (T,9), nrm (T,3)) of a symmetric wedge extruded along
    z with a seeded triangulation of its faces (uclbs_geom_wedge)."""
    lib = load_library()
    nt = ctypes.c_int64()
    args = (float(apex[0]), float(apex[1]), float(length), float(half_angle_deg),
            float(z_span[0]), float(z_span[1]), int(n_len), int(n_span), int(seed))
    rc = lib.uclbs_geom_wedge(*args, None, None, 0, ctypes.byref(nt))
    if rc:
        raise UclbsError(f"wedge size query failed ({rc})")
    tri = np.empty((nt.value, 9))
    nrm = np.empty((nt.value, 3))
    rc = lib.uclbs_geom_wedge(*args, _dptr(tri), _dptr(nrm), nt.value, ctypes.byref(nt))
    if rc:
        raise UclbsError(f"wedge failed ({rc})")
    return tri, nrm


def copy_bandwidth(nbytes: int, reps: int = 10) -> float:
    """GB/s of a plain device copy of nbytes (read + write bytes / time)."""
    lib = load_library()
    g = ctypes.c_double()
    rc = lib.uclbs_copy_bandwidth(int(nbytes), reps, ctypes.byref(g))
    if rc:
        raise UclbsError(f"copy_bandwidth failed ({rc})")
    return g.value


def cluster_tags(tags: np.ndarray, lo=(0, 0, 0), efficiency: float = 0.7, blocking: int = 1,
                 buffer: int = 0, max_size: int = 0):
    """uclbs_cluster_tags: boxes [(lo, hi), ...] (inclusive) over the nonzero
    entries of tags [z, y, x], whose node (0, 0, 0) sits at position lo."""
    lib = load_library()
    t = np.ascontiguousarray(tags, dtype=np.uint8)
    nz, ny, nx = t.shape
    lo3 = np.ascontiguousarray(lo, np.int32)
    hi3 = np.ascontiguousarray([lo[0] + nx - 1, lo[1] + ny - 1, lo[2] + nz - 1], np.int32)
    cap = 64
    while True:
        blo = np.zeros(3 * cap, np.int32)
        bhi = np.zeros(3 * cap, np.int32)
        nb = ctypes.c_int()
        rc = lib.uclbs_cluster_tags(t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _iptr(lo3),
                                    _iptr(hi3), float(efficiency), int(blocking), int(buffer),
                                    int(max_size), cap, _iptr(blo), _iptr(bhi), ctypes.byref(nb))
        if rc == 0:
            break
        if nb.value <= cap:
            raise UclbsError(f"cluster_tags failed ({rc}): {lib.uclbs_last_error().decode()}")
        cap = nb.value
    return [(tuple(int(v) for v in blo[3 * b:3 * b + 3]), tuple(int(v) for v in bhi[3 * b:3 * b + 3]))
            for b in range(nb.value)]


class Solver:
    """One library context (one device stream, resident level state)."""

    def __init__(self, lattice: str = "d3q27", precision: str = "fp32s",
                 alpha_mode: str = "newton", nu: float = 1e-3, alpha_tol: float = 1e-10,
                 alpha_maxit: int = 30):
        self.lib = load_library()
        cfg = Config()
        cfg.lattice = LATTICE[lattice]
        cfg.precision = PRECISION[precision]
        cfg.alpha_mode = ALPHA_MODE[alpha_mode]
        cfg.alpha_maxit = alpha_maxit
        cfg.alpha_tol = alpha_tol
        cfg.nu = nu
        self.q = cfg.lattice
        self.ctx = ctypes.c_void_p()
        self._check(self.lib.uclbs_create(ctypes.byref(cfg), ctypes.byref(self.ctx)))
        self.shapes = {}

    def _check(self, rc: int):
        if rc != 0:
            msg = self.lib.uclbs_last_error().decode(errors="replace")
            raise UclbsError(f"uclbs call failed ({rc}): {msg}")

    def close(self):
        if self.ctx:
            self.lib.uclbs_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def define_level(self, lvl: int, shape: Sequence[int], periodic=(1, 1, 1), lo=(0, 0, 0)):
        """shape: valid nodes; a refined level (lvl >= 1, periodic (0,0,0)) is
        stored with a ghost layer (one node on D3Q27, four on D3Q39), so its
        arrays have shape + 2 (D3Q27) or shape + 8 (D3Q39)."""
        lo3 = (ctypes.c_int32 * 3)(*lo)
        hi3 = (ctypes.c_int32 * 3)(*[lo[d] + shape[d] - 1 for d in range(3)])
        per = (ctypes.c_int32 * 3)(*[int(p) for p in periodic])
        self._check(self.lib.uclbs_define_level(self.ctx, lvl, lo3, hi3, per))
        g = (8 if self.q == 39 else 2) if lvl > 0 else 0
        self.shapes[lvl] = tuple(int(s) + g for s in shape)
        self.places = getattr(self, "places", {})
        self.places[lvl] = (tuple(int(v) for v in lo), tuple(int(v) for v in shape),
                            tuple(int(p) for p in periodic))

    def define_level_boxes(self, lvl: int, boxes):
        """boxes: [(lo, hi), ...] valid nodes inclusive in the level's index
        space.  The level's arrays are in stored order: each box with a one-node
        ghost shell, boxes one after another (see box_layout)."""
        lo = np.ascontiguousarray([b[0] for b in boxes], np.int32).reshape(-1)
        hi = np.ascontiguousarray([b[1] for b in boxes], np.int32).reshape(-1)
        pi = ctypes.POINTER(ctypes.c_int32)
        self._check(self.lib.uclbs_define_level_boxes(self.ctx, lvl, len(boxes),
                                                      lo.ctypes.data_as(pi), hi.ctypes.data_as(pi)))
        self._set_boxes(lvl, boxes)

    def _set_boxes(self, lvl: int, boxes):
        self.boxes = getattr(self, "boxes", {})
        self.boxes[lvl] = [(tuple(int(v) for v in b[0]), tuple(int(v) for v in b[1])) for b in boxes]
        n = sum(int(np.prod([h - l + 3 for l, h in zip(*b)])) for b in self.boxes[lvl])
        self.shapes[lvl] = (n, 1, 1)
        getattr(self, "places", {}).pop(lvl, None)

    # ---- regrid ----------------------------------------------------------------
    def regrid_level(self, lvl: int, boxes) -> int:
        """uclbs_regrid_level: level lvl remade on boxes [(lo, hi), ...] (AMReX
        remake); returns the number of nodes taken from the old level.  Finer
        levels become stale until regridded or cleared; a level that had a link
        table needs build_links again before it steps."""
        lo = np.ascontiguousarray([b[0] for b in boxes], np.int32).reshape(-1)
        hi = np.ascontiguousarray([b[1] for b in boxes], np.int32).reshape(-1)
        pi = ctypes.POINTER(ctypes.c_int32)
        self._check(self.lib.uclbs_regrid_level(self.ctx, lvl, len(boxes), lo.ctypes.data_as(pi),
                                                hi.ctypes.data_as(pi)))
        self._set_boxes(lvl, boxes)
        n = ctypes.c_int64()
        self._check(self.lib.uclbs_regrid_copied(self.ctx, ctypes.byref(n)))
        return n.value

    def clear_level(self, lvl: int):
        """uclbs_clear_level: drop level lvl and every finer level."""
        self._check(self.lib.uclbs_clear_level(self.ctx, lvl))
        for d in (self.shapes, getattr(self, "boxes", {}), getattr(self, "places", {})):
            for k in [k for k in d if k >= lvl]:
                del d[k]

    def tag_ggm(self, lvl: int, quantity: int = 1, scheme: int = 0, limit: float = 0.04):
        """uclbs_tag_ggm on a one-box level: (ggm [n], tags [n] uint8, count),
        both in the level's stored order."""
        n = self.n_nodes(lvl)
        ggm = np.empty(n)
        tags = np.empty(n, np.uint8)
        cnt = ctypes.c_int64()
        self._check(self.lib.uclbs_tag_ggm(self.ctx, lvl, quantity, scheme, float(limit), _dptr(ggm),
                                           tags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                           ctypes.byref(cnt)))
        return ggm, tags, cnt.value

    def refine_by_ggm(self, lvl: int, quantity: int = 1, scheme: int = 0, limit: float = 0.04,
                      efficiency: float = 0.7, blocking: int = 2, buffer: int = 2,
                      max_size: int = 32):
        """One regrid of level lvl (>= 1) from the GGM tags of level lvl-1
        (regrid_question_mark.f95:38-184 through AMReX's regrid): tag, cluster
        where the new level's ghost shell nests in level lvl-1's valid nodes,
        refine the boxes by 2, then regrid level lvl (define and fill it from
        level lvl-1 when it does not exist yet).  Level lvl-1 of one box:
        clustered within its valid nodes one node in from its non-periodic
        faces.  Level lvl-1 of many boxes (refined): tags kept on valid nodes
        whose 26 neighbours are valid nodes too, and a clustered box that
        leaves them is clustered again at efficiency 1 (boxes of tags only;
        the buffer grows the tags before they are kept to those nodes).
        No tags: level lvl and finer are cleared.  Returns the new boxes
        (level lvl's index space)."""
        places, mboxes = getattr(self, "places", {}), getattr(self, "boxes", {})
        _, tags, cnt = self.tag_ggm(lvl - 1, quantity, scheme, limit)
        if lvl - 1 in places:
            clo, cshape, cper = places[lvl - 1]
            g = 2 if lvl - 1 > 0 else 0
            st = tags.reshape(tuple(v + g for v in cshape)[::-1])
            if g:
                st = st[1:-1, 1:-1, 1:-1]
            inset = [0 if cper[d] else 1 for d in range(3)]
            rlo = [clo[d] + inset[d] for d in range(3)]
            region = np.ascontiguousarray(st[inset[2]:cshape[2] - inset[2],
                                             inset[1]:cshape[1] - inset[1],
                                             inset[0]:cshape[0] - inset[0]])
            boxes = cluster_tags(region, rlo, efficiency=efficiency, blocking=blocking,
                                 buffer=buffer, max_size=max_size) if region.any() else []
        elif lvl - 1 in mboxes and lvl - 1 >= 1:
            vlo = np.min([b[0] for b in mboxes[lvl - 1]], axis=0)
            vhi = np.max([b[1] for b in mboxes[lvl - 1]], axis=0)
            ext = vhi - vlo + 1
            valid = np.zeros(tuple(ext[::-1]), bool)
            t = np.zeros(tuple(ext[::-1]), np.uint8)
            for off, e, lo in self.box_layout(lvl - 1):
                blk = tags[off:off + e[0] * e[1] * e[2]].reshape(e[2], e[1], e[0])[1:-1, 1:-1, 1:-1]
                z0, y0, x0 = lo[2] - vlo[2], lo[1] - vlo[1], lo[0] - vlo[0]
                sl = (slice(z0, z0 + e[2] - 2), slice(y0, y0 + e[1] - 2), slice(x0, x0 + e[0] - 2))
                valid[sl] = True
                t[sl] = blk
            inner = valid.copy()
            pad = np.pad(valid, 1)
            for dz in (0, 1, 2):
                for dy in (0, 1, 2):
                    for dx in (0, 1, 2):
                        inner &= pad[dz:dz + ext[2], dy:dy + ext[1], dx:dx + ext[0]]
            # the buffer grows the tags (a cube) before they are kept to `inner`
            for ax in range(3):
                grown = t.astype(bool)
                for o in range(1, buffer + 1):
                    lo_sl = [slice(None)] * 3
                    hi_sl = [slice(None)] * 3
                    lo_sl[2 - ax], hi_sl[2 - ax] = slice(o, None), slice(None, -o)
                    grown[tuple(lo_sl)] |= t[tuple(hi_sl)].astype(bool)
                    grown[tuple(hi_sl)] |= t[tuple(lo_sl)].astype(bool)
                t = grown.astype(np.uint8)
            t = np.ascontiguousarray(t * inner)
            boxes = []
            if t.any():
                for lo, hi in cluster_tags(t, tuple(int(v) for v in vlo), efficiency=efficiency,
                                           blocking=blocking, buffer=0, max_size=max_size):
                    sl = tuple(slice(lo[d] - vlo[d], hi[d] - vlo[d] + 1) for d in (2, 1, 0))
                    if inner[sl].all():
                        boxes.append((lo, hi))
                    else:
                        boxes += cluster_tags(np.ascontiguousarray(t[sl]), lo, efficiency=1.0)
        else:
            raise UclbsError(f"refine_by_ggm: level {lvl - 1} is neither a one-box level nor a "
                             f"refined level of many boxes")
        if not boxes:
            if lvl in self.shapes:
                self.clear_level(lvl)
            return []
        fine = [(tuple(2 * v for v in lo), tuple(2 * v for v in hi)) for lo, hi in boxes]
        if lvl in self.shapes:
            self.regrid_level(lvl, fine)
        else:
            self.define_level_boxes(lvl, fine)
            self.init_from_coarse(lvl)
        return fine

    def box_layout(self, lvl: int):
        """(offset, stored extents (x, y, z), lo of the valid region) per box."""
        out, off = [], 0
        for lo, hi in self.boxes[lvl]:
            e = tuple(h - l + 3 for l, h in zip(lo, hi))
            out.append((off, e, lo))
            off += e[0] * e[1] * e[2]
        return out

    def dense_to_stored(self, lvl: int, a: np.ndarray, origin=(0, 0, 0)):
        """a[..., z, y, x] over level positions origin + (x, y, z) -> stored
        order (valid nodes taken from a, ghost-shell nodes from the wrapped
        position, so a periodic level-0 array fills every stored node)."""
        a = np.asarray(a)
        lead = a.shape[:-3]
        out = np.zeros(lead + (self.shapes[lvl][0],), a.dtype)
        nz, ny, nx = a.shape[-3:]
        for off, e, lo in self.box_layout(lvl):
            zz = (np.arange(e[2]) + lo[2] - 1 - origin[2]) % nz
            yy = (np.arange(e[1]) + lo[1] - 1 - origin[1]) % ny
            xx = (np.arange(e[0]) + lo[0] - 1 - origin[0]) % nx
            blk = a[..., zz[:, None, None], yy[None, :, None], xx[None, None, :]]
            out[..., off:off + e[0] * e[1] * e[2]] = blk.reshape(lead + (-1,))
        return out

    def stored_to_dense(self, lvl: int, v: np.ndarray, shape, origin=(0, 0, 0), fill=np.nan):
        """Valid nodes of stored-order v into a dense [..., z, y, x] array of
        the given shape over positions origin + (x, y, z); other entries fill."""
        v = np.asarray(v)
        lead = v.shape[:-1]
        out = np.full(lead + tuple(shape), fill, dtype=np.result_type(v.dtype, np.float64))
        for off, e, lo in self.box_layout(lvl):
            blk = v[..., off:off + e[0] * e[1] * e[2]].reshape(lead + (e[2], e[1], e[0]))
            z0, y0, x0 = lo[2] - origin[2], lo[1] - origin[1], lo[0] - origin[0]
            out[..., z0:z0 + e[2] - 2, y0:y0 + e[1] - 2, x0:x0 + e[0] - 2] = blk[..., 1:-1, 1:-1, 1:-1]
        return out

    def advance_native(self, ncoarse: int, graph: bool = False):
        """uclbs_advance: the same schedule as advance(), in the library, with
        the launches of two coarse steps replayed as a hipGraph when graph."""
        self._check(self.lib.uclbs_advance(self.ctx, ncoarse, int(graph)))

    def autotune(self, lvl: int) -> int:
        t = ctypes.c_int()
        self._check(self.lib.uclbs_autotune(self.ctx, lvl, ctypes.byref(t)))
        return t.value

    def interlevel(self, lvl: int, theta: float):
        self._check(self.lib.uclbs_interlevel(self.ctx, lvl, theta))

    def init_from_coarse(self, lvl: int):
        self._check(self.lib.uclbs_init_from_coarse(self.ctx, lvl))

    def fine_to_coarse(self, lvl: int):
        self._check(self.lib.uclbs_fine_to_coarse(self.ctx, lvl))

    def advance(self, nsteps: int, lvl: int = 0):
        """Level lvl one step, each finer level two substeps per step (ghosts
        at the coarse level's previous state, then half way), finer levels
        injected back after their second substep; repeated nsteps times."""
        for _ in range(nsteps):
            self.step(lvl, 1)
            if lvl + 1 in self.shapes:
                for theta in (0.0, 0.5):
                    self.interlevel(lvl + 1, theta)
                    self.advance(1, lvl + 1)
                self.fine_to_coarse(lvl + 1)

    def n_nodes(self, lvl: int) -> int:
        s = self.shapes[lvl]
        return s[0] * s[1] * s[2]

    def upload_f(self, lvl: int, f: np.ndarray):
        n = self.n_nodes(lvl)
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(self.q, n)
        self._check(self.lib.uclbs_upload_f(self.ctx, lvl, _dptr(f), n))

    # ---- D3Q39 ---------------------------------------------------------------
    def set_gas(self, gama=1.4, temperature=300.0, pressure=101325.0, mach=0.3, dx=1e-3,
                startup=25, turbulent=True, shifted=-1, lm_maxit=50):
        self._gas = dict(gama=gama, temperature=temperature, pressure=pressure, mach=mach)
        g = Gas(gama=gama, temperature=temperature, pressure=pressure, mach=mach, dx=dx,
                startup_timesteps=startup, turbulent=int(turbulent), shifted=shifted,
                lm_maxit=lm_maxit)
        self._check(self.lib.uclbs_set_gas(self.ctx, ctypes.byref(g)))

    def init_fields(self, lvl: int, rho, u, v, w, temp):
        n = self.n_nodes(lvl)
        arrs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (n,)))
                for a in (rho, u, v, w, temp)]
        self._check(self.lib.uclbs_init_fields(self.ctx, lvl, *[_dptr(a) for a in arrs]))

    def upload_state(self, lvl: int, f, g, lm):
        n = self.n_nodes(lvl)
        f = np.ascontiguousarray(f, np.float64).reshape(39, n)
        g = np.ascontiguousarray(g, np.float64).reshape(39, n)
        lm = np.ascontiguousarray(lm, np.float64).reshape(13, n)
        self._check(self.lib.uclbs_upload_state(self.ctx, lvl, _dptr(f), _dptr(g), _dptr(lm), n))

    def download_state(self, lvl: int):
        n = self.n_nodes(lvl)
        f = np.empty((39, n))
        g = np.empty((39, n))
        lm = np.empty((13, n))
        self._check(self.lib.uclbs_download_state(self.ctx, lvl, _dptr(f), _dptr(g), _dptr(lm), n))
        return f, g, lm

    # ---- learned alpha ---------------------------------------------------------
    def set_bc_states39(self, inlets=(), outlets=()):
        """D3Q39 inlets / outlets: sequences of (rho, u, v, w, T, type)."""
        def pack(items):
            st = np.ascontiguousarray([list(t[:5]) for t in items] or [[0.0] * 5], dtype=np.float64)
            ty = np.ascontiguousarray([int(t[5]) for t in items] or [0], dtype=np.int32)
            return st, ty
        si, ti = pack(inlets)
        so, to = pack(outlets)
        self._check(self.lib.uclbs_set_bc_states39(self.ctx, len(inlets), _dptr(si), _iptr(ti),
                                                   len(outlets), _dptr(so), _iptr(to)))

    FIELD = {"f": 0, "g": 1, "lm": 2}

    def import_fab(self, lvl: int, field: str, device_ptr: int, ghost: int, comp_stride: int):
        """Level state from an AMReX FArrayBox in device memory (address device_ptr,
        FP64, component-major, ghost width, component stride in elements)."""
        self._check(self.lib.uclbs_import_fab(self.ctx, lvl, self.FIELD[field],
                                              ctypes.c_void_p(device_ptr), ghost, comp_stride))

    def export_fab(self, lvl: int, field: str, device_ptr: int, ghost: int, comp_stride: int):
        """Level state into the valid region of a device FArrayBox."""
        self._check(self.lib.uclbs_export_fab(self.ctx, lvl, self.FIELD[field],
                                              ctypes.c_void_p(device_ptr), ghost, comp_stride))

    def checkpoint_write(self, path: str):
        """Versioned dump of every level's device state (uclbs_checkpoint_write)."""
        self._check(self.lib.uclbs_checkpoint_write(self.ctx, os.fsencode(path)))

    def checkpoint_read(self, path: str):
        """Restore a dump into this context, set up as the saved one was."""
        self._check(self.lib.uclbs_checkpoint_read(self.ctx, os.fsencode(path)))

    # ---- output: Plot3D and microphones -------------------------------------------
    def set_placement(self, problo, probhi):
        """amrex_problo / amrex_probhi of the level-0 domain."""
        lo = np.ascontiguousarray(problo, np.float64)
        hi = np.ascontiguousarray(probhi, np.float64)
        self._check(self.lib.uclbs_set_placement(self.ctx, _dptr(lo), _dptr(hi)))

    def plot3d(self, directory: str, timestep: int, time: float = 0.0, max_level: int = 31,
               wait: bool = True, max_subrecord: int = 0, **values):
        """Write uclbs_grid_<timestep>.x / uclbs_data_<timestep>.q into directory
        (plot3d_output.f95:76-305).  values: the header's mach, reynolds,
        gamma, t_ref, stag_temperature, cv, cp, gas_r (mach and gamma default
        to the gas of set_gas, the rest to 0).  wait=False returns once the
        copy is enqueued; output_wait() then waits for the files."""
        gas = getattr(self, "_gas", {})
        h = Plot3DHeader(timestep=int(timestep), max_level=int(max_level),
                         max_subrecord=int(max_subrecord), time=float(time),
                         mach=float(values.pop("mach", gas.get("mach", 0.0))),
                         gamma=float(values.pop("gamma", gas.get("gama", 0.0))))
        for k, v in values.items():
            if k not in ("reynolds", "t_ref", "stag_temperature", "cv", "cp", "gas_r"):
                raise TypeError(f"unknown header value {k}")
            setattr(h, k, float(v))
        self._check(self.lib.uclbs_plot3d_write(self.ctx, os.fsencode(directory), ctypes.byref(h)))
        if wait:
            self.output_wait()

    def output_wait(self):
        self._check(self.lib.uclbs_output_wait(self.ctx))

    def mic_setup(self, directory: str, locations, types, levels) -> np.ndarray:
        """Microphones at physical locations (n x 3) recording the fields of
        types (bits 32 p, 16 T, 8 rho, 4 u, 2 v, 1 w) on levels; returns the
        validity per microphone (0 invalid, 1 sampled, 2 outside the level's
        boxes)."""
        loc = np.ascontiguousarray(locations, np.float64).reshape(-1, 3)
        n = loc.shape[0]
        ty = np.ascontiguousarray(np.broadcast_to(np.asarray(types, np.int32), (n,)))
        lv = np.ascontiguousarray(np.broadcast_to(np.asarray(levels, np.int32), (n,)))
        valid = np.zeros(max(n, 1), np.int32)
        self._check(self.lib.uclbs_mic_setup(self.ctx, os.fsencode(directory), n, _dptr(loc),
                                             _iptr(ty), _iptr(lv), _iptr(valid)))
        return valid[:n]

    def mic_sample(self):
        self._check(self.lib.uclbs_mic_sample(self.ctx))

    def mic_flush(self):
        self._check(self.lib.uclbs_mic_flush(self.ctx))

    def mic_weights(self, i: int):
        """(corner positions 8 x 3, weights 8, stored nodes 8) of microphone i."""
        pos = np.zeros(24, np.int32)
        w = np.zeros(8)
        node = np.zeros(8, np.int64)
        self._check(self.lib.uclbs_mic_weights(self.ctx, i, _iptr(pos), _dptr(w),
                                               node.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return pos.reshape(8, 3), w, node

    def macro39(self, lvl: int):
        """(rho, u, v, w, T, press) of a D3Q39 level, computed on the device
        (press = rho (T + |u|^2 / 2) as calculate_flow_variables.f95:221)."""
        n = self.n_nodes(lvl)
        arrs = [np.empty(n, dtype=np.float64) for _ in range(6)]
        self._check(self.lib.uclbs_download_macro39(self.ctx, lvl, *[_dptr(a) for a in arrs], n))
        return arrs

    def mlp_init(self, seed: int):
        self._check(self.lib.uclbs_mlp_init(self.ctx, seed))

    def mlp_set_params(self, params, y_mean: float, y_std: float):
        p = np.ascontiguousarray(params, np.float32)
        self._check(self.lib.uclbs_mlp_set_params(self.ctx, _fptr(p), y_mean, y_std))

    def mlp_get_params(self):
        p = np.empty(MLP_PARAMS, np.float32)
        m, s = ctypes.c_float(), ctypes.c_float()
        self._check(self.lib.uclbs_mlp_get_params(self.ctx, _fptr(p), ctypes.byref(m), ctypes.byref(s)))
        return p, m.value, s.value

    def mlp_loss_grad(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = ctypes.c_double()
        self._check(self.lib.uclbs_mlp_loss_grad(self.ctx, _fptr(X), _fptr(y), len(y), _fptr(g),
                                                 ctypes.byref(loss)))
        return loss.value, g

    def mlp_forward(self, X, mfma: bool):
        X = np.ascontiguousarray(X, np.float32)
        out = np.empty(X.shape[0], np.float32)
        self._check(self.lib.uclbs_mlp_forward(self.ctx, _fptr(X), X.shape[0], int(mfma), _fptr(out)))
        return out

    def mlp_forward_noalloc(self, X):
        """BF16-MFMA forward through the library's fixed staging area (no
        device allocation; at most 65,536 rows)."""
        X = np.ascontiguousarray(X, np.float32)
        out = np.empty(X.shape[0], np.float32)
        self._check(self.lib.uclbs_mlp_forward_noalloc(self.ctx, _fptr(X), X.shape[0], _fptr(out)))
        return out

    def mlp_reserve(self, capacity: int):
        self._check(self.lib.uclbs_mlp_reserve(self.ctx, int(capacity)))

    def mlp_collect(self, lvl: int, stride: int) -> int:
        n = ctypes.c_int64()
        self._check(self.lib.uclbs_mlp_collect(self.ctx, lvl, stride, ctypes.byref(n)))
        return n.value

    def mlp_samples(self, capacity: int):
        X = np.empty((capacity, 12), np.float32)
        y = np.empty(capacity, np.float32)
        n = ctypes.c_int64()
        self._check(self.lib.uclbs_mlp_download_samples(self.ctx, _fptr(X), _fptr(y), capacity,
                                                        ctypes.byref(n)))
        k = min(n.value, capacity)
        return X[:k], y[:k]

    def mlp_sample_nodes(self, capacity: int) -> np.ndarray:
        """Node index of each held pair, in the order of mlp_samples."""
        node = np.empty(capacity, np.uint32)
        n = ctypes.c_int64()
        self._check(self.lib.uclbs_mlp_download_sample_nodes(
            self.ctx, node.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), capacity,
            ctypes.byref(n)))
        return node[:min(n.value, capacity)]

    def mlp_train(self, steps: int, batch: int, lr: float, seed: int):
        l0, l1 = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.uclbs_mlp_train(self.ctx, int(steps), batch, lr, seed,
                                             ctypes.byref(l0), ctypes.byref(l1)))
        return l0.value, l1.value

    def set_timestep(self, lvl: int, step: int):
        self._check(self.lib.uclbs_set_timestep(self.ctx, lvl, int(step)))

    def set_faces(self, lvl: int, codes):
        c6 = (ctypes.c_int32 * 6)(*[int(v) for v in codes])
        self._check(self.lib.uclbs_set_faces(self.ctx, lvl, c6))

    def set_bc_states(self, inlets=(), outlets=(), freestream=None):
        ins = np.ascontiguousarray(np.array(inlets, dtype=np.float64).reshape(-1, 4))
        outs = np.ascontiguousarray(np.array(outlets, dtype=np.float64).reshape(-1, 4))
        fs = None if freestream is None else np.ascontiguousarray(freestream, dtype=np.float64)
        self._check(self.lib.uclbs_set_bc_states(
            self.ctx, len(ins), _dptr(ins) if len(ins) else None, len(outs),
            _dptr(outs) if len(outs) else None, _dptr(fs) if fs is not None else None))

    def build_links(self, lvl: int, tri: np.ndarray, nrm: np.ndarray, origin=(0.0, 0.0, 0.0),
                    dx: float = 1.0):
        tri = np.ascontiguousarray(tri, dtype=np.float64)
        nrm = np.ascontiguousarray(nrm, dtype=np.float64)
        o3 = np.ascontiguousarray(origin, dtype=np.float64)
        ns, nt = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.uclbs_build_links(
            self.ctx, lvl, _dptr(o3), dx, len(tri), _dptr(tri) if len(tri) else None,
            _dptr(nrm) if len(nrm) else None, ctypes.byref(ns), ctypes.byref(nt)))
        return ns.value, nt.value

    def set_links(self, lvl: int, state: np.ndarray):
        n = self.n_nodes(lvl)
        st = np.ascontiguousarray(state, dtype=np.int32).reshape(self.q, n)
        self._check(self.lib.uclbs_set_links(self.ctx, lvl, _iptr(st), n))

    def download_links(self, lvl: int) -> np.ndarray:
        n = self.n_nodes(lvl)
        out = np.empty((self.q, n), dtype=np.int32)
        self._check(self.lib.uclbs_download_links(self.ctx, lvl, _iptr(out), n))
        return out

    def link_counts(self, lvl: int):
        nb, ns = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.uclbs_link_counts(self.ctx, lvl, ctypes.byref(nb), ctypes.byref(ns)))
        return nb.value, ns.value

    def init_uniform(self, lvl: int, rho: float, u):
        self._check(self.lib.uclbs_init_uniform(self.ctx, lvl, rho, *[float(v) for v in u]))

    def download_f(self, lvl: int) -> np.ndarray:
        n = self.n_nodes(lvl)
        out = np.empty((self.q, n), dtype=np.float64)
        self._check(self.lib.uclbs_download_f(self.ctx, lvl, _dptr(out), n))
        return out

    def macro(self, lvl: int):
        n = self.n_nodes(lvl)
        arrs = [np.empty(n, dtype=np.float64) for _ in range(4)]
        self._check(self.lib.uclbs_download_macro(self.ctx, lvl, *[_dptr(a) for a in arrs]))
        return arrs

    def keep_alpha(self, keep: bool = True):
        self._check(self.lib.uclbs_set_keep_alpha(self.ctx, int(keep)))

    def alpha(self, lvl: int) -> np.ndarray:
        out = np.empty(self.n_nodes(lvl), dtype=np.float64)
        self._check(self.lib.uclbs_download_alpha(self.ctx, lvl, _dptr(out)))
        return out

    def init_tgv(self, lvl: int, u0: float, noise_amp: float, seed: int):
        self._check(self.lib.uclbs_init_tgv(self.ctx, lvl, u0, noise_amp, seed))

    def step(self, lvl: int, nsteps: int = 1):
        self._check(self.lib.uclbs_step_level(self.ctx, lvl, nsteps))

    def step_async(self, lvl: int, nsteps: int = 1):
        self._check(self.lib.uclbs_step_level_async(self.ctx, lvl, nsteps))

    def time_steps(self, lvl: int, nsteps: int) -> float:
        ms = ctypes.c_double()
        self._check(self.lib.uclbs_time_steps(self.ctx, lvl, nsteps, ctypes.byref(ms)))
        return ms.value

    def time_kernels(self, lvl: int, nsteps: int):
        """(total ms of nsteps steps, summed ms of the streaming kernel launches)"""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.uclbs_time_kernels(self.ctx, lvl, nsteps, ctypes.byref(a),
                                                ctypes.byref(b)))
        return a.value, b.value

    def sync(self):
        self._check(self.lib.uclbs_sync(self.ctx))

    def stats(self) -> dict:
        s = Stats()
        self._check(self.lib.uclbs_get_stats(self.ctx, ctypes.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        self._check(self.lib.uclbs_reset_stats(self.ctx))
