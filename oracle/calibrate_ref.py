"""Per-call cost of the oracle's restatements against the reference's own
compiled routines on this machine's CPU.  TEST INFRASTRUCTURE ONLY (it times
the oracle, it is not part of the product).

The reference's LM Newton (calculate_lagrangian_mults) and alpha estimate
need AMReX and cannot be built here; the routines that can -- lu_decomp (one
13x13 LU-solve Newton update) and quick_calcs (Sutherland viscosity and heat
conductivity) -- are timed against the oracle's bit-exact restatements on the
same inputs, one core, many repetitions inside the compiled code (no Python
call overhead per repetition).  The oracle's whole per-node costs (multiplier
solve, full D3Q39 step, D3Q27 step) are timed alongside, for the share the
calibrated parts take.

Run from the repository root where /root/reference exists:
    python3 oracle/calibrate_ref.py
"""
import ctypes
import os
import subprocess
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import oracle as o  # noqa: E402


def best_of(fn, reps=5):
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        v = fn()
        t.append(time.perf_counter() - t0)
    return min(t), v


def main():
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    here = os.path.dirname(os.path.abspath(__file__))
    o.build()
    subprocess.run(["make", "-C", os.path.join(here, "ref")], check=True, stdout=subprocess.DEVNULL)
    ref = o.ref_lib()
    if ref is None:
        raise SystemExit("the reference routines are not built (oracle/ref/Makefile)")
    L = o.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    ref.ref_lu_decomp_batch.argtypes = [ctypes.c_int, dp, dp, dp]
    ref.ref_lu_decomp_batch.restype = ctypes.c_double
    ref.ref_quick_batch.argtypes = [ctypes.c_int, dp, dp]
    ref.ref_quick_batch.restype = ctypes.c_double
    L.o39_lu_batch.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_int]
    L.o39_lu_batch.restype = ctypes.c_double
    L.o39_quick_batch.argtypes = [ctypes.c_int, dp, dp, ctypes.POINTER(o.Consts39)]
    L.o39_quick_batch.restype = ctypes.c_double

    def p(a):
        return a.ctypes.data_as(dp)

    rng = np.random.default_rng(20260105)
    M = rng.normal(0, 1, (13, 13))
    J = np.ascontiguousarray(M @ M.T + 13 * np.eye(13))
    Jf = np.asfortranarray(J)
    mom = rng.normal(0, 0.3, 13)
    mom[0] = 1.0
    lm0 = rng.normal(0, 0.2, 13)
    nrep = 2_000_000
    t_ref, s_ref = best_of(lambda: ref.ref_lu_decomp_batch(nrep, p(Jf), p(mom), p(lm0)))
    t_orc, s_orc = best_of(lambda: L.o39_lu_batch(nrep, p(J), p(mom), p(lm0), 1))
    t_exa, _ = best_of(lambda: L.o39_lu_batch(nrep, p(J), p(mom), p(lm0), 0))
    assert s_ref == s_orc, (s_ref, s_orc)  # the same bits, call for call

    c = o.consts39(mach=0.3)
    ref.ref_set_constants(c.gama, c.t_ref, c.mu_ref, c.rho_ref, c.kappa_ref, c.R)
    n = 4_000_000
    T = rng.uniform(0.5, 1.5, n) / c.gama
    rho = rng.uniform(0.8, 1.2, n)
    tq_ref, q_ref = best_of(lambda: ref.ref_quick_batch(n, p(rho), p(T)))
    tq_orc, q_orc = best_of(lambda: L.o39_quick_batch(n, p(rho), p(T), ctypes.byref(c)))
    assert q_ref == q_orc, (q_ref, q_orc)

    # the oracle's whole per-node costs (one core)
    shape = (24, 20, 16)
    nn = int(np.prod(shape))
    c39 = o.consts39(mach=0.3, shifted=False)
    f0, g0, lm = o.init39(shape, 1.0, 0.3, 0.0, 0.0, 1 / 1.4, c39)
    f1, g1, lm1, *_ = o.step_periodic39(shape, f0, g0, lm, c39, 0, 30)
    t39, _ = best_of(lambda: o.step_periodic39(shape, f1, g1, lm1, c39, 30, 2), reps=3)
    f27 = o.init_tgv27((48, 48, 48), 0.05, 1e-3, 1)
    t27, _ = best_of(lambda: o.step_periodic27((48, 48, 48), f27, 0.995, 2, nthreads=1), reps=3)

    rows = [
        ("lu_decomp, one 13x13 LU-solve Newton update (per call)", t_ref / nrep * 1e6,
         t_orc / nrep * 1e6, "equal bits (reference mode, its two index slips); exact mode %.3f us"
         % (t_exa / nrep * 1e6)),
        ("quick_sutherlands + quick_therm_cond (per node)", tq_ref / n * 1e6, tq_orc / n * 1e6,
         "equal bits"),
    ]
    print("| Routine | Reference, compiled (us) | Oracle restatement (us) | oracle / reference | Check |")
    print("|---|---|---|---|---|")
    for name, a, b, chk in rows:
        print(f"| {name} | {a:.4f} | {b:.4f} | {b / a:.2f} | {chk} |")
    print()
    print(f"Oracle, whole per-node costs on one core: D3Q39 plain-lattice step (Mach 0.3, "
          f"periodic {shape}, steps 30-31) {t39 / (2 * nn) * 1e6:.2f} us per node update; "
          f"D3Q27 entropic step (TGV 48^3) {t27 / (2 * 48 ** 3) * 1e6:.3f} us per node update.")


if __name__ == "__main__":
    main()
