"""D3Q27 wall links on the GPU (linear interpolated bounce-back) in a
pressure-driven channel between flat walls cut at fraction q of the links,
through the C ABI with a reference-form link table (uclbs_set_links):
  FP64 storage against the FP64 oracle: populations relative L2 <= 1e-12
  FP32-shifted storage: the developed mid-channel profile is the parabola
    vanishing on the walls to within 1e-2 (oracle at the same size: 5e-3)."""
import numpy as np
import pytest

from test_oracle_box27 import _channel_table

pytestmark = pytest.mark.gpu
PER = 2147483647
FACES = [-101, -10, PER, -201, -10, PER]


NZ = 3  # the library's boxes span at least 3 nodes per axis; z is periodic


def _channel(pkg, orc, L, H, q, U, nu, precision, mode):
    st = np.concatenate([_channel_table(orc, L, H, q, q)] * NZ, axis=1)
    s = pkg.Solver(precision=precision, nu=nu, alpha_mode=mode, alpha_tol=1e-12)
    s.define_level(0, (L, H, NZ), periodic=(0, 0, 1))
    s.set_faces(0, FACES)
    s.set_bc_states(inlets=[(1.0, U, 0.0, 0.0)], outlets=[(1.0, U, 0.0, 0.0)])
    s.set_links(0, st)
    s.init_uniform(0, 1.0, (U, 0.0, 0.0))
    return s, st


@pytest.mark.parametrize("q", [0.2, 0.7])
def test_channel_fp64_matches_oracle(pkg, orc, q):
    L, H, U, nu = 32, 8, 0.02, 0.05
    s, st = _channel(pkg, orc, L, H, q, U, nu, "fp64", "newton")
    f0 = orc.uniform27(st, 1.0, (U, 0.0, 0.0))
    s.upload_f(0, f0)
    s.step(0, 150)
    f_o, _, _, bst = orc.step_box27((L, H, NZ), FACES, st, f0, 1 / (6 * nu + 1), nsteps=150,
                                    inlets=[(1.0, U, 0.0, 0.0)], outlets=[(1.0, U, 0.0, 0.0)],
                                    tol=1e-12)
    f_g = s.download_f(0)
    e = float(np.linalg.norm(f_g - f_o) / np.linalg.norm(f_o))
    print(f"channel q {q}: FP64 GPU vs oracle relL2 {e:.2e}, wall links per step {bst['n_ibb_links'] // 150}")
    assert e < 1e-12
    assert s.stats()["n_ibb_links"] == bst["n_ibb_links"]


def test_channel_fp32_poiseuille_profile(pkg, orc):
    L, H, q, U, nu = 64, 16, 0.2, 0.01, 1 / 6
    s, _ = _channel(pkg, orc, L, H, q, U, nu, "fp32s", "fixed2")
    s.step(0, int(3 * H * H / nu))
    ux = s.macro(0)[1].reshape(NZ, H, L)[0, :, L // 2]
    y = np.arange(H)
    p = (y + q) * (H - 1 + q - y)
    A = ux @ p / (p @ p)
    e = float(np.linalg.norm(ux - A * p) / np.linalg.norm(ux))
    print(f"FP32 channel H {H} q {q}: profile error {e:.2e}")
    assert e < 1e-2
