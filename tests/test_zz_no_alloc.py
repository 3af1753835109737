"""uclbs_mlp_forward_noalloc (staging through the library's fixed device area,
no allocation) against uclbs_mlp_forward with the MFMA path: the same outputs
bit for bit, on two calls in a row, the second with a smaller batch, and from
a second context on the same device; a batch beyond the area is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _net(pkg, seed):
    s = pkg.Solver()
    s.define_level(0, (8, 8, 8))
    s.mlp_init(seed)
    return s


def test_noalloc_forward_matches_forward(pkg):
    s = _net(pkg, 31)
    rng = np.random.default_rng(7)
    for B in (65536, 50000, 777):
        X = rng.uniform(-1, 1, (B, 12)).astype(np.float32)
        ref = s.mlp_forward(X, mfma=True)
        out = s.mlp_forward_noalloc(X)
        print(f"noalloc forward, {B} rows: max |diff| {np.abs(out - ref).max():.1e}", flush=True)
        assert np.array_equal(out, ref)
    # a second context shares the device's staging area
    s2 = _net(pkg, 32)
    X = rng.uniform(-1, 1, (1000, 12)).astype(np.float32)
    assert np.array_equal(s2.mlp_forward_noalloc(X), s2.mlp_forward(X, mfma=True))
    assert np.array_equal(s.mlp_forward_noalloc(X), s.mlp_forward(X, mfma=True))
    with pytest.raises(pkg.UclbsError, match="staging area"):
        s.mlp_forward_noalloc(np.zeros((65537, 12), np.float32))
