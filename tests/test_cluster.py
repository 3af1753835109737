"""Clustering of tagged nodes into boxes (uclbs_cluster_tags: host code in the
library, no device).  The reference takes it from AMReX's regrid (AMReX
>= 22.06, not in the reference tree), so the exact box lists are not pinned
against it; what any valid clustering must give is:
  - every grown tag is covered, boxes are disjoint and inside the region
  - each box is at least `efficiency` tagged on the blocking grid, unless it
    is a single block
  - faces on multiples of the blocking factor except where clipped to the
    region; no box longer than max_size on any axis
  - known answers: two separated blocks give exactly their bounding boxes;
    a full box gives itself; no tags give no boxes; a buffer grows a single
    tag to a cube clipped at the region's faces; an L of two slabs splits
    at its inflection into the two slabs."""
import itertools

import numpy as np
import pytest


def grow(t, buf):
    g = t.astype(bool).copy()
    for ax in range(3):
        h = g.copy()
        for o in range(1, buf + 1):
            h[(slice(None),) * ax + (slice(o, None),)] |= g[(slice(None),) * ax + (slice(None, -o),)]
            h[(slice(None),) * ax + (slice(None, -o),)] |= g[(slice(None),) * ax + (slice(o, None),)]
        g = h
    return g


def check(boxes, tags, lo, eff, bf, buf, max_size):
    nz, ny, nx = tags.shape
    hi = (lo[0] + nx - 1, lo[1] + ny - 1, lo[2] + nz - 1)
    cover = np.zeros(tags.shape, np.int32)
    g = grow(tags, buf)
    for blo, bhi in boxes:
        for d in range(3):
            assert lo[d] <= blo[d] <= bhi[d] <= hi[d]
            if blo[d] != lo[d]:
                assert blo[d] % bf == 0
            if bhi[d] != hi[d]:
                assert (bhi[d] + 1) % bf == 0
            if max_size:
                assert bhi[d] - blo[d] + 1 <= max_size
        cover[blo[2] - lo[2]:bhi[2] - lo[2] + 1, blo[1] - lo[1]:bhi[1] - lo[1] + 1,
              blo[0] - lo[0]:bhi[0] - lo[0] + 1] += 1
    assert cover.max() <= 1, "boxes overlap"
    assert cover[g].min() >= 1 if g.any() else not boxes
    return cover


def blob_tags(shape, seed, n=3):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx]
    t = np.zeros(shape, bool)
    for _ in range(n):
        c = rng.uniform(0, 1, 3) * (nx, ny, nz)
        r = rng.uniform(2, 6)
        t |= (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 < r * r
    t &= rng.random(shape) < 0.9
    return t.astype(np.uint8)


@pytest.mark.parametrize("seed,eff,bf,buf,max_size", list(itertools.product(
    (1, 2, 3), (0.5, 0.7, 0.9), (1, 2, 4), (0, 2), (0, 8))))
def test_invariants(pkg, seed, eff, bf, buf, max_size):
    tags = blob_tags((22, 27, 31), seed)
    lo = (3, -5, 7)
    boxes = pkg.cluster_tags(tags, lo, efficiency=eff, blocking=bf, buffer=buf, max_size=max_size)
    check(boxes, tags, lo, eff, bf, buf, max_size)
    if bf == 1 and buf == 0 and max_size == 0:
        # each box reaches the efficiency unless it is a single node
        for blo, bhi in boxes:
            sub = tags[blo[2] - lo[2]:bhi[2] - lo[2] + 1, blo[1] - lo[1]:bhi[1] - lo[1] + 1,
                       blo[0] - lo[0]:bhi[0] - lo[0] + 1]
            assert sub.sum() >= eff * sub.size or sub.size == 1


def test_known_answers(pkg):
    t = np.zeros((20, 20, 30), np.uint8)
    t[2:6, 3:9, 4:10] = 1
    t[12:18, 10:15, 20:27] = 1
    boxes = sorted(pkg.cluster_tags(t, (0, 0, 0), efficiency=0.7))
    assert boxes == [((4, 3, 2), (9, 8, 5)), ((20, 10, 12), (26, 14, 17))]
    full = np.ones((4, 5, 6), np.uint8)
    assert pkg.cluster_tags(full, (1, 2, 3)) == [((1, 2, 3), (6, 6, 6))]
    assert pkg.cluster_tags(np.zeros((4, 5, 6), np.uint8)) == []
    one = np.zeros((10, 10, 10), np.uint8)
    one[1, 5, 8] = 1
    assert pkg.cluster_tags(one, (0, 0, 0), buffer=2) == [((6, 3, 0), (9, 7, 3))]
    # an L: a long slab along x and a tall one along y meeting at a corner
    ell = np.zeros((4, 24, 24), np.uint8)
    ell[:, 0:4, :] = 1
    ell[:, :, 0:4] = 1
    boxes = sorted(pkg.cluster_tags(ell, (0, 0, 0), efficiency=0.9))
    cov = check(boxes, ell, (0, 0, 0), 0.9, 1, 0, 0)
    assert len(boxes) == 2 and cov.sum() == int(ell.sum())


def test_bad_arguments(pkg):
    t = np.ones((3, 3, 3), np.uint8)
    for kw in ({"efficiency": 0.0}, {"blocking": 0}, {"buffer": -1}, {"blocking": 4, "max_size": 6}):
        with pytest.raises(pkg.UclbsError):
            pkg.cluster_tags(t, (0, 0, 0), **kw)
