"""The oracle's geodesic backtracking, pinned without the GPU.

* The capped distance to the seeds equals a brute-force shortest-path
  distance on the 3^d-neighbour graph with edge costs round(5 |offset|),
  values above the target tolerance dropped (the min-plus iteration's fixed
  point within niter steps).
* On an exact radial field (values |x - s|, flow pointing to the seed, flow
  weight sum 1 off the seed and 0 on it) every geodesic is the segment from
  its tip to the seed: it stops AtSeed, stays within a fraction of a pixel
  of the segment, and has |tip - s| / step points.
* A flow pointing away from the seed stops PastSeed near the seed and
  leaves the grid (InWall) farther away; a zero flow stops at once
  (VanishingFlow); a tip outside the grid is InWall; a flow circling the
  seed never reaches it.
"""
import heapq
import itertools

import numpy as np
import pytest

import geodesic_oracle as G  # noqa: E402


def brute_eucl(seeds, bound, mult=5):
    d = seeds.ndim
    dist = np.full(seeds.shape, np.inf)
    pq = []
    for p in zip(*np.nonzero(seeds)):
        dist[p] = 0
        pq.append((0, p))
    heapq.heapify(pq)
    offs = [o for o in itertools.product((-1, 0, 1), repeat=d) if any(o)]
    while pq:
        dv, p = heapq.heappop(pq)
        if dv > dist[p]:
            continue
        for o in offs:
            q = tuple(p[k] + o[k] for k in range(d))
            if any(q[k] < 0 or q[k] >= seeds.shape[k] for k in range(d)):
                continue
            nd = dv + int(np.round(mult * np.sqrt(sum(abs(x) for x in o))))
            if nd < dist[q]:
                dist[q] = nd
                heapq.heappush(pq, (nd, q))
    out = np.where(dist > bound, 255, dist)
    return out.astype(np.uint8)


@pytest.mark.parametrize("d", [2, 3])
def test_eucl_matches_shortest_paths(d):
    rng = np.random.default_rng(3 + d)
    shape = (23, 17) if d == 2 else (11, 9, 10)
    seeds = rng.random(shape) < (0.01 if d == 2 else 0.008)
    seeds.flat[5] = True
    e = G.eucl_distance(seeds, niter=6, mult=5, bound=30)
    assert np.array_equal(e, brute_eucl(seeds, 30))
    # a plausible mistake (one min-plus pass too few) changes the field
    assert not np.array_equal(G.eucl_distance(seeds, niter=5, mult=5, bound=30), e)


def radial_field(shape, s):
    X = np.stack(np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing='ij'))
    diff = np.asarray(s, dtype=np.float64).reshape((-1,) + (1,) * len(shape)) - X
    r = np.sqrt((diff ** 2).sum(0))
    flow = np.where(r > 0, diff / np.where(r > 0, r, 1), 0.)
    wsum = np.where(r > 0, 1., 0.)
    seeds = r == 0
    return r, flow, wsum, seeds


@pytest.mark.parametrize("d", [2, 3])
def test_radial_geodesics_are_segments(d):
    shape = (41, 37) if d == 2 else (21, 19, 23)
    s = tuple(n // 2 for n in shape)
    u, flow, wsum, seeds = radial_field(shape, s)
    rng = np.random.default_rng(7)
    tips = np.stack([rng.uniform(1, n - 2, 12) for n in shape], -1)
    geos, stops = G.geodesics(u, flow, wsum, seeds, tips)
    sv = np.array(s, dtype=np.float64)
    for tip, geo, stop in zip(tips, geos, stops):
        assert stop == 'AtSeed'
        seg = sv - tip
        L = np.linalg.norm(seg)
        rel = geo - tip
        t = rel @ seg / L ** 2
        perp = np.linalg.norm(rel - t[:, None] * seg, axis=1)
        assert perp.max() < 0.2
        assert np.linalg.norm(geo[-1] - sv) < 1.5          # ends in a cell touching the seed
        assert abs(len(geo) - L / 0.25) <= 8               # |tip - s| / step points


def test_stopping_codes():
    shape = (30, 30)
    s = (15, 15)
    u, flow, wsum, seeds = radial_field(shape, s)
    tip = np.array([[20.3, 12.7]])
    # flow reversed: from a tip within the target tolerance (6 pixels) the
    # distance to the seed grows (PastSeed); from farther away the geodesic
    # walks to the boundary (InWall)
    _, stops = G.geodesics(u, -flow, wsum, seeds, np.array([[18.3, 13.7]]))
    assert stops == ['PastSeed']
    _, stops = G.geodesics(u, -flow, wsum, seeds, np.array([[25.3, 26.7]]))
    assert stops == ['InWall']
    # no flow at all
    _, stops = G.geodesics(u, 0 * flow, wsum, seeds, tip)
    assert stops == ['VanishingFlow']
    # tip outside the grid
    _, stops = G.geodesics(u, flow, wsum, seeds, np.array([[-5., 3.]]))
    assert stops == ['InWall']
    # a flow circling the seed never reaches it within max_len
    rot = np.stack([-flow[1], flow[0]])
    _, stops = G.geodesics(u, rot, wsum, seeds, tip, max_len=200, typical_len=50)
    assert stops[0] in ('Stationnary', 'Stopping criterion', 'PastSeed')
