"""The oracle's value pasting between local charts (ChartPaste.h:71-167) and
the solve that continues from pasted values, pinned by what they must give:

* multilinear interpolation reproduces affine functions exactly, at any
  mapped point, over one or two mapped axes (broadcast over the first);
* the boundary rules: points within 1e-4 of the last layer are moved onto
  it, points beyond, NaN mappings and zero-weight infinite corners paste
  nothing; a mapping onto the cell itself changes nothing;
* fast marching from tentative values u0 equals the decreasing Jacobi
  iteration started from u0;
* a band glued to the opposite side by a whole-cell shift (the circular band
  of Notebooks_Repro/Chart_GPU.ipynb, section 3) solves, after the paste
  loop, to exactly the periodic solution on the fundamental domain.
"""
import numpy as np
import pytest


def test_paste_interpolates_affine_functions_exactly(O):
    rng = np.random.default_rng(1)
    shape = (9, 12, 10)
    X = np.indices(shape).astype(np.float64)
    a = np.array([0.3, -1.2, 0.7])
    u = 50. + np.tensordot(a, X, axes=1)
    # two mapped axes (the last two), broadcast over the first
    m = np.stack([rng.uniform(0, 11, size=shape[1:]), rng.uniform(0, 9, size=shape[1:])])
    new, low = O.chart_paste(u, m)
    at_m = 50. + a[0] * X[0] + a[1] * m[0][None] + a[2] * m[1][None]
    np.testing.assert_array_equal(low, at_m < u - 1e-9)
    np.testing.assert_allclose(new, np.minimum(u, at_m), rtol=0, atol=1e-12)
    assert 0.2 < low.mean() < 0.8
    # one mapped axis
    m1 = rng.uniform(0, 9, size=(1, shape[2]))
    new, low = O.chart_paste(u, m1)
    at_m = 50. + a[0] * X[0] + a[1] * X[1] + a[2] * m1[0][None, None]
    np.testing.assert_allclose(new, np.minimum(u, at_m), atol=1e-12)


def test_paste_boundary_rules(O):
    rows = np.arange(4, dtype=np.float64)[:, None]
    u = 100. + rows + np.zeros((4, 6))
    u[:, 0] = 5. + rows[:, 0]
    u[:, 5] = 7. + rows[:, 0]
    m = np.array([[np.nan,       # no mapping
                   5 + 2e-4,     # beyond the last layer by more than the tolerance: nothing
                   -5e-5,        # just below the first layer: moved onto it
                   4.5,          # inside
                   5 + 5e-5,     # just beyond the last layer: moved onto it
                   np.nan]])
    new, low = O.chart_paste(u, m)
    np.testing.assert_array_equal(low[0], [False, False, True, True, True, False])
    np.testing.assert_array_equal(new[:, 2], u[:, 0])
    np.testing.assert_allclose(new[:, 3], 0.5 * u[:, 4] + 0.5 * u[:, 5], rtol=1e-15)
    np.testing.assert_array_equal(new[:, 4], u[:, 5])
    # identity: nothing is lower than itself
    assert not np.any(O.chart_paste(u, np.arange(6, dtype=np.float64)[None])[1])
    # a zero-weight corner that is infinite gives NaN: no paste
    v = u.copy()
    v[:, 4] = 1. + rows[:, 0]
    v[:, 5] = np.inf
    m2 = np.full((1, 6), np.nan)
    m2[0, 1] = 4.0            # q = 4, r = 0: corner 5 has weight 0 and value inf
    m2[0, 2] = 3.5            # both corners finite
    new, low = O.chart_paste(v, m2)
    assert not np.any(low[:, 1]) and np.all(low[:, 2])


def test_fast_marching_from_tentative_values_equals_iteration(O):
    rng = np.random.default_rng(2)
    n = 40
    P = O.make_problem('Isotropic', (n, n), (1. / n, 1. / n), cost=1. + rng.uniform(size=(n, n)),
                       seeds_idx=[[5, 7]])
    u = P.solve_fm()
    u0 = np.full((n, n), np.inf)
    u0[30:35, 20:26] = u[30:35, 20:26] * 0.5        # tentative values below the solution
    a = P.solve_fm(u0=u0)
    b, _ = P.solve_jacobi(1e-14, 0., u0=u0)
    np.testing.assert_allclose(a, b, rtol=1e-12)
    assert np.all(a <= u + 1e-15) and np.any(a < u - 1e-3)


def chart_band(O, n=100, band=5):
    """Isotropic unit cost on [-1,1] x [-1-b,1+b] with the y-band glued to the
    opposite side: y -> y - 2 (y > 0) or y + 2 (y < 0), that is 100 cells."""
    h = 2. / n
    ny = n + 2 * band
    y = -1 - band * h + (np.arange(ny) + 0.5) * h
    mapping = np.where(y > 0, y - 2, y + 2)
    # continuous index along y: whole cells (rounded, so that no paste mixes
    # two cells and the loop ends exactly)
    cont = np.round((mapping - (-1 - band * h)) / h - 0.5)[None]
    seed = (int(0.75 * n), band + int(0.65 * n))
    P = O.make_problem('Isotropic', (n, ny), (h, h), seeds_idx=[seed])
    return P, cont, seed


def test_band_glued_by_charts_is_periodic(O):
    n, band = 100, 5
    P, cont, seed = chart_band(O, n, band)
    u = P.solve_fm()
    for it in range(40):
        new, low = O.chart_paste(u, cont)
        if not low.any():
            break
        u = P.solve_fm(u0=new)
    assert not low.any()
    Pp = O.make_problem('Isotropic', (n, n), (2. / n, 2. / n), seeds_idx=[(seed[0], seed[1] - band)],
                        periodic=[0, 1])
    up = Pp.solve_fm()
    core = u[:, band:band + n]
    np.testing.assert_allclose(core, up, rtol=1e-13, atol=0)
    # and not without the charts
    assert np.max(np.abs(P.solve_fm()[:, band:band + n] - up)) > 0.1


def test_projective_plane_exact_distance(O):
    """Notebooks_Repro/Chart_GPU.ipynb section 1 on the oracle: the unit
    sphere with antipodes identified, parametrised from the equator (cost
    2 / (1 + |x|^2), mapping x -> -x / |x|^2); after the paste loop the
    values lie within the notebook's 0.02 of the exact spherical distance,
    and without the charts they do not."""
    r, n = 1.1, 251
    h = 2 * r / n
    x = -r + (np.arange(n) + 0.5) * h
    X = np.stack(np.meshgrid(x, x, indexing='ij'))
    e2 = np.sum(X ** 2, axis=0)
    seed = [[int(round((0.5 + r) / h - 0.5)), int(round((0.7 + r) / h - 0.5))]]
    P = O.make_problem('Isotropic', (n, n), (h, h), cost=2. / (1. + e2), seeds_idx=seed)
    with np.errstate(divide='ignore', invalid='ignore'):
        cont = (-X / e2 + r) / h - 0.5
    u0 = u = P.solve_fm()
    for _ in range(5):
        new, low = O.chart_paste(u, cont)
        if not low.any():
            break
        u = P.solve_fm(u0=new)

    def to_sphere(y):
        s2 = np.sum(y ** 2, axis=0)
        return np.concatenate([(1. - s2)[None], 2 * y]) / (1. + s2)

    s = to_sphere(X[:, seed[0][0], seed[0][1]].reshape(2, 1, 1))
    XS = to_sphere(X)
    dist = lambda p, q: 2 * np.arcsin(np.minimum(1., np.linalg.norm(p - q, axis=0) / 2))
    exact = np.minimum(dist(s, XS), dist(-s, XS))
    assert np.max(np.abs(u - exact)) < 0.02
    assert np.max(np.abs(u0 - exact)) > 0.3


def test_geodesic_jumps_across_the_glued_band(O):
    """The geodesic oracle's chart jump on the band glued by a whole-cell
    shift: a tip whose shortest path crosses the seam reaches the seed, jumps
    by the shift once, and has the length of the geodesic of the periodic
    solve."""
    import geodesic_oracle as G
    n, band = 100, 5
    P, cont, seed = chart_band(O, n, band)
    u = P.solve_fm()
    for _ in range(40):
        new, low = O.chart_paste(u, cont)
        if not low.any():
            break
        u = P.solve_fm(u0=new)
    fl = P.flow(u)
    seeds = np.zeros(u.shape, bool)
    seeds[seed] = True
    jump = np.zeros(u.shape, bool)
    jump[:, :band - 2] = jump[:, n + band + 2:] = True      # deep in the bands
    tip = np.array([[float(seed[0]), float(seed[1] - 60)]])   # closer across the seam
    geos, stops = G.geodesics(u, fl['vector'], fl['weightsum'], seeds, tip, chart_mapping=cont,
                              chart_jump=jump)
    g = geos[0]
    steps = np.abs(np.diff(g[:, 1]))
    assert stops[0] == 'AtSeed'
    assert np.sum(steps > n / 2) == 1                           # one jump, by the period
    Pp = O.make_problem('Isotropic', (n, n), (2. / n, 2. / n), seeds_idx=[(seed[0], seed[1] - band)],
                        periodic=[0, 1])
    up = Pp.solve_fm()
    flp = Pp.flow(up)
    seeds_p = np.zeros(up.shape, bool)
    seeds_p[seed[0], seed[1] - band] = True
    gp, sp_ = G.geodesics(up, flp['vector'], flp['weightsum'], seeds_p, tip - [0, band], periodic=[0, 1])
    assert sp_[0] == 'AtSeed'
    assert abs(len(g) - len(gp[0])) <= 3
