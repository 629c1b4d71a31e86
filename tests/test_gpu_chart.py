"""Manifolds described by local charts on the GPU (ChartPaste.h, the paste loop
of _StopCriterion.py:60-80, the geodesic chart jump of GeodesicODE_Opt.h:43-99):
the paste kernel against the oracle's restatement through the C ABI, and the
reference's own notebook checks (Notebooks_Repro/Chart_GPU.ipynb: the
projective plane, the sphere from two charts, the circular band) against
their exact distances, plus the band against the periodic solve."""
import ctypes

import numpy as np
import pytest

from adaptivegriddiscretizations_amd import Eikonal, _lib
from gpu_helpers import log

pytestmark = pytest.mark.gpu


def tiled_paste(u, mapping, shape_i, scalar):
    """agd_chart_paste on cell-major u (expanded to the tiled layout and
    squeezed back): the new values, the lowered mask and the count."""
    import torch
    from adaptivegriddiscretizations_amd.Eikonal._device import device_tensor, empty, zeros
    from adaptivegriddiscretizations_amd.Eikonal._device import ptr
    dt = np.float32 if scalar == _lib.F32 else np.float64
    shape = u.shape
    d = len(shape)
    desc = _lib.EikonalDesc()
    desc.ndim, desc.scalar, desc.model = d, scalar, _lib.ISOTROPIC
    for k in range(3):
        desc.shape[k] = shape[k] if k < d else 1
        desc.shape_i[k] = shape_i[k] if k < d else 1
    so = [-(-shape[k] // shape_i[k]) for k in range(d)]
    size_tot = int(np.prod(so)) * int(np.prod(shape_i))
    u_c = device_tensor(np.ascontiguousarray(u, dtype=dt), dt)
    u_t = empty((size_tot,), dt)
    _lib.check(_lib.lib.agd_block_expand(d, _lib.i64(shape), _lib.i32(shape_i), 1, np.dtype(dt).itemsize,
                                         ptr(u_c), 1, ptr(u_t), float('inf'), None))
    m = device_tensor(np.ascontiguousarray(mapping, dtype=dt), dt)
    scratch = empty((size_tot,), dt)
    trig = zeros(((size_tot + 31) // 32,), np.int32)
    cnt = zeros((1,), np.int32)
    _lib.check(_lib.lib.agd_chart_paste(desc, ptr(u_t), ptr(m), mapping.shape[0], 0., 0., ptr(scratch),
                                        ptr(trig), ptr(cnt), None))
    out = empty(shape, dt)
    _lib.check(_lib.lib.agd_block_squeeze(d, _lib.i64(shape), _lib.i32(shape_i), 1, np.dtype(dt).itemsize,
                                          ptr(u_t), ptr(out), None))
    bits = trig.cpu().numpy().view(np.uint32)
    lowered_t = ((bits[np.arange(size_tot) >> 5] >> (np.arange(size_tot) & 31)) & 1).astype(np.uint8)
    low_t = device_tensor(lowered_t, np.uint8)
    low = empty(shape, np.uint8)
    _lib.check(_lib.lib.agd_block_squeeze(d, _lib.i64(shape), _lib.i32(shape_i), 1, 1, ptr(low_t), ptr(low),
                                          None))
    torch.cuda.synchronize()
    return out.cpu().numpy().astype(np.float64), low.cpu().numpy().astype(bool), int(cnt.item())


@pytest.mark.parametrize("scalar", [_lib.F64, _lib.F32])
def test_chart_paste_matches_oracle(O, scalar):
    rng = np.random.default_rng(4)
    shape, shape_i = (7, 13, 11), (4, 4, 4)
    X = np.indices(shape).astype(np.float64)
    u = 3. + np.sin(X[0] * 0.7) + np.cos(X[1] * 0.4 + X[2] * 0.3) + rng.uniform(0, 0.1, size=shape)
    u[2, 5, :] = np.inf                                   # unreached cells
    m = np.stack([rng.uniform(-1.5, 13.5, size=shape[1:]), rng.uniform(-1.5, 11.5, size=shape[1:])])
    m[0, 3, 4] = np.nan
    m[1, 0, :3] = [10 + 5e-5, -5e-5, 10 + 3e-4]            # boundary tolerance cases
    m[0, 1, :4] = [5.0, 5.0, 4.0, 3.0]                     # corners of zero weight
    if scalar == _lib.F32:
        u, m = u.astype(np.float32).astype(np.float64), m.astype(np.float32).astype(np.float64)
    new, low, cnt = tiled_paste(u, m, shape_i, scalar)
    ref, low_ref = O.chart_paste(u, m)
    log(f"chart paste ({'f64' if scalar == _lib.F64 else 'f32'}): {int(low_ref.sum())} of {u.size} cells lowered, "
        f"device {cnt}")
    if scalar == _lib.F64:
        np.testing.assert_array_equal(low, low_ref)
        np.testing.assert_allclose(new, ref, rtol=1e-14)
    else:
        # FP32 interpolation: cells whose pasted value ties the old one within
        # rounding may differ in the mask only
        tie = np.abs(np.where(np.isfinite(ref), ref, 0) - np.where(np.isfinite(u), u, 0)) <= 1e-5 * np.abs(u)
        assert np.all((low == low_ref) | tie)
        fin = np.isfinite(ref)
        np.testing.assert_allclose(new[fin], ref[fin], rtol=1e-5)
    assert cnt == int(low.sum())


def band_input(n=100, r=1.1, **extra):
    hfm = Eikonal.dictIn({'model': 'Isotropic2', 'seed': [0.5, 0.3], 'exportValues': True, 'cost': 1,
                          'factoringRadius': 10, 'verbosity': 0, **extra})
    hfm.SetRect([[-1, 1], [-r, r]], dimx=n)
    return hfm


def test_circular_band(O):
    """Chart_GPU.ipynb section 3: periodicity along y from a chart mapping
    broadcast over x; exact distance within 0.02, and the periodic solve on
    the fundamental domain."""
    r = 1.1
    hfm = band_input(r=r)
    hfm.SetUniformTips((4, 4))
    _, aY = hfm.Axes()
    aY = np.asarray(aY, dtype=np.float64)
    hfm['chart_mapping'] = np.expand_dims(np.where(aY > 0, aY - 2, aY + 2), axis=0)
    hfm['chart_jump'] = np.abs(aY) > (1 + r) / 2
    out = hfm.Run()
    u = out['values'].cpu().numpy().astype(np.float64)
    X = np.stack(np.meshgrid(*[np.asarray(a, dtype=np.float64) for a in hfm.Axes()], indexing='ij'))
    seed = np.array([0.5, 0.3])
    exact = np.minimum(np.linalg.norm(X - seed[:, None, None], axis=0),
                       np.linalg.norm(X - (seed - [0, 2])[:, None, None], axis=0))
    err = np.max(np.abs(exact - u))
    stops = out['geodesic_stopping_criteria']
    log(f"circular band: pastes {out['stats']['eikonal'].get('chart_pastes')}, max error {err:.3e}, "
        f"geodesic stops {sorted(set(stops))}")
    assert err < 0.02
    # against the periodic solve of the fundamental domain [-1, 1)
    per = Eikonal.dictIn({'model': 'Isotropic2', 'seed': [0.5, 0.3], 'exportValues': True, 'cost': 1,
                          'factoringRadius': 10, 'verbosity': 0, 'periodic': [False, True]})
    per.SetRect([[-1, 1], [-1, 1]], dimx=100)
    up = per.Run()['values'].cpu().numpy().astype(np.float64)
    core = u[:, 5:105]
    gap = np.max(np.abs(core - up))
    log(f"circular band vs periodic: max gap {gap:.3e}")
    assert gap < 2e-3
    assert all(s == 'AtSeed' for s in stops)


def test_projective_plane():
    """Chart_GPU.ipynb section 1: the projective plane as the unit sphere
    with antipodes identified, parametrised from the equator; exact spherical
    distance within 0.02; the geodesics jump and reach the seed."""
    n, r = 201, 1.1
    hfm = Eikonal.dictIn({'model': 'Isotropic2', 'seed': [0.5, 0.7], 'exportValues': True,
                          'factoringRadius': 10, 'verbosity': 0})
    hfm.SetRect([[-r, r], [-r, r]], dimx=n)
    X = np.stack(np.meshgrid(*[np.asarray(a, dtype=np.float64) for a in hfm.Axes()], indexing='ij'))
    e2 = np.sum(X ** 2, axis=0)
    hfm['cost'] = 2. / (1. + e2)
    hfm.SetUniformTips((4, 4))
    with np.errstate(divide='ignore', invalid='ignore'):
        hfm['chart_mapping'] = -X / e2             # plane_from_sphere(-sphere_from_plane(x))
    hfm['chart_jump'] = np.sqrt(e2) > (1 + r) / 2
    out = hfm.Run()
    u = out['values'].cpu().numpy().astype(np.float64)

    def to_sphere(x):
        s2 = np.sum(x ** 2, axis=0)
        return np.concatenate([(1. - s2)[None], 2 * x]) / (1. + s2)

    seed = to_sphere(np.array([0.5, 0.7]).reshape(2, 1, 1))
    XS = to_sphere(X)
    dist = lambda p, q: 2 * np.arcsin(np.linalg.norm(p - q, axis=0) / 2)
    exact = np.minimum(dist(seed, XS), dist(-seed, XS))
    err = np.max(np.abs(exact - u))
    stops = out['geodesic_stopping_criteria']
    log(f"projective plane {n}^2: pastes {out['stats']['eikonal'].get('chart_pastes')}, max error {err:.3e}, "
        f"geodesics {len(stops)}, AtSeed {stops.count('AtSeed')}")
    assert err < 0.02
    assert stops.count('AtSeed') >= len(stops) - 1


def test_sphere_from_two_charts():
    """Chart_GPU.ipynb section 2: two charts side by side separated by a
    wall, each the stereographic image of a hemisphere; exact spherical
    distance within 0.03 outside the wall."""
    r = 1.1
    hfm = Eikonal.dictIn({'model': 'Isotropic2', 'seed': [0.5 - r, 0.7], 'exportValues': True,
                          'factoringRadius': 10, 'verbosity': 0})
    hfm.SetRect([[-2 * r, 2 * r], [-r, r]], dimx=401)
    X = np.stack(np.meshgrid(*[np.asarray(a, dtype=np.float64) for a in hfm.Axes()], indexing='ij'))

    def local(x):
        x = x.copy()
        s = np.sign(x[0])
        x[0] -= s * r
        return x, s

    xl, s = local(X)
    e2 = np.sum(xl ** 2, axis=0)
    hfm['cost'] = 2. / (1. + e2)
    with np.errstate(divide='ignore', invalid='ignore'):
        y = -xl / e2
    y[:, np.abs(y[0]) > r] = np.nan
    y[0] -= s * r
    walls = np.abs(X[0]) < 0.02
    hfm['walls'] = walls
    hfm['chart_mapping'] = np.where(walls, np.nan, y)
    hfm['chart_nitermax'] = 10
    out = hfm.Run()
    u = out['values'].cpu().numpy().astype(np.float64)

    def to_sphere(x):
        xl, s = local(x)
        s2 = np.sum(xl ** 2, axis=0)
        return s * np.concatenate([(1. - s2)[None], 2 * xl]) / (1. + s2)

    seed = to_sphere(np.array([0.5 - r, 0.7]).reshape(2, 1, 1))
    dist = 2 * np.arcsin(np.linalg.norm(to_sphere(X) - seed, axis=0) / 2)
    err = np.max(np.abs(np.where(walls, 0., u - dist)))
    log(f"sphere from two charts: pastes {out['stats']['eikonal'].get('chart_pastes')}, max error {err:.3e}")
    assert err < 0.03


def test_band_geodesics_match_oracle():
    """Geodesics jumping across the glued band: the device against the
    geodesic oracle's chart jump on the same values and flow (same stopping
    criteria, lengths within 2 points, points within 2e-2 pixel)."""
    import geodesic_oracle as G
    from test_gpu_geodesic import oracle_inputs
    r = 1.1
    hfm = band_input(r=r)
    del hfm['factoringRadius']
    _, aY = hfm.Axes()
    aY = np.asarray(aY, dtype=np.float64)
    hfm['chart_mapping'] = np.expand_dims(np.where(aY > 0, aY - 2, aY + 2), axis=0)
    hfm['chart_jump'] = np.abs(aY) > (1 + r) / 2
    hfm['tips'] = np.array([[0.5, -0.8], [-0.6, -0.9], [0.2, 0.95], [0.9, -0.3]])
    hfm['retself'] = True
    out, iface = hfm.Run()
    u, fv, ws, seeds = oracle_inputs(hfm, iface)
    mapping = iface.chart_mapping_t.cpu().numpy().astype(np.float64)
    jump = np.broadcast_to(np.asarray(hfm['chart_jump'], bool), u.shape)
    tips = hfm.PointFromIndex(np.asarray(hfm['tips']).reshape(-1, 2), to=True)
    geos_o, stops_o = G.geodesics(u, fv, ws, seeds, tips, chart_mapping=mapping, chart_jump=jump)
    worst, dlen, jumps = 0., 0, 0
    for geo, go, so, sg in zip(out['geodesics'], geos_o, stops_o, out['geodesic_stopping_criteria']):
        gi = hfm.PointFromIndex(geo.T, to=True)
        assert sg == so
        m = min(len(gi), len(go))
        dlen = max(dlen, abs(len(gi) - len(go)))
        worst = max(worst, float(np.abs(gi[:m] - go[:m]).max()))
        jumps += int(np.sum(np.abs(np.diff(gi[:, 1])) > 50))
    log(f"band geodesics: stops {stops_o}, jumps {jumps}, max point difference {worst:.2e} px, "
        f"max length difference {dlen}")
    assert jumps >= 2
    assert dlen <= 2
    assert worst <= 2e-2
