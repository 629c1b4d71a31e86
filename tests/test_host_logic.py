"""Host-side logic that needs no device: grid helpers and seed rounding of
dictIn, tiled layout helpers, metric classes."""
import numpy as np
import pytest

from adaptivegriddiscretizations_amd import Eikonal, Metrics
from adaptivegriddiscretizations_amd import FiniteDifferences as fd


def test_setrect_dims_cell_centred(O):
    hfmIn = Eikonal.dictIn({'model': 'Isotropic2'})
    hfmIn.SetRect([[-1, 1], [-1, 1]], dims=[512, 512])
    assert hfmIn.shape == (512, 512)
    h, origin = O.grid_h([[-1, 1], [-1, 1]], [512, 512])
    np.testing.assert_allclose(hfmIn.gridScales, h)
    np.testing.assert_allclose(hfmIn['origin'], origin)
    X = hfmIn.Grid()
    np.testing.assert_allclose(X[0][:, 0], -1 + (np.arange(512) + 0.5) * h[0], rtol=1e-6)


def test_seed_on_rounding_tie_goes_to_even(O):
    """Seed (0,0) on 512 cells over [-1,1]: continuous index 255.5 -> 256."""
    hfmIn = Eikonal.dictIn({'model': 'Isotropic2'})
    hfmIn.SetRect([[-1, 1], [-1, 1]], dims=[512, 512])
    cont = hfmIn.PointFromIndex(np.zeros((1, 2), dtype=np.float32), to=True)
    assert cont.tolist() == [[255.5, 255.5]]
    assert np.round(cont).astype(int).tolist() == [[256, 256]]
    h, origin = O.grid_h([[-1, 1], [-1, 1]], [512, 512])
    assert O.seed_indices([[0., 0.]], origin, h).tolist() == [[256, 256]]


def test_setrect_sample_boundary_and_round_trip():
    hfmIn = Eikonal.dictIn({'model': 'Riemann2'})
    hfmIn.SetRect([[0, 1], [0, 2]], dimx=51, sampleBoundary=True)
    assert hfmIn.shape == (51, 101)
    p = hfmIn.PointFromIndex(np.array([[3, 7]]))
    np.testing.assert_allclose(p, [[3 * 0.02, 7 * 0.02]], atol=1e-6)
    idx, err = hfmIn.IndexFromPoint(p)
    assert idx.tolist() == [[3, 7]]
    assert np.max(np.abs(err)) < 1e-4


def test_single_key_aliases_and_float_cast():
    hfmIn = Eikonal.dictIn({'model': 'Isotropic2', 'seed': [0.5, 0.25], 'cost': 2})
    assert hfmIn['seeds'].shape == (1, 2) and hfmIn['seeds'].dtype == np.float32
    assert hfmIn['seed'].tolist() == [0.5, 0.25]
    assert hfmIn['cost'].dtype == np.float32
    with pytest.raises(ValueError):
        hfmIn['mode'] = 'cpu'


def test_cpu_modes_are_not_provided():
    hfmIn = Eikonal.dictIn({'model': 'Isotropic2', 'mode': 'cpu'})
    hfmIn.SetRect([[0, 1], [0, 1]], dims=[8, 8])
    with pytest.raises(NotImplementedError):
        hfmIn.Run()


@pytest.mark.parametrize("shape,shape_i", [((10, 13), (4, 4)), ((9, 5, 7), (4, 4, 4))])
def test_tiled_layout_and_tiled_index(O, shape, shape_i):
    rng = np.random.default_rng(0)
    a = rng.normal(size=(2,) + shape)
    b = O.tile_expand(a, shape_i, np.inf)
    shape_o = tuple(fd.round_up_ratio(shape, shape_i))
    assert b.shape == (2,) + shape_o + tuple(shape_i)
    np.testing.assert_array_equal(O.tile_squeeze(b, shape), a)
    # the padding holds the pad value and nothing else
    assert np.isinf(b).sum() == 2 * (int(np.prod(b.shape[1:])) - int(np.prod(shape)))
    idx = np.stack(np.unravel_index(np.arange(int(np.prod(shape))), shape), axis=-1)
    t = fd.tiled_index(idx, shape_i, shape_o)
    np.testing.assert_array_equal(b[0].reshape(-1)[t], a[0].reshape(-1))
    # closed form: cell x sits in tile x // shape_i at position x % shape_i
    x = idx[len(idx) // 2]
    tile, pos = tuple(x // np.array(shape_i)), tuple(x % np.array(shape_i))
    assert b[(0,) + tile + pos] == a[(0,) + tuple(x)]


def test_oracle_set_rect_closed_forms(O):
    """The oracle's SetRect arithmetic on closed-form cases: [-1,1]^2 with
    512 cells has h = 2^-8 and origin -1; [0,1] x [0,2] sampled with 51
    points on the first axis, faces included, has 51 x 101 points, h = 1/50,
    and the box of its cells starts half a cell below 0."""
    r = O.set_rect([[-1, 1], [-1, 1]], dims=[512, 512])
    assert r['gridScales'].tolist() == [2. ** -8] * 2
    assert r['dims'].tolist() == [512, 512] and r['origin'].tolist() == [-1., -1.]
    r = O.set_rect([[0, 1], [0, 2]], sample_boundary=True, dimx=51)
    assert r['gridScale'] == np.float32(1 / 50) and r['dims'].tolist() == [51, 101]
    assert np.all(np.abs(r['origin'] + np.float32(0.01)) <= 1e-8)
    r = O.set_rect([[0, 3]], gridScale=0.5)
    assert r['dims'].tolist() == [6] and r['origin'].tolist() == [0.]
    # a width that is not a whole number of cells: the grid is re-centred
    r = O.set_rect([[0, 1.3125]], gridScale=0.25)     # 5.25 cells
    assert r['dims'].tolist() == [5] and r['origin'].tolist() == [0.03125]
    r = O.set_rect([[0, 1.375]], gridScale=0.25)      # 5.5 cells: a tie, rounded to even
    assert r['dims'].tolist() == [6] and r['origin'].tolist() == [-0.0625]
    # cell centres of [a, b]
    np.testing.assert_array_equal(O.centred_axis(0., 1., 4), [0.125, 0.375, 0.625, 0.875])
    assert O.point_from_index([-1., -1.], [2. ** -8] * 2, [[255.5, 255.5]]).tolist() == [[0., 0.]]


def test_set_rect_and_point_maps_match_oracle_bitwise(O):
    """dictIn's SetRect / PointFromIndex / Axes reproduce the float32
    arithmetic of DictIn.py:97-107,404-453 to the bit, on random boxes and
    every way of giving the grid scale."""
    rng = np.random.default_rng(5)
    checked = 0
    for trial in range(600):
        d = int(rng.integers(1, 4))
        lo = rng.uniform(-5, 5, d).round(int(rng.integers(0, 4)))
        hi = lo + rng.uniform(0.1, 10, d).round(int(rng.integers(0, 4)))
        sides = np.stack([lo, hi], 1).tolist()
        sb = bool(rng.integers(0, 2))
        kind = int(rng.integers(0, 4))
        kw = [{'gridScale': float(rng.uniform(0.001, 0.5))},
              {'gridScales': rng.uniform(0.001, 0.5, d).tolist()},
              {'dimx': int(rng.integers(2, 3000))},
              {'dims': rng.integers(2, 3000, d).tolist()}][kind]
        hf = Eikonal.dictIn({'model': f'Isotropic{d}'})
        with np.errstate(all='ignore'):
            try:
                hf.SetRect(sides, sampleBoundary=sb, **kw)
            except ValueError:
                continue
        ref = O.set_rect(sides, np.float32, sb, **kw)
        for k, v in ref.items():
            got = np.asarray(hf[k])
            assert got.dtype == v.dtype and np.array_equal(got, v), (k, got, v)
        idx = rng.integers(0, 50, (4, d))
        p = hf.PointFromIndex(idx)
        np.testing.assert_array_equal(p, O.point_from_index(hf['origin'], hf.gridScales, idx))
        np.testing.assert_array_equal(hf.PointFromIndex(p, to=True),
                                      O.point_from_index(hf['origin'], hf.gridScales, p, to=True))
        if trial % 10 == 0:
            lo_c, hi_c = hf.corners
            for ax, a, b, n in zip(hf.Axes(), lo_c, hi_c, hf.shape):
                np.testing.assert_array_equal(ax, np.asarray(O.centred_axis(a, b, n), np.float32))
        checked += 1
    assert checked > 500


def test_aliases_copy_and_run_join():
    hfmIn = Eikonal.dictIn({'model': 'Isotropic2', 'tip': [0.1, 0.2]})
    assert hfmIn['tips'].shape == (1, 2) and 'tip' not in hfmIn.store
    del hfmIn['tip']
    assert 'tips' not in hfmIn
    hfmIn['seeds'] = [[0., 0.], [1., 1.]]
    with pytest.raises(ValueError):
        hfmIn['seed']
    c = hfmIn.copy()
    c['model'] = 'Isotropic3'
    assert hfmIn['model'] == 'Isotropic2'
    with pytest.raises(ValueError):
        Eikonal.dictIn({'mode': 'gpu', 'float_t': np.float32})['float_t'] = np.float64
    out = Eikonal.dictOut({'geodesics': [np.zeros((2, 3))]})
    assert out['geodesic'].shape == (2, 3)
    out['geodesic'] = np.ones(2)
    assert len(out['geodesics']) == 1


def test_riemann_dual_and_with_costs():
    rng = np.random.default_rng(1)
    A = rng.normal(size=(3, 3, 4))
    m = np.einsum('ki...,kj...->ij...', A, A) + np.eye(3)[:, :, None]
    R = Metrics.Riemann(m)
    D = R.dual()
    np.testing.assert_allclose(np.einsum('ij...,jk...->ik...', m, D.m), np.eye(3)[:, :, None] * np.ones(4),
                               atol=1e-10)
    h = np.array([0.1, 0.2, 0.3])
    Rh = R.with_costs(h)
    np.testing.assert_allclose(Rh.m, m * np.outer(h, h)[:, :, None])
    flat = R.flatten()
    assert flat.shape == (6, 4)
    np.testing.assert_allclose(Metrics.Riemann.expand(flat).m, m)


def test_rander_zermelo_round_trip():
    M = np.array([[2., 0.3], [0.3, 1.]])
    drift = np.array([0.2, -0.1])
    r = Metrics.Rander.from_Zermelo(M, drift)
    # the dual of the dual is the metric itself
    rr = r.dual().dual()
    np.testing.assert_allclose(rr.m, r.m, atol=1e-12)
    np.testing.assert_allclose(rr.w, r.w, atol=1e-12)
    # unit speed + drift: moving along the drift costs less than against it
    v = np.array([1., 0.])
    assert r.norm(v) < r.norm(-v)


def test_batch_bit_arrays_concatenate_at_bit_offsets():
    """Batched solves with 16-cell tiles: the problems' seed/trigger bit
    arrays (32 cells per word) concatenate at bit offsets that are not word
    aligned; the packed result holds every problem's bits at its offset."""
    import types
    import torch
    from adaptivegriddiscretizations_amd.Eikonal import batch
    rng = np.random.default_rng(3)
    sizes = [48, 80, 16, 64]
    bits = [rng.integers(0, 2, size=n) for n in sizes]

    def pack(b):
        b = np.concatenate([b, np.zeros((-len(b)) % 32, dtype=b.dtype)]).reshape(-1, 32)
        w = (b.astype(np.uint64) << np.arange(32, dtype=np.uint64)).sum(axis=1)
        return torch.from_numpy(w.astype(np.uint32).view(np.int32).copy())

    fake = types.SimpleNamespace(sizes=sizes,
                                 ifaces=[types.SimpleNamespace(seeds_t=pack(b)) for b in bits])
    got = batch.Batch._cat_bits(fake, 'seeds_t').numpy().view(np.uint32)
    ref = pack(np.concatenate(bits)).numpy().view(np.uint32)
    assert np.array_equal(got, ref)
    # word-aligned sizes take the plain concatenation
    fake.sizes = [64, 32]
    fake.ifaces = fake.ifaces[:2]
    fake.ifaces[0].seeds_t = pack(bits[1][:64])
    fake.ifaces[1].seeds_t = pack(bits[0][:32])
    got = batch.Batch._cat_bits(fake, 'seeds_t').numpy().view(np.uint32)
    assert np.array_equal(got, pack(np.concatenate([bits[1][:64], bits[0][:32]])).numpy().view(np.uint32))


def test_subwave_tile_shapes_are_batched_only():
    from adaptivegriddiscretizations_amd import _lib
    for shape in ((4, 4), (4, 8), (8, 4)):
        assert _lib.lib.agd_batch_tile_supported(_lib.ISOTROPIC, 2, _lib.i32(shape)) == 1
        assert _lib.lib.agd_tile_supported(_lib.ISOTROPIC, 2, _lib.i32(shape)) == 0
    assert _lib.lib.agd_batch_tile_supported(_lib.ISOTROPIC, 2, _lib.i32((8, 8))) == 1
    assert _lib.lib.agd_tile_supported(_lib.ISOTROPIC, 2, _lib.i32((8, 8))) == 1
    assert _lib.lib.agd_batch_tile_supported(_lib.ISOTROPIC, 2, _lib.i32((2, 2))) == 0


def test_curvature_grid_geometry():
    """SE grids (DictIn.py:323-372, 464-482): gridScales (h, h, 2 pi / nTheta),
    the angle axis starting half a cell below 0, so that cell k sits at
    angle 2 pi k / nTheta; OrientedPoints pairs each point with every angle."""
    from adaptivegriddiscretizations_amd.Eikonal import dictIn
    hfm = dictIn({'model': 'ReedsShepp2'})
    hfm.SetRect([[-1, 1], [0, 1]], dimx=20)
    hfm.nTheta = 12
    assert hfm.shape == (20, 10, 12) and hfm.vdim == 3 and hfm.nTheta == 12
    np.testing.assert_allclose(hfm.gridScales, [0.1, 0.1, 2 * np.pi / 12], rtol=1e-6)
    th = np.asarray(hfm.Axes()[2], dtype=np.float64)
    np.testing.assert_allclose(th, 2 * np.pi * np.arange(12) / 12, atol=1e-6)
    pts = hfm.OrientedPoints([[0.3, 0.4], [-0.5, 0.2]])
    assert pts.shape == (12, 2, 3)
    np.testing.assert_allclose(pts[5, 1], [-0.5, 0.2, th[5]], rtol=1e-6)
    idx = hfm.PointFromIndex(pts.reshape(-1, 3), to=True)
    np.testing.assert_allclose(idx[:, 2].reshape(12, 2), np.arange(12)[:, None] + 0 * idx[:2, 2], atol=1e-5)
    with pytest.raises(ValueError):
        dictIn({'model': 'Riemann2'}).nTheta


def test_curvature_traits_and_refusals():
    from adaptivegriddiscretizations_amd import _lib
    from adaptivegriddiscretizations_amd.Eikonal import HFM_GPU
    for model, fim in (('ReedsShepp2', 5), ('ReedsSheppForward2', 5), ('Elastica2', 4), ('Dubins2', 4)):
        traits, f = HFM_GPU.default_traits(model)
        assert traits['shape_i'] == (4, 4, 4) and f == fim
        mid = HFM_GPU._MODELS[model[:-1]]
        assert _lib.lib.agd_tile_supported(mid, 3, _lib.i32((4, 4, 4))) == 1
        assert _lib.lib.agd_tile_supported(mid, 2, _lib.i32((8, 8))) == 0
    assert HFM_GPU.scheme_slots('Elastica', 3) == 30 and HFM_GPU.scheme_slots('ReedsSheppForward', 3) == 7
