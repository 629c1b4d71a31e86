"""Bit-exact checks of the device building blocks through the C ABI:
Selling stencils (weights and integer offsets) against the oracle, layout
conversions against numpy, geometry preparation against the numpy
restatement."""
import ctypes

import numpy as np
import pytest

from gpu_helpers import log

pytestmark = pytest.mark.gpu


def _lib():
    from adaptivegriddiscretizations_amd import _lib
    return _lib


def random_spd_flat(rng, d, n, cond):
    A = rng.normal(size=(n, d, d))
    Q, _ = np.linalg.qr(A)
    lam = np.exp(rng.uniform(0, np.log(cond), (n, d)))
    D = np.einsum('nij,nj,nkj->nik', Q, lam, Q)
    return np.stack([D[:, i, j] for i in range(d) for j in range(i + 1)], axis=1)


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_selling_stencils_bit_exact(O, d, dtype):
    import torch
    L = _lib()
    rng = np.random.default_rng(100 + d)
    # moderate, strong and extreme anisotropy: the last group reaches the
    # reduction's iteration cap and superbase entries beyond 2^11 (the
    # integer path of the 3D reduction)
    n = 21000
    m = np.concatenate([random_spd_flat(rng, d, n // 3, c) for c in (10., 1e4, 1e7)])
    m = m.astype(dtype)
    geom = torch.from_numpy(np.ascontiguousarray(m.T)).cuda()  # component-major
    sd = d * (d + 1) // 2
    w = torch.empty((n, sd), dtype=geom.dtype, device='cuda')
    e = torch.empty((n, sd, d), dtype=torch.int32, device='cuda')
    L.check(L.lib.agd_stencil_export(L.RIEMANN, d, L.F32 if dtype == np.float32 else L.F64, n,
                                     geom.data_ptr(), None, w.data_ptr(), e.data_ptr(), None))
    torch.cuda.synchronize()
    w_ref, e_ref = O.selling(m.astype(np.float64), use_float=(dtype == np.float32))
    e_gpu = e.cpu().numpy()
    w_gpu = w.cpu().numpy().astype(np.float64)
    ndiff = int(np.sum(np.any(e_gpu != e_ref, axis=(1, 2))))
    log(f"selling d={d} {np.dtype(dtype).name}: cells with differing offsets {ndiff}/{n}")
    assert ndiff == 0
    assert np.array_equal(w_gpu, w_ref)


@pytest.mark.parametrize("shape,shape_i", [((37, 50), (8, 8)), ((64, 64), (16, 16)),
                                           ((9, 13, 10), (4, 4, 4)), ((20, 18, 7), (8, 8, 4))])
def test_block_expand_squeeze(O, shape, shape_i):
    import torch
    from adaptivegriddiscretizations_amd import FiniteDifferences as fd
    L = _lib()
    rng = np.random.default_rng(0)
    a = rng.normal(size=(3,) + shape).astype(np.float32)
    src = torch.from_numpy(a).cuda()
    ref = O.tile_expand(a, shape_i, np.inf)
    dst = torch.empty(ref.size, dtype=torch.float32, device='cuda')
    L.check(L.lib.agd_block_expand(len(shape), L.i64(shape), L.i32(shape_i), 3, 4, src.data_ptr(), 1,
                                   dst.data_ptr(), float('inf'), None))
    torch.cuda.synchronize()
    assert np.array_equal(dst.cpu().numpy(), ref.reshape(-1))
    back = torch.empty_like(src)
    L.check(L.lib.agd_block_squeeze(len(shape), L.i64(shape), L.i32(shape_i), 3, 4, dst.data_ptr(),
                                    back.data_ptr(), None))
    torch.cuda.synchronize()
    assert np.array_equal(back.cpu().numpy(), a)
    # tiled_index agrees with the device layout
    idx = np.stack(np.unravel_index(np.arange(int(np.prod(shape))), shape), axis=-1)
    t = fd.tiled_index(idx, shape_i, fd.round_up_ratio(shape, shape_i))
    np.testing.assert_array_equal(dst.cpu().numpy()[t], a[0].reshape(-1))


@pytest.mark.parametrize("d,model", [(2, 'Riemann'), (3, 'Riemann'), (2, 'Rander'), (3, 'Rander')])
def test_prepare_geometry(O, d, model):
    import torch
    L = _lib()
    rng = np.random.default_rng(d)
    shape = (11, 7) if d == 2 else (5, 6, 7)
    shape_i = (8, 8) if d == 2 else (4, 4, 4)
    A = rng.normal(size=(d, d) + shape)
    m = (np.einsum('ki...,kj...->ij...', A, A) + 0.5 * np.eye(d).reshape((d, d) + (1,) * d)).astype(np.float32)
    w = (0.2 * rng.normal(size=(d,) + shape)).astype(np.float32)
    h = np.array([0.1, 0.2, 0.05][:d])
    P = O.make_problem(model, shape, h, metric=m.astype(np.float64), drift=w.astype(np.float64))
    ref = O.tile_expand(P.geom.reshape((-1,) + shape), shape_i, np.inf)
    ng = ref.shape[0]
    out = torch.empty(ref.size, dtype=torch.float32, device='cuda')
    md, wd = torch.from_numpy(m).cuda(), torch.from_numpy(w).cuda()
    hh = (ctypes.c_double * 3)(*(list(h) + [1.] * (3 - d)))
    L.check(L.lib.agd_prepare_geometry(L.RIEMANN if model == 'Riemann' else L.RANDER, d, L.i64(shape),
                                       L.i32(shape_i), L.F32, hh, md.data_ptr(), wd.data_ptr(), 1,
                                       out.data_ptr(), None))
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(ng, -1)
    ref = ref.reshape(ng, -1)
    np.testing.assert_allclose(got, ref.astype(np.float32), rtol=2e-6, atol=1e-7 * np.nanmax(np.abs(ref[np.isfinite(ref)])))


@pytest.mark.parametrize("which", ["riemann2", "riemann3"])
def test_precomputed_scheme_codes(O, which):
    """agd_eikonal_scheme stores, per tiled cell, the Selling weights
    (bit-identical to agd_stencil_export on the same geometry) and the
    neighbour codes x -/+ e_k resolved on the host from the exported offsets:
    tiled index outside the tile, -2 - slot inside it, -1 outside the grid."""
    import torch
    from adaptivegriddiscretizations_amd import workloads
    from adaptivegriddiscretizations_amd.Eikonal import HFM_GPU
    L = _lib()
    if which == "riemann2":
        wl = workloads.riemann2_aniso(n=45, seed=3, sigma=3., amax=10.)
    else:
        wl = workloads.riemann3_aniso(n=14, seed=3, sigma=2.)
    iface = HFM_GPU.Interface(wl.dict_in(precompute_scheme=True))
    iface.SetKernelTraits()
    iface.SetGeometry()
    iface.SetArgs()
    torch.cuda.synchronize()
    d, N, si = iface.ndim, iface.size_tot, np.array(iface.shape_i)
    sd = d * (d + 1) // 2
    w = torch.empty((N, sd), dtype=torch.float32, device='cuda')
    e = torch.empty((N, sd, d), dtype=torch.int32, device='cuda')
    L.check(L.lib.agd_stencil_export(L.RIEMANN, d, L.F32, N, iface.geom.data_ptr(), None,
                                     w.data_ptr(), e.data_ptr(), None))
    torch.cuda.synchronize()
    w, e = w.cpu().numpy(), e.cpu().numpy()
    sw = iface.scheme_w_t.cpu().numpy().T
    sc = iface.scheme_c_t.cpu().numpy().T
    shape, so = np.array(iface.shape), np.array(iface.shape_o)
    NI = int(np.prod(si))
    n_t = np.arange(N)
    xo = np.stack(np.unravel_index(n_t // NI, tuple(so)), -1)
    xi = np.stack(np.unravel_index(n_t % NI, tuple(si)), -1)
    xt = xo * si + xi
    dom = np.all(xt < shape, axis=1)
    assert np.array_equal(sw[dom], w[dom])
    for k in range(sd):
        for s, sg in ((0, -1), (1, 1)):
            y = xt + sg * e[:, k, :]
            inside = np.all((y >= 0) & (y < shape), axis=1)
            yc = np.clip(y, 0, shape - 1)
            yo, yi = yc // si, yc % si
            local = np.all(yo == xo, axis=1)
            slot = np.ravel_multi_index(tuple(yi.T), tuple(si))
            tiled = np.ravel_multi_index(tuple(yo.T), tuple(so)) * NI + slot
            ref = np.where(~inside, -1, np.where(local, -2 - slot, tiled))
            got = sc[:, 2 * k + s]
            nbad = int(np.sum(got[dom] != ref[dom]))
            assert nbad == 0, f"component {k} side {sg}: {nbad} cells differ"
    log(f"{which}: precomputed scheme of {int(dom.sum())} cells matches")
