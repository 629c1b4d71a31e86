"""Pins the oracle's Selling decomposition (Geometry2.h / Geometry3.h) against
the reference's stored outputs and against what the mathematics fixes."""
import itertools
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "selling_kat.json")))


def flat(D):
    D = np.asarray(D, dtype=np.float64)
    d = len(D)
    return np.array([D[i, j] for i in range(d) for j in range(i + 1)])


def canonical(weights, offsets, tol=0.):
    """Decomposition as a sorted set of (offset up to sign, weight)."""
    out = []
    for w, e in zip(weights, offsets):
        e = np.asarray(e, dtype=int)
        nz = np.flatnonzero(e)
        if len(nz) and e[nz[0]] < 0:
            e = -e
        if w > tol:
            out.append((tuple(e.tolist()), float(w)))
    return sorted(out)


def test_ordered_2d_kat(O):
    k = KAT["ordered_2d"]
    w, e = O.selling(flat(k["D"])[None])
    assert e[0].tolist() == k["offsets"]
    np.testing.assert_allclose(w[0], k["weights"], rtol=0, atol=1e-15)
    wf, ef = O.selling(flat(k["D"])[None], use_float=True)
    assert ef[0].tolist() == k["offsets"]
    np.testing.assert_allclose(wf[0], k["weights"], rtol=1e-6)


@pytest.mark.parametrize("name", ["notebook_2d", "notebook_3d"])
def test_notebook_kats(O, name):
    k = KAT[name]
    w, e = O.selling(flat(k["D"])[None])
    got, ref = canonical(w[0], e[0]), canonical(k["weights"], k["offsets"])
    assert [g[0] for g in got] == [r[0] for r in ref]
    np.testing.assert_allclose([g[1] for g in got], [r[1] for r in ref], rtol=1e-6)


def test_rotating_max_offset_norm(O):
    k = KAT["rotating_2d"]
    t = np.linspace(0, np.pi / 2, k["n"])
    u = np.array([np.cos(t), np.sin(t)])
    v = np.array([-np.sin(t), np.cos(t)])
    D = k["mu"] ** 2 * u[:, None] * u[None] + v[:, None] * v[None]
    m = np.stack([D[0, 0], D[1, 0], D[1, 1]], axis=1)
    w, e = O.selling(m)
    norms = np.linalg.norm(e, axis=-1)
    assert np.max(norms) == pytest.approx(k["max_offset_norm"], rel=1e-12)


def random_spd(rng, d, n, cond=100.):
    A = rng.normal(size=(n, d, d))
    Q, _ = np.linalg.qr(A)
    lam = np.exp(rng.uniform(0, np.log(cond), (n, d)))
    D = np.einsum('nij,nj,nkj->nik', Q, lam, Q)
    return np.stack([D[:, i, j] for i in range(d) for j in range(i + 1)], axis=1), D


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("cond", [2., 100., 1e4])
def test_reconstruction_and_positivity(O, d, cond):
    rng = np.random.default_rng(7 + d)
    m, D = random_spd(rng, d, 500, cond)
    w, e = O.selling(m)
    assert np.all(w >= 0)
    rec = np.einsum('nk,nki,nkj->nij', w, e, e)
    np.testing.assert_allclose(rec, D, rtol=1e-9, atol=1e-9 * np.abs(D).max())
    # offsets are primitive integer vectors spanning the lattice
    if d == 2:
        det = e[:, 0, 0] * e[:, 1, 1] - e[:, 0, 1] * e[:, 1, 0]
        assert np.all(np.abs(det) == 1)
    else:
        det = np.linalg.det(e[:, :3, :].astype(float))
        assert np.all(np.abs(np.round(det)) >= 1)


def brute_force_selling2(D, R=6):
    """Obtuse superbase by exhaustive search over small integer bases."""
    best = None
    rng = range(-R, R + 1)
    for b0 in itertools.product(rng, rng):
        for b1 in itertools.product(rng, rng):
            if b0[0] * b1[1] - b0[1] * b1[0] != 1:
                continue
            b = np.array([b0, b1, (-b0[0] - b1[0], -b0[1] - b1[1])], dtype=float)
            g = b @ D @ b.T
            if g[0, 1] <= 1e-12 and g[1, 2] <= 1e-12 and g[0, 2] <= 1e-12:
                best = b
                break
        if best is not None:
            break
    assert best is not None
    out = []
    for r in range(3):
        i, j, k = r, (r + 1) % 3, (r + 2) % 3
        w = -best[i] @ D @ best[j]
        e = np.array([-best[k][1], best[k][0]])
        out.append((w, e))
    return [o[0] for o in out], [o[1] for o in out]


def test_matches_exhaustive_obtuse_superbase_2d(O):
    rng = np.random.default_rng(3)
    m, D = random_spd(rng, 2, 25, cond=30.)
    w, e = O.selling(m)
    for n in range(len(m)):
        bw, be = brute_force_selling2(D[n])
        got, ref = canonical(w[n], e[n], 1e-12), canonical(bw, be, 1e-12)
        assert [g[0] for g in got] == [r[0] for r in ref]
        np.testing.assert_allclose([g[1] for g in got], [r[1] for r in ref], rtol=1e-9)


def test_float_and_double_agree_on_generic_matrices(O):
    rng = np.random.default_rng(11)
    for d in (2, 3):
        m, _ = random_spd(rng, d, 2000, cond=50.)
        m32 = m.astype(np.float32).astype(np.float64)
        w64, e64 = O.selling(m32)
        w32, e32 = O.selling(m32, use_float=True)
        same = np.all(e64 == e32, axis=(1, 2))
        # only near-degenerate ties may pick another (equivalent) superbase
        assert same.mean() > 0.99
        np.testing.assert_allclose(w32[same], w64[same], rtol=1e-4, atol=1e-6 * np.abs(m).max())
