"""Pins of the oracle's pair distribution g(r) and structure factor S(k)
(tdo_gr_grid / tdo_kshells / tdo_observables, restating
BosonsBulk::CalculateAdditionalSystemProperties, src/PhysicalSystems/BosonsBulk.cpp:474-520):

- golden vectors written by the compiled reference plug-in (tests/golden/obs_*.npz);
- the reference's own wave-vector shell files (tests/golden/kshells.json);
- closed form of S(k) for particles on a simple cubic lattice;
- g(r) times the bin volumes sums to the number of pairs inside the grid.
"""
import glob
import json
import os

import numpy as np
import pytest

from oracle import oracle as o

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_kshells_match_reference_files():
    with open(os.path.join(GOLDEN, "kshells.json")) as fh:
        ref = json.load(fh)["shells"]
    for D in (1, 2, 3):
        vec, cnt, n2 = o.kshells(D, 50)
        shells = ref[str(D)]["shells"]
        off = 0
        for s in range(50):
            got = vec[off:off + cnt[s]].tolist()
            assert got == shells[s], (D, s)
            off += cnt[s]
        # kNorm<D>D.csv holds |n| printed to ~16 digits
        np.testing.assert_allclose(np.sqrt(n2.astype(float)), ref[str(D)]["norms"], rtol=3e-15, atol=0)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "obs_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_observables_vs_reference(path):
    g = np.load(path)
    D, N, L, P, bins = int(g["D"]), int(g["N"]), float(g["L"]), int(g["P"]), int(g["bins"])
    gr, sk = o.observables(g["R"], L, P, L / 2, bins)
    assert gr.shape == g["gr"].shape
    np.testing.assert_allclose(gr, g["gr"], rtol=1e-14, atol=0)
    np.testing.assert_allclose(sk, g["sk"], rtol=1e-13, atol=1e-15)
    _, _, cen, _ = o.gr_grid(D, P, L / 2, bins)
    np.testing.assert_array_equal(cen, g["gr_grid"])
    _, _, n2 = o.kshells(D, 50)
    np.testing.assert_allclose(np.sqrt(n2.astype(float)) * 2 * np.pi / L, g["sk_grid"], rtol=3e-15)


@pytest.mark.parametrize("shift", [(0.0, 0.0, 0.0), (0.37, -1.21, 2.05)])
def test_sk_simple_cubic_lattice(shift):
    # N = m^3 particles on a lattice of spacing L/m, shifted by s: the sum over
    # particles of exp(i k.r) is N exp(i k.s) when every component of n is a
    # multiple of m and 0 otherwise.  The shell sums add these before squaring,
    # so S = N |sum_{n in shell, n = 0 mod m} exp(i k.s)|^2 / n_k.
    m, L = 4, 6.0
    g = np.arange(m) * (L / m)
    R = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + np.array(shift)
    N = len(R)
    _, sk = o.observables(R[None], L, 10, L / 2, 30)
    vec, cnt, _ = o.kshells(3, 50)
    off = 0
    for s in range(50):
        v = vec[off:off + cnt[s]]
        off += cnt[s]
        z = np.exp(1j * (v[np.all(v % m == 0, axis=1)] * (2 * np.pi / L)) @ np.array(shift)).sum()
        assert abs(sk[0, s] - N * abs(z) ** 2 / cnt[s]) < 1e-9 * N, (s, sk[0, s], z)


@pytest.mark.parametrize("D,N,L,P,bins,count", [(1, 17, 9.0, 12, 40, 40), (2, 30, 7.0, 14, 25, 24),
                                                 (3, 50, 5.0, 20, 64, 64)])
def test_gr_pair_count(D, N, L, P, bins, count):
    # Grid.cpp:22 truncates (gmax - 0) / (gmax / bins): for 3.5 / 25 that is
    # 24.999999999999996, so the grid has 24 bins, and the reference would
    # write the pairs of [24 spacing, gmax) past its histogram; they are dropped.
    rng = np.random.default_rng(D)
    R = rng.uniform(-L, L, (1, N, D))
    gr, _ = o.observables(R, L, P, L / 2, bins)
    gmax, sp, _, scl = o.gr_grid(D, P, L / 2, bins)
    assert len(scl) == count
    d = R[0][:, None, :] - R[0][None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d * d).sum(-1))[np.tril_indices(N, -1)]
    inside = (r < gmax) & (np.floor(r / sp) < count)
    total = np.sum(gr[0] * scl) * (N - 1) / D
    assert abs(total - inside.sum()) < 1e-9 * inside.sum()
    # and bin by bin
    counts = np.bincount(np.floor(r[inside] / sp).astype(int), minlength=len(scl))[:len(scl)]
    np.testing.assert_allclose(gr[0] * scl * (N - 1) / D, counts, rtol=1e-12, atol=1e-12)
