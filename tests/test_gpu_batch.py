"""Batched instances and ensembles: every instance of a batched launch follows
the same trajectory as the same grid run alone (and as the oracle), the graph
replay equals the stream path, and the launch-shape tuner leaves the state alone."""
import numpy as np
import pytest

from ps3d_amd import ics

pytestmark = pytest.mark.gpu

PI = np.pi
TOL = 1.0e-12


@pytest.fixture(scope="module")
def G():
    from ps3d_amd import gpu
    gpu.lib()
    return gpu


def kw(**extra):
    d = dict(filtering="2/3-rule", stepper="impl-diff-rk4", nnu=3, prediss=30.0, pretype="vorch",
             length_scale="Kolmogorov", alpha=0.1, dt_fixed=0.0, limit=1.0e30)
    d.update(extra)
    return d


def grid(n, nz=None):
    return dict(nx=n, ny=n, nz=nz or n, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI))


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def ic(n, nz, seed):
    return ics.random_phase_isotropic(n, n, nz, k0=2.0 + (seed % 5), seed=seed)


@pytest.mark.parametrize("n,nz,nb", [(32, 32, 3), (16, 128, 2)])
def test_batch_instances_match_single_runs(G, n, nz, nb):
    p = G.make_params(**grid(n, nz), **kw())
    bat = G.Ps3dBatch(p, nb)
    singles = []
    for b in range(nb):
        vor = ic(n, nz, 10 + b)
        bat.set_vorticity(b, vor)
        s = G.Ps3dGpu(p)
        s.set_vorticity(vor)
        singles.append(s)
    print("batch", n, nz, nb, flush=True)
    diag = bat.advance_n(4, series=True)
    for b, s in enumerate(singles):
        ds = s.advance_n(4)
        assert abs(bat.t[b] - s.t) < 1e-15
        assert rel(bat.get_svor(b), s.get(G.SVOR)) < 1e-14
        assert np.abs(diag[b] - ds).max() <= 1e-14 * np.abs(ds).max()
        s.close()
    bat.close()


def test_batch_instance_matches_oracle(oracle_mod, G):
    g = grid(32)
    o = oracle_mod.Oracle(oracle_mod.make_params(**g, **kw()))
    bat = G.Ps3dBatch(G.make_params(**g, **kw()), 2)
    vor0, vor1 = ic(32, 32, 21), ic(32, 32, 22)
    bat.set_vorticity(0, vor0)
    bat.set_vorticity(1, vor1)
    o.set_vorticity(vor1)
    bat.advance_n(5)
    o.advance_n(5)
    assert rel(bat.get_svor(1), o.get(oracle_mod.SVOR)) < TOL
    bat.close()
    o.close()


def make_ensemble(G, sizes, counts):
    ens = G.Ps3dEnsemble([G.make_params(**grid(n), **kw()) for n in sizes], counts)
    vors = {}
    for gi, (n, c) in enumerate(zip(sizes, counts)):
        for b in range(c):
            v = ic(n, n, 100 * gi + b)
            ens.groups[gi].set_vorticity(b, v)
            vors[(gi, b)] = v
    return ens, vors


def test_ensemble_graph_equals_streams_and_singles(G):
    sizes, counts = [32, 48], [3, 2]
    e1, vors = make_ensemble(G, sizes, counts)
    e2, _ = make_ensemble(G, sizes, counts)
    print("ensemble graph", flush=True)
    e1.advance(3, graph=True)
    e2.advance(3, graph=False)
    for gi, n in enumerate(sizes):
        p = G.make_params(**grid(n), **kw())
        for b in range(counts[gi]):
            a = e1.groups[gi].get_svor(b)
            assert rel(a, e2.groups[gi].get_svor(b)) == 0.0
            s = G.Ps3dGpu(p)
            s.set_vorticity(vors[(gi, b)])
            s.advance_n(3)
            assert rel(a, s.get(G.SVOR)) < 1e-14
            s.close()
    e1.close()
    e2.close()


def test_ensemble_tuner_restores_state(G):
    sizes, counts = [32, 48], [2, 2]
    e1, _ = make_ensemble(G, sizes, counts)
    e2, _ = make_ensemble(G, sizes, counts)
    ch = e1.tune(steps=1)
    print("tuned shapes", ch, flush=True)
    assert all(c in (1, 2, 4, 8, 100, 101, 102) for c in ch)
    e1.advance(2)
    e2.advance(2)
    for gi in range(len(sizes)):
        for b in range(counts[gi]):
            assert rel(e1.groups[gi].get_svor(b), e2.groups[gi].get_svor(b)) < 1e-14
    e1.close()
    e2.close()
