"""The principal-subspace screen of propagation (option pca_screen).

Projection on orthonormal directions never lengthens a distance, so the
distance of two points' projections, less their rounding radii and divided by
a bound of the basis' largest singular value, is a lower bound of their exact
FP32 key.  A candidate is dropped only when that bound is above the current
k-th key, where the exact comparison would refuse it too: graphs with the
screen on and off must be identical bit for bit, on data where the screen is
active (near a 16- or 32-dimensional subspace) and where it switches itself off.
The projections are kept in FP32 or FP16 (option pca_fp16); in FP16 each
point's radius includes its projection's rounding, and the graphs are the same.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _runs(aknn, x, k, trees, leaf, props, **opts):
    """Screen off, on with FP32 projections, on with FP16 projections."""
    out = []
    for pca, fp16 in ((0, 1), (1, 0), (1, 1)):
        g = aknn.Graph()
        g.initialize(x, k, trees, leaf, props, 0, pca_screen=pca, pca_fp16=fp16, **opts)
        g.calculate()
        out.append((g.graph(), g.candidate_counts("pca"), g.candidate_counts("propagate")))
        g.close()
    return out


def _same(a, b):
    return (a[0] == b[0]).all() and (bits(a[1]) == bits(b[1])).all()


@pytest.mark.parametrize("name,opts", [("C5", {}), ("C5", {"split": 1}), ("C5", {"sketch": 1, "leaf_mfma": 1}),
                                       ("C1", {}), ("C2", {})])
def test_pca_screen_identical_graphs(aknn, datasets, name, opts):
    x = datasets.generate(name, n=20000)
    if name == "C5":
        x[100:180] = x[7]  # duplicates: zero distances at the screen
    print(f"[gpu] pca screen {name} {opts}", flush=True)
    (g0, _, c0), (g1, pca, c1), (g2, _, c2) = _runs(aknn, x, 10, 3, 32, 3, **opts)
    assert _same(g0, g1) and _same(g0, g2)
    rank, retained = pca
    print(f"[gpu] rank {rank} retained16 {retained:.3f} examined {c1[0]:.0f} rescored {c1[1]:.0f} "
          f"(FP16 projections: {c2[1]:.0f})", flush=True)
    if name == "C5":
        # a 16-dimensional mixture: the screen is on and drops most candidates
        assert rank == 16 and retained > 0.97 and c1[1] < 0.5 * c1[0] and c2[1] < 0.5 * c2[0]
    if name == "C2":
        # uniform in 128 dimensions: no small subspace holds the data
        assert rank == 0


@pytest.mark.parametrize("d", [16, 24, 40])
def test_pca_screen_edge_widths(aknn, datasets, d):
    """D = 16 (the basis is a full rotation), D = 24 (no 32-wide basis), D = 40."""
    rng = np.random.default_rng(d)
    z = rng.normal(size=(8000, 12)).astype(np.float32)
    w = rng.normal(size=(12, d)).astype(np.float32)
    x = z @ w + 0.001 * rng.normal(size=(8000, d)).astype(np.float32)
    print(f"[gpu] pca edge d={d}", flush=True)
    (g0, _, _), (g1, pca, c1), (g2, _, _) = _runs(aknn, x, 8, 2, 24, 2)
    assert _same(g0, g1) and _same(g0, g2)
    assert pca[0] == 16
