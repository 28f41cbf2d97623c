"""The learned router: its FP64 oracle pinned by finite differences, and
(marked gpu) the device's forward/backward kernels against that oracle, the
training loop, and what routed trees gain over the median forest without them.

The router has no counterpart in the reference (euclideandistancematrixgpu.cu
draws split directions from random pivot pairs, :2138-2164), so its oracle is
pinned by mathematics: central differences of the FP64 loss give the
gradient the backward pass must produce.
"""
import numpy as np
import pytest

D, DP, H, NB = 5, 16, 8, 4


def _params(rng, d=D, dp=DP, h=H, nb=NB):
    w1 = np.zeros((h, dp), np.float32)
    w1[:, :d] = rng.normal(0, 0.7, (h, d))
    parts = [w1.ravel(), rng.normal(0, 0.1, h), rng.normal(0, 0.4, h * h), rng.normal(0, 0.1, h),
             rng.normal(0, 0.4, nb * h), rng.normal(0, 0.1, nb)]
    return np.concatenate([np.asarray(p, np.float32).ravel() for p in parts]).astype(np.float32)


def test_oracle_gradient_matches_finite_differences(oracle):
    rng = np.random.default_rng(0)
    p = _params(rng)
    x = rng.normal(size=(6, D)).astype(np.float32)
    mu = rng.normal(0, 0.1, D).astype(np.float32)
    lab = np.array([0, 3, 1, -1, 2, 3], np.int32)
    loss, g, _ = oracle.mlp(x, (D, DP, H, NB), p, mu, 1.3, lab, round_bf16=False)
    assert np.isfinite(loss) and loss > 0
    # parameters are float32: perturb through float64 copies is impossible, so
    # use steps large enough to be exact in float32 and a second-order check
    idx = [i for i in range(len(p)) if not (i < H * DP and (i % DP) >= D)]
    idx = rng.choice(idx, 60, replace=False)
    for i in idx:
        eps = np.float32(2.0 ** -10)
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp = oracle.mlp(x, (D, DP, H, NB), pp, mu, 1.3, lab, round_bf16=False, want_grads=False)[0]
        lm = oracle.mlp(x, (D, DP, H, NB), pm, mu, 1.3, lab, round_bf16=False, want_grads=False)[0]
        fd = (lp - lm) / (2 * float(eps))
        assert abs(fd - g[i]) <= 2e-4 + 2e-3 * abs(fd), (i, fd, g[i])


def test_oracle_gradient_structure(oracle):
    rng = np.random.default_rng(1)
    p = _params(rng)
    x = rng.normal(size=(8, D)).astype(np.float32)
    mu = np.zeros(D, np.float32)
    lab = np.array([0, 1, 2, 3, 0, 1, -1, -1], np.int32)
    loss, g, z = oracle.mlp(x, (D, DP, H, NB), p, mu, 1.0, lab, round_bf16=False)
    # padding columns of W1 get no gradient
    assert (g[:H * DP].reshape(H, DP)[:, D:] == 0).all()
    # unlabelled rows contribute nothing
    l2, g2, _ = oracle.mlp(x[:6], (D, DP, H, NB), p, mu, 1.0, lab[:6], round_bf16=False)
    assert abs(l2 - loss) < 1e-12 and np.allclose(g2, g, rtol=1e-6, atol=1e-9)
    # the b3 gradient is the mean of softmax - onehot: it sums to zero
    assert abs(g[-NB:].sum()) < 1e-6
    # the BF16-rounded mode stays close to FP64
    lb, gb, zb = oracle.mlp(x, (D, DP, H, NB), p, mu, 1.0, lab, round_bf16=True)
    assert abs(lb - loss) < 0.05 * loss and np.linalg.norm(gb - g) < 0.05 * np.linalg.norm(g)


# ---- the device kernels -------------------------------------------------------

def _rel(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(b.astype(np.float64)), 1e-30))


@pytest.mark.gpu
@pytest.mark.parametrize("d,hidden,bin_bits,m", [(128, 256, 8, 300), (20, 64, 5, 64), (37, 32, 6, 1000)])
def test_gpu_router_gradients_match_oracle(aknn, oracle, d, hidden, bin_bits, m):
    """Forward logits and every gradient tensor of one batch against the FP64
    oracle with the device's BF16 rounding points (tolerance 2e-3 relative,
    Frobenius, per tensor; FP32 accumulation order is the only difference)."""
    print(f"[gpu] router grads d={d} h={hidden} bits={bin_bits} m={m}", flush=True)
    rng = np.random.default_rng(d)
    r = aknn.Router(d, hidden, bin_bits, seed=3)
    p, mu, inv = r.get_params()
    mu = rng.normal(0, 0.5, d).astype(np.float32)
    r.set_params(p, mu, 0.8)
    x = rng.normal(0, 1.2, (m, d)).astype(np.float32)
    lab = rng.integers(0, 1 << bin_bits, m).astype(np.int32)
    lab[::7] = -1
    g, loss = r.gradients(x, lab)
    z = r.logits(x)
    p, mu, inv = r.get_params()
    ol, og, oz = oracle.mlp(x, (d, r.dp, hidden, 1 << bin_bits), p, mu, inv, lab, round_bf16=True)
    assert _rel(z, oz) < 2e-3
    assert abs(loss - ol) <= 1e-4 * abs(ol)
    for name, a, b in zip(("W1", "b1", "W2", "b2", "W3", "b3"), r.split_params(g), r.split_params(og)):
        assert _rel(a, b) < 2e-3, name
    r.close()


@pytest.mark.gpu
def test_gpu_router_save_load_roundtrip(aknn, tmp_path):
    r = aknn.Router(16, 32, 5, seed=9)
    x = np.random.default_rng(2).normal(size=(50, 16)).astype(np.float32)
    z0 = r.logits(x)
    r.save(str(tmp_path / "r.bin"))
    r2 = aknn.Router(16, 32, 5, seed=10)
    assert not np.array_equal(r2.logits(x), z0)
    r2.load(str(tmp_path / "r.bin"))
    assert np.array_equal(r2.logits(x), z0)
    with pytest.raises(aknn.AknnError):
        aknn.Router(17, 32, 5).load(str(tmp_path / "r.bin"))


@pytest.mark.gpu
def test_gpu_router_trains_and_routed_trees_gain(aknn, oracle, datasets):
    """Train routers on a graph the method built (C5 distribution, N=100,000,
    its own sample seed), then on a fresh sample compare median forests of
    the same size with and without routed tops.

    Each routed tree takes its own router (routers started from different
    trees' leaf orders), so the trees keep different top partitions.  The
    labels must cut fewer graph edges than the initial partition, the router
    must learn them well above chance, and the routed forest's initial graph
    (no propagation, where the forest alone decides) must reach a higher
    recall than the unrouted one."""
    k, leaf, trees = 20, 128, 4
    xt = datasets.generate("C5", n=100000, sample_seed=105)
    gt = aknn.Graph()
    gt.initialize(xt, k, 8, leaf, 3, 0, split=1)
    gt.calculate()
    routers = []
    for seed in range(1, trees + 1):
        r = aknn.Router(xt.shape[1], 256, 8, seed=seed)
        lab, cut = r.labels(gt, sweeps=20)
        assert cut[1] < 0.8 * cut[0]
        assert np.bincount(lab, minlength=256).max() <= int(np.ceil(1.05 * len(xt) / 256))
        st = r.fit(gt, lab, epochs=10, batch=1024, lr=2e-3)
        print(f"[gpu] router {seed}: cut {cut} fit {st}", flush=True)
        assert st["accuracy"] > 0.5 and st["loss"] < np.log(256) / 2
        routers.append(r)
    gt.close()

    x = datasets.generate("C5", n=100000, sample_seed=205)
    q = np.sort(np.random.default_rng(7).choice(len(x), 3000, replace=False))
    rec = {}
    for routed in (0, trees):
        g = aknn.Graph()
        g.initialize(x, k, trees, leaf, 0, 0, split=1, router_trees=routed)
        for r in routers:
            g.add_router(r)
        g.calculate()
        rec[routed] = g.recall(q)
        g.set_router(None)
        g.close()
    print(f"[gpu] initial-graph recall: unrouted {rec[0]:.4f} routed {rec[trees]:.4f}", flush=True)
    assert rec[trees] > rec[0] + 0.01
    for r in routers:
        r.close()


@pytest.mark.gpu
@pytest.mark.parametrize("d,hidden,bin_bits", [(128, 256, 8), (20, 256, 5), (96, 256, 6), (256, 256, 7),
                                               (300, 256, 8), (37, 64, 6)])
def test_gpu_fused_bins_equal_layerwise(aknn, d, hidden, bin_bits):
    """The one-kernel inference calculate() uses gives the layer-by-layer
    path's bins, which are the argmax of its logits (lowest bin on ties);
    shapes outside the fused kernel fall back to the layer-by-layer path."""
    print(f"[gpu] fused bins d={d} h={hidden} bits={bin_bits}", flush=True)
    rng = np.random.default_rng(d + bin_bits)
    r = aknn.Router(d, hidden, bin_bits, seed=5)
    p, mu, inv = r.get_params()
    r.set_params(p, rng.normal(0, 0.3, d).astype(np.float32), 0.7)
    m = 5000 + d  # not a multiple of the 128-row tile
    x = rng.normal(0, 1.5, (m, d)).astype(np.float32)
    x[7] = x[3]
    bf = r.bins_of(x, fused=True)
    bl = r.bins_of(x, fused=False)
    z = r.logits(x)
    assert (bf == bl).all()
    assert (bl == np.argmax(z, axis=1)).all()
    assert 0 <= bf.min() and bf.max() < (1 << bin_bits)
    # fewer rows than one tile
    assert (r.bins_of(x[:77], fused=True) == bl[:77]).all()
    r.close()
