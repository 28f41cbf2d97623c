"""Seeded synthetic inputs of the benchmark configurations.

C1  N=20,000, D=32: Gaussian mixture of 10 equal clusters, centres ~ N(0, 25 I),
    within-cluster std 1, integer cluster label column.
C2  N=1,000,000, D=128: uniform integers 0..255 cast to float.
C3  N=1,000,000, D=784: 64 clusters with Zipf(1.2) sizes; each cluster fixes a
    549-of-784 zero mask (70% exact zeros), non-zeros uniform in (0, 1].
C4  N=2,000,000, D=96: 1,000 clusters with Zipf(1.1) sizes, centres ~ N(0, 25 I),
    x = centre + N(0, I), then x / ||x||.
C5  N=4,000,000, D=128: z from a 16-D mixture of 64 equal clusters (centres
    ~ N(0, 25 I), covariance A A^T with A ~ N(0, 1/16)), x = W z + N(0, 0.01^2),
    W ~ N(0, 1/16) of shape 128 x 16.

Every generator takes n so that parity tests can use the same distribution at
small sizes.  Rows are shuffled by a seeded permutation.
"""
import numpy as np

CONFIGS = {
    "C1": dict(n=20_000, d=32, k=10, trees=10, leaf=32, props=3, metric="EUCLIDEAN"),
    "C2": dict(n=1_000_000, d=128, k=20, trees=16, leaf=256, props=5, metric="EUCLIDEAN"),
    "C3": dict(n=1_000_000, d=784, k=20, trees=12, leaf=512, props=4, metric="COSINE"),
    "C4": dict(n=2_000_000, d=96, k=30, trees=24, leaf=384, props=6, metric="COSINE"),
    "C5": dict(n=4_000_000, d=128, k=20, trees=6, leaf=256, props=2, metric="EUCLIDEAN"),
}


def _zipf_sizes(n, clusters, s, rng):
    w = 1.0 / np.arange(1, clusters + 1) ** s
    sizes = np.floor(w / w.sum() * n).astype(np.int64)
    sizes[: n - sizes.sum()] += 1
    return rng.permutation(sizes)


def gmm_c1(n=20_000, d=32, clusters=10, seed=1):
    """Returns (x float32 [n][d], labels int32 [n])."""
    rng = np.random.default_rng(seed)
    centres = rng.normal(0.0, 5.0, (clusters, d))
    labels = np.repeat(np.arange(clusters), -(-n // clusters))[:n]
    x = centres[labels] + rng.normal(0.0, 1.0, (n, d))
    p = rng.permutation(n)
    return x[p].astype(np.float32), labels[p].astype(np.int32)


def uniform_c2(n=1_000_000, d=128, seed=2):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 256, size=(n, d), dtype=np.int32).astype(np.float32)


def sparse_c3(n=1_000_000, d=784, clusters=64, seed=3):
    rng = np.random.default_rng(seed)
    sizes = _zipf_sizes(n, clusters, 1.2, rng)
    nnz = d - int(round(0.7 * d))
    x = np.zeros((n, d), np.float32)
    row = 0
    for c in range(clusters):
        cols = np.sort(rng.choice(d, nnz, replace=False))
        m = int(sizes[c])
        x[row:row + m][:, cols] = (1.0 - rng.random((m, nnz))).astype(np.float32)
        row += m
    return x[rng.permutation(n)]


def sphere_c4(n=2_000_000, d=96, clusters=1000, seed=4):
    rng = np.random.default_rng(seed)
    sizes = _zipf_sizes(n, clusters, 1.1, rng)
    labels = np.repeat(np.arange(clusters), sizes)
    centres = rng.normal(0.0, 5.0, (clusters, d)).astype(np.float32)
    x = centres[labels] + rng.normal(0.0, 1.0, (n, d)).astype(np.float32)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    return x[rng.permutation(n)].astype(np.float32)


def embedded_c5(n=4_000_000, d=128, latent=16, clusters=64, seed=5, sample_seed=None):
    """seed fixes the distribution (centres, covariances, embedding) and, when
    sample_seed is None, also the sample; a sample_seed draws another sample
    of the same distribution (the router's training and test sets)."""
    rng = np.random.default_rng(seed)
    centres = rng.normal(0.0, 5.0, (clusters, latent))
    a = rng.normal(0.0, 0.25, (clusters, latent, latent))
    w = rng.normal(0.0, 0.25, (d, latent)).astype(np.float32)
    if sample_seed is not None:
        rng = np.random.default_rng(sample_seed)
    labels = rng.integers(0, clusters, n)
    z = np.empty((n, latent), np.float32)
    for c in range(clusters):
        idx = np.nonzero(labels == c)[0]
        z[idx] = centres[c] + rng.normal(0.0, 1.0, (len(idx), latent)) @ a[c].T
    x = z @ w.T + rng.normal(0.0, 0.01, (n, d)).astype(np.float32)
    return x.astype(np.float32)


def generate(config, n=None, seed=None, sample_seed=None):
    """Points of configuration `config` ("C1".."C5"), optionally at another
    size; sample_seed (C5) draws another sample of the same distribution."""
    c = CONFIGS[config]
    n = c["n"] if n is None else n
    kw = {} if seed is None else {"seed": seed}
    if sample_seed is not None:
        if config != "C5":
            raise ValueError("sample_seed applies to C5")
        kw["sample_seed"] = sample_seed
    if config == "C1":
        return gmm_c1(n, c["d"], **kw)[0]
    if config == "C2":
        return uniform_c2(n, c["d"], **kw)
    if config == "C3":
        return sparse_c3(n, c["d"], **kw)
    if config == "C4":
        return sphere_c4(n, c["d"], **kw)
    return embedded_c5(n, c["d"], **kw)


def write_csv(path, x, labels=None):
    """The reference's input format (README.md:32-44): N, D, then rows."""
    n, d = x.shape
    with open(path, "w") as f:
        f.write(f"{n}\n{d}\n")
        for i in range(n):
            row = ", ".join(repr(float(v)) for v in x[i])
            if labels is not None:
                row += f", {int(labels[i])}"
            f.write(row + "\n")
