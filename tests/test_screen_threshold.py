"""The subspace screen's per-point FP32 thresholds (propagate.hip, the
principal-subspace screen of propagation).

The lower bound of a candidate's key is a nondecreasing function of its
projection distance p, evaluated in FP64; the kernel decides a candidate with
one FP32 comparison when p lies above p_hi (the bound certainly exceeds the
k-th key thk: dropped) or below p_lo (certainly not: kept), and evaluates the
bound itself in the band between.  The thresholds come from inverting the
bound in exact arithmetic, widened by 2^-13 relative and rounded outward to
floats.  Here both are restated in numpy (float64 steps, float32 p) and every
certain decision is checked against the FP64 bound, at p on and around the
inverted boundary (and each threshold's float neighbours), over random and
edge-case parameters.  A mistake in the inversion (a dropped tiny, k1 on the
wrong side, a missing sigma) moves the thresholds across the boundary and
fails here.
"""
import numpy as np

F32_MAX = float(np.finfo(np.float32).max)


def bound_drops(p, thk, ei, ec, sigma, kap, tiny):
    """The kernel's FP64 bound: drop iff lb > thk."""
    k1 = 1.0 - kap
    sp = max(0.0, float(p) * k1 - tiny)
    r = max(0.0, np.sqrt(sp) - float(ei) - float(ec)) / sigma
    lb = np.sqrt(max(0.0, r * r * k1 - tiny)) * k1
    return lb > float(thk)


def f32_down(v):
    f = np.float32(v)
    return np.nextafter(f, np.float32(-np.inf)) if float(f) > v else f


def f32_up(v):
    f = np.float32(v)
    return np.nextafter(f, np.float32(np.inf)) if float(f) < v else f


def thresholds(thk, ei, ec, sigma, kap, tiny):
    k1 = 1.0 - kap
    if not (thk >= 0 and np.isfinite(thk) and ei >= 0 and np.isfinite(ei) and ec >= 0):
        return np.float32(np.inf), np.float32(-np.inf), None
    t = float(thk) / k1
    rr = np.sqrt((t * t + tiny) / k1)
    ss = float(ei) + float(ec) + sigma * rr
    pp = (ss * ss + tiny) / k1
    if not pp < 1e300:
        return np.float32(np.inf), np.float32(-np.inf), None
    return f32_up(pp * (1.0 + 1.0 / 8192.0)), f32_down(pp * (1.0 - 1.0 / 8192.0)), pp


def test_certain_decisions_agree_with_the_fp64_bound():
    rng = np.random.default_rng(5)
    checked = decided = 0
    cases = []
    for _ in range(400):
        d = int(rng.integers(16, 1024))
        kap = 2.0 * (d + d // 256 + 8) * 2.0 ** -24
        tiny = (d + 16) * 1.401298464324817e-45
        thk = np.float32(10.0 ** rng.uniform(-6, 3))
        ei = np.float32(10.0 ** rng.uniform(-8, 1))
        ec = np.float32(10.0 ** rng.uniform(-8, 1))
        sigma = 10.0 ** rng.uniform(-0.5, 0.5)
        cases.append((thk, ei, ec, sigma, kap, tiny))
    # edge cases: zero key and radii (the boundary made of tiny alone), a
    # zero key with radii, large keys
    k_e = 2.0 * (128 + 8) * 2.0 ** -24
    t_e = 144 * 1.401298464324817e-45
    cases += [(np.float32(0.0), np.float32(0.0), np.float32(0.0), 1.0, k_e, t_e),
              (np.float32(0.0), np.float32(0.5), np.float32(0.25), 1.3, k_e, t_e),
              (np.float32(1e18), np.float32(1.0), np.float32(1.0), 0.7, k_e, t_e),
              (np.float32(3.0), np.float32(0.0), np.float32(0.0), 1.0, k_e, t_e)]
    for thk, ei, ec, sigma, kap, tiny in cases:
        p_hi, p_lo, pp = thresholds(thk, ei, ec, sigma, kap, tiny)
        assert pp is not None and p_lo <= p_hi
        probes = [p_hi, p_lo, np.nextafter(p_hi, np.float32(np.inf)), np.nextafter(p_lo, np.float32(-np.inf)),
                  np.float32(0.0)]
        for rel in (1e-12, 1e-9, 1e-7, 1e-5, 2e-4, 1e-3, 1e-1):
            for sgn in (-1.0, 1.0):
                v = pp * (1.0 + sgn * rel)
                if v < F32_MAX:
                    probes.append(np.float32(v))
        if pp < F32_MAX:
            probes.append(np.float32(pp))
        for p in probes:
            drop = bound_drops(p, thk, ei, ec, sigma, kap, tiny)
            checked += 1
            if p > p_hi:
                decided += 1
                assert drop, (p, p_hi, thk, ei, ec, sigma)
            elif p < p_lo:
                decided += 1
                assert not drop, (p, p_lo, thk, ei, ec, sigma)
    # about half the probes lie outside the 2^-13 band and are decided
    assert decided > 0.4 * checked


def test_degenerate_points_take_the_bound():
    """Negative, infinite or NaN keys and radii leave every candidate to the
    FP64 bound (no certain decisions)."""
    k_e = 2.0 * (128 + 8) * 2.0 ** -24
    t_e = 144 * 1.401298464324817e-45
    for thk, ei, ec in ((np.float32(np.inf), 0.1, 0.1), (np.float32(np.nan), 0.1, 0.1), (np.float32(-1.0), 0.1, 0.1),
                        (np.float32(1.0), np.float32(np.inf), 0.1), (np.float32(1.0), np.float32(-0.5), 0.1),
                        (np.float32(1.0), 0.1, -1.0)):
        p_hi, p_lo, _ = thresholds(thk, ei, ec, 1.0, k_e, t_e)
        assert p_hi == np.inf and p_lo == -np.inf
