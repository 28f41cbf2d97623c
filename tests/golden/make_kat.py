"""Writes tests/golden/readme_kat.json: known-answer vectors for the README
example of the reference (README.md:38-44: N=3, D=2, rows 0.1,0.2 / 0.3,0.4 /
0.5,0.6; launched with k=2 as in README.md:25).

Everything here is computed with exact rational arithmetic and explicit
IEEE-754 round-to-nearest-even to binary32/binary64, so the vectors do not
depend on the oracle, on numpy or on any compiler.  The operation sequence is
the reference's own:

  * input: atof (decimal -> double) then double -> float (.cu:2746)
  * normalization: sequential FP32 mean, centre, max-abs scale (.cu:19-85)
  * distances: EUCLIDEAN sqrt(fma chain of t*t), TAXICAB sqrt(sum |t|),
    COSINE float(1 - (double)(ab / (sqrt(aa) * sqrt(bb)))) (.cu:243-273,
    :540-570, :840-881)
  * initial pass: with N=3 and k=2 every start node is the root, so cand(i) is
    the other two points; chunks hold dimIdx - k = min(k*k+k, N) - k = 1
    candidate, so COSINE's store-time flip (.cu:1007) runs twice
  * propagation: with N=3, k=2 the candidate set is empty, rows are unchanged
  * output file: "N;k;8\\n", long 0x01020304, {int32 id, float32 value} sorted
    ascending (COSINE descending) (.cu:2908-2978)

Run:  python tests/golden/make_kat.py
"""
import json
import math
import os
import struct
from fractions import Fraction as F


def _round(q, mant_bits, emin):
    if q == 0:
        return F(0)
    s = -1 if q < 0 else 1
    a = abs(q)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if F(2) ** e > a:
        e -= 1
    while F(2) ** (e + 1) <= a:
        e += 1
    e = max(e, emin)
    ulp = F(2) ** (e - (mant_bits - 1))
    m = a / ulp
    fl = m.numerator // m.denominator
    rem = m - fl
    if rem > F(1, 2) or (rem == F(1, 2) and fl % 2 == 1):
        fl += 1
    return s * fl * ulp


def f32(q):
    return _round(F(q), 24, -126)


def f64(q):
    return _round(F(q), 53, -1022)


def f32_bits(q):
    return struct.unpack("<I", struct.pack("<f", float(q)))[0]


def f32_next(q, up):
    b = f32_bits(q)
    if q == 0:
        b = 1 if up else 0x80000001
    elif (q > 0) == up:
        b += 1
    else:
        b -= 1
    return F(struct.unpack("<f", struct.pack("<I", b))[0])


def sqrt32(x):
    """Correctly rounded binary32 square root of a binary32 x >= 0."""
    if x == 0:
        return F(0)
    r = f32(F(math.sqrt(float(x))))
    while True:
        lo, hi = f32_next(r, False), f32_next(r, True)
        mid_lo, mid_hi = (lo + r) / 2, (r + hi) / 2
        if mid_lo * mid_lo > x:
            r = lo
        elif mid_hi * mid_hi < x:
            r = hi
        else:
            return r


def fma32(a, b, c):
    return f32(a * b + c)


def distance(a, b, metric):
    d = len(a)
    if metric == "COSINE":
        ab = aa = bb = F(0)
        for c0 in range(0, d, 256):
            pab = paa = pbb = F(0)
            for i in range(c0, min(c0 + 256, d)):
                pab = fma32(a[i], b[i], pab)
                paa = fma32(a[i], a[i], paa)
                pbb = fma32(b[i], b[i], pbb)
            ab, aa, bb = f32(ab + pab), f32(aa + paa), f32(bb + pbb)
        c = f32(ab / f32(sqrt32(aa) * sqrt32(bb)))
        return f32(f64(1 - c))
    tot = F(0)
    for c0 in range(0, d, 256):
        p = F(0)
        for i in range(c0, min(c0 + 256, d)):
            t = f32(a[i] - b[i])
            p = fma32(t, t, p) if metric == "EUCLIDEAN" else f32(p + abs(t))
        tot = f32(tot + p)
    return sqrt32(tot)


def normalize(rows):
    n, d = len(rows), len(rows[0])
    x = [list(r) for r in rows]
    for j in range(d):
        m = F(0)
        for i in range(n):
            m = f32(m + x[i][j])
        m = f32(m / n)
        for i in range(n):
            x[i][j] = f32(x[i][j] - m)
    dm = min(d, 256)
    g = max(abs(x[i][j]) for i in range(n) for j in range(dm))
    for i in range(n):
        for j in range(dm):
            x[i][j] = f32(x[i][j] / g)
    return x


def initial(x, k, metric):
    n = len(x)
    cap = min(k * k + k, n) - k
    out = []
    for i in range(n):
        cand = [j for j in range(n) if j != i]
        lst = [(F(100000), -1)] * k
        for c0 in range(0, len(cand), cap):
            pool = list(lst) + [(distance(x[i], x[j], metric), j) for j in cand[c0:c0 + cap]]
            sel = sorted(pool)[:k]
            if metric == "COSINE":
                sel = [(f32(f64(1 - key)), j) for key, j in sel]
            lst = sel
        out.append(sorted(lst))
    return out


def result_file(n, k, metric, rows):
    blob = b"%d;%d;8\n" % (n, k) + struct.pack("<q", 0x01020304)
    for row in rows:
        key = (lambda e: (-e[0], e[1])) if metric == "COSINE" else (lambda e: (e[0], e[1]))
        for v, j in sorted(row, key=key):
            blob += struct.pack("<i", j) + struct.pack("<I", f32_bits(v))
    return blob


def main():
    csv = "3\n2\n0.1, 0.2\n0.3, 0.4\n0.5, 0.6\n"
    raw = [[f32(F(float(v.strip()))) for v in line.split(",")] for line in csv.splitlines()[2:]]
    x = normalize(raw)
    k = 2
    kat = {
        "source": "README.md:38-44 example, k=2 (README.md:25)",
        "csv": csv,
        "n": 3, "d": 2, "k": k,
        "raw_bits": [[f32_bits(v) for v in r] for r in raw],
        "normalized_bits": [[f32_bits(v) for v in r] for r in x],
        "graphs": {},
        "distance_kats": [],
    }
    for metric in ("EUCLIDEAN", "TAXICAB", "COSINE"):
        rows = initial(x, k, metric)
        kat["graphs"][metric] = {
            "ids": [[j for _, j in r] for r in rows],
            "val_bits": [[f32_bits(v) for v, _ in r] for r in rows],
            "result_file_hex": result_file(3, k, metric, rows).hex(),
        }
    pairs = [
        ([0.0, 0.0], [3.0, 4.0]),
        ([1.0, 0.0], [0.0, 1.0]),
        ([1.0, 2.0, 3.0], [-2.0, 0.5, 7.25]),
        ([0.1] * 300, [0.2] * 300),
    ]
    for a, b in pairs:
        fa = [f32(F(v)) for v in a]
        fb = [f32(F(v)) for v in b]
        kat["distance_kats"].append({
            "a_bits": [f32_bits(v) for v in fa],
            "b_bits": [f32_bits(v) for v in fb],
            "bits": {m: f32_bits(distance(fa, fb, m)) for m in ("EUCLIDEAN", "TAXICAB", "COSINE")
                     if m != "COSINE" or (any(fa) and any(fb))},
        })
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "readme_kat.json")
    with open(out, "w") as f:
        json.dump(kat, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main()
