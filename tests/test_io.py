"""Input loaders without a device: the multi-threaded CSV parser returns what
the sequential one (the reference's loadData semantics,
euclideandistancematrixgpu.cu:2712-2757: strtok on ',', trim spaces, atof,
atoi label) returns, value for value and error for error; *.fbin binary input."""
import numpy as np
import pytest


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)
    return str(path)


def test_parallel_csv_matches_sequential(aknn, datasets, tmp_path):
    x, lab = datasets.gmm_c1(5000, 12, clusters=4, seed=3)
    p = str(tmp_path / "in.csv")
    datasets.write_csv(p, x, lab)
    a, la = aknn.read_points(p, 5000, 12, klaster=True, threads=1)
    for t in (0, 2, 7, 16):
        b, lb = aknn.read_points(p, 5000, 12, klaster=True, threads=t)
        assert (a.view(np.uint32) == b.view(np.uint32)).all() and (la == lb).all()
    # repr() of float32 values round-trips exactly
    assert (a == x).all() and (la == lab).all()


def test_csv_quirks_match(aknn, tmp_path):
    """Empty fields vanish (strtok), spaces trim, tabs and CR pass to atof,
    an extra text tail after N rows is ignored, and the label is the last field."""
    text = ("3\n2\n"
            " 1.5 ,, 2.25 ,7\n"
            "\t-3e2,4\t,\t8\r\n"
            "5,,,6 ,  9  ,\n"
            "garbage line that is never read\n")
    p = _write(tmp_path / "q.csv", text)
    for t in (1, 0, 3):
        x, lab = aknn.read_points(p, 3, 2, klaster=True, threads=t)
        assert x.tolist() == [[1.5, 2.25], [-300.0, 4.0], [5.0, 6.0]]
        assert lab.tolist() == [7, 8, 9]


@pytest.mark.parametrize("text,msg", [("2\n2\n1,2\n3\n", "data row 1 has 1 fields"),
                                      ("3\n2\n1,2\n3,4\n", "holds 2 data rows"),
                                      ("2\n2\n1,2\n\n", "data row 1 has 0 fields")])
def test_csv_errors_match(aknn, tmp_path, text, msg):
    p = _write(tmp_path / "e.csv", text)
    n = int(text.split("\n")[0])
    for t in (1, 0, 4):
        with pytest.raises(aknn.AknnError, match=msg):
            aknn.read_points(p, n, 2, threads=t)


def test_fbin_input(aknn, tmp_path):
    x = np.random.default_rng(1).normal(size=(100, 7)).astype(np.float32)
    p = str(tmp_path / "pts.fbin")
    with open(p, "wb") as f:
        np.array([100, 7], np.uint32).tofile(f)
        x.tofile(f)
    got, _ = aknn.read_points(p, 100, 7)
    assert (got == x).all()
    got, _ = aknn.read_points(p, 60, 7)
    assert (got == x[:60]).all()
    with pytest.raises(aknn.AknnError, match="expected"):
        aknn.read_points(p, 101, 7)
    with pytest.raises(aknn.AknnError, match="expected"):
        aknn.read_points(p, 10, 8)
