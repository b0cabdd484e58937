"""Host-side logic of the shipped library (no GPU needed): the layout reader,
bounds, error codes, the Green table, the generators and the result writer."""
import os

import numpy as np
import pytest

import frw
import oracle_ffi as O
from conftest import FIXTURES, layout_path


@pytest.mark.parametrize("name", [f for f, _ in FIXTURES])
def test_reader_matches_oracle(oracle_lib, name):
    lay = frw.Layout.load(layout_path(name))
    n, r, nl = O.parse(layout_path(name))
    assert len(lay) == n
    assert np.array_equal(lay.rects().astype(np.float64), r)
    assert tuple(lay.bounds().astype(float)) == O.bounds(r)
    assert frw.Layout.layer_count(layout_path(name)) == nl


def _write(tmp_path, text, name="l.txt"):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


def test_reader_line_rules(tmp_path, oracle_lib):
    text = ("magic\n<< a >>\nrect 0 0 10 10\nrect 20 0 30 5\n<< b >>\n"
            "  direct 1 2 3 4\nnotes without the keyword\n<< c >>\nrect 5 5 6 6\n")
    p = _write(tmp_path, text)
    assert frw.Layout.layer_count(p) == 2                   # rects after the last "<<" dropped
    assert frw.Layout.load(p, 0).rects().tolist() == [[0, 0, 10, 10], [20, 0, 30, 5]]
    assert frw.Layout.load(p, 1).rects().tolist() == [[1, 2, 3, 4]]   # "rect" anywhere in a line
    for layer in (0, 1):
        n, r, nl = O.parse(p, layer)
        assert np.array_equal(frw.Layout.load(p, layer).rects().astype(float), r)
    # a line longer than 49 characters ends the read (istream::getline failbit)
    long = "<< a >>\nrect 1 1 2 2\n<< a >>\nrect 1 1 2 2" + " " * 60 + "\n<< a >>\n"
    p2 = _write(tmp_path, long, "long.txt")
    assert frw.Layout.layer_count(p2) == 1
    assert O.parse(p2, 0)[2] == 1
    # CRLF line ends and a missing final newline
    p3 = _write(tmp_path, "<< x >>\r\nrect 1 2 3 4\r\n<< x >>", "crlf.txt")
    assert frw.Layout.load(p3).rects().tolist() == [[1, 2, 3, 4]]


def test_reader_errors(tmp_path):
    with pytest.raises(frw.FrwError) as e:
        frw.Layout.load(str(tmp_path / "missing.txt"))
    assert e.value.code == frw.E_IO
    p = _write(tmp_path, "<< a >>\nrect 1 2 3\n<< a >>\n")
    with pytest.raises(frw.FrwError) as e:
        frw.Layout.load(p)
    assert e.value.code == frw.E_PARSE
    p = _write(tmp_path, "<< a >>\nrect 1 2 3.5 4\n<< a >>\n", "frac.txt")
    with pytest.raises(frw.FrwError) as e:
        frw.Layout.load(p)
    assert e.value.code == frw.E_PARSE
    p = _write(tmp_path, "<< a >>\nrect 1 2 3 4\n<< a >>\n", "ok.txt")
    with pytest.raises(frw.FrwError) as e:
        frw.Layout.load(p, 3)
    assert e.value.code == frw.E_LAYER
    with pytest.raises(frw.FrwError) as e:
        frw.Layout.from_rects([[5, 5, 1, 1]])
    assert e.value.code == frw.E_PARSE


def test_bounds_padding():
    lay = frw.Layout.from_rects([[0, 0, 10, 10], [990, 500, 1000, 510]])
    assert lay.bounds().tolist() == [-10, -10, 1010, 520]   # int(0.01 * width) on every side
    small = frw.Layout.from_rects([[3, 4, 50, 60]])
    assert small.bounds().tolist() == [3, 4, 50, 60]        # width < 100: no padding


def test_green_table_matches_oracle(oracle_lib):
    assert np.array_equal(frw.green_thresholds(), O.green_thresholds())


def test_weight_tables_match_oracle(oracle_lib):
    """The weighted estimator's first-hop tables equal the oracle's to the bit."""
    gx, gy, pw, qx, qy, tx, ty = frw.weight_tables()
    ogx, ogy = O.green_grad()
    assert np.array_equal(gx, ogx) and np.array_equal(gy, ogy)
    T = O.green_thresholds().astype(np.float64)
    assert np.array_equal(pw, (np.append(T[1:], 2.0 ** 32) - T) / 2.0 ** 32)
    otx, oqx = O.importance_table(ogx)
    oty, oqy = O.importance_table(ogy)
    assert np.array_equal(tx, otx) and np.array_equal(ty, oty)
    assert np.array_equal(qx, oqx) and np.array_equal(qy, oqy)


def _overlaps(r):
    """Closed-interval overlap (touching counts) by sweep over x."""
    order = np.argsort(r[:, 0], kind="stable")
    rs = r[order]
    for i in range(len(rs)):
        j = i + 1
        while j < len(rs) and rs[j, 0] <= rs[i, 2]:
            if rs[j, 1] <= rs[i, 3] and rs[i, 1] <= rs[j, 3]:
                return True
            j += 1
    return False


def test_generator_uniform():
    a = frw.Layout.generate(100, 10_000, 10, 200, seed=1)
    b = frw.Layout.generate(100, 10_000, 10, 200, seed=1)
    c = frw.Layout.generate(100, 10_000, 10, 200, seed=2)
    ra = a.rects()
    assert np.array_equal(ra, b.rects()) and not np.array_equal(ra, c.rects())
    assert len(ra) == 100 and not _overlaps(ra)
    w, h = ra[:, 2] - ra[:, 0], ra[:, 3] - ra[:, 1]
    assert w.min() >= 10 and w.max() <= 200 and h.min() >= 10 and h.max() <= 200
    assert ra.min() >= 0 and ra[:, 2:].max() <= 10_000


def test_generator_clustered():
    lay = frw.Layout.generate(5000, 1_000_000, 5, 300, seed=1, cluster_frac=0.2, sigma=10_000,
                              n_clusters=10)
    r = lay.rects()
    assert len(r) == 5000 and not _overlaps(r)
    assert r.min() >= 0 and r[:, 2:].max() <= 1_000_000


def test_save_load_roundtrip(tmp_path):
    a = frw.Layout.generate(300, 50_000, 10, 500, seed=4)
    p = str(tmp_path / "gen.txt")
    a.save(p)
    b = frw.Layout.load(p)
    assert np.array_equal(a.rects(), b.rects())
    assert np.array_equal(a.bounds(), b.bounds())


def test_results_file(tmp_path):
    lay = frw.Layout.from_rects([[0, 0, 10, 10], [20, 0, 30, 10], [1000000, 0, 1000010, 10]])
    counts = np.array([3, 0, 1, 4, 0], dtype=np.int64)
    p = str(tmp_path / "wyniki.txt")
    lay.write_results(counts, 8, p)
    lines = open(p, encoding="utf-8").read().splitlines()
    assert lines[0] == "[GPU]"
    assert lines[1] == "Ile scieżek wyszło poza granice warstwy: 50%"
    assert lines[2] == "0 0 10 10: 37.5%"
    assert lines[3] == "1e+06 0 1.00001e+06 10: 12.5%"   # std::ostream default formatting


def test_reader_parallel_chunks_match_oracle(tmp_path, oracle_lib):
    """Files above 1 MB are classified in parallel chunks; every layer, the
    stopping line and the errors must equal the sequential rules (oracle)."""
    rng = np.random.default_rng(8)
    parts = ["header line\n<< start >>\n"]
    for layer in range(5):
        for i in range(12_000):
            x, y = rng.integers(0, 10 ** 6, 2)
            w, h = rng.integers(1, 500, 2)
            kind = rng.integers(0, 20)
            if kind == 0:
                parts.append(f"  comment {i}\n")
            elif kind == 1:
                parts.append(f"direct {x} {y} {x + w} {y + h}\r\n")
            else:
                parts.append(f"rect {x} {y} {x + w} {y + h}\n")
        parts.append(f"<< layer {layer} >>\n")
    text = "".join(parts)
    assert len(text) > 1 << 20  # several 256 KB chunks
    p = _write(tmp_path, text, "big.txt")
    n_layers = frw.Layout.layer_count(p)
    assert n_layers == O.parse(p, 0)[2] == 5
    for layer in range(n_layers):
        n, r, _ = O.parse(p, layer)
        assert np.array_equal(frw.Layout.load(p, layer).rects().astype(np.float64), r), layer
    # a line of 50 characters late in the file ends the read there
    cut = text.index("<< layer 3 >>")
    p2 = _write(tmp_path, text[:cut] + "x" * 50 + "\n" + text[cut:], "long.txt")
    assert frw.Layout.layer_count(p2) == O.parse(p2, 0)[2] == 3
    # a malformed rect in the last chunk is reported with its line number
    cut = text.index("<< layer 4 >>")
    p3 = _write(tmp_path, text[:cut] + "rect 1 2 3\n" + text[cut:], "bad.txt")
    with pytest.raises(frw.FrwError) as e:
        frw.Layout.load(p3)
    assert e.value.code == frw.E_PARSE
    assert f"line {text[:cut].count(chr(10)) + 1} " in str(e.value)
