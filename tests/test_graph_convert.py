"""graph-convert subset (galois_amd/bin/graph-convert, host only): the
reference's own converter tests (tools/graph-convert/CMakeLists.txt:32-34,
fixtures copied as data under tests/golden/graph_convert/) and numpy
statements of the symmetrise / transpose / cleanup conversions."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
TOOL = ROOT / "galois_amd" / "bin" / "graph-convert"
FIX = ROOT / "tests" / "golden" / "graph_convert"
sys.path.insert(0, str(ROOT))


def conv(*args):
    r = subprocess.run([str(TOOL), *map(str, args)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("kind,src,expected", [
    ("-edgelist2gr", "with-blank-lines.edgelist", "with-blank-lines.edgelist.expected"),
    ("-csv2gr", "sample.csv", "with-blank-lines.edgelist.expected"),
    ("-edgelist2gr", "with-comments.edgelist", "with-comments.edgelist.expected"),
])
def test_reference_converter_cases(tmp_path, kind, src, expected):
    """create -> convert back -> compare_files, as the reference's CMake tests."""
    conv(kind, FIX / src, tmp_path / "g.test")
    conv("-gr2edgelist", tmp_path / "g.test", tmp_path / "g.compare")
    assert (tmp_path / "g.compare").read_bytes() == (FIX / expected).read_bytes()


def random_graph(rng, n=60, m=400):
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    w = rng.integers(1, 1000, m).astype(np.uint32)
    order = np.argsort(src, kind="stable")
    src, dst, w = src[order], dst[order], w[order]
    row = np.zeros(n + 1, np.int64)
    np.add.at(row, src + 1, 1)
    return np.cumsum(row).astype(np.int32), dst.astype(np.int32), w


def test_symmetric_transpose_cleanup(tmp_path):
    import galois_amd as gx
    rng = np.random.default_rng(5)
    row, col, w = random_graph(rng)
    n = len(row) - 1
    gx.write_gr(str(tmp_path / "in.gr"), row, col, w, 4)
    src = np.repeat(np.arange(n), np.diff(row))
    # -gr2sgr: every edge plus its reverse unless a self-loop, grouped by source
    # in the order the edges are visited.
    conv("-gr2sgr", "-edgeType=uint32", tmp_path / "in.gr", tmp_path / "s.gr")
    srow, scol, sdat, _ = gx.read_gr(str(tmp_path / "s.gr"))
    pairs = []
    for u, v, x in zip(src, col, w):
        pairs.append((u, v, x))
        if u != v:
            pairs.append((v, u, x))
    exp = sorted(range(len(pairs)), key=lambda i: (pairs[i][0], i))
    assert list(scol) == [pairs[i][1] for i in exp]
    assert list(sdat.view(np.uint32)) == [pairs[i][2] for i in exp]
    # -gr2tgr: in-edges in ascending source order.
    conv("-gr2tgr", "-edgeType=uint32", tmp_path / "in.gr", tmp_path / "t.gr")
    trow, tcol, tdat, _ = gx.read_gr(str(tmp_path / "t.gr"))
    order = np.lexsort((src, col))
    assert (tcol == src[order]).all() and (tdat.view(np.uint32) == w[order]).all()
    assert (np.diff(trow) == np.bincount(col, minlength=n)).all()
    # -gr2cgr: self-loops and duplicates gone, first duplicate (input order) kept.
    conv("-gr2cgr", "-edgeType=uint32", tmp_path / "in.gr", tmp_path / "c.gr")
    crow, ccol, cdat, _ = gx.read_gr(str(tmp_path / "c.gr"))
    keep = {}
    for i, (u, v) in enumerate(zip(src, col)):
        if u != v and (u, v) not in keep:
            keep[(u, v)] = i
    exp = sorted(keep.items())
    assert list(ccol) == [v for (u, v), _ in exp]
    assert list(cdat.view(np.uint32)) == [w[i] for _, i in exp]
    # Already clean: copied unchanged.
    out = conv("-gr2cgr", "-edgeType=uint32", tmp_path / "c.gr", tmp_path / "c2.gr")
    assert "already simplified" in out
    assert (tmp_path / "c2.gr").read_bytes() == (tmp_path / "c.gr").read_bytes()


def test_random_weights(tmp_path):
    import galois_amd as gx
    rng = np.random.default_rng(6)
    row, col, _ = random_graph(rng, n=300, m=20000)
    gx.write_gr(str(tmp_path / "in.gr"), row, col)
    conv("-gr2randomweightgr", "-edgeType=uint32", "-minValue=1", "-maxValue=255", tmp_path / "in.gr",
         tmp_path / "w.gr")
    conv("-gr2randomweightgr", "-edgeType=uint32", "-minValue=1", "-maxValue=255", tmp_path / "in.gr",
         tmp_path / "w2.gr")
    _, wcol, wdat, esize = gx.read_gr(str(tmp_path / "w.gr"))
    vals = wdat.view(np.uint32)
    assert esize == 4 and (wcol == col).all()
    assert vals.min() >= 1 and vals.max() <= 255 and abs(vals.mean() - 128) < 3
    assert (tmp_path / "w.gr").read_bytes() == (tmp_path / "w2.gr").read_bytes()  # fixed default seed


def _mt19937_weights(tmp_path, m, kind, lo, hi):
    """m draws of std::mt19937 (default seed) through uniform_int_distribution
    <uint32_t>(lo, hi) or uniform_real_distribution<float>(lo, hi), the recipe
    of RandomizeEdgeWeights (graph-convert.cpp:1144-1192), from a standalone
    build of exactly that standard-library code."""
    dist = ("std::uniform_int_distribution<uint32_t>" if kind == "uint32"
            else "std::uniform_real_distribution<float>")
    src = f"""
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
int main() {{
  std::mt19937 gen;
  {dist} d({lo}, {hi});
  for (long i = 0; i < {m}L; ++i) {{
    auto v = d(gen);
    uint32_t bits;
    std::memcpy(&bits, &v, 4);
    std::printf("%u\\n", bits);
  }}
}}
"""
    cpp, exe = tmp_path / f"mt_{kind}.cpp", tmp_path / f"mt_{kind}"
    cpp.write_text(src)
    subprocess.run(["g++", "-O1", str(cpp), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    return np.array([int(x) for x in out], np.uint64).astype(np.uint32)


@pytest.mark.parametrize("kind,lo,hi", [("uint32", 1, 255), ("uint32", 1, 1000), ("float32", 0, 1)])
def test_random_weights_follow_mt19937(tmp_path, kind, lo, hi):
    """The weights equal the standard-library draws of the reference's recipe
    in CSR edge order, bit for bit."""
    import galois_amd as gx
    rng = np.random.default_rng(8)
    row, col, _ = random_graph(rng, n=200, m=5000)
    gx.write_gr(str(tmp_path / "in.gr"), row, col)
    conv("-gr2randomweightgr", f"-edgeType={kind}", f"-minValue={lo}", f"-maxValue={hi}", tmp_path / "in.gr",
         tmp_path / "w.gr")
    _, wcol, wdat, esize = gx.read_gr(str(tmp_path / "w.gr"))
    assert esize == 4 and (wcol == col).all()
    np.testing.assert_array_equal(wdat.view(np.uint32), _mt19937_weights(tmp_path, len(col), kind, lo, hi))


def test_converter_errors(tmp_path):
    r = subprocess.run([str(TOOL), "-gr2edgelist", str(tmp_path / "missing.gr"), str(tmp_path / "x")],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "ERROR" in r.stderr
    r = subprocess.run([str(TOOL), str(tmp_path / "a"), str(tmp_path / "b")], capture_output=True, text=True)
    assert r.returncode == 1 and "conversion" in r.stderr
