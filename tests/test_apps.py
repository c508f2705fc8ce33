"""The lonestar-compatible app executables (galois_amd/bin/*): command line,
output lines and -statFile table of the reference apps, with the printed
results checked against the oracle on the same .gr inputs."""
from __future__ import annotations

import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "galois_amd" / "bin"
APPS = ["bfs", "sssp", "connected-components", "k-core", "pagerank-pull", "betweennesscentrality",
        "pagerank-push", "bfs-directionopt"]
INF = 2147483646


def run(app, *args, timeout=120):
    print(f"app {app} {' '.join(map(str, args))}", flush=True)
    return subprocess.run([str(BIN / app), *map(str, args)], capture_output=True, text=True, timeout=timeout)


def write_gr(path, row, col, w=None):
    sys.path.insert(0, str(ROOT))
    import galois_amd as gx
    gx.write_gr(str(path), row, col, None if w is None else np.ascontiguousarray(w, np.uint32), 4 if w is not None
                else 0)


def line_value(out, prefix):
    for ln in out.splitlines():
        if ln.startswith(prefix):
            return ln[len(prefix):].strip()
    raise AssertionError(f"no line starting with {prefix!r} in:\n{out}")


def stat_rows(text):
    rows = [ln.split(", ") for ln in text.strip().splitlines()]
    assert rows[0] == ["STAT_TYPE", "REGION", "CATEGORY", "TOTAL_TYPE", "TOTAL"]
    return rows[1:]


# ------------------------------------------------------------------ CPU side
@pytest.mark.parametrize("app", APPS)
def test_app_built_and_help(app):
    assert (BIN / app).exists(), "build the library and apps first"
    r = run(app, "-help")
    assert r.returncode == 0
    assert "USAGE:" in r.stdout and "-statFile" in r.stdout and "-noverify" in r.stdout


def test_app_rejects_unknown_flag_and_missing_input(tmp_path):
    r = run("bfs", "-noSuchFlag", "x.gr")
    assert r.returncode == 1 and "Unknown command line argument '-noSuchFlag'" in r.stderr
    r = run("bfs")
    assert r.returncode == 1 and "positional" in r.stderr
    r = run("k-core", "-symmetricGraph", "x.gr")
    assert r.returncode == 1 and "kcore" in r.stderr


# ------------------------------------------------------------------ GPU side
@pytest.fixture(scope="module")
def inputs(tmp_path_factory, gx):
    d = tmp_path_factory.mktemp("apps")
    G = gx.Graph.rmat(12, 16, symmetric=True, weights="uniform", wlo=1, whi=255)
    row, col, w = G.to_host()
    write_gr(d / "rmat12.sgr", row, col, w)
    D = gx.Graph.rmat(12, 8, symmetric=False)
    drow, dcol, _ = D.to_host()
    return d, (row, col, w), (drow, dcol)


@pytest.mark.gpu
def test_bfs_app(inputs, oracle):
    d, (row, col, w), _ = inputs
    src = int(np.argmax(np.diff(row)))
    r = run("bfs", f"-startNode={src}", "-reportNode", 7, f"-statFile={d / 'bfs.stats'}", d / "rmat12.sgr")
    assert r.returncode == 0, r.stderr
    lv = oracle.bfs(row, col, src)
    reached = lv[lv != INF].astype(np.int64)
    assert line_value(r.stdout, "Node 7 has distance ") == str(lv[7])
    assert line_value(r.stdout, "INFO: # visited nodes is ") == str(len(reached))
    assert line_value(r.stdout, "INFO: Max distance is ") == str(reached.max())
    assert line_value(r.stdout, "INFO: Sum of visited distances is ") == str(reached.sum())
    assert "Verification successful." in r.stdout
    rows = stat_rows((d / "bfs.stats").read_text())
    assert any(x[:4] == ["STAT", "(NULL)", "Timer_0", "TMAX"] and x[4].isdigit() for x in rows)
    assert any(x[:3] == ["PARAM", "(NULL)", "CommandLine"] for x in rows)


@pytest.mark.gpu
def test_sssp_app(inputs, oracle):
    d, (row, col, w), _ = inputs
    r = run("sssp", "-delta=6", "-algo=deltaStep", "-startNode=0", "-reportNode=5", d / "rmat12.sgr")
    assert r.returncode == 0, r.stderr
    dist = oracle.sssp(row, col, w, 0, 6)
    reached = dist[dist != INF].astype(np.int64)
    assert "INFO: Using delta-step of 64" in r.stdout
    assert line_value(r.stdout, "Node 5 has distance ") == str(dist[5])
    assert line_value(r.stdout, "INFO: # visited nodes is ") == str(len(reached))
    assert line_value(r.stdout, "INFO: Sum of visited distances is ") == str(reached.sum())
    assert "Verification successful." in r.stdout
    r = run("sssp", d / "rmat12.sgr")  # Auto: reports its choice
    assert r.returncode == 0 and "INFO: Choosing " in r.stdout


@pytest.mark.gpu
def test_cc_and_kcore_apps(inputs, oracle):
    d, (row, col, w), _ = inputs
    r = run("connected-components", "-symmetricGraph", d / "rmat12.sgr")
    assert r.returncode == 0, r.stderr
    lab = oracle.cc(row, col)
    sizes = np.bincount(lab, minlength=len(lab))
    comps = int((sizes > 0).sum())
    nontrivial = int((sizes > 1).sum())
    assert line_value(r.stdout, "Total components: ") == str(comps)
    m = re.search(r"Number of non-trivial components: (\d+) \(largest size: (\d+) ", r.stdout)
    assert m and int(m.group(1)) == nontrivial and int(m.group(2)) == int(sizes.max())
    r = run("connected-components", d / "rmat12.sgr")
    assert r.returncode == 1 and "symmetric" in r.stderr
    r = run("k-core", "-kcore=10", "-symmetricGraph", d / "rmat12.sgr")
    assert r.returncode == 0, r.stderr
    alive = int((oracle.kcore(row, col, 10) >= 10).sum())
    assert f"Number of nodes in the 10-core is {alive}" in r.stdout


@pytest.mark.gpu
def test_pagerank_pull_app(inputs, oracle):
    d, _, (drow, dcol) = inputs
    trow, tcol, _ = oracle.transpose(drow, dcol)
    write_gr(d / "rmat12.tgr", trow, tcol)
    r = run("pagerank-pull", "-transposedGraph", "-tolerance=1e-4", d / "rmat12.tgr")
    assert r.returncode == 0, r.stderr
    x, _ = oracle.pagerank_residual(drow, dcol, tol=1e-4)
    lines = r.stdout.splitlines()
    top = lines[lines.index("Rank PageRank Id") + 1:][:20]
    order = sorted(range(len(x)), key=lambda v: (-float(x[v]), v))[:20]
    for i, ln in enumerate(top):
        rank, val, vid = ln.split()
        assert rank == f"{i + 1}:" and int(vid) == order[i]
        np.testing.assert_allclose(float(val), x[order[i]], rtol=1e-5)
    r = run("pagerank-pull", d / "rmat12.tgr")
    assert r.returncode == 1 and "transposed" in r.stderr


@pytest.mark.gpu
def test_bc_app(inputs, oracle):
    d, (row, col, w), _ = inputs
    r = run("betweennesscentrality", "-algo=Level", "-numOfSources=4", "-output", d / "rmat12.sgr")
    assert r.returncode == 0, r.stderr
    exp = oracle.bc(row, col, np.arange(4, dtype=np.uint32))
    vals = {}
    for ln in r.stdout.splitlines():
        p = ln.split()
        if len(p) == 2 and p[0].isdigit():
            vals[int(p[0])] = float(p[1])
    assert len(vals) == len(exp)
    got = np.array([vals[v] for v in range(len(exp))])
    np.testing.assert_allclose(got, exp.astype(np.float32), rtol=1e-6, atol=1e-6)
    assert line_value(r.stdout, "Max BC is ") and "INFO: Running level BC" in r.stdout


def node_values(path, kind=int):
    """The -o file: one 'node value' line per node, in node order."""
    lines = Path(path).read_text().splitlines()
    ids = [int(ln.split()[0]) for ln in lines]
    assert ids == list(range(len(lines)))
    return np.array([kind(ln.split()[1]) for ln in lines])


@pytest.mark.gpu
def test_node_output_files(inputs, oracle):
    """-o writes the LonestarGPU per-node output (node value lines; unreached
    distances as 1073741823) that scripts/result_checker.py compares."""
    d, (row, col, w), (drow, dcol) = inputs
    src = int(np.argmax(np.diff(row)))
    assert run("bfs", f"-startNode={src}", "-o", d / "bfs.out", d / "rmat12.sgr").returncode == 0
    lv = oracle.bfs(row, col, src).astype(np.int64)
    np.testing.assert_array_equal(node_values(d / "bfs.out"), np.where(lv == INF, 1073741823, lv))
    assert run("sssp", "-delta=6", "-startNode=3", f"-o={d / 'sssp.out'}", d / "rmat12.sgr").returncode == 0
    dist = oracle.sssp(row, col, w, 3, 6).astype(np.int64)
    np.testing.assert_array_equal(node_values(d / "sssp.out"), np.where(dist == INF, 1073741823, dist))
    assert run("connected-components", "-symmetricGraph", "-o", d / "cc.out", d / "rmat12.sgr").returncode == 0
    np.testing.assert_array_equal(node_values(d / "cc.out"), oracle.cc(row, col))
    assert run("k-core", "-kcore=10", "-symmetricGraph", "-o", d / "kc.out", d / "rmat12.sgr").returncode == 0
    np.testing.assert_array_equal(node_values(d / "kc.out"), (oracle.kcore(row, col, 10) >= 10).astype(int))
    assert run("betweennesscentrality", "-numOfSources=4", "-o", d / "bc.out", d / "rmat12.sgr").returncode == 0
    np.testing.assert_allclose(node_values(d / "bc.out", float),
                               oracle.bc(row, col, np.arange(4, dtype=np.uint32)), rtol=1e-6, atol=1e-6)
    trow, tcol, _ = oracle.transpose(drow, dcol)
    write_gr(d / "rmat12o.tgr", trow, tcol)
    r = run("pagerank-pull", "-transposedGraph", "-tolerance=1e-4", "-o", d / "pr.out", d / "rmat12o.tgr")
    assert r.returncode == 0, r.stderr
    x, _ = oracle.pagerank_residual(drow, dcol, tol=1e-4)
    np.testing.assert_allclose(node_values(d / "pr.out", float), x, rtol=1e-7)
    r = run("bfs", "-o", "-", d / "rmat12.sgr")  # "-" is stdout
    assert r.returncode == 0 and "0 0" in r.stdout.splitlines()


@pytest.mark.gpu
def test_pagerank_push_app(inputs, oracle):
    d, _, (drow, dcol) = inputs
    write_gr(d / "rmat12d.gr", drow, dcol)
    r = run("pagerank-push", "-tolerance=1e-7", "-o", d / "push.out", d / "rmat12d.gr")
    assert r.returncode == 0, r.stderr
    assert "tolerance:1e-07, maxIterations:1000" in r.stdout and "Running Edge Async push version," in r.stdout
    got = node_values(d / "push.out", float)
    exp, _ = oracle.pagerank_push(drow, dcol, tol=1e-7)
    np.testing.assert_allclose(got, exp, atol=2e-5)  # FP32 atomics: order-dependent last bits
    lines = r.stdout.splitlines()
    top = [ln.split() for ln in lines[lines.index("Rank PageRank Id") + 1:][:20]]
    assert [int(t[2]) for t in top[:5]] == list(np.argsort(-exp, kind="stable")[:5])
    r = run("pagerank-push", "-algo=Sync", "-tolerance=1e-3", d / "rmat12d.gr")
    assert r.returncode == 0 and "Running Edge Sync push version," in r.stdout


@pytest.mark.gpu
def test_bfs_directionopt_app(inputs, oracle):
    d, (row, col, w), _ = inputs
    src = int(np.argmax(np.diff(row)))
    r = run("bfs-directionopt", f"-startNode={src}", "-reportNode=9", "-numRuns=2", "-o", d / "par.out",
            d / "rmat12.sgr")
    assert r.returncode == 0, r.stderr
    lv = oracle.bfs(row, col, src)
    par = node_values(d / "par.out")
    assert par[src] == 0  # the reference's source keeps its initial 0
    for v in range(len(lv)):
        if v == src:
            continue
        if lv[v] == INF:
            assert par[v] == INF
        else:
            p = par[v]
            assert lv[p] == lv[v] - 1 and v in col[row[p]:row[p + 1]]
    assert line_value(r.stdout, "Node 9 has parent ") == str(par[9])
    assert "parent[0] : " in r.stdout and "BFS::go run 1 called" in r.stdout


@pytest.mark.gpu
def test_autoalgo_choice(inputs, gx, tmp_path):
    """The Auto branches of sssp (SSSP.cpp:410-419), bfs-directionopt
    (bfsDirectionOpt.cpp:469-478) and BC (BetweennessCentrality.cpp:110-122)
    choose by the degree test: the power-law RMAT input picks deltaStep /
    SyncDO / Async, the road-like grid (average degree below 10) picks
    deltaStepBarrier / Async / Level."""
    d, (row, col, w), _ = inputs
    G = gx.Graph.grid(64, 50000, 10000, 7, 1, 1000)
    grow, gcol, gw = G.to_host()
    write_gr(tmp_path / "grid.sgr", grow, gcol, gw)
    assert gx.degree_power_law(row) and not gx.degree_power_law(grow)
    cases = [("sssp", ["-delta=6"], "deltaStep", "deltaStepBarrier"),
             ("bfs-directionopt", [], "SyncDO", "Async"),
             ("betweennesscentrality", ["-numOfSources=2"], "Async", "Level")]
    for app, extra, pl, other in cases:
        for path, exp in ((d / "rmat12.sgr", pl), (tmp_path / "grid.sgr", other)):
            r = run(app, *extra, path)
            assert r.returncode == 0, r.stderr
            assert f"INFO: Choosing {exp} algorithm" in r.stdout, (app, path)
