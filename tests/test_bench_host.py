"""bench.py's host logic (CPU only): the per-workload defaults, the
algorithmic-bytes formula, and the lookup of committed rocprof traffic by
configuration (a summary is used only when its configuration, the
nontemporal setting included, matches the run)."""
import sys

import pytest

import bench
from cuda_fot_amd import capi


def _parse(argv, monkeypatch):
    monkeypatch.setattr(sys, "argv", ["bench.py"] + argv)
    return bench.parse()


@pytest.mark.parametrize("workload,part,bs,order", [
    ("res_calc", "points", 512, "owner"),
    ("volna-batch", "points", 512, "owner"),
    ("hex", "bisect", 512, "two-phase"),
    ("mini_aero", "bisect", 128, "two-phase"),
    ("lulesh", "bisect", 512, "two-phase"),
    ("getacc", "bisect", 512, "two-phase"),
    ("mine", "bisect", 512, "two-phase"),
])
def test_workload_defaults(monkeypatch, workload, part, bs, order):
    a = _parse(["--workload", workload], monkeypatch)
    assert (a.partition, a.block_size, a.order) == (part, bs, order)


def test_explicit_options_win_over_defaults(monkeypatch):
    a = _parse(["--workload", "hex", "--partition", "gps", "--block-size", "256",
                "--order", "launch"], monkeypatch)
    assert (a.partition, a.block_size, a.order) == ("gps", 256, "launch")


def test_unique_bytes_is_the_papers_formula():
    # scripts/visualise_paper.py:19-22 on the 1200x600 airfoil: 76.1 B/edge
    nodes, cells, edges = 1201 * 601, 1200 * 600, 1438200
    assert bench.unique_bytes(nodes, cells, edges) == 109_440_016


def test_nontemporal_flags():
    assert bench.nt_flags("auto") == 0
    assert bench.nt_flags("on") == capi.LOOP_NONTEMPORAL
    assert bench.nt_flags("off") == capi.LOOP_CACHED
    assert capi.LOOP_NONTEMPORAL & capi.LOOP_CACHED == 0


def test_workload_traffic_needs_the_same_nontemporal_setting():
    # hex's committed summary was taken with the nontemporal force loads,
    # mine's without (its loop has none)
    assert bench.load_workload_traffic("hex", "bisect_b512_two-phase", 1) > 0
    assert bench.load_workload_traffic("hex", "bisect_b512_two-phase", 0) is None
    assert bench.load_workload_traffic("mine", "bisect_b512_two-phase", 0) > 0
    assert bench.load_workload_traffic("mine", "bisect_b512_two-phase", 1) is None
    assert bench.load_workload_traffic("hex", "no_such_variant", 1) is None


def test_res_calc_traffic_matches_configuration(monkeypatch):
    a = _parse(["--tune", "none"], monkeypatch)
    got = bench.load_traffic(a, {"nontemporal": 0})
    assert got is not None and 0.98 < got / 109_440_016 < 1.0
    # the 1200x600 summary is of the cached kernel
    assert bench.load_traffic(a, {"nontemporal": 1}) is None
    # 2400x1200: the summary is of the nontemporal kernel
    b = _parse(["--tune", "none", "--nx", "2400", "--ny", "1200"], monkeypatch)
    assert bench.load_traffic(b, {"nontemporal": 1}) is not None
    assert bench.load_traffic(b, {"nontemporal": 0}) is None
    # another part size has no summary
    c = _parse(["--tune", "none", "--points-per-block", "200"], monkeypatch)
    assert bench.load_traffic(c, {"nontemporal": 0}) is None


def test_share_mesh_covers_the_mesh_once_per_owner():
    # the shares of a 4-way cell partition: owned cells partition the mesh,
    # every edge is in the share of each of its cells' owners (twice across
    # the cut, once inside), halos are the other cells those edges touch,
    # and every submesh id is in range
    import numpy as np
    from cuda_fot_amd import meshgen, plan
    mesh = meshgen.airfoil(60, 40, seed=1, shuffle=7)
    ind, end = plan.graph_csr(mesh["pecell"], mesh["num_cells"])
    part, n = plan.bisect_partition(ind, end, -(-mesh["num_cells"] // 4))
    assert n == 4
    pe = mesh["pecell"]
    cut = int((part[pe[:, 0]] != part[pe[:, 1]]).sum())
    owned_total, edges_total = 0, 0
    for s in range(n):
        sub, owned, halo = bench.share_mesh(mesh, part, s)
        owned_total += owned
        edges_total += sub["num_edges"]
        assert sub["pecell"].max() < sub["num_cells"] == owned + halo
        assert sub["pedge"].max() < sub["num_nodes"]
        # every edge of the share touches an owned cell (numbered first)
        assert np.all((sub["pecell"] < owned).any(axis=1))
        # the halo cells are exactly the non-owned cells its edges read
        assert np.array_equal(np.unique(sub["pecell"][sub["pecell"] >= owned]),
                              np.arange(owned, owned + halo))
        # data follow the renumbering
        mine = np.nonzero(part == s)[0]
        assert np.array_equal(sub["q"][:owned], mesh["q"][mine])
    assert owned_total == mesh["num_cells"]
    assert edges_total == mesh["num_edges"] + cut
