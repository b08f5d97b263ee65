"""Host-side logic of the engine (no GPU): topology indexing, compression,
tip encoding, ratio transform, JSON parsing, and the C ABI library itself."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import data as odata
from torchtree_amd import alignment as al
from torchtree_amd import tree as ht

ROOT = Path(__file__).resolve().parents[1]


def test_topology_matches_oracle(tiny_newick, flu_newick, flu_fasta):
    for nwk, taxa in ((tiny_newick, None), (flu_newick, [t for t, _ in odata.read_fasta(flu_fasta)])):
        a = ht.Topology(nwk, taxa)
        b = odata.Tree(nwk, taxa)
        assert a.taxa == b.taxa
        assert [tuple(map(int, r)) for r in a.postorder] == b.postorder
        assert np.allclose(a.edge_lengths(), b.edge_lengths())
        assert np.allclose(a.unrooted_lengths(), b.unrooted_branch_lengths())


def test_heights_and_ratio_roundtrip(flu_newick, flu_fasta):
    taxa = [t for t, _ in odata.read_fasta(flu_fasta)]
    dates = [float(t.split("_")[-1]) for t in taxa]
    topo = ht.Topology(flu_newick, taxa)
    lh = ht.leaf_heights_from_dates(dates)
    h = topo.heights_from_lengths(lh)
    ot = odata.Tree(flu_newick, taxa)
    ot.set_dates(dates)
    assert np.allclose(h, ot.heights_from_branch_lengths().numpy())
    bounds = ht.ratio_bounds(topo, lh)
    ht_ = torch.tensor(h, dtype=torch.float64)
    r = ht.heights_to_ratios(topo, bounds, ht_)
    assert torch.all((r[:-1] > 0) & (r[:-1] < 1))
    back = ht.ratios_to_heights(topo, bounds, r)
    assert torch.allclose(back, ht_, rtol=1e-12)
    bl = ht.time_branch_lengths(topo, torch.tensor(lh), ht_)
    assert torch.allclose(bl, ot.time_branch_lengths(ht_))


def test_compress_matches_oracle_order_and_weights(tiny_fasta):
    seqs = [s for _, s in odata.read_fasta(tiny_fasta)]
    a, wa = al.compress(seqs)
    b, wb = odata.compress(seqs)
    assert [list(x) for x in a] == [list(x) for x in b]
    assert torch.equal(wa, wb)


def test_reference_site_pattern_fixture():
    """test/test_site_pattern.py:9-20: weights [2, 1], four taxa, two patterns."""
    cols, w = al.compress_bytes(["AAG", "AAC", "AAC", "AAT"])
    assert w.tolist() == [2, 1]
    codes, tbl = al.encode_tips(cols, al.NucleotideType())
    assert codes.shape == (4, 2) and tbl.shape[1] == 4


@pytest.mark.parametrize("dt,chars", [("aa", "ACDEFGHIKLMNPQRSTVWYBZX-?J"), ("codon", None)])
def test_encodings_match_oracle(dt, chars):
    rng = np.random.default_rng(0)
    if dt == "aa":
        seqs = ["".join(rng.choice(list(chars), 60)) for _ in range(5)]
        cols, w = al.compress_bytes(seqs)
        codes, tbl = al.encode_tips(cols, al.AminoAcidType(), use_ambiguities=True)
        ref, wref = odata.compress_alignment(seqs, odata.amino_acid_partial)
    else:
        trip = ["".join(t) for t in __import__("itertools").product("ACGT", repeat=3)] + ["---", "NNN"]
        seqs = ["".join(rng.choice(trip, 40)) for _ in range(5)]
        cols, w = al.compress_bytes(seqs, 3)
        ct = al.CodonType()
        codes, tbl = al.encode_tips(cols, ct)
        ref, wref = odata.compress_alignment(
            seqs, lambda c: odata.codon_partial(c, odata_table(), 61), size=3)
    assert torch.equal(w, wref)
    for t in range(len(seqs)):
        assert torch.equal(tbl[codes[t].long()].t(), ref[t])


def odata_table():
    from oracle.models import UNIVERSAL
    return UNIVERSAL


def test_reference_datatype_fixtures():
    """test/test_datatype.py:58-89."""
    nt = al.NucleotideType()
    assert nt.state("A") == 0 and nt.state("T") == 3 and nt.state("U") == 3
    assert nt.partial("-") == (1.0,) * 4
    aa = al.AminoAcidType()
    assert aa.state("A") == 0 and aa.state("Y") == 19
    assert aa.partial("/") == (1.0,) * 20
    ct = al.CodonType()
    assert ct.state_count == 61
    assert ct.state("AAA") == 0 and ct.state("TTT") == 60
    assert ct.state("TAC") == 48 and ct.state("TAT") == 49
    assert ct.state("TGC") == 54 and ct.state("TGG") == 55
    assert ct.partial("---") == (1.0,) * 61


def test_library_exports_every_header_symbol():
    from torchtree_amd import _lib

    header = (ROOT / "include" / "hiptree.h").read_text()
    declared = set(re.findall(r"\b(hx_[a-z0-9_]+)\s*\(", header))
    lib = _lib.load()
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(_lib.EXPORTED)


def test_plan_rejects_bad_topology():
    from torchtree_amd import _lib

    lib = _lib.load()
    h = ctypes.c_void_p()
    bad = np.array([[3, 0, 1], [4, 3, 1]], dtype=np.int32)  # node 1 has two parents
    assert lib.hx_plan_create(3, bad.ctypes.data, 0, ctypes.byref(h)) == 5
    assert b"repeated" in lib.hx_last_error()
    wrong_root = np.array([[3, 0, 1], [3, 2, 0]], dtype=np.int32)
    assert lib.hx_plan_create(3, wrong_root.ctypes.data, 0, ctypes.byref(h)) == 5
    assert lib.hx_error_string(5) == b"bad topology"


def test_json_objects_parse(tiny_newick, tiny_fasta):
    from torchtree_amd import treelikelihood_model as tm

    dic = {}
    taxa = tm.process_object({"id": "taxa", "type": "Taxa", "taxa": [
        {"id": t, "type": "Taxon", "attributes": {"date": 0.0}} for t, _ in odata.read_fasta(tiny_fasta)]}, dic)
    tree = tm.process_object({"id": "tree", "type": "UnRootedTreeModel", "newick": tiny_newick,
                              "keep_branch_lengths": True, "taxa": "taxa",
                              "branch_lengths": {"id": "bl", "type": "torchtree.Parameter", "tensor": [0.0]}},
                             dic)
    assert tree.branch_lengths().shape == (2 * 10 - 3,)
    gtr = tm.process_object({"id": "gtr", "type": "torchtree.evolution.substitution_model.GTR",
                             "rates": {"id": "r", "type": "Parameter", "tensor": [1.0, 2, 1, 1, 2, 1], "dtype": "torch.float64"},
                             "frequencies": {"id": "f", "type": "Parameter", "full": [4], "value": 0.25, "dtype": "torch.float64"}}, dic)
    q = gtr.q_norm()
    assert torch.allclose(-(torch.diagonal(q) * gtr.frequencies()).sum(), torch.tensor(1.0, dtype=torch.float64))
    w = tm.process_object({"id": "sm", "type": "WeibullSiteModel", "categories": 4,
                           "shape": {"id": "shape", "type": "Parameter", "tensor": [1.0], "dtype": "torch.float64"}}, dic)
    r, p = w.rates_probs()
    assert torch.allclose(r, torch.tensor([0.1457844, 0.5131316, 1.0708310, 2.2702530], dtype=torch.float64), rtol=1e-6)
    with pytest.raises(tm.JSONParseError):
        tm.process_object("missing", dic)


def test_integration_install_roundtrip():
    """install() swaps exactly the six pruning functions and uninstall() restores them."""
    import types

    from torchtree_amd import integration, likelihood

    mod = types.ModuleType("tree_likelihood_standin")
    originals = {}
    for n in integration.NAMES:
        f = (lambda name: (lambda *a, **k: name))(n)
        setattr(mod, n, f)
        originals[n] = f
    mod.other = 1
    saved = integration.install(mod)
    assert saved == originals
    for n in integration.NAMES:
        assert getattr(mod, n) is getattr(likelihood, n)
    assert mod.other == 1
    integration.uninstall(mod, saved)
    for n in integration.NAMES:
        assert getattr(mod, n) is originals[n]


def test_integration_install_rejects_other_modules():
    import types

    import pytest

    from torchtree_amd import integration

    mod = types.ModuleType("unrelated")
    mod.calculate_treelikelihood = 1
    with pytest.raises(AttributeError):
        integration.install(mod)
    assert mod.calculate_treelikelihood == 1


def test_slice_for_fits_the_stream_workspace():
    """Server.slice_for picks a pattern slice whose job workspace, as the library computes it on
    the host (hx_job_workspace_bytes_for), fits one stream: four-state, amino-acid and codon
    jobs, both precisions (the FP32 codon gradient adds padded P copies that do not shrink
    with the slice)."""
    from types import SimpleNamespace

    from torchtree_amd import _lib, serving

    lib = _lib.load()
    assert lib.hx_job_workspace_bytes_for(1, 4, 10, 4, 1, _lib.HX_F32, 1) == 0
    assert lib.hx_job_workspace_bytes_for(16, 4, 10, 4, 1, 7, 1) == 0
    for T, N, C, S, dt in ((512, 100000, 4, 4, torch.float32), (17, 150, 4, 4, torch.float32),
                           (256, 20000, 4, 20, torch.float32), (1024, 100000, 4, 61, torch.float32),
                           (1024, 100000, 4, 61, torch.float64)):
        code = _lib.HX_F64 if dt == torch.float64 else _lib.HX_F32
        full = lib.hx_job_workspace_bytes_for(T, S, N, C, 32, code, 1)
        assert full > lib.hx_job_workspace_bytes_for(T, S, N, C, 32, code, 0) > 0
        one = lib.hx_job_workspace_bytes_for(T, S, 1, C, 32, code, 1)
        for frac in (0.3, 0.7, 1.0):
            srv = SimpleNamespace(max_request_bytes=int(full * frac))
            if srv.max_request_bytes < one:
                with pytest.raises(serving.BudgetError):
                    serving.Server.slice_for(srv, T, N, C, S, dt)
                continue
            sl = serving.Server.slice_for(srv, T, N, C, S, dt)
            assert 1 <= sl <= N
            assert lib.hx_job_workspace_bytes_for(T, S, sl, C, 32, code, 1) <= srv.max_request_bytes
            if frac == 1.0:
                assert sl == N


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asked for more it exits with status 2 and says so, before
    touching a device."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_tip_cache_misses_on_new_tensors():
    """A fresh tip tensor that reuses a freed tensor's id and address is a cache miss."""
    import torch

    from torchtree_amd import likelihood as hl

    hits = 0
    for i in range(40):
        tips = [torch.full((4, 30), float(i % 7 + 1), dtype=torch.float64) for _ in range(3)]
        tips[0][i % 4, 0] = 0.0
        td = hl.get_tips(tips, "partials", 4, "cpu")
        dense = td.table[td.tips.long()] if td.kind == 0 else td.tips
        want = torch.stack(tips).permute(0, 2, 1)
        assert torch.equal(dense, want)
        hits += hl.get_tips(tips, "partials", 4, "cpu") is td
        tips[1].mul_(2.0)  # an in-place write is a miss too
        td2 = hl.get_tips(tips, "partials", 4, "cpu")
        assert td2 is not td
        del tips, td, td2
    assert hits == 40


def test_tip_codes_outside_table_rejected():
    import torch

    from torchtree_amd import likelihood as hl
    from torchtree_amd._lib import HipTreeError

    with pytest.raises(HipTreeError):
        hl.TipData(0, torch.tensor([[0, 3]], dtype=torch.uint8), torch.ones(3, 4), 4, 2, 1)


def test_gamma_site_model_batched_shape():
    """A batched shape [K, 1] gives each sample its own gamma rates (not the first sample's)."""
    import torch

    from oracle import models as omodels
    from torchtree_amd import treelikelihood_model as tm

    dic = {}
    site = tm.process_object({"type": "GammaSiteModel", "id": "sm", "categories": 4,
                              "shape": {"type": "Parameter", "id": "a",
                                        "tensor": [[0.5], [2.0], [7.0]]}}, dic)
    assert tuple(site.sample_shape) == (3,)
    rates, probs = site.rates_probs()
    assert rates.shape == (3, 4)
    for k, a in enumerate((0.5, 2.0, 7.0)):
        want, p = omodels.gamma_rates_probs(a, 4)
        assert torch.allclose(rates[k], want, rtol=1e-14)
    assert torch.allclose(probs, p)


def test_plugin_discovery_and_retarget():
    """Replays torchtree's PluginManager.load_plugins (cli/plugin_manager.py:10-23) and the CLI's
    load_arguments / process_tree_likelihood calls (cli/cli.py:59-60, cli/evolution.py:697-698)."""
    import argparse
    import importlib
    import pkgutil

    loaded = {}
    for _finder, name, _ispkg in pkgutil.iter_modules([str(ROOT)]):
        if name.startswith("torchtree_"):
            m = importlib.import_module(name)
            full = f"{name}.{m.__plugin__}".split(".")
            loaded[m.__plugin__] = getattr(importlib.import_module(".".join(full[:-1])), full[-1])()
    assert list(loaded) == ["plugin.HipTreePlugin"]
    plugin = loaded["plugin.HipTreePlugin"]

    parser = argparse.ArgumentParser()
    sub = parser.add_subparsers()
    for cmd in ("advi", "map", "mcmc", "hmc"):
        sub.add_parser(cmd)
    plugin.load_arguments(sub)
    like = {"id": "like", "type": "TreeLikelihoodModel", "tree_model": "tree"}
    plugin.process_tree_likelihood(parser.parse_args(["advi"]), like)
    assert like["type"] == "TreeLikelihoodModel" and "precision" not in like
    plugin.process_tree_likelihood(
        parser.parse_args(["map", "--hiptree", "--hiptree_precision", "float32"]), like)
    assert like["type"] == "torchtree_amd.treelikelihood_model.TreeLikelihoodModel"
    assert like["precision"] == "float32"
    # the type resolves the way torchtree's get_class does (core/utils.py:116-125)
    parts = like["type"].split(".")
    klass = getattr(importlib.import_module(".".join(parts[:-1])), parts[-1])
    assert hasattr(klass, "from_json_safe") and hasattr(klass, "from_json")

    # --engine torchtree_amd: cli/evolution.py imports <engine>.cli.evolution
    evol = importlib.import_module("torchtree_amd.cli.evolution")
    like2 = {"id": "like", "type": "TreeLikelihoodModel"}
    evol.process_tree_likelihood(argparse.Namespace(hiptree_precision=None), like2)
    assert like2["type"].endswith("treelikelihood_model.TreeLikelihoodModel")
    with pytest.raises(ValueError):
        evol.process_tree_likelihood(argparse.Namespace(), {"id": "x", "type": "JC69"})


def test_site_pattern_indices_match_oracle():
    """SitePattern ``indices`` (site_pattern.py:69-97, core/utils.py:404-419): the selected
    columns, in the order given and with repeats, compressed like the whole alignment."""
    import numpy as np

    from oracle import data as odata
    from torchtree_amd import treelikelihood_model as tm

    idx = tm.parse_indices("3:40,7,0:60:3,100:,::17")
    assert idx == [slice(3, 40), 7, slice(0, 60, 3), slice(100, None), slice(None, None, 17)]
    rng = np.random.default_rng(8)
    seqs = ["".join(rng.choice(list("ACGT-"), 130)) for _ in range(6)]
    taxa = [tm.Taxon(f"t{i}", {}) for i in range(6)]
    sp = tm.SitePattern(tm.Alignment([(f"t{i}", s) for i, s in enumerate(seqs)], taxa,
                                     al.NucleotideType()), idx)
    cols, w = sp.compressed()
    ref, wref = odata.compress(seqs, indices=idx)
    assert torch.equal(w, wref)
    got = [["".join(chr(c) for c in cols[t, :, j]) for j in range(cols.shape[2])]
           for t in range(6)]
    assert [list(p) for p in ref] == got


def test_bench_byte_accounting():
    """bench.py's roofline bytes: the level-by-level per-update figures reproduce the worked
    numbers for the 512-taxon, 100,000-pattern, K=32 step (216 GB post-order, 837 GB
    pre-order); the walks' own minimum drops one partial per unstored cherry from each pass."""
    import bench

    canon = bench.canonical_bytes(512, 100_000, 4, 32, 4, 4)
    assert round(canon["postorder"] / 1e9) == 216 and round(canon["preorder"] / 1e9) == 837
    part = 32 * 100_000 * 4 * 4 * 4
    post0, pre0 = bench.algorithmic_bytes(512, 100_000, 4, 32, 4, 4, 391, 0)
    post1, pre1 = bench.algorithmic_bytes(512, 100_000, 4, 32, 4, 4, 391, 171)
    assert post0 - post1 == 171 * part
    assert pre0 - pre1 == 171 * part - 32 * 171 * 100_000 * 2  # the cherry code pairs are read
    assert bench.walk_drops(512, [], None, 20, 4, 100_000, 32, 4) == 0  # dense states drop none
