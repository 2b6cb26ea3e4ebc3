"""CPU tests of libvdock's host side: the C ABI surface, the FP64 tables, PDBQT flattening,
heuristics, clustering and error handling -- all without a GPU (host-only context, device -1)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as orc
from conftest import ROOT, golden


def test_library_exports_every_declared_symbol():
    import vdock
    lib = vdock.load()
    header = open(os.path.join(ROOT, "include", "vdock.h")).read()
    declared = set(re.findall(r"\b(vd_[a-z_]+)\s*\(", header))
    assert declared == set(vdock.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    assert vdock.version().startswith("vdock")


def test_no_device_is_an_error_not_a_fallback(host_ctx):
    import vdock
    box = vdock.box_from_center([0, 0, 0], [4, 4, 4])
    with pytest.raises(vdock.VdockError) as ei:
        host_ctx.grid_populate(box)
    assert ei.value.code == vdock.VD_ERR_NO_DEVICE


def test_lut_matches_oracle(host_ctx):
    lut = orc.Lut()
    worst = 0.0
    for tp in range(153):
        for a, b in zip(lut.tables(tp), host_ctx.lut(tp)):
            worst = max(worst, float(np.abs(a - b).max()))
    assert worst < 1e-11


def test_weights_change_tables(host_ctx):
    import vdock
    w = list(vdock.DEFAULT_WEIGHTS)
    w[2] *= 2.0
    host_ctx.set_weights(w)
    try:
        lut = orc.Lut(tuple(w[:5]))
        for tp in (0, 40, 152):
            assert np.allclose(lut.tables(tp)[2], host_ctx.lut(tp)[2], atol=1e-11)
    finally:
        host_ctx.set_weights(vdock.DEFAULT_WEIGHTS)


def flat_equal(host_ctx, text):
    lid = host_ctx.add_ligand(text)
    ex = host_ctx.ligand_export(lid)
    L = orc.Ligand(text)
    a, nd = L.atoms(), L.nodes()
    assert np.array_equal(ex["xs"], a["xs"])
    assert np.array_equal(ex["el"] == 0, a["el"] == 0)
    assert np.array_equal(ex["node"], a["node"])
    assert np.array_equal(ex["parent"], nd["parent"])
    assert np.array_equal(ex["begin"], nd["begin"]) and np.array_equal(ex["end"], nd["end"])
    assert np.array_equal(ex["pairs"], L.pairs())
    assert np.abs(ex["local"] - a["local"]).max() < 1e-5
    assert np.abs(ex["rel_origin"] - nd["rel_origin"]).max() < 1e-5
    assert np.abs(ex["rel_axis"] - nd["rel_axis"]).max() < 1e-6
    return host_ctx.ligand_info(lid)


@pytest.mark.parametrize("name", ["1u1b", "2bm2", "2er7"])
def test_flattening_matches_oracle_reference_ligands(host_ctx, name):
    inf = flat_equal(host_ctx, golden(f"{name}_ligand.pdbqt"))
    assert inf.n_atoms == orc.Ligand(golden(f"{name}_ligand.pdbqt")).n_atoms


def test_flattening_matches_oracle_synthetic(host_ctx):
    import vdock
    for k in range(12):
        n_heavy = 10 + 6 * k
        text = vdock.synth_ligand(n_heavy, min(16, n_heavy // 3), 1000 + k)
        inf = flat_equal(host_ctx, text)
        assert inf.n_heavy == n_heavy
        assert inf.n_torsions == min(16, n_heavy // 3) or inf.n_torsions < min(16, n_heavy // 3)


def test_receptor_typing_matches_oracle(host_ctx):
    text = golden("1u1b_protein.pdbqt")
    n = host_ctx.set_receptor_pdbqt(text)
    xyz, xs = host_ctx.receptor()
    xyz_o, xs_o, _ = orc.parse_receptor(text)
    assert n == len(xs_o) == 1192 + (len(xs_o) - 1192)
    assert np.array_equal(xs, xs_o)
    assert np.abs(xyz - xyz_o).max() < 1e-4


def test_box_matches_vina_grid_dims():
    import vdock
    b = vdock.box_from_center([-30.661, -29.540, 17.010], [24.0, 26.25, 22.5], 0.375)
    ob, oe, on = orc.grid_dims([-30.661, -29.540, 17.010], [24.0, 26.25, 22.5], 0.375)
    assert list(b.n) == on == [64, 70, 60]
    assert np.allclose(list(b.begin), ob) and np.allclose(list(b.end), oe)
    # 65 x 71 x 61 points, as the reference log reports for the 1u1b box
    assert b.dims == (65, 71, 61)


def test_heuristics_match_reference_log_and_formula():
    import vdock
    # V11/main/main.cpp:354-391; values reported in QVW/input_file_example/auto_test/log_1u1b.txt
    assert vdock.heuristics(19, 7) == (6, 16)
    assert vdock.heuristics(41, 11) == (19, 23)
    assert vdock.heuristics(44, 12) == (21, 24)
    assert vdock.heuristics(33, 11) == (15, 21)
    assert vdock.heuristics(7, 2) == (1, 12)
    assert vdock.heuristics(32, 8) == (13, 20)
    assert vdock.heuristics(32, 8, rilc=False)[1] == (25 + 32) // 3


def test_cluster_matches_oracle():
    import vdock
    rng = np.random.default_rng(2)
    for trial in range(5):
        n, h = 300, 12
        centers = rng.normal(scale=3.0, size=(40, 1, 3))
        heavy = centers[rng.integers(0, 40, n)] + rng.normal(scale=0.3, size=(n, h, 3))
        e = rng.normal(size=n)
        a = vdock.cluster(e.astype(np.float32), heavy.astype(np.float32), 1.0, 20)
        b = orc.cluster(e.astype(np.float32).astype(np.float64), heavy.astype(np.float32).astype(np.float64), 1.0, 20)
        assert list(a) == list(b)


def test_parse_errors_are_reported(host_ctx):
    import vdock
    with pytest.raises(vdock.VdockError) as ei:
        host_ctx.add_ligand("ROOT\nATOM      1  C   UNL     1       0.000   0.000   0.000  0.00  0.00    +0.000 Xx\nENDROOT\nTORSDOF 0\n")
    assert ei.value.code == vdock.VD_ERR_PARSE
    with pytest.raises(vdock.VdockError):
        host_ctx.add_ligand("ROOT\nENDROOT\n")
    with pytest.raises(vdock.VdockError) as ei:
        host_ctx.add_ligand(golden("1u1b_ligand.pdbqt").replace("TORSDOF", "REMARK"))
    assert ei.value.code == vdock.VD_ERR_PARSE


def test_capacity_errors_are_reported(host_ctx):
    import vdock
    big = vdock.synth_ligand(260, 10, 5)     # more atoms than the 8-bit atom indices of a ligand record address
    with pytest.raises(vdock.VdockError):
        host_ctx.add_ligand(big)
    # beyond the reference's 128-atom cap (MAX_NUM_OF_ATOMS) but within the record's
    lid = host_ctx.add_ligand(vdock.synth_ligand(200, 20, 5))
    assert host_ctx.ligand_info(lid).n_atoms >= 200
    n_before = host_ctx.lib.vd_ligand_count(host_ctx.h)
    ok = host_ctx.add_ligand(vdock.synth_ligand(12, 2, 5))
    assert ok == n_before   # a rejected ligand does not consume an id


def test_invalid_arguments(host_ctx):
    import vdock
    lib = vdock.load()
    assert lib.vd_create(0, None) == vdock.VD_ERR_ARG
    b = vdock.Box()
    assert lib.vd_box_from_center(None, None, 0.375, ctypes.byref(b)) == vdock.VD_ERR_ARG
    assert lib.vd_set_graphs(None, 1) == vdock.VD_ERR_ARG
    assert lib.vd_set_graphs(host_ctx.h, 2) == vdock.VD_ERR_ARG
    assert lib.vd_set_graphs(host_ctx.h, 0) == vdock.VD_OK
    # result fetches take the caller's capacity: with no batch docked there is nothing to copy
    assert lib.vd_batch_best(None, 0, None) == vdock.VD_ERR_ARG
    assert lib.vd_batch_best(host_ctx.h, -1, None) == vdock.VD_ERR_ARG
    assert lib.vd_batch_best(host_ctx.h, 1, None) == vdock.VD_ERR_ARG
    assert lib.vd_batch_best(host_ctx.h, 0, None) == 0
    with pytest.raises(vdock.VdockError):
        host_ctx.set_receptor(np.array([[np.nan, 0, 0]], dtype=np.float32), np.array([0], dtype=np.int32))
    with pytest.raises(vdock.VdockError):
        host_ctx.ligand_info(10 ** 6)
    # parameters are checked before the device: the QuickVina2 and QuickVina-W policies run with RILC-BFGS only
    lid = host_ctx.add_ligand(vdock.synth_ligand(12, 2, 1))
    for bad in (dict(policy=1, local_search=0), dict(policy=2, local_search=0), dict(policy=3), dict(chains=0), dict(chains_per_wave=3), dict(waves_per_simd=5)):
        with pytest.raises(vdock.VdockError) as ei:
            host_ctx.dock(lid, vdock.default_params(**bad))
        assert ei.value.code == vdock.VD_ERR_ARG, bad


def test_synthetic_receptor_recipe():
    import vdock
    xyz, xs = vdock.synth_receptor(2000, 36.0, 8.0, 42)
    assert len(xs) == 2000
    d = np.linalg.norm(xyz[:, None] - xyz[None], axis=-1) + np.eye(2000) * 10
    assert d.min() >= 1.2 - 1e-5
    assert np.linalg.norm(xyz, axis=1).min() >= 8.0 - 1e-5
    frac = np.bincount(xs, minlength=17) / 2000.0
    assert abs(frac[0] - 0.35) < 0.05 and abs(frac[1] - 0.28) < 0.05 and abs(frac[8] - 0.12) < 0.04


@pytest.mark.parametrize("lanes", [16, 32, 64])
def test_pair_chunks_cover_every_pair_once_without_conflicts(host_ctx, lanes):
    """Device pair layout: every intramolecular pair exactly once (either way round, same type
    pair), and within a chunk no atom is the first atom of two pairs or the second atom of two
    pairs -- the condition under which the kernels scatter pair forces without atomics."""
    import vdock
    from conftest import golden
    texts = [golden(f"{n}_ligand.pdbqt") for n in ("1u1b", "2bm2", "2er7")]
    texts += [vdock.synth_ligand(h, min(16, h // 3), 300 + h) for h in (10, 25, 40, 60)]
    texts += [vdock.synth_ligand(h, 24, 300 + h) for h in (140, 225)]      # beyond 128 atoms
    for text in texts:
        lid = host_ctx.add_ligand(text)
        pairs = host_ctx.ligand_export(lid)["pairs"]          # (tp, a, b)
        ch = host_ctx.pair_chunks(lid, lanes)
        assert ch.shape[1] == lanes
        words = ch.ravel()
        used = words[words != 0xFFFFFFFF]
        a, b, tp = used & 255, (used >> 8) & 255, used >> 16
        got = sorted(zip(tp.tolist(), np.minimum(a, b).tolist(), np.maximum(a, b).tolist()))
        want = sorted(zip(pairs[:, 0].tolist(), np.minimum(pairs[:, 1], pairs[:, 2]).tolist(),
                          np.maximum(pairs[:, 1], pairs[:, 2]).tolist()))
        assert got == want
        for row in ch:
            r = row[row != 0xFFFFFFFF]
            assert len(set((r & 255).tolist())) == len(r)
            assert len(set(((r >> 8) & 255).tolist())) == len(r)
            if host_ctx.ligand_info(lid).n_atoms <= lanes:     # each atom has a lane: slot = first atom
                slots = np.nonzero(row != 0xFFFFFFFF)[0]
                assert np.array_equal(slots, (row[slots] & 255))
        # a chunk holds at most `lanes` pairs and each atom in at most two of them, so at least
        # max(ceil(P / lanes), ceil(max degree / 2)) chunks are needed; the edge colouring
        # (Euler orientation + equitable bipartite colouring) reaches that bound
        deg = np.bincount(np.concatenate([pairs[:, 1], pairs[:, 2]]), minlength=1)
        assert ch.shape[0] == max(-(-len(pairs) // lanes), -(-int(deg.max()) // 2))


def test_cpp_consumer_of_the_header():
    """vina-cuda_amd/examples/vdock_dock, a C++ program built against include/vdock.h and linked
    with libvdock, reports bad options with status 2 and, on a host without a GPU, the missing
    device with status 3 (its run on a GPU is test_gpu_example.py)."""
    import subprocess
    import vdock
    exe = os.path.join(ROOT, "vina-cuda_amd", "examples", "vdock_dock")
    assert os.path.exists(exe), "build with __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "missing --receptor" in r.stderr
    args = [exe, "--receptor", os.path.join(ROOT, "tests", "golden", "1u1b_protein.pdbqt"), "--ligand_directory",
            os.path.join(ROOT, "tests", "golden", "qvw_auto"), "--output_directory", "/nonexistent/out",
            "--center_x", "0", "--center_y", "0", "--center_z", "0", "--size_x", "10", "--size_y", "10",
            "--size_z", "10", "--thread", "999", "--seed", "1"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "thread must be at least 1000" in r.stderr
    h = ctypes.c_void_p()
    if vdock.load().vd_create(0, ctypes.byref(h)) == vdock.VD_OK:
        vdock.load().vd_destroy(h)
        pytest.skip("a GPU is visible here")
    args[args.index("999")] = "8000"
    r = subprocess.run(args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 3 and "no usable GPU" in r.stderr
