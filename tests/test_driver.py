"""The command-line flow's host part (CPU only): config files, option checks and output names
as V11/main/main.cpp:630-930 defines them."""
import os

import pytest

from conftest import GOLDEN


def test_reads_the_reference_example_config():
    from vdock import driver
    with open(os.path.join(GOLDEN, "1u1b_config.txt")) as f:
        raw = driver.read_config(f.read())
    assert raw["center_x"] == "-30.661" and raw["size_y"] == "26.25" and raw["thread"] == "8192"
    assert raw["receptor"] == "./input_file_example/1u1b_protein.pdbqt"
    opts = driver.resolve(os.path.join(GOLDEN, "1u1b_config.txt"), seed=5)
    assert opts["thread"] == 8192 and opts["size_z"] == 22.5 and opts["seed"] == 5
    assert opts["num_modes"] == 9 and opts["energy_range"] == 3.0 and opts["rilc_bfgs"] == 1 and opts["search_depth"] == 0
    assert opts["weight_rot"] == 0.05846


def test_option_checks_follow_the_reference(tmp_path):
    from vdock import driver
    base = dict(receptor="r.pdbqt", ligand="l.pdbqt", thread=8192, center_x=0, center_y=0, center_z=0,
                size_x=20, size_y=20, size_z=20)
    assert driver.resolve(**base)["seed"] is not None          # auto seed
    for change, msg in [(dict(thread=999), "at least 1000"), (dict(thread=None), "Missing thread"),
                        (dict(receptor=None), "Missing receptor"), (dict(ligand=None), "Missing ligand"),
                        (dict(ligand_directory="d"), "at the same time"), (dict(size_y=0), "positive"),
                        (dict(center_z=None), "center_z"), (dict(num_modes=0), "num_modes"),
                        (dict(thread="many"), "bad value")]:
        kw = dict(base)
        kw.update(change)
        with pytest.raises(driver.UsageError, match=msg):
            driver.resolve(**{k: v for k, v in kw.items() if v is not None})
    cfg = tmp_path / "c.txt"
    cfg.write_text("receptor = r.pdbqt  # comment\nligand=l.pdbqt\n\nthread = 2000\ncenter_x=1\ncenter_y=2\n"
                   "center_z=3\nsize_x=4\nsize_y=5\nsize_z=6\n")
    o = driver.resolve(str(cfg), thread=4000)
    assert o["thread"] == 4000 and o["center_y"] == 2.0 and o["receptor"] == "r.pdbqt"
    cfg.write_text("recptor = r.pdbqt\n")
    with pytest.raises(driver.UsageError, match="unrecognised"):
        driver.resolve(str(cfg))


def test_output_names(tmp_path):
    from vdock import driver
    assert driver.default_output("a/b/lig.pdbqt") == "a/b/lig_out.pdbqt"
    assert driver.ligand_jobs(dict(ligand="x.pdbqt", ligand_directory=None, out=None)) == [("x.pdbqt", "x_out.pdbqt")]
    assert driver.ligand_jobs(dict(ligand="x.pdbqt", ligand_directory=None, out="y.pdbqt")) == [("x.pdbqt", "y.pdbqt")]
    d = tmp_path / "ligs"
    d.mkdir()
    for n in ("b.pdbqt", "a.pdbqt"):
        (d / n).write_text("")
    jobs = driver.ligand_jobs(dict(ligand=None, ligand_directory=str(d), output_directory=None))
    assert jobs == [(str(d / "a.pdbqt"), str(d) + "_out/a_out.pdbqt"), (str(d / "b.pdbqt"), str(d) + "_out/b_out.pdbqt")]
    jobs = driver.ligand_jobs(dict(ligand=None, ligand_directory=str(d), output_directory=str(tmp_path / "o")))
    assert jobs[0][1] == str(tmp_path / "o") + "/a_out.pdbqt"


def test_mode_table_format():
    import vdock
    from vdock import driver
    t = driver.mode_table([vdock.Mode(-7.93, 0.0, 0.0, 0), vdock.Mode(-7.616, 5.6041, 9.4939, 3)])
    lines = t.splitlines()
    assert lines[0] == "mode |   affinity | dist from best mode"
    assert lines[3] == "   1         -7.9      0.000      0.000"
    assert lines[4] == "   2         -7.6      5.604      9.494"
