"""Writes the QuickVina-W auto_test fixtures of tests/golden/ (run from the repository root, in a
checkout that has the reference beside it at /root/reference):

  qvw_auto/<ligand>.pdbqt   the ligand inputs of QVW/input_file_example/auto_test/{1u1b,2bm2,2er7}/
                            that are not already golden files (the 2bm2 and 2er7 folders hold the
                            same ten inputs, byte for byte; 1u1b_ligand.pdbqt is the V11 example's)
  qvw_auto_test.npz         for each of the 31 docked outputs (<dir>/<ligand>_out.pdbqt): the
                            atom serials, the coordinates of the nine modes in milli-Angstrom
                            (exactly as printed, %8.3f) and the REMARK VINA RESULT triples
  qvw_auto_confs.npz        the conformations (position, orientation, torsions; rounded to FP32)
                            that put each ligand's atoms on those modes, fitted with
                            tests/test_oracle_rescore.py's conf_from_coords (Kabsch fit of the
                            root, then a scan per torsion), for the device post-processing test

The receptors (auto_test/{1u1b,2bm2,2er7}_protein.pdbqt) are byte-identical to the golden
1u1b/2bm2/2er7_protein.pdbqt.  Nothing here runs reference code: the files are read as text.
`--confs` writes only qvw_auto_confs.npz, from the committed fixtures.
"""
import os
import shutil
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
GOLDEN = os.path.join(ROOT, "tests", "golden")
SRC = "/root/reference/QuickVina-W-CUDA/input_file_example/auto_test"
DIRS = ("1u1b", "2bm2", "2er7")


def parse_out(text):
    """MODEL blocks -> (serials [n], milli-A coordinates [models, n, 3], remarks [models, 3])."""
    models, remarks, cur = [], [], None
    for ln in text.splitlines():
        if ln.startswith("MODEL"):
            cur = []
        elif ln.startswith("REMARK VINA RESULT:"):
            remarks.append([float(x) for x in ln.split()[3:6]])
        elif ln.startswith(("ATOM", "HETATM")) and cur is not None:
            cur.append((int(ln[6:11]), [round(float(ln[c:c + 8]) * 1000) for c in (30, 38, 46)]))
        elif ln.startswith("ENDMDL"):
            models.append(cur)
            cur = None
    ser = [s for s, _ in models[0]]
    assert all([s for s, _ in m] == ser for m in models)
    xyz = np.array([[p for _, p in m] for m in models], dtype=np.int32)
    return np.array(ser, dtype=np.int32), xyz, np.array(remarks, dtype=np.float64)


def main():
    out_dir = os.path.join(GOLDEN, "qvw_auto")
    os.makedirs(out_dir, exist_ok=True)
    arrays, recs, ligs = {}, [], []
    for d in DIRS:
        names = sorted(f for f in os.listdir(os.path.join(SRC, d)) if f.endswith("_out.pdbqt"))
        for f in names:
            lig = f[:-len("_out.pdbqt")]
            src_in = os.path.join(SRC, d, lig + ".pdbqt")
            if lig != "1u1b_ligand" and not os.path.exists(os.path.join(out_dir, lig + ".pdbqt")):
                shutil.copyfile(src_in, os.path.join(out_dir, lig + ".pdbqt"))
            with open(os.path.join(out_dir if lig != "1u1b_ligand" else GOLDEN, lig + ".pdbqt"), "rb") as fh, \
                    open(src_in, "rb") as fs:
                assert fh.read() == fs.read(), ("inputs of the same name differ", d, lig)
            with open(os.path.join(SRC, d, f)) as fh:
                ser, xyz, rem = parse_out(fh.read())
            k = len(recs)
            arrays[f"serial_{k}"], arrays[f"xyz_{k}"], arrays[f"remark_{k}"] = ser, xyz, rem
            recs.append(d)
            ligs.append(lig)
    np.savez_compressed(os.path.join(GOLDEN, "qvw_auto_test.npz"), receptor=np.array(recs),
                        ligand=np.array(ligs), **arrays)
    print(len(recs), "outputs")


def fit_confs():
    sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]
    import oracle as orc
    from test_oracle_rescore import conf_from_coords
    from test_oracle_rescore_qvw_auto import CASES, ligand_text, modes
    out, worst = {}, 0.0
    for _, name, k in CASES:
        lig = orc.Ligand(ligand_text(name))
        X = modes(k, lig)
        c = np.array([conf_from_coords(lig, x) for x in X]).astype(np.float32)
        worst = max(worst, max(np.abs(lig.set_conf(ci.astype(np.float64)) - x).max() for ci, x in zip(c, X)))
        out[f"conf_{k}"] = c
    np.savez_compressed(os.path.join(GOLDEN, "qvw_auto_confs.npz"), **out)
    print(len(out), "outputs fitted, worst atom distance", worst)


if __name__ == "__main__":
    if "--confs" in sys.argv:
        fit_confs()
    else:
        main()
