"""Vina-CUDA's command-line flow over the C ABI.

Restates what `main()` and `main_procedure` do around the docking path (V11/main/main.cpp:
600-970 and 330-501): read the options (a config file of `name = value` lines and/or keyword
overrides, the reference's option names and defaults), check them as the reference does, read
the receptor and the ligand -- or every file of a ligand directory (virtual screening mode) --,
dock all ligands in one batch, run the step after the search (refinement, final scores, modes)
and write one output PDBQT per ligand under the reference's output names.  Ligands that do not
parse are reported and skipped, as the reference's catch does.

Everything runs through libvdock (`vdock.Context`); there is no CPU path.
"""
import os
import random

from . import (VD_ERR_PARSE, VD_LS_BFGS, VD_LS_RILC_BFGS, Context, VdockError, box_from_center,
               default_params, post_params)

# option -> (type, default); V11/main/main.cpp:630-700 (None = no default)
OPTIONS = {
    "receptor": (str, None), "ligand": (str, None), "ligand_directory": (str, None),
    "output_directory": (str, None), "out": (str, None), "log": (str, None),
    "thread": (int, None), "search_depth": (int, 0), "rilc_bfgs": (int, 1),
    "center_x": (float, None), "center_y": (float, None), "center_z": (float, None),
    "size_x": (float, None), "size_y": (float, None), "size_z": (float, None),
    "seed": (int, None), "num_modes": (int, 9), "energy_range": (float, 3.0),
    "weight_gauss1": (float, -0.035579), "weight_gauss2": (float, -0.005156),
    "weight_repulsion": (float, 0.840245), "weight_hydrophobic": (float, -0.035069),
    "weight_hydrogen": (float, -0.587439), "weight_rot": (float, 0.05846),
    "exhaustiveness": (int, 8), "cpu": (int, 0),
}
BOX = ("center_x", "center_y", "center_z", "size_x", "size_y", "size_z")
GRANULARITY = 0.375


class UsageError(ValueError):
    """An option problem the reference reports before any work (main.cpp:745-860)."""


def read_config(text):
    """`name = value` lines as boost's parse_config_file reads them: '#' starts a comment,
    blank lines are skipped, unknown names are an error."""
    out = {}
    for n, raw in enumerate(text.splitlines(), 1):
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        if "=" not in line:
            raise UsageError(f"configuration line {n}: expected name = value")
        name, value = (s.strip() for s in line.split("=", 1))
        if name not in OPTIONS:
            raise UsageError(f"configuration line {n}: unrecognised option '{name}'")
        out[name] = value
    return out


def resolve(config=None, **overrides):
    """Options from a config file (path) and keyword overrides (these win, like the command
    line over the config file), typed, with the reference's defaults and checks."""
    raw = {}
    if config is not None:   # relative paths inside it are taken from the working directory, as the reference does
        with open(config) as f:
            raw.update(read_config(f.read()))
    raw.update({k: v for k, v in overrides.items() if v is not None})
    opts = {}
    for name, (typ, default) in OPTIONS.items():
        if name in raw:
            try:
                opts[name] = typ(raw[name])
            except ValueError:
                raise UsageError(f"option '{name}': bad value {raw[name]!r}") from None
        else:
            opts[name] = default
    for name in raw:
        if name not in OPTIONS:
            raise UsageError(f"unrecognised option '{name}'")
    if opts["receptor"] is None:
        raise UsageError("Missing receptor.")
    if opts["thread"] is None:
        raise UsageError("Missing thread, the recommend thread value is 8000")
    if opts["thread"] < 1000:
        raise UsageError("thread must be at least 1000")
    if opts["ligand"] is not None and opts["ligand_directory"] is not None:
        raise UsageError("Can not define ligand and ligand_directory at the same time")
    if opts["ligand"] is None and opts["ligand_directory"] is None:
        raise UsageError("Missing ligand")
    missing = [b for b in BOX if opts[b] is None]
    if missing:
        raise UsageError("the search space needs " + ", ".join(missing))
    if min(opts["size_x"], opts["size_y"], opts["size_z"]) <= 0:
        raise UsageError("Search space dimensions should be positive")
    if opts["exhaustiveness"] < 1:
        raise UsageError("exhaustiveness must be 1 or greater")
    if opts["num_modes"] < 1:
        raise UsageError("num_modes must be 1 or greater")
    if opts["seed"] is None:
        opts["seed"] = random.SystemRandom().randrange(1 << 31)
    return opts


def default_output(ligand):
    """main.cpp:81-87: strip '.pdbqt', append '_out.pdbqt'."""
    return (ligand[:-6] if ligand.endswith(".pdbqt") else ligand) + "_out.pdbqt"


def ligand_jobs(opts):
    """(ligand path, output path) pairs (main.cpp:898-926).  A directory's entries are taken in
    sorted order; each output is the entry's path with the directory part replaced by the output
    directory (default: the ligand directory + '_out') and '.pdbqt' by '_out.pdbqt'."""
    if opts["ligand_directory"] is None:
        return [(opts["ligand"], opts["out"] or default_output(opts["ligand"]))]
    d = opts["ligand_directory"]
    out_dir = opts["output_directory"] or d + "_out"
    jobs = []
    for entry in sorted(os.listdir(d)):
        p = os.path.join(d, entry)
        stem = p[: p.find(".pdbqt")] if ".pdbqt" in p else p
        jobs.append((p, out_dir + (stem + "_out.pdbqt")[len(d):]))
    return jobs


def mode_table(modes):
    """The reference's result table (main.cpp:472-490)."""
    lines = ["mode |   affinity | dist from best mode", "     | (kcal/mol) | rmsd l.b.| rmsd u.b.",
             "-----+------------+----------+----------"]
    for i, m in enumerate(modes):
        lines.append(f"{i + 1:4d}    {m.e:9.1f}  {m.rmsd_lb:9.3f}  {m.rmsd_ub:9.3f}")
    return "\n".join(lines)


def run(opts, device=0, log=print, ctx=None):
    """Docks every ligand of `opts` (from resolve()) and writes the outputs.  Returns one dict
    per docked ligand: ligand, out, modes (list of vdock.Mode), text (the PDBQT written)."""
    own = ctx is None
    ctx = ctx or Context(device)
    try:
        ctx.set_weights([opts["weight_gauss1"], opts["weight_gauss2"], opts["weight_repulsion"],
                         opts["weight_hydrophobic"], opts["weight_hydrogen"], opts["weight_rot"]])
        with open(opts["receptor"]) as f:
            ctx.set_receptor_pdbqt(f.read())
        box = box_from_center([opts["center_x"], opts["center_y"], opts["center_z"]],
                              [opts["size_x"], opts["size_y"], opts["size_z"]], GRANULARITY)
        ctx.grid_populate(box)
        jobs = ligand_jobs(opts)
        if opts["ligand_directory"] is not None:
            out_dir = opts["output_directory"] or opts["ligand_directory"] + "_out"
            os.makedirs(out_dir, exist_ok=True)
            log(f"Output will be in the directory {out_dir}")
        ids, valid = [], []
        for lig, out in jobs:
            try:
                with open(lig) as f:
                    ids.append(ctx.add_ligand(f.read()))
                valid.append((lig, out))
            except VdockError as e:
                if e.code != VD_ERR_PARSE:
                    raise
                log(f"Parse error in file \"{lig}\": {e}")
        if not ids:
            raise UsageError("No valid ligands in the input directory")
        ls = VD_LS_RILC_BFGS if opts["rilc_bfgs"] == 1 else VD_LS_BFGS
        params = default_params(chains=opts["thread"], seed=opts["seed"] & 0xFFFFFFFF, local_search=ls,
                                search_depth=opts["search_depth"])
        ctx.dock_batch(ids, params)
        post = post_params(num_modes=opts["num_modes"], energy_range=opts["energy_range"], local_search=ls)
        ctx.postprocess_docked(20, post)
        results = []
        for k, (lig, out) in enumerate(valid):
            modes, _, coords = ctx.post_results(k, opts["num_modes"])
            text = ctx.write_pdbqt(ids[k], modes, coords)
            with open(out, "w") as f:
                f.write(text)
            log(f"Refining ligand {os.path.basename(lig)} results...")
            log(mode_table(modes))
            if not modes:
                log("WARNING: Could not find any conformations completely within the search space.")
            results.append(dict(ligand=lig, out=out, modes=modes, text=text))
        return results
    finally:
        if own:
            ctx.close()
