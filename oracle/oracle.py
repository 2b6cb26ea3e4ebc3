"""ctypes front end of the FP64 CPU oracle (oracle/vina_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the checker.  The product (vina-cuda_amd/) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIBRARY selects another build of the oracle (its sanitizer build)
LIB_PATH = os.environ.get("ORACLE_LIBRARY") or os.path.join(HERE, "_build", "liboracle.so")

# Weights of the five Vina terms (V11/main/main.cpp:649-654)
VINA_WEIGHTS = (-0.035579, -0.005156, 0.840245, -0.035069, -0.587439)
VINA_WEIGHTS_ROT = 0.05846  # weight_rot, V11/main/main.cpp:654
N_XS = 17
N_PAIRS = 153


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    sig = {
        "orc_last_error": (C.c_char_p, []),
        "orc_philox": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "orc_term_eval": (C.c_double, [C.c_int, C.c_int, C.c_double, dp]),
        "orc_set_terms": (None, [C.c_int, C.c_int]),
        "orc_lut_create": (vp, [dp]),
        "orc_lut_destroy": (None, [vp]),
        "orc_lut_size": (C.c_int, []),
        "orc_lut_get": (C.c_int, [vp, C.c_int, dp, dp, dp]),
        "orc_lut_eval_fast": (C.c_double, [vp, C.c_int, C.c_double]),
        "orc_lut_eval_deriv": (C.c_int, [vp, C.c_int, C.c_double, dp, dp]),
        "orc_receptor_parse": (C.c_int, [C.c_char_p, C.c_int, dp, ip, ip]),
        "orc_grid_create": (vp, [vp, dp, dp, ip]),
        "orc_grid_destroy": (None, [vp]),
        "orc_grid_set_receptor": (C.c_int, [vp, C.c_int, dp, ip]),
        "orc_grid_populate": (C.c_int, [vp, C.c_int]),
        "orc_grid_populate_slack": (C.c_int, [vp, C.c_int, C.c_double, dp]),
        "orc_grid_get_maps": (C.c_int, [vp, dp]),
        "orc_grid_set_maps": (C.c_int, [vp, dp]),
        "orc_grid_cell_candidates": (C.c_int, [vp, dp, ip, C.c_int]),
        "orc_grid_eval": (C.c_double, [vp, C.c_int, dp, C.c_double, C.c_double, dp]),
        "orc_ligand_parse": (vp, [C.c_char_p]),
        "orc_ligand_destroy": (None, [vp]),
        "orc_ligand_sizes": (C.c_int, [vp, ip]),
        "orc_ligand_atoms": (C.c_int, [vp, ip, ip, ip, ip, dp, dp]),
        "orc_ligand_nodes": (C.c_int, [vp, ip, ip, ip, dp, dp, dp]),
        "orc_ligand_pairs": (C.c_int, [vp, ip, ip, ip]),
        "orc_ligand_bonds": (C.c_int, [vp, ip, ip, C.c_int]),
        "orc_set_conf": (C.c_int, [vp, dp, dp]),
        "orc_eval_deriv": (C.c_double, [vp, vp, dp, C.c_double, C.c_double, C.c_double, dp, dp, dp]),
        "orc_conf_increment": (C.c_int, [vp, dp, dp, C.c_double]),
        "orc_local_search": (C.c_double, [vp, vp, dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, ip]),
        "orc_opt_test": (C.c_double, [C.c_int, C.c_int, dp, dp, C.c_int, C.c_int, ip]),
        "orc_opt_trace": (C.c_double, [C.c_int, C.c_int, dp, dp, C.c_int, C.c_int, C.c_int, dp, dp, ip]),
        "orc_mc_trace": (C.c_int, [vp, vp, vp, C.c_int, ip, dp, dp, dp, ip, dp, dp, dp]),
        "orc_mc_run": (C.c_int, [vp, vp, vp, dp, dp, dp, C.POINTER(C.c_longlong)]),
        "orc_mc_run_from": (C.c_int, [vp, vp, vp, dp, dp, dp, dp, C.POINTER(C.c_longlong)]),
        "orc_initial_conf": (C.c_int, [vp, vp, C.c_int, dp]),
        "orc_cluster": (C.c_int, [C.c_int, C.c_int, dp, dp, C.c_double, C.c_int, ip]),
        "orc_visited_feed": (C.c_int, [C.c_int, C.c_int, dp, dp, dp, C.c_int, ip]),
        "orc_qvw_feed": (C.c_int, [C.c_int, C.c_int, dp, dp, dp, ip, ip, dp, dp, C.c_int, ip]),
        "orc_qvw_levels": (C.c_int, [C.c_longlong, C.c_longlong]),
        "orc_ligand_serials": (C.c_int, [vp, ip]),
        "orc_num_tors": (C.c_double, [vp]),
        "orc_conf_independent": (C.c_double, [vp, C.c_double, C.c_double]),
        "orc_noncache_eval": (C.c_double, [vp, vp, dp, C.c_double, C.c_double, dp]),
        "orc_intra_fast": (C.c_double, [vp, vp, dp, C.c_double]),
        "orc_within": (C.c_int, [vp, vp, dp]),
        "orc_rmsd_bounds": (C.c_int, [vp, dp, dp, dp, dp]),
        "orc_noncache_eval_deriv": (C.c_double, [vp, vp, dp, C.c_double, C.c_double, dp]),
        "orc_refine": (C.c_double, [vp, vp, dp, C.c_int, C.c_double, dp]),
        "orc_postprocess": (C.c_int, [vp, vp, C.c_int, dp, C.c_int, C.c_double, C.c_int, C.c_double,
                                      C.c_double, C.c_double, dp, dp, dp, dp, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _err():
    return lib().orc_last_error().decode()


def tri_index(t1, t2):
    if t1 > t2:
        t1, t2 = t2, t1
    return t1 + t2 * (t2 + 1) // 2


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox(c, k, o)
    return list(o)


def set_terms(inter=True, intra=True):
    lib().orc_set_terms(1 if inter else 0, 1 if intra else 0)


def term_eval(t1, t2, r, w=VINA_WEIGHTS):
    wa = np.array(w, dtype=np.float64)
    return lib().orc_term_eval(t1, t2, r, _d(wa))


class Lut:
    def __init__(self, w=VINA_WEIGHTS):
        self._w = np.array(w, dtype=np.float64)
        self.h = lib().orc_lut_create(_d(self._w))
        self.n = lib().orc_lut_size()

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_lut_destroy(self.h)

    def tables(self, tp):
        se, sd, fa = (np.zeros(self.n) for _ in range(3))
        assert lib().orc_lut_get(self.h, tp, _d(se), _d(sd), _d(fa)) == 0
        return se, sd, fa

    def eval_fast(self, tp, r2):
        return lib().orc_lut_eval_fast(self.h, tp, r2)

    def eval_deriv(self, tp, r2):
        e, d = C.c_double(), C.c_double()
        lib().orc_lut_eval_deriv(self.h, tp, r2, C.byref(e), C.byref(d))
        return e.value, d.value


def parse_receptor(text, max_atoms=200000):
    xyz = np.zeros((max_atoms, 3))
    xs = np.zeros(max_atoms, dtype=np.int32)
    ad = np.zeros(max_atoms, dtype=np.int32)
    n = lib().orc_receptor_parse(text.encode(), max_atoms, _d(xyz), _i(xs), _i(ad))
    if n < 0:
        raise ValueError(_err())
    return xyz[:n].copy(), xs[:n].copy(), ad[:n].copy()


def grid_dims(center, size, granularity=0.375):
    """grid_dims construction of V11/main/main.cpp:864-876 -> (begin, end, n)."""
    import math
    begin, end, n = [], [], []
    for c, s in zip(center, size):
        ni = int(math.ceil(s / granularity))
        span = granularity * ni
        b = c - span / 2
        begin.append(b)
        end.append(b + span)
        n.append(ni)
    return begin, end, n


class Grid:
    def __init__(self, lut, begin, end, n):
        self.lut = lut
        self.begin = np.array(begin, dtype=np.float64)
        self.end = np.array(end, dtype=np.float64)
        self.n = np.array(n, dtype=np.int32)
        self.h = lib().orc_grid_create(lut.h, _d(self.begin), _d(self.end), _i(self.n))
        self.dims = tuple(int(x) + 1 for x in n)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_grid_destroy(self.h)

    def set_receptor(self, xyz, xs):
        self._xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self._xs = np.ascontiguousarray(xs, dtype=np.int32)
        lib().orc_grid_set_receptor(self.h, len(self._xs), _d(self._xyz), _i(self._xs))

    def populate(self, use_cells=True):
        lib().orc_grid_populate(self.h, 1 if use_cells else 0)

    def populate_with_slack(self, eps=0.01, use_cells=True):
        nx, ny, nz = self.dims
        slack = np.zeros((N_XS, nz, ny, nx))
        lib().orc_grid_populate_slack(self.h, 1 if use_cells else 0, eps, _d(slack))
        return slack

    def maps(self):
        nx, ny, nz = self.dims
        out = np.zeros((N_XS, nz, ny, nx))
        lib().orc_grid_get_maps(self.h, _d(out))
        return out

    def set_maps(self, maps):
        m = np.ascontiguousarray(maps, dtype=np.float64)
        lib().orc_grid_set_maps(self.h, _d(m))

    def cell_candidates(self, p):
        pa = np.array(p, dtype=np.float64)
        out = np.zeros(200000, dtype=np.int32)
        k = lib().orc_grid_cell_candidates(self.h, _d(pa), _i(out), len(out))
        return out[:k].copy()

    def eval(self, t, xyz, slope=1e6, v=10.0):
        p = np.array(xyz, dtype=np.float64)
        d = np.zeros(3)
        e = lib().orc_grid_eval(self.h, t, _d(p), slope, v, _d(d))
        return e, d


class Ligand:
    def __init__(self, text):
        self.h = lib().orc_ligand_parse(text.encode())
        if not self.h:
            raise ValueError(_err())
        s = np.zeros(5, dtype=np.int32)
        lib().orc_ligand_sizes(self.h, _i(s))
        self.n_atoms, self.n_nodes, self.n_pairs, self.n_heavy, self.torsdof = (int(x) for x in s)
        self.T = self.n_nodes - 1

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ligand_destroy(self.h)

    def atoms(self):
        n = self.n_atoms
        xs, el, ad, node = (np.zeros(n, dtype=np.int32) for _ in range(4))
        local, c0 = np.zeros((n, 3)), np.zeros((n, 3))
        lib().orc_ligand_atoms(self.h, _i(xs), _i(el), _i(ad), _i(node), _d(local), _d(c0))
        return dict(xs=xs, el=el, ad=ad, node=node, local=local, coords0=c0)

    def nodes(self):
        n = self.n_nodes
        parent, begin, end = (np.zeros(n, dtype=np.int32) for _ in range(3))
        ro, ra, o0 = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        lib().orc_ligand_nodes(self.h, _i(parent), _i(begin), _i(end), _d(ro), _d(ra), _d(o0))
        return dict(parent=parent, begin=begin, end=end, rel_origin=ro, rel_axis=ra, origin0=o0)

    def pairs(self):
        n = self.n_pairs
        tp, a, b = (np.zeros(max(n, 1), dtype=np.int32) for _ in range(3))
        lib().orc_ligand_pairs(self.h, _i(tp), _i(a), _i(b))
        return np.stack([tp[:n], a[:n], b[:n]], axis=1)

    def bonds(self):
        bi = np.zeros(4096, dtype=np.int32)
        bj = np.zeros(4096, dtype=np.int32)
        k = lib().orc_ligand_bonds(self.h, _i(bi), _i(bj), 4096)
        return list(zip(bi[:k].tolist(), bj[:k].tolist()))

    def serials(self):
        out = np.zeros(self.n_atoms, dtype=np.int32)
        lib().orc_ligand_serials(self.h, _i(out))
        return out

    def num_tors(self):
        return lib().orc_num_tors(self.h)

    def conf_independent(self, e, w_rot=VINA_WEIGHTS_ROT):
        return lib().orc_conf_independent(self.h, e, w_rot)

    def rmsd_bounds(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        lb, ub = C.c_double(), C.c_double()
        lib().orc_rmsd_bounds(self.h, _d(x), _d(y), C.byref(lb), C.byref(ub))
        return lb.value, ub.value

    def identity_conf(self):
        c = np.zeros(7 + self.T)
        c[:3] = self.nodes()["origin0"][0]
        c[3] = 1.0
        return c

    def set_conf(self, conf):
        c = np.ascontiguousarray(conf, dtype=np.float64)
        out = np.zeros((self.n_atoms, 3))
        lib().orc_set_conf(self.h, _d(c), _d(out))
        return out

    def increment(self, conf, d, factor):
        c = np.array(conf, dtype=np.float64)
        dd = np.ascontiguousarray(d, dtype=np.float64)
        lib().orc_conf_increment(self.h, _d(c), _d(dd), factor)
        return c


def eval_deriv(grid, lig, conf, v_intra=10.0, v_inter=10.0, slope=1e6):
    c = np.ascontiguousarray(conf, dtype=np.float64)
    g = np.zeros(6 + lig.T)
    coords = np.zeros((lig.n_atoms, 3))
    e_inter = C.c_double()
    e = lib().orc_eval_deriv(grid.h, lig.h, _d(c), v_intra, v_inter, slope, _d(g), _d(coords), C.byref(e_inter))
    return e, g, coords, e_inter.value


def local_search(grid, lig, conf, max_steps, mode=1, v_intra=10.0, v_inter=10.0, slope=1e6):
    c = np.array(conf, dtype=np.float64)
    ne = C.c_int()
    e = lib().orc_local_search(grid.h, lig.h, _d(c), max_steps, mode, v_intra, v_inter, slope, C.byref(ne))
    return e, c, ne.value


def opt_test(kind, x0, coef=None, max_steps=100, mode=1):
    x = np.array(x0, dtype=np.float64)
    cf = np.ascontiguousarray(coef if coef is not None else np.ones(len(x)), dtype=np.float64)
    ne = C.c_int()
    f = lib().orc_opt_test(kind, len(x), _d(x), _d(cf), max_steps, mode, C.byref(ne))
    return f, x, ne.value


def opt_trace(kind, x0, coef=None, max_steps=100, mode=1, max_pts=256):
    """opt_test, also returning every evaluated point (rows) and its value, in evaluation order."""
    x = np.array(x0, dtype=np.float64)
    cf = np.ascontiguousarray(coef if coef is not None else np.ones(len(x)), dtype=np.float64)
    pts = np.zeros((max_pts, len(x)))
    fs = np.zeros(max_pts)
    ne = C.c_int()
    f = lib().orc_opt_trace(kind, len(x), _d(x), _d(cf), max_steps, mode, max_pts, _d(pts), _d(fs), C.byref(ne))
    k = min(ne.value, max_pts)
    return f, x, pts[:k], fs[:k]


class MCParams(C.Structure):
    _fields_ = [
        ("chains", C.c_int), ("search_depth", C.c_int), ("bfgs_steps", C.c_int), ("mode", C.c_int),
        ("mutation_quirk", C.c_int), ("seed", C.c_uint32), ("stream", C.c_uint32),
        ("box_lo", C.c_double * 3), ("box_hi", C.c_double * 3),
        ("v_intra", C.c_double), ("v_inter", C.c_double), ("slope", C.c_double),
        ("temperature", C.c_double), ("amplitude", C.c_double), ("threads", C.c_int), ("policy", C.c_int),
    ]


def mc_params(chains, depth, bfgs_steps, box_lo, box_hi, seed=42, stream=0, mode=1, quirk=0, threads=0, policy=0):
    """policy: 0 Vina, 1 QuickVina2, 2 QuickVina-W."""
    p = MCParams()
    p.chains, p.search_depth, p.bfgs_steps, p.mode, p.mutation_quirk = chains, depth, bfgs_steps, mode, quirk
    p.seed, p.stream = seed, stream
    for i in range(3):
        p.box_lo[i] = box_lo[i]
        p.box_hi[i] = box_hi[i]
    p.v_intra, p.v_inter, p.slope, p.temperature, p.amplitude = 10.0, 10.0, 1e6, 1.2, 2.0
    p.threads = threads
    p.policy = policy
    return p


def mc_run(grid, lig, params, starts=None):
    """Monte Carlo of params.chains chains -> (e, conf, heavy, evaluations); starts[chains][7+T]
    gives the starting conformations (else drawn from the shared random stream)."""
    n = params.chains
    e = np.zeros(n)
    conf = np.zeros((n, 7 + lig.T))
    heavy = np.zeros((n, lig.n_heavy, 3))
    ev = C.c_longlong()
    if starts is not None:
        st = np.ascontiguousarray(starts, dtype=np.float64)
        assert st.shape == (n, 7 + lig.T)
        rc = lib().orc_mc_run_from(grid.h, lig.h, C.byref(params), _d(st), _d(e), _d(conf), _d(heavy), C.byref(ev))
    else:
        rc = lib().orc_mc_run(grid.h, lig.h, C.byref(params), _d(e), _d(conf), _d(heavy), C.byref(ev))
    if rc != 0:
        raise RuntimeError(_err())
    return e, conf, heavy, ev.value


def mc_trace(grid, lig, params, chain):
    """One chain's MC steps: dict of which, u, e_prev, e_cand, accept, rg (per step) and the
    conformations before and after each mutation (rows of 7+T)."""
    d, nc = params.search_depth, 7 + lig.T
    which, accept = np.zeros(d, dtype=np.int32), np.zeros(d, dtype=np.int32)
    u, e_prev, e_cand, rg = (np.zeros(d) for _ in range(4))
    before, mutated = np.zeros((d, nc)), np.zeros((d, nc))
    rc = lib().orc_mc_trace(grid.h, lig.h, C.byref(params), chain, _i(which), _d(u), _d(e_prev), _d(e_cand),
                            _i(accept), _d(rg), _d(before), _d(mutated))
    if rc != 0:
        raise RuntimeError(_err())
    return dict(which=which, u=u, e_prev=e_prev, e_cand=e_cand, accept=accept.astype(bool), rg=rg, before=before,
                mutated=mutated)


def initial_conf(lig, params, chain):
    c = np.zeros(7 + lig.T)
    lib().orc_initial_conf(lig.h, C.byref(params), chain, _d(c))
    return c


def cluster(e, heavy, min_rmsd=1.0, max_size=20):
    e = np.ascontiguousarray(e, dtype=np.float64)
    h = np.ascontiguousarray(heavy, dtype=np.float64)
    kept = np.zeros(max_size, dtype=np.int32)
    k = lib().orc_cluster(len(e), h.shape[1], _d(e), _d(h), min_rmsd, max_size, _i(kept))
    return kept[:k].copy()


def visited_feed(T, x, f, g, mode=1):
    """Gate answers for points fed in order into one QuickVina2 history (mode 1: add the
    interesting ones, mode 2: add all)."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 7 + T)
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    g = np.ascontiguousarray(g, dtype=np.float64).reshape(-1, 6 + T)
    out = np.zeros(len(f), dtype=np.int32)
    if lib().orc_visited_feed(T, len(f), _d(x), _d(f), _d(g), mode, _i(out)) != 0:
        raise RuntimeError(_err())
    return out.astype(bool)


# --- after the search: refinement against the receptor atoms and rescoring ----------------

AUTHENTIC_V = 1000.0


def noncache_eval(grid, lig, coords, v=AUTHENTIC_V, slope=1e6, deriv=False):
    """non_cache::eval (fast table values) or, with deriv, ::eval_deriv -> (e, minus_forces)."""
    x = np.ascontiguousarray(coords, dtype=np.float64)
    if not deriv:
        return lib().orc_noncache_eval(grid.h, lig.h, _d(x), v, slope, None)
    f = np.zeros((lig.n_atoms, 3))
    return lib().orc_noncache_eval(grid.h, lig.h, _d(x), v, slope, _d(f)), f


def noncache_eval_deriv(grid, lig, conf, v=AUTHENTIC_V, slope=1e6):
    """model::eval_deriv over non_cache: inter on the receptor atoms + intra, both capped at v."""
    c = np.ascontiguousarray(conf, dtype=np.float64)
    g = np.zeros(6 + lig.T)
    return lib().orc_noncache_eval_deriv(grid.h, lig.h, _d(c), v, slope, _d(g)), g


def intra_fast(grid, lig, coords, v=AUTHENTIC_V):
    x = np.ascontiguousarray(coords, dtype=np.float64)
    return lib().orc_intra_fast(grid.h, lig.h, _d(x), v)


def within(grid, lig, coords):
    x = np.ascontiguousarray(coords, dtype=np.float64)
    return bool(lib().orc_within(grid.h, lig.h, _d(x)))


def refine(grid, lig, conf, max_steps, v=AUTHENTIC_V):
    """refine_structure: returns (e, refined conf, all-atom coordinates)."""
    c = np.array(conf, dtype=np.float64)
    xyz = np.zeros((lig.n_atoms, 3))
    e = lib().orc_refine(grid.h, lig.h, _d(c), max_steps, v, _d(xyz))
    return e, c, xyz


def refine_steps(lig, rilc_bfgs=True):
    """BFGS step budget of the refinement (V11/main/main.cpp:383-391)."""
    return ((30 if rilc_bfgs else 25) + lig.n_atoms) // 3


def postprocess(grid, lig, confs, max_steps, w_rot=VINA_WEIGHTS_ROT, num_modes=9, energy_range=3.0,
                min_rmsd=1.0, slope=1e6):
    """Per-ligand post-processing of V11/main/main.cpp:425-470 -> list of modes
    (conf, score, rmsd lb, rmsd ub, refined energy)."""
    confs = np.ascontiguousarray(confs, dtype=np.float64).reshape(-1, 7 + lig.T)
    n = len(confs)
    co = np.zeros_like(confs)
    e, lb, ub, rf = (np.zeros(max(n, 1)) for _ in range(4))
    k = lib().orc_postprocess(grid.h, lig.h, n, _d(confs), max_steps, w_rot, num_modes, energy_range,
                              min_rmsd, slope, _d(co), _d(e), _d(lb), _d(ub), _d(rf))
    return [dict(conf=co[i].copy(), e=e[i], lb=lb[i], ub=ub[i], refined=rf[i]) for i in range(k)]


def qvw_feed(T, ops, box_lo, box_hi, levels):
    """Runs QuickVina-W gate operations: ops = [(op, chain, x, f, g)] with op 0 gate query,
    1 individual add, 2 global add, 3 G-check alone; returns the per-op outputs."""
    m = len(ops)
    x = np.ascontiguousarray([o[2] for o in ops], dtype=np.float64).reshape(m, 7 + T)
    f = np.ascontiguousarray([o[3] for o in ops], dtype=np.float64)
    g = np.ascontiguousarray([o[4] for o in ops], dtype=np.float64).reshape(m, 6 + T)
    ch = np.ascontiguousarray([o[1] for o in ops], dtype=np.int32)
    op = np.ascontiguousarray([o[0] for o in ops], dtype=np.int32)
    lo = np.ascontiguousarray(box_lo, dtype=np.float64)
    hi = np.ascontiguousarray(box_hi, dtype=np.float64)
    out = np.zeros(max(m, 1), dtype=np.int32)
    if lib().orc_qvw_feed(T, m, _d(x), _d(f), _d(g), _i(ch), _i(op), _d(lo), _d(hi), levels, _i(out)) != 0:
        raise RuntimeError(_err())
    return out[:m].tolist()


def qvw_levels(chains, depth):
    return lib().orc_qvw_levels(chains, depth)
