"""vdock: Python front end of libvdock (include/vdock.h), the MI355X docking hot path.

The compute runs in libvdock.so (HIP kernels for gfx950 behind a C ABI).  This module only
marshals arrays through ctypes; it has no fallback: if the library is missing or no GPU is
visible, the calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# VDOCK_LIBRARY selects another build of the same library (the sanitizer build of the host code)
LIB_PATH = os.environ.get("VDOCK_LIBRARY") or os.path.join(HERE, "libvdock.so")

VD_OK, VD_ERR_ARG, VD_ERR_CAPACITY, VD_ERR_HIP, VD_ERR_PARSE, VD_ERR_NO_DEVICE = 0, -1, -2, -3, -4, -5
VD_LS_BFGS, VD_LS_RILC_BFGS = 0, 1
VD_MC_VINA, VD_MC_QVINA2, VD_MC_QVINAW = 0, 1, 2
N_XS = 17
N_TYPE_PAIRS = 153
LUT_BINS = 2051
DEFAULT_WEIGHTS = (-0.035579, -0.005156, 0.840245, -0.035069, -0.587439, 0.05846)

# Every symbol the header declares (checked by the CPU test suite).
EXPORTS = (
    "vd_version", "vd_create", "vd_destroy", "vd_last_error", "vd_set_weights", "vd_get_lut",
    "vd_set_receptor", "vd_set_receptor_pdbqt", "vd_get_receptor", "vd_box_from_center",
    "vd_grid_populate", "vd_grid_last_ms", "vd_grid_download", "vd_grid_upload",
    "vd_ligand_add_pdbqt", "vd_ligand_info_get", "vd_ligand_export", "vd_ligand_count",
    "vd_eval", "vd_eval_last_ms", "vd_dock_params_default", "vd_dock_heuristics", "vd_dock",
    "vd_dock_batch", "vd_dock_batch_plain", "vd_batch_results", "vd_batch_best", "vd_cluster",
    "vd_synth_receptor", "vd_synth_ligand", "vd_device_alloc", "vd_device_free", "vd_copy", "vd_random_confs",
    "vd_last_launches", "vd_set_class_overlap", "vd_set_graphs", "vd_set_autotune", "vd_ligand_pair_chunks",
    "vd_post_params_default", "vd_refine", "vd_select_modes", "vd_postprocess_batch", "vd_postprocess_docked",
    "vd_post_results", "vd_write_pdbqt", "vd_local_search", "vd_dock_noalloc",
)


class VdockError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vdock error {code}: {msg}")
        self.code = code


class Box(C.Structure):
    _fields_ = [("begin", C.c_double * 3), ("end", C.c_double * 3), ("n", C.c_int * 3)]

    @property
    def dims(self):
        return tuple(int(self.n[i]) + 1 for i in range(3))

    def as_tuple(self):
        return list(self.begin), list(self.end), list(self.n)


class LigandInfo(C.Structure):
    _fields_ = [("n_atoms", C.c_int), ("n_torsions", C.c_int), ("n_pairs", C.c_int), ("n_heavy", C.c_int),
                ("torsdof", C.c_int), ("tree_depth", C.c_int), ("num_tors", C.c_float)]


class DockParams(C.Structure):
    _fields_ = [("chains", C.c_int), ("search_depth", C.c_int), ("bfgs_steps", C.c_int), ("local_search", C.c_int),
                ("policy", C.c_int), ("mutation_quirk", C.c_int), ("chains_per_wave", C.c_int), ("waves_per_simd", C.c_int),
                ("seed", C.c_uint32),
                ("temperature", C.c_float), ("amplitude", C.c_float), ("v_intra", C.c_float), ("v_inter", C.c_float),
                ("slope", C.c_float)]


class Stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("evals", C.c_ulonglong), ("chains_per_wave", C.c_int), ("blocks", C.c_int)]


class PostParams(C.Structure):
    _fields_ = [("max_steps", C.c_int), ("local_search", C.c_int), ("num_modes", C.c_int), ("energy_range", C.c_float),
                ("min_rmsd", C.c_float), ("v", C.c_float), ("slope", C.c_float)]


class Refined(C.Structure):
    _fields_ = [("e", C.c_float), ("e_inter", C.c_float), ("e_intra", C.c_float), ("inside", C.c_int)]


class Mode(C.Structure):
    _fields_ = [("e", C.c_float), ("rmsd_lb", C.c_float), ("rmsd_ub", C.c_float), ("pose", C.c_int)]


class Launch(C.Structure):
    _fields_ = [("lanes_per_chain", C.c_int), ("mode", C.c_int), ("waves_per_simd", C.c_int), ("blocks", C.c_int),
                ("items", C.c_int), ("ligands", C.c_int), ("persistent", C.c_int), ("kernel_ms", C.c_double),
                ("evals", C.c_ulonglong), ("eval_bytes", C.c_double), ("start_ms", C.c_double)]

    @property
    def kernel(self):
        """Name of the launched kernel as a profiler lists it (template arguments)."""
        return "vd_mc_kernel<%d, %d, %d>" % (self.lanes_per_chain, self.mode, self.waves_per_simd)


_lib = None


def load():
    """Loads libvdock.so (built by __graft_entry__.build()); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libvdock.so not found at {LIB_PATH}; build it with __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    vp, dp, fp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_int)
    sig = {
        "vd_version": (C.c_char_p, []),
        "vd_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
        "vd_destroy": (None, [vp]),
        "vd_last_error": (C.c_char_p, [vp]),
        "vd_set_weights": (C.c_int, [vp, dp]),
        "vd_get_lut": (C.c_int, [vp, C.c_int, dp, dp, dp]),
        "vd_set_receptor": (C.c_int, [vp, C.c_int, fp, ip]),
        "vd_set_receptor_pdbqt": (C.c_int, [vp, C.c_char_p, ip]),
        "vd_get_receptor": (C.c_int, [vp, C.c_int, fp, ip]),
        "vd_box_from_center": (C.c_int, [dp, dp, C.c_double, C.POINTER(Box)]),
        "vd_grid_populate": (C.c_int, [vp, C.POINTER(Box)]),
        "vd_grid_last_ms": (C.c_double, [vp]),
        "vd_grid_download": (C.c_int, [vp, fp]),
        "vd_grid_upload": (C.c_int, [vp, C.POINTER(Box), fp]),
        "vd_ligand_add_pdbqt": (C.c_int, [vp, C.c_char_p, ip]),
        "vd_ligand_info_get": (C.c_int, [vp, C.c_int, C.POINTER(LigandInfo)]),
        "vd_ligand_export": (C.c_int, [vp, C.c_int, ip, ip, ip, fp, ip, ip, ip, fp, fp, ip]),
        "vd_ligand_count": (C.c_int, [vp]),
        "vd_eval": (C.c_int, [vp, C.c_int, C.c_longlong, vp, vp, vp, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int]),
        "vd_eval_last_ms": (C.c_double, [vp]),
        "vd_dock_params_default": (C.c_int, [C.POINTER(DockParams)]),
        "vd_dock_heuristics": (C.c_int, [C.c_int, C.c_int, C.c_int, ip, ip]),
        "vd_dock": (C.c_int, [vp, C.c_int, C.POINTER(DockParams), fp, fp, fp, fp, C.POINTER(Stats)]),
        "vd_local_search": (C.c_int, [vp, C.c_int, C.c_int, fp, C.POINTER(DockParams), fp, fp, C.POINTER(Stats)]),
        "vd_dock_noalloc": (C.c_int, [vp, C.c_int, C.POINTER(DockParams), vp, vp, vp, vp, C.POINTER(Stats)]),
        "vd_dock_batch": (C.c_int, [vp, C.c_int, ip, C.POINTER(DockParams), C.POINTER(Stats)]),
        "vd_dock_batch_plain": (C.c_int, [vp, C.c_int, ip, C.POINTER(DockParams), C.POINTER(Stats)]),
        "vd_batch_results": (C.c_int, [vp, C.c_int, C.c_int, fp, fp, fp]),
        "vd_batch_best": (C.c_int, [vp, C.c_int, fp]),
        "vd_last_launches": (C.c_int, [vp, C.c_int, C.POINTER(Launch)]),
        "vd_set_class_overlap": (C.c_int, [vp, C.c_int]),
        "vd_set_autotune": (C.c_int, [vp, C.c_int]),
        "vd_set_graphs": (C.c_int, [vp, C.c_int]),
        "vd_ligand_pair_chunks": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint32)]),
        "vd_cluster": (C.c_int, [C.c_int, C.c_int, fp, fp, C.c_double, C.c_int, ip]),
        "vd_device_alloc": (C.c_int, [vp, C.c_size_t, C.POINTER(C.c_void_p)]),
        "vd_device_free": (C.c_int, [vp, vp]),
        "vd_copy": (C.c_int, [vp, vp, vp, C.c_size_t, C.c_int]),
        "vd_random_confs": (C.c_int, [vp, C.c_int, C.c_longlong, C.c_uint32, vp]),
        "vd_synth_receptor": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_uint64, fp, ip]),
        "vd_synth_ligand": (C.c_int, [C.c_int, C.c_int, C.c_uint64, C.c_char_p, C.c_int]),
        "vd_post_params_default": (C.c_int, [C.POINTER(PostParams)]),
        "vd_refine": (C.c_int, [vp, C.c_int, C.c_int, fp, C.POINTER(PostParams), fp, fp, C.POINTER(Refined)]),
        "vd_select_modes": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(Refined), fp, C.POINTER(PostParams),
                                      C.POINTER(Mode)]),
        "vd_postprocess_batch": (C.c_int, [vp, C.c_int, ip, ip, fp, C.POINTER(PostParams), C.POINTER(Stats)]),
        "vd_postprocess_docked": (C.c_int, [vp, C.c_int, C.POINTER(PostParams), C.POINTER(Stats)]),
        "vd_post_results": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(Mode), fp, fp]),
        "vd_write_pdbqt": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(Mode), fp, C.c_char_p, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def default_params(**kw):
    p = DockParams()
    load().vd_dock_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def heuristics(n_atoms, n_torsions, rilc=True):
    d, b = C.c_int(), C.c_int()
    load().vd_dock_heuristics(n_atoms, n_torsions, 1 if rilc else 0, C.byref(d), C.byref(b))
    return d.value, b.value


def box_from_center(center, size, granularity=0.375):
    b = Box()
    c = np.array(center, dtype=np.float64)
    s = np.array(size, dtype=np.float64)
    rc = load().vd_box_from_center(_d(c), _d(s), granularity, C.byref(b))
    if rc != VD_OK:
        raise VdockError(rc, "bad box")
    return b


def synth_receptor(n_atoms, cube, cavity_radius, seed):
    xyz = np.zeros((n_atoms, 3), dtype=np.float32)
    xs = np.zeros(n_atoms, dtype=np.int32)
    n = load().vd_synth_receptor(n_atoms, cube, cavity_radius, seed, _f(xyz), _i(xs))
    if n < 0:
        raise VdockError(n, "synth_receptor")
    return xyz[:n].copy(), xs[:n].copy()


def synth_ligand(n_heavy, n_rot, seed):
    need = load().vd_synth_ligand(n_heavy, n_rot, seed, None, 0)
    buf = C.create_string_buffer(need + 1)
    load().vd_synth_ligand(n_heavy, n_rot, seed, buf, need + 1)
    return buf.value.decode()


def cluster(e, heavy, min_rmsd=1.0, max_size=20):
    e = np.ascontiguousarray(e, dtype=np.float32)
    h = np.ascontiguousarray(heavy, dtype=np.float32)
    kept = np.zeros(max_size, dtype=np.int32)
    k = load().vd_cluster(len(e), h.shape[1] if h.ndim == 3 else 0, _f(e), _f(h), min_rmsd, max_size, _i(kept))
    if k < 0:
        raise VdockError(k, "cluster")
    return kept[:k].copy()


class Context:
    """One HIP device context (one stream).  Mirrors the C ABI call by call."""

    def __init__(self, device=0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.vd_create(device, C.byref(h))
        if rc != VD_OK:
            raise VdockError(rc, "vd_create failed (no usable GPU?)")
        self.h = h
        self.box = None

    def close(self):
        if getattr(self, "h", None):
            self.lib.vd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != VD_OK:
            raise VdockError(rc, self.lib.vd_last_error(self.h).decode())

    # -- scoring function
    def set_weights(self, w):
        a = np.array(w, dtype=np.float64)
        self._check(self.lib.vd_set_weights(self.h, _d(a)))

    def lut(self, tp):
        se, sd, fa = (np.zeros(LUT_BINS) for _ in range(3))
        self._check(self.lib.vd_get_lut(self.h, tp, _d(se), _d(sd), _d(fa)))
        return se, sd, fa

    # -- receptor and maps
    def set_receptor(self, xyz, xs):
        x = np.ascontiguousarray(xyz, dtype=np.float32)
        t = np.ascontiguousarray(xs, dtype=np.int32)
        self._check(self.lib.vd_set_receptor(self.h, len(t), _f(x), _i(t)))

    def set_receptor_pdbqt(self, text):
        n = C.c_int()
        self._check(self.lib.vd_set_receptor_pdbqt(self.h, text.encode(), C.byref(n)))
        return n.value

    def receptor(self):
        n = self.lib.vd_get_receptor(self.h, 0, None, None)
        xyz = np.zeros((max(n, 1), 3), dtype=np.float32)
        xs = np.zeros(max(n, 1), dtype=np.int32)
        self.lib.vd_get_receptor(self.h, n, _f(xyz), _i(xs))
        return xyz[:n], xs[:n]

    def grid_populate(self, box):
        self._check(self.lib.vd_grid_populate(self.h, C.byref(box)))
        self.box = box
        return self.lib.vd_grid_last_ms(self.h)

    def grid_download(self):
        nx, ny, nz = self.box.dims
        out = np.zeros((N_XS, nz, ny, nx), dtype=np.float32)
        self._check(self.lib.vd_grid_download(self.h, _f(out)))
        return out

    def grid_upload(self, box, maps):
        m = np.ascontiguousarray(maps, dtype=np.float32)
        self._check(self.lib.vd_grid_upload(self.h, C.byref(box), _f(m)))
        self.box = box

    # -- ligands
    def add_ligand(self, text):
        i = C.c_int()
        self._check(self.lib.vd_ligand_add_pdbqt(self.h, text.encode(), C.byref(i)))
        return i.value

    def ligand_info(self, lid):
        info = LigandInfo()
        self._check(self.lib.vd_ligand_info_get(self.h, lid, C.byref(info)))
        return info

    def ligand_export(self, lid):
        inf = self.ligand_info(lid)
        A, N, P = inf.n_atoms, inf.n_torsions + 1, inf.n_pairs
        xs, el, node = (np.zeros(A, dtype=np.int32) for _ in range(3))
        local = np.zeros((A, 3), dtype=np.float32)
        parent, begin, end = (np.zeros(N, dtype=np.int32) for _ in range(3))
        ro, ra = np.zeros((N, 3), dtype=np.float32), np.zeros((N, 3), dtype=np.float32)
        pairs = np.zeros((max(P, 1), 3), dtype=np.int32)
        self._check(self.lib.vd_ligand_export(self.h, lid, _i(xs), _i(el), _i(node), _f(local), _i(parent), _i(begin),
                                              _i(end), _f(ro), _f(ra), _i(pairs)))
        return dict(xs=xs, el=el, node=node, local=local, parent=parent, begin=begin, end=end, rel_origin=ro,
                    rel_axis=ra, pairs=pairs[:P])

    # -- energy and gradient
    def eval(self, lid, conf, v_intra=10.0, v_inter=10.0, slope=1e6, chains_per_wave=0):
        conf = np.ascontiguousarray(conf, dtype=np.float32)
        n, m = conf.shape
        T = m - 7
        e = np.zeros(n, dtype=np.float32)
        g = np.zeros((n, 6 + T), dtype=np.float32)
        self._check(self.lib.vd_eval(self.h, lid, n, conf.ctypes.data, e.ctypes.data, g.ctypes.data, v_intra, v_inter,
                                     slope, chains_per_wave, 0))
        return e, g

    def eval_device(self, lid, conf_ptr, e_ptr, g_ptr, n, v_intra=10.0, v_inter=10.0, slope=1e6, chains_per_wave=0):
        """Device-pointer variant: pointers are integers (e.g. torch tensor .data_ptr())."""
        self._check(self.lib.vd_eval(self.h, lid, n, conf_ptr, e_ptr, g_ptr, v_intra, v_inter, slope,
                                     chains_per_wave, 1))
        return self.lib.vd_eval_last_ms(self.h)

    # -- device memory (inputs resident in HBM)
    def alloc(self, nbytes):
        ptr = C.c_void_p()
        self._check(self.lib.vd_device_alloc(self.h, nbytes, C.byref(ptr)))
        return ptr.value

    def free(self, ptr):
        self._check(self.lib.vd_device_free(self.h, ptr))

    def copy_device(self, dst, src, nbytes):
        """Device-to-device copy of nbytes (synchronous)."""
        self._check(self.lib.vd_copy(self.h, dst, src, nbytes, 2))

    def to_host(self, ptr, shape, dtype=np.float32):
        out = np.empty(shape, dtype=dtype)
        self._check(self.lib.vd_copy(self.h, out.ctypes.data, ptr, out.nbytes, 1))
        return out

    def random_confs(self, lid, n, seed, ptr):
        self._check(self.lib.vd_random_confs(self.h, lid, n, seed, ptr))

    def eval_last_ms(self):
        return self.lib.vd_eval_last_ms(self.h)

    # -- docking
    def dock(self, lid, params, init_conf=None):
        """Plain per-ligand launch -> (e, conf, heavy, stats); init_conf [chains][7+T] gives each
        chain's starting conformation (else drawn from its random stream)."""
        inf = self.ligand_info(lid)
        n = params.chains
        e = np.zeros(n, dtype=np.float32)
        conf = np.zeros((n, 7 + inf.n_torsions), dtype=np.float32)
        heavy = np.zeros((n, inf.n_heavy, 3), dtype=np.float32)
        st = Stats()
        init = None
        if init_conf is not None:
            init = np.ascontiguousarray(init_conf, dtype=np.float32)
            if init.shape != (n, 7 + inf.n_torsions):
                raise ValueError("init_conf must have shape (chains, 7 + T)")
        self._check(self.lib.vd_dock(self.h, lid, C.byref(params), _f(init) if init is not None else None, _f(e), _f(conf),
                                     _f(heavy), C.byref(st)))
        return e, conf, heavy, st

    def local_search(self, lid, conf, params):
        """One local search (params.local_search, params.bfgs_steps) from every row of conf
        -> (e, conf_out, stats)."""
        inf = self.ligand_info(lid)
        c = np.ascontiguousarray(conf, dtype=np.float32).reshape(-1, 7 + inf.n_torsions)
        e = np.zeros(len(c), dtype=np.float32)
        out = np.zeros_like(c)
        st = Stats()
        self._check(self.lib.vd_local_search(self.h, lid, len(c), _f(c), C.byref(params), _f(out), _f(e), C.byref(st)))
        return e, out, st

    def dock_noalloc(self, lid, params, d_e, d_conf, d_heavy, d_init=None):
        """The no-allocation docking call on device pointers (integers) the caller owns."""
        st = Stats()
        self._check(self.lib.vd_dock_noalloc(self.h, lid, C.byref(params), d_init, d_e, d_conf, d_heavy, C.byref(st)))
        return st

    def dock_batch(self, ids, params, plain=False):
        a = np.ascontiguousarray(ids, dtype=np.int32)
        st = Stats()
        f = self.lib.vd_dock_batch_plain if plain else self.lib.vd_dock_batch
        self._check(f(self.h, len(a), _i(a), C.byref(params), C.byref(st)))
        self._batch_ids = list(a)
        return st

    def batch_results(self, k, params=None):
        """Per-chain results of ligand k of the last batch -> (e, conf, heavy), sized by the chain
        count the ligand was docked with (params is accepted for older callers and ignored)."""
        n = self.lib.vd_batch_results(self.h, k, 0, None, None, None)
        self._check(min(n, 0))
        inf = self.ligand_info(self._batch_ids[k])
        e = np.zeros(n, dtype=np.float32)
        conf = np.zeros((n, 7 + inf.n_torsions), dtype=np.float32)
        heavy = np.zeros((n, inf.n_heavy, 3), dtype=np.float32)
        m = self.lib.vd_batch_results(self.h, k, n, _f(e), _f(conf), _f(heavy))
        self._check(min(m, 0))
        return e, conf, heavy

    def batch_best(self):
        n = self.lib.vd_batch_best(self.h, 0, None)
        self._check(min(n, 0))
        out = np.zeros(n, dtype=np.float32)
        self._check(min(self.lib.vd_batch_best(self.h, n, _f(out)), 0))
        return out

    def pair_chunks(self, lid, lanes):
        """Device pair layout for chain groups of `lanes` lanes: array (chunks, lanes) of pair words."""
        n = self.lib.vd_ligand_pair_chunks(self.h, lid, lanes, 0, None)
        self._check(min(n, 0))
        out = np.zeros(max(n, 1), dtype=np.uint32)
        self._check(min(self.lib.vd_ligand_pair_chunks(self.h, lid, lanes, n, out.ctypes.data_as(C.POINTER(C.c_uint32))), 0))
        return out[:n].reshape(-1, lanes)

    def set_class_overlap(self, on):
        """Run the lanes-per-chain classes of a batch at the same time (True, the default) or one
        after the other (False)."""
        self._check(self.lib.vd_set_class_overlap(self.h, 1 if on else 0))

    def set_graphs(self, on):
        """Run persistent batches as captured hipGraphs, replayed when the same ligands are docked
        again with the same parameters (any seed) (True), or eagerly (False, the default)."""
        self._check(self.lib.vd_set_graphs(self.h, 1 if on else 0))

    def set_autotune(self, on):
        """Time each size bucket's (lanes per chain, waves per SIMD) pair on first use (True) or
        take the fixed per-bucket pair, reproducible across processes (False, the default)."""
        self._check(self.lib.vd_set_autotune(self.h, 1 if on else 0))

    def last_launches(self):
        """MC kernel launches of the last docking call (list of Launch)."""
        n = self.lib.vd_last_launches(self.h, 0, None)
        self._check(min(n, 0))
        arr = (Launch * max(n, 1))()
        self._check(min(self.lib.vd_last_launches(self.h, n, arr), 0))
        return list(arr[:n])

    # -- after the search: refinement, final scores, output
    def refine(self, lid, conf, params=None):
        """refine_structure of every row of conf -> (conf, all-atom coords, list of Refined)."""
        p = params or post_params()
        inf = self.ligand_info(lid)
        c = np.ascontiguousarray(conf, dtype=np.float32).reshape(-1, 7 + inf.n_torsions)
        n = len(c)
        co = np.zeros_like(c)
        xyz = np.zeros((n, inf.n_atoms, 3), dtype=np.float32)
        res = (Refined * max(n, 1))()
        self._check(self.lib.vd_refine(self.h, lid, n, _f(c), C.byref(p), _f(co), _f(xyz), res))
        return co, xyz, list(res[:n])

    def select_modes(self, lid, refined, coords, params=None):
        """Mode selection over refined poses (host only) -> list of Mode."""
        p = params or post_params()
        n = len(refined)
        r = (Refined * max(n, 1))(*refined)
        xyz = np.ascontiguousarray(coords, dtype=np.float32)
        modes = (Mode * max(p.num_modes, 1))()
        k = self.lib.vd_select_modes(self.h, lid, n, r, _f(xyz), C.byref(p), modes)
        self._check(min(k, 0))
        return list(modes[:k])

    def postprocess(self, ids, confs, params=None):
        """Refinement + selection for several ligands in one launch; confs[k] = poses of ids[k]."""
        p = params or post_params()
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        counts = np.array([len(np.atleast_2d(c)) if len(c) else 0 for c in confs], dtype=np.int32)
        flat = np.concatenate([np.asarray(c, dtype=np.float32).reshape(-1) for c in confs]) if len(confs) else \
            np.zeros(1, dtype=np.float32)
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        st = Stats()
        self._check(self.lib.vd_postprocess_batch(self.h, len(ids), _i(ids), _i(counts), _f(flat), C.byref(p), C.byref(st)))
        self._post_ids = list(ids)
        return st

    def postprocess_docked(self, max_poses=20, params=None):
        p = params or post_params()
        st = Stats()
        self._check(self.lib.vd_postprocess_docked(self.h, max_poses, C.byref(p), C.byref(st)))
        self._post_ids = list(self._batch_ids)
        return st

    def post_results(self, k, num_modes=None):
        """Modes of ligand k of the last post-processing call -> (modes, conf, coords): all of them,
        or at most num_modes."""
        n = self.lib.vd_post_results(self.h, k, 0, None, None, None)
        self._check(min(n, 0))
        if num_modes is not None:
            n = min(n, num_modes)
        inf = self.ligand_info(self._post_ids[k])
        modes = (Mode * max(n, 1))()
        conf = np.zeros((n, 7 + inf.n_torsions), dtype=np.float32)
        xyz = np.zeros((n, inf.n_atoms, 3), dtype=np.float32)
        m = self.lib.vd_post_results(self.h, k, n, modes, _f(conf), _f(xyz))
        self._check(min(m, 0))
        return list(modes[:n]), conf, xyz

    def write_pdbqt(self, lid, modes, coords):
        """PDBQT text of the given modes (coords[m][n_atoms][3])."""
        n = len(modes)
        arr = (Mode * max(n, 1))(*modes)
        xyz = np.ascontiguousarray(coords, dtype=np.float32)
        size = self.lib.vd_write_pdbqt(self.h, lid, n, arr, _f(xyz), None, 0)
        self._check(min(size, 0))
        buf = C.create_string_buffer(size + 1)
        self._check(min(self.lib.vd_write_pdbqt(self.h, lid, n, arr, _f(xyz), buf, size + 1), 0))
        return buf.value.decode()


def post_params(**kw):
    p = PostParams()
    load().vd_post_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def version():
    return load().vd_version().decode()
