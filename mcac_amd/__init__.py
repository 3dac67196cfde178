"""MI355X-native MCAC aggregate-move / collision-search path.

Python view of the two native libraries built in-tree by ``mcac_amd.build``:

* ``libmcac_gpu.so`` — HIP kernels behind the C-ABI of ``include/mcac_gpu.h``
  (one aggregation box resident in HBM: contact search, translate, merge).
* ``libmcac_sim.so`` — the host driver of ``include/mcac_sim.h``: MCAC parameter
  file, physical model, glibc random stream, time-step weighted pick and the
  reference MC loop (``src/calcul.cpp``), calling the device for the collision path.

There is no CPU fallback: importing works without a GPU, but every entry point
that needs the libraries raises if they are not built.
"""
from __future__ import annotations

import ctypes
import math
import pathlib

import numpy as np

__all__ = [
    "physics_scalars", "swept_sphere_batch", "neighbour_order", "arvo_part", "LIB_DIR", "load_gpu", "load_sim", "Simulation", "Box", "STEP_DTYPE", "MERGE_DTYPE",
    "CONTACT_DTYPE", "MERGE_RESULT_DTYPE", "GPU_SYMBOLS", "SIM_SYMBOLS",
]

LIB_DIR = pathlib.Path(__file__).resolve().parent / "lib"

STEP_DTYPE = np.dtype([
    ("step", "<i8"), ("agg", "<i8"), ("label", "<i8"), ("dir", "<f8", (3,)), ("lpm", "<f8"),
    ("dist", "<f8"), ("moving_sphere", "<i8"), ("other_sphere", "<i8"), ("other_agg", "<i8"),
    ("contact", "<i4"), ("merged", "<i4"), ("dt", "<f8"), ("time", "<f8"),
])
MERGE_DTYPE = np.dtype([
    ("step", "<i8"), ("keeper", "<i8"), ("removed", "<i8"), ("n", "<i8"),
    ("volume", "<f8"), ("surface", "<f8"), ("rg", "<f8"), ("rmax", "<f8"), ("dp", "<f8"),
    ("f_agg", "<f8"), ("lpm", "<f8"), ("time_step", "<f8"), ("d_m", "<f8"), ("proper_time", "<f8"),
    ("overlapping", "<f8"), ("coordination", "<f8"), ("com", "<f8", (3,)), ("rcm", "<f8", (3,)),
])
CONTACT_DTYPE = np.dtype([
    ("distance", "<f8"), ("move", "<f8"), ("moving_sphere", "<i8"), ("other_sphere", "<i8"),
    ("moving_agg", "<i8"), ("other_agg", "<i8"), ("contact", "<i4"), ("status", "<i4"),
    ("n_query", "<i8"), ("n_filtered", "<i8"), ("n_pairs", "<i8"),
])
MERGE_RESULT_DTYPE = np.dtype([
    ("keeper", "<i8"), ("removed", "<i8"), ("n", "<i8"), ("merged", "<i4"), ("status", "<i4"),
    ("volume", "<f8"), ("surface", "<f8"), ("rg", "<f8"), ("rmax", "<f8"), ("radius_sum", "<f8"),
    ("com", "<f8", (3,)), ("rcm", "<f8", (3,)), ("overlapping", "<f8"), ("coordination", "<f8"),
    ("sphere_volume", "<f8"),
])

# Every symbol declared in include/mcac_gpu.h and include/mcac_sim.h.
GPU_SYMBOLS = (
    "mcac_box_create", "mcac_box_destroy", "mcac_box_upload", "mcac_next_contact", "mcac_translate",
    "mcac_move", "mcac_merge", "mcac_move_enqueue", "mcac_move_result", "mcac_release", "mcac_box_set_auto_merge",
    "mcac_box_download_spheres", "mcac_box_sphere_owners", "mcac_box_download_aggs", "mcac_box_members",
    "mcac_box_set_timing", "mcac_box_timing", "mcac_box_stats", "mcac_box_prune_stats", "mcac_swept_sphere_batch",
    "mcac_version", "mcac_group_create", "mcac_group_destroy", "mcac_group_add", "mcac_move_enqueue_group",
    "mcac_group_stats", "mcac_group_times", "mcac_box_spare_moves", "mcac_box_block_steps", "mcac_neighbour_order", "mcac_arvo_part", "mcac_box_device_bytes",
    "mcac_swept_sphere_batch_noalloc", "mcac_box_set_engine", "mcac_box_engine_stats", "mcac_box_grow",
    "mcac_box_grow_ids", "mcac_device_memory",
    "mcac_box_download_radii", "mcac_box_download_members",
)
SIM_SYMBOLS = (
    "mcac_sim_create", "mcac_sim_destroy", "mcac_sim_init", "mcac_sim_run", "mcac_sim_finished",
    "mcac_sim_set_trace", "mcac_sim_n_steps", "mcac_sim_get_steps", "mcac_sim_n_merges",
    "mcac_sim_get_merges", "mcac_sim_scalar", "mcac_sim_n_aggs", "mcac_sim_export_aggs", "mcac_sim_box",
    "mcac_physics_scalars", "mcac_pick_table_selftest", "mcac_pick_table_selftest_ties", "mcac_batch_run",
    "mcac_write_test_snapshots", "mcac_batch_run_groups", "mcac_sim_device_bytes",
)

_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_I64 = ctypes.c_int64
_DP = ctypes.POINTER(ctypes.c_double)
_I64P = ctypes.POINTER(ctypes.c_int64)

_gpu = None
_sim = None


class MCACError(RuntimeError):
    """A native call returned a non-zero status (reference ErrorCodes or MCAC_E_*)."""


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise MCACError(f"{what} failed with status {rc}")


def _require(path: pathlib.Path) -> None:
    if not path.exists():
        raise RuntimeError(
            f"{path.name} is not built; run `python -m mcac_amd.build` (or __graft_entry__.build()) first")


def load_gpu():
    """Loads libmcac_gpu.so (HIP kernels + C-ABI)."""
    global _gpu
    if _gpu is None:
        path = LIB_DIR / "libmcac_gpu.so"
        _require(path)
        L = ctypes.CDLL(str(path))
        L.mcac_box_create.argtypes = [_P, ctypes.POINTER(_P)]
        L.mcac_box_device_bytes.restype = _I64
        L.mcac_box_device_bytes.argtypes = [_P]
        L.mcac_box_destroy.argtypes = [_P]
        L.mcac_box_upload.argtypes = [_P, _I64] + [_DP] * 9 + [_I64] + [_DP] * 5 + [_I64P, _I64P]
        L.mcac_next_contact.argtypes = [_P, _I64, _DP, _D, _D, _P]
        L.mcac_move.argtypes = [_P, _I64, _DP, _D, _D, _P]
        L.mcac_translate.argtypes = [_P, _I64, _DP]
        L.mcac_merge.argtypes = [_P, _P, _P]
        L.mcac_move_enqueue.argtypes = [_P, _I64, _DP, _D, _D, ctypes.POINTER(ctypes.c_uint64)]
        L.mcac_move_result.argtypes = [_P, ctypes.c_uint64, _P, ctypes.POINTER(_I)]
        L.mcac_release.argtypes = [_P]
        L.mcac_box_set_auto_merge.argtypes = [_P, _I]
        L.mcac_box_download_spheres.argtypes = [_P] + [_DP] * 6
        L.mcac_box_download_aggs.argtypes = [_P] + [_DP] * 6 + [_I64P]
        L.mcac_box_members.argtypes = [_P, _I64, _I64P, _I64, _I64P]
        L.mcac_box_sphere_owners.argtypes = [_P, _I64P]
        L.mcac_box_set_timing.argtypes = [_P, _I]
        L.mcac_box_set_engine.argtypes = [_P, _I]
        L.mcac_box_grow.argtypes = [_P, _D] + [_DP] * 5
        L.mcac_box_download_radii.argtypes = [_P] + [_DP] * 3
        L.mcac_box_download_members.argtypes = [_P, _I64P, _I64P]
        L.mcac_box_engine_stats.argtypes = [_P, _DP]
        L.mcac_box_timing.argtypes = [_P, _DP]
        L.mcac_box_stats.argtypes = [_P, _I64P]
        L.mcac_box_prune_stats.argtypes = [_P, _I64P]
        L.mcac_swept_sphere_batch.argtypes = [_I64, _DP, _DP, _DP, _DP, _DP, _DP, _D, _DP, _I]
        L.mcac_swept_sphere_batch_noalloc.argtypes = [_I64, _DP, _DP, _DP, _DP, _DP, _DP, _D, _DP, _I]
        L.mcac_device_memory.argtypes = [_I, _I64P, _I64P]
        L.mcac_group_create.argtypes = [_I, _I, ctypes.POINTER(_P)]
        L.mcac_group_destroy.argtypes = [_P]
        L.mcac_group_add.argtypes = [_P, _P]
        L.mcac_move_enqueue_group.argtypes = [_P, ctypes.POINTER(_P), _I64P, _DP, _DP, _DP, _I,
                                              ctypes.POINTER(ctypes.c_uint64)]
        L.mcac_group_stats.argtypes = [_P, _I64P]
        L.mcac_group_times.argtypes = [_P, ctypes.POINTER(ctypes.c_double)]
        L.mcac_box_spare_moves.argtypes = [_P]
        L.mcac_box_block_steps.argtypes = [_P, _D]
        L.mcac_neighbour_order.argtypes = [_I, _I64P, _I, _I64P]
        L.mcac_arvo_part.argtypes = [_I, _DP, _DP, _DP, _DP, _DP]
        L.mcac_version.restype = ctypes.c_char_p
        _gpu = L
    return _gpu


def load_sim():
    """Loads libmcac_sim.so (host driver; loads libmcac_gpu.so)."""
    global _sim
    if _sim is None:
        load_gpu()
        path = LIB_DIR / "libmcac_sim.so"
        _require(path)
        L = ctypes.CDLL(str(path))
        L.mcac_sim_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _I, ctypes.POINTER(_P), ctypes.c_char_p, _I]
        L.mcac_sim_destroy.argtypes = [_P]
        L.mcac_sim_init.argtypes = [_P]
        L.mcac_sim_run.argtypes = [_P, _I64, _I64P]
        L.mcac_sim_finished.argtypes = [_P]
        L.mcac_sim_set_trace.argtypes = [_P, _I, _I]
        for f in ("mcac_sim_n_steps", "mcac_sim_n_merges", "mcac_sim_n_aggs"):
            getattr(L, f).restype = _I64
            getattr(L, f).argtypes = [_P]
        L.mcac_sim_get_steps.restype = _I64
        L.mcac_sim_get_steps.argtypes = [_P, _P, _I64]
        L.mcac_sim_get_merges.restype = _I64
        L.mcac_sim_get_merges.argtypes = [_P, _P, _I64]
        L.mcac_sim_scalar.restype = _D
        L.mcac_sim_scalar.argtypes = [_P, ctypes.c_char_p]
        L.mcac_sim_export_aggs.argtypes = [_P, _I64P] + [_DP] * 6 + [_I64P]
        L.mcac_sim_box.restype = _P
        L.mcac_sim_box.argtypes = [_P]
        L.mcac_physics_scalars.argtypes = [ctypes.c_char_p, _DP]
        L.mcac_pick_table_selftest.restype = _I64
        L.mcac_pick_table_selftest.argtypes = [_I64, _I64, _I64, ctypes.c_uint64, _I]
        L.mcac_pick_table_selftest_ties.restype = _I64
        L.mcac_write_test_snapshots.argtypes = [ctypes.c_char_p, _I64, _I64, _I64]
        L.mcac_batch_run.argtypes = [ctypes.POINTER(ctypes.c_char_p), _DP, _I64, _I, _I, _I64, _I64, _DP, _DP,
                                     _I64P, ctypes.POINTER(_I), _I64P]
        L.mcac_batch_run_groups.argtypes = [ctypes.POINTER(ctypes.c_char_p), _DP, _I64, _I, _I, _I, _I, _I64, _I64,
                                            _DP, _DP, _I64P, ctypes.POINTER(_I), _DP, _I64P]
        L.mcac_sim_device_bytes.restype = _I64
        L.mcac_sim_device_bytes.argtypes = [ctypes.c_char_p]
        _sim = L
    return _sim


def _dp(a):
    return a.ctypes.data_as(_DP)


def _ip(a):
    return a.ctypes.data_as(_I64P)


def _v3(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(3))


PHYSICS_KEYS = ("box_length", "lambda", "viscosity", "rpeqmass", "gamma", "u_sg", "box_volume",
                "aggregate_concentration")


def physics_scalars(ini_text: str) -> dict:
    """PhysicalModel derived scalars of a parameter file (host code, no GPU needed)."""
    out = np.zeros(8)
    _check(load_sim().mcac_physics_scalars(ini_text.encode(), _dp(out)), "mcac_physics_scalars")
    return dict(zip(PHYSICS_KEYS, out.tolist()))


def neighbour_order(prev: int, keys) -> np.ndarray:
    """Walk order of one sphere's neighbour map after an update (host-only
    check, mcac_neighbour_order): keys ascending, prev the map's previous size."""
    k = np.ascontiguousarray(keys, dtype=np.int64)
    out = np.zeros_like(k)
    _check(load_gpu().mcac_neighbour_order(int(prev), _ip(k), len(k), _ip(out)), "mcac_neighbour_order")
    return out


def arvo_part(balls) -> tuple:
    """(volume, area) part of ball 0 among balls (rows x, y, z, r in nm): host-only check (mcac_arvo_part)."""
    b = np.ascontiguousarray(balls, dtype=np.float64).reshape(-1, 4)
    cols = [np.ascontiguousarray(b[:, q]) for q in range(4)]
    out = np.zeros(2)
    _check(load_gpu().mcac_arvo_part(len(b) - 1, *(_dp(c) for c in cols), _dp(out)), "mcac_arvo_part")
    return float(out[0]), float(out[1])


def device_memory(device: int = 0) -> tuple:
    """(free, total) bytes of device memory (mcac_device_memory)."""
    f, t = np.zeros(1, np.int64), np.zeros(1, np.int64)
    _check(load_gpu().mcac_device_memory(device, _ip(f), _ip(t)), "mcac_device_memory")
    return int(f[0]), int(t[0])


def swept_sphere_batch(p1, r1, p2, r2, direction, dist, box: float, device: int = 0, noalloc: bool = False):
    """Device evaluation of distance_to_contact(Sphere, Sphere) for n inputs
    (noalloc: through mcac_swept_sphere_batch_noalloc, n <= 65,536)."""
    p1 = np.ascontiguousarray(p1, np.float64).reshape(-1, 3)
    n = p1.shape[0]
    p2 = np.ascontiguousarray(p2, np.float64).reshape(n, 3)
    d = np.ascontiguousarray(direction, np.float64).reshape(n, 3)
    r1 = np.ascontiguousarray(np.broadcast_to(r1, (n,)), np.float64)
    r2 = np.ascontiguousarray(np.broadcast_to(r2, (n,)), np.float64)
    dist = np.ascontiguousarray(np.broadcast_to(dist, (n,)), np.float64)
    out = np.zeros(n)
    f = load_gpu().mcac_swept_sphere_batch_noalloc if noalloc else load_gpu().mcac_swept_sphere_batch
    _check(f(n, _dp(p1), _dp(r1), _dp(p2), _dp(r2), _dp(d), _dp(dist), box, _dp(out), device), f.__name__)
    return out


class Box:
    """One aggregation box in HBM driven through the C-ABI (include/mcac_gpu.h)."""

    PARAMS = np.dtype([("box_length", "<f8"), ("n_div", "<i8"), ("max_spheres", "<i8"),
                       ("max_aggregates", "<i8"), ("candidate_capacity", "<i8"), ("member_pool", "<i8"),
                       ("device", "<i4"), ("precision", "<i4"), ("volsurf", "<i4"), ("reserved", "<i4")])

    def __init__(self, box_length: float, n_div: int, max_spheres: int, max_aggregates: int,
                 device: int = 0, precision: int = 0, candidate_capacity: int = 0, member_pool: int = 0,
                 volsurf: int = 0):
        L = load_gpu()
        p = np.zeros(1, self.PARAMS)
        p["box_length"], p["n_div"], p["max_spheres"], p["max_aggregates"] = box_length, n_div, max_spheres, max_aggregates
        p["candidate_capacity"], p["member_pool"], p["device"], p["precision"] = (
            candidate_capacity, member_pool, device, precision)
        p["volsurf"] = volsurf
        h = _P()
        _check(L.mcac_box_create(p.ctypes.data, ctypes.byref(h)), "mcac_box_create")
        self.h = h
        self.n_spheres = 0
        self.n_aggs = 0

    def close(self):
        if getattr(self, "h", None):
            load_gpu().mcac_box_destroy(self.h)
            self.h = None

    __del__ = close

    def upload(self, spheres: dict, aggs: dict):
        """spheres: x y z r vol surf rx ry rz (index order); aggs: pos (n,3), rcm (n,3), rmax,
        volume, surface, count, members (concatenated, stable-id order)."""
        s = {k: np.ascontiguousarray(spheres[k], np.float64) for k in ("x", "y", "z", "r", "vol", "surf", "rx", "ry", "rz")}
        a = {k: np.ascontiguousarray(aggs[k], np.float64) for k in ("pos", "rcm", "rmax", "volume", "surface")}
        count = np.ascontiguousarray(aggs["count"], np.int64)
        members = np.ascontiguousarray(aggs["members"], np.int64)
        ns, na = s["x"].size, count.size
        _check(load_gpu().mcac_box_upload(self.h, ns, *(_dp(s[k]) for k in ("x", "y", "z", "r", "vol", "surf", "rx", "ry", "rz")),
                                          na, _dp(a["pos"]), _dp(a["rcm"]), _dp(a["rmax"]), _dp(a["volume"]),
                                          _dp(a["surface"]), _ip(count), _ip(members)), "mcac_box_upload")
        self.n_spheres, self.n_aggs = ns, na

    def next_contact(self, agg: int, direction, lpm: float, maxradius: float, move: bool = False):
        out = np.zeros(1, CONTACT_DTYPE)
        f = load_gpu().mcac_move if move else load_gpu().mcac_next_contact
        _check(f(self.h, agg, _dp(_v3(direction)), lpm, maxradius, out.ctypes.data),
               "mcac_move" if move else "mcac_next_contact")
        return out[0]

    def move_enqueue(self, agg: int, direction, lpm: float, maxradius: float) -> int:
        """Queues a gated move; returns its ticket (results are read in ticket order)."""
        t = ctypes.c_uint64()
        _check(load_gpu().mcac_move_enqueue(self.h, agg, _dp(_v3(direction)), lpm, maxradius, ctypes.byref(t)),
               "mcac_move_enqueue")
        return t.value

    def move_result(self, ticket: int):
        """(contact record, skipped) of a queued move: skipped 0 ran, 1 did nothing
        behind a closed gate, 2 ran while the moves queued after it were skipped."""
        out = np.zeros(1, CONTACT_DTYPE)
        sk = ctypes.c_int()
        _check(load_gpu().mcac_move_result(self.h, ticket, out.ctypes.data, ctypes.byref(sk)), "mcac_move_result")
        return out[0], int(sk.value)

    def release(self):
        """Reopens the gate closed by a queued move that found a contact."""
        _check(load_gpu().mcac_release(self.h), "mcac_release")

    def translate(self, agg: int, vec):
        _check(load_gpu().mcac_translate(self.h, agg, _dp(_v3(vec))), "mcac_translate")

    def merge(self, contact):
        c = np.zeros(1, CONTACT_DTYPE)
        c[0] = contact
        out = np.zeros(1, MERGE_RESULT_DTYPE)
        _check(load_gpu().mcac_merge(self.h, c.ctypes.data, out.ctypes.data), "mcac_merge")
        return out[0]

    def spheres(self):
        n = self.n_spheres
        arrs = [np.zeros(n) for _ in range(6)]
        _check(load_gpu().mcac_box_download_spheres(self.h, *(_dp(a) for a in arrs)), "download_spheres")
        return dict(zip(("x", "y", "z", "rx", "ry", "rz"), arrs))

    def radii(self):
        """Radius, volume and surface of every sphere (they change with surface growth)."""
        n = self.n_spheres
        r, v, s = np.zeros(n), np.zeros(n), np.zeros(n)
        _check(load_gpu().mcac_box_download_radii(self.h, _dp(r), _dp(v), _dp(s)), "mcac_box_download_radii")
        return dict(r=r, vol=v, surf=s)

    def aggregates(self):
        n = self.n_aggs
        pos, rcm = np.zeros((n, 3)), np.zeros((n, 3))
        rmax, vol, surf, rg = (np.zeros(n) for _ in range(4))
        count = np.zeros(n, np.int64)
        _check(load_gpu().mcac_box_download_aggs(self.h, _dp(pos), _dp(rcm), _dp(rmax), _dp(vol), _dp(surf),
                                                 _dp(rg), _ip(count)), "download_aggs")
        return dict(pos=pos, rcm=rcm, rmax=rmax, volume=vol, surface=surf, rg=rg, count=count)

    def members(self, agg: int):
        n = np.zeros(1, np.int64)
        _check(load_gpu().mcac_box_members(self.h, agg, None, 0, _ip(n)), "mcac_box_members")
        out = np.zeros(int(n[0]), np.int64)
        _check(load_gpu().mcac_box_members(self.h, agg, _ip(out), out.size, _ip(n)), "mcac_box_members")
        return out

    def set_timing(self, stride: int):
        """HIP events around every `stride`-th search (and every merge); 0 turns timing off."""
        _check(load_gpu().mcac_box_set_timing(self.h, int(stride)), "mcac_box_set_timing")

    def timing(self):
        out = np.zeros(8)
        _check(load_gpu().mcac_box_timing(self.h, _dp(out)), "mcac_box_timing")
        keys = ("search_ms", "translate_ms", "merge_ms", "n_search", "n_translate", "n_merge", "filter_ms", "solve_ms")
        return dict(zip(keys, out.tolist()))


    def prune_stats(self):
        """FP32 broad phase: entries tested, passed to FP64, dropped-but-FP64-kept (precision 2)."""
        out = np.zeros(3, np.int64)
        _check(load_gpu().mcac_box_prune_stats(self.h, _ip(out)), "mcac_box_prune_stats")
        return dict(zip(("tested", "kept", "missed"), (int(v) for v in out)))

    def set_engine(self, blocks: int):
        """Queued one-launch steps go to a resident grid of `blocks` workgroups (0: one launch each)."""
        _check(load_gpu().mcac_box_set_engine(self.h, int(blocks)), "mcac_box_set_engine")

    def engine_stats(self):
        out = np.zeros(8)
        _check(load_gpu().mcac_box_engine_stats(self.h, _dp(out)), "mcac_box_engine_stats")
        keys = ("launches", "posted", "steps", "busy_ms", "relaunches", "end_reason", "idle_ms", "grid")
        return dict(zip(keys, out.tolist()))

    def set_auto_merge(self, enable: bool):
        """Queued moves that find a contact merge in the same launch (small aggregates)."""
        _check(load_gpu().mcac_box_set_auto_merge(self.h, int(enable)), "mcac_box_set_auto_merge")

    def stats(self):
        out = np.zeros(8, np.int64)
        _check(load_gpu().mcac_box_stats(self.h, _ip(out)), "mcac_box_stats")
        keys = ("one_launch", "two_launch", "handover", "cell_rebuilds", "pool_compactions", "pool_used",
                "inline_merges", "launched_merges")
        return dict(zip(keys, (int(v) for v in out)))


class _BorrowedBox(Box):
    """The box owned by a Simulation (not destroyed by Python)."""

    def __init__(self, handle, n_spheres, n_aggs):  # noqa: D401 - no create call
        self.h = handle
        self.n_spheres = n_spheres
        self.n_aggs = n_aggs

    def close(self):
        self.h = None

    __del__ = close


class Simulation:
    """One MCAC run on the MI355X path (include/mcac_sim.h)."""

    def __init__(self, ini_text: str, output_dir: str | None = None, device: int = 0,
                 trace_steps: bool = False, trace_merges: bool = False):
        L = load_sim()
        err = ctypes.create_string_buffer(512)
        h = _P()
        rc = L.mcac_sim_create(ini_text.encode(), output_dir.encode() if output_dir else None, device,
                               ctypes.byref(h), err, 512)
        if rc != 0:
            raise MCACError(f"mcac_sim_create: {err.value.decode()} (status {rc})")
        self.h = h
        _check(L.mcac_sim_set_trace(self.h, int(trace_steps), int(trace_merges)), "mcac_sim_set_trace")
        _check(L.mcac_sim_init(self.h), "mcac_sim_init")

    def close(self):
        if getattr(self, "h", None):
            load_sim().mcac_sim_destroy(self.h)
            self.h = None

    __del__ = close

    def run(self, max_steps: int = 2**62) -> int:
        done = np.zeros(1, np.int64)
        _check(load_sim().mcac_sim_run(self.h, max_steps, _ip(done)), "mcac_sim_run")
        return int(done[0])

    def finished(self) -> bool:
        return bool(load_sim().mcac_sim_finished(self.h))

    def scalar(self, name: str) -> float:
        v = load_sim().mcac_sim_scalar(self.h, name.encode())
        if math.isnan(v) and name not in ("",):
            pass
        return v

    def steps(self):
        L = load_sim()
        n = L.mcac_sim_n_steps(self.h)
        out = np.zeros(n, STEP_DTYPE)
        L.mcac_sim_get_steps(self.h, out.ctypes.data, n)
        return out

    def merges(self):
        L = load_sim()
        n = L.mcac_sim_n_merges(self.h)
        out = np.zeros(n, MERGE_DTYPE)
        L.mcac_sim_get_merges(self.h, out.ctypes.data, n)
        return out

    def aggregates(self):
        L = load_sim()
        n = L.mcac_sim_n_aggs(self.h)
        uid, nsph = np.zeros(n, np.int64), np.zeros(n, np.int64)
        lpm, ts, pt, vol, rg, rmax = (np.zeros(n) for _ in range(6))
        _check(L.mcac_sim_export_aggs(self.h, _ip(uid), _dp(lpm), _dp(ts), _dp(pt), _dp(vol), _dp(rg), _dp(rmax),
                                      _ip(nsph)), "mcac_sim_export_aggs")
        return dict(uid=uid, lpm=lpm, time_step=ts, proper_time=pt, volume=vol, rg=rg, rmax=rmax, n=nsph)

    def box(self) -> Box:
        n_sph = int(self.scalar("n_sph"))
        return _BorrowedBox(load_sim().mcac_sim_box(self.h), n_sph, n_sph)
