"""ctypes binding of the oracle (oracle/libmcac_oracle.so) — test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's CPU legs (cpu_baseline, ensemble_vs_cpu) use this.
"""
import ctypes
import pathlib
import subprocess

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
LIB_PATH = ROOT / "oracle" / "libmcac_oracle.so"

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
assert STEP_DTYPE.itemsize == 112 and MERGE_DTYPE.itemsize == 176

_lib = None
_P = ctypes.c_void_p
_D = ctypes.c_double
_I64 = ctypes.c_int64
_DP = ctypes.POINTER(ctypes.c_double)
_I64P = ctypes.POINTER(ctypes.c_int64)


def build():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        L.ora_create.restype = _P
        L.ora_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.ora_destroy.argtypes = [_P]
        L.ora_init.argtypes = [_P, ctypes.c_int]
        L.ora_run.restype = _I64
        L.ora_run.argtypes = [_P, _I64]
        L.ora_finished.argtypes = [_P]
        L.ora_set_trace.argtypes = [_P, ctypes.c_int, ctypes.c_int]
        for f in ("ora_n_steps", "ora_n_merges", "ora_n_spheres", "ora_n_aggs"):
            getattr(L, f).restype = _I64
            getattr(L, f).argtypes = [_P]
        L.ora_get_steps.restype = _I64
        L.ora_get_steps.argtypes = [_P, _P, _I64]
        L.ora_get_merges.restype = _I64
        L.ora_get_merges.argtypes = [_P, _P, _I64]
        L.ora_get_scalar.restype = _D
        L.ora_get_scalar.argtypes = [_P, ctypes.c_char_p]
        L.ora_export_spheres.argtypes = [_P] + [_DP] * 9 + [_I64P]
        L.ora_export_aggs.argtypes = [_P] + [_DP] * 9 + [_I64P] * 4
        L.ora_next_contact.restype = _D
        L.ora_next_contact.argtypes = [_P, _I64, _DP, _D, _I64P, _I64P, _I64P, _I64P, _I64P, _I64P, _I64P]
        L.ora_swept_sphere.restype = _D
        L.ora_swept_sphere.argtypes = [_DP, _D, _DP, _D, _DP, _D, _D]
        L.ora_contact.argtypes = [_DP, _D, _DP, _D, _D]
        L.ora_contact_components.restype = _I64
        L.ora_contact_components.argtypes = [_I64] + [_DP] * 4 + [_D, _I64P, _I64P]
        L.ora_rng_create.restype = _P
        L.ora_rng_create.argtypes = [ctypes.c_int]
        L.ora_rng_destroy.argtypes = [_P]
        L.ora_rng_uniform.restype = _D
        L.ora_rng_uniform.argtypes = [_P]
        L.ora_rng_direction.argtypes = [_P, _DP]
        L.ora_inverf.restype = _D
        L.ora_inverf.argtypes = [_D]
        L.ora_inverfc.restype = _D
        L.ora_inverfc.argtypes = [_D]
        L.ora_physics.argtypes = [ctypes.c_char_p, _DP]
        L.ora_verlet_query.restype = _I64
        L.ora_verlet_query.argtypes = [_I64, _D, _I64P, _I64, _DP, _DP, _D, _I64P, _I64]
        L.ora_caps_intersection.argtypes = [_D] * 7 + [_DP]
        L.ora_caps_intersection.restype = None
        L.ora_merge_geometry.argtypes = [ctypes.c_int, _I64, _I64] + [_DP] * 6 + [_DP]
        L.ora_neighbour_order.argtypes = [ctypes.c_int, _I64P, ctypes.c_int, _I64P]
        L.ora_arvo_part.argtypes = [ctypes.c_int, _DP, _DP, _DP, _DP, _DP]
        L.ora_arvo_part.restype = None
        _lib = L
    return _lib


def dp(a):
    return a.ctypes.data_as(_DP)


def ip(a):
    return a.ctypes.data_as(_I64P)


def v3(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(3))


class OracleSim:
    """One oracle run (AggregatList + calcul loop)."""

    def __init__(self, ini_text: str, fast_init: bool = False, trace=True):
        L = lib()
        err = ctypes.create_string_buffer(512)
        self.h = L.ora_create(ini_text.encode(), err, 512)
        if not self.h:
            raise ValueError(err.value.decode())
        L.ora_set_trace(self.h, int(trace), int(trace))
        rc = L.ora_init(self.h, int(fast_init))
        if rc != 0:
            raise RuntimeError(f"oracle init failed with code {rc}")

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_destroy(self.h)
            self.h = None

    def run(self, max_steps: int) -> int:
        n = lib().ora_run(self.h, max_steps)
        if n < 0:
            raise RuntimeError(f"oracle run failed with code {-n}")
        return n

    def finished(self) -> bool:
        return bool(lib().ora_finished(self.h))

    def scalar(self, name: str) -> float:
        return lib().ora_get_scalar(self.h, name.encode())

    def steps(self):
        L = lib()
        n = L.ora_n_steps(self.h)
        out = np.zeros(n, STEP_DTYPE)
        L.ora_get_steps(self.h, out.ctypes.data, n)
        return out

    def merges(self):
        L = lib()
        n = L.ora_n_merges(self.h)
        out = np.zeros(n, MERGE_DTYPE)
        L.ora_get_merges(self.h, out.ctypes.data, n)
        return out

    def export(self):
        """Current state: spheres (sphere-index order) and aggregates (label order)."""
        L = lib()
        ns, na = L.ora_n_spheres(self.h), L.ora_n_aggs(self.h)
        sph = {k: np.zeros(ns) for k in ("x", "y", "z", "r", "vol", "surf", "rx", "ry", "rz")}
        label = np.zeros(ns, np.int64)
        L.ora_export_spheres(self.h, *(dp(sph[k]) for k in ("x", "y", "z", "r", "vol", "surf", "rx", "ry", "rz")),
                             ip(label))
        sph["label"] = label
        agg = {"pos": np.zeros((na, 3)), "rcm": np.zeros((na, 3))}
        for k in ("rmax", "volume", "surface", "rg", "lpm", "time_step", "proper_time"):
            agg[k] = np.zeros(na)
        for k in ("n", "uid", "offset"):
            agg[k] = np.zeros(na, np.int64)
        members = np.zeros(ns, np.int64)
        L.ora_export_aggs(self.h, dp(agg["pos"]), dp(agg["rcm"]), dp(agg["rmax"]), dp(agg["volume"]),
                          dp(agg["surface"]), dp(agg["rg"]), dp(agg["lpm"]), dp(agg["time_step"]),
                          dp(agg["proper_time"]), ip(agg["n"]), ip(agg["uid"]), ip(agg["offset"]), ip(members))
        agg["members"] = members
        return sph, agg

    def next_contact(self, label: int, direction, lpm: float):
        L = lib()
        d = v3(direction)
        outs = [np.zeros(1, np.int64) for _ in range(7)]
        dist = L.ora_next_contact(self.h, label, dp(d), lpm, *(ip(o) for o in outs))
        ms, os_, ol, nq, nf, nsv, npairs = (int(o[0]) for o in outs)
        return dist, ms, os_, ol, dict(query=nq, filtered=nf, solved=nsv, pairs=npairs)


def swept_sphere(p1, r1, p2, r2, direction, dist, box):
    return lib().ora_swept_sphere(dp(v3(p1)), r1, dp(v3(p2)), r2, dp(v3(direction)), dist, box)


def contact(p1, r1, p2, r2, box):
    return bool(lib().ora_contact(dp(v3(p1)), r1, dp(v3(p2)), r2, box))


def contact_components(pos, r, box):
    """The oracle's split components (aggregat.cpp:553-577) of spheres at pos (n x 3) with radii r."""
    pos = np.ascontiguousarray(pos, np.float64)
    r = np.ascontiguousarray(r, np.float64)
    n = len(r)
    x, y, z = (np.ascontiguousarray(pos[:, k]) for k in range(3))
    order = np.zeros(max(n, 1), np.int64)
    starts = np.zeros(n + 1, np.int64)
    k = lib().ora_contact_components(n, dp(x), dp(y), dp(z), dp(r), box, ip(order), ip(starts))
    return [order[starts[c]:starts[c + 1]].tolist() for c in range(int(k))]


def caps_intersection(r1, vol1, surf1, r2, vol2, surf2, dist):
    """(volume_1, surface_1, volume_2, surface_2) of sphere_intersection.cpp:29-66."""
    out = np.zeros(4)
    lib().ora_caps_intersection(r1, vol1, surf1, r2, vol2, surf2, dist, dp(out))
    return tuple(out.tolist())


def merge_geometry(volsurf, nk, rel, r, vol, surf):
    """Aggregate::update geometry of the merged list (first nk spheres: the keeper).
    Returns dict(volume, surface, rg, rmax, rcm, overlapping, coordination, sphere_volume, radius_sum)."""
    rel = np.ascontiguousarray(rel, dtype=np.float64)
    cols = [np.ascontiguousarray(rel[:, k]) for k in range(3)]
    r, vol, surf = (np.ascontiguousarray(a, dtype=np.float64) for a in (r, vol, surf))
    out = np.zeros(11)
    rc = lib().ora_merge_geometry(int(volsurf), int(nk), len(r), *(dp(c) for c in cols), dp(r), dp(vol), dp(surf), dp(out))
    if rc != 0:
        raise RuntimeError(f"ora_merge_geometry: {rc}")
    return dict(volume=out[0], surface=out[1], rg=out[2], rmax=out[3], rcm=out[4:7].copy(), overlapping=out[7],
                coordination=out[8], sphere_volume=out[9], radius_sum=out[10])


def arvo_part(balls):
    """(volume, area) part of ball 0 among balls (rows x, y, z, r in nm), restated ARVO."""
    b = np.ascontiguousarray(balls, dtype=np.float64).reshape(-1, 4)
    cols = [np.ascontiguousarray(b[:, q]) for q in range(4)]
    out = np.zeros(2)
    lib().ora_arvo_part(len(b) - 1, *(dp(c) for c in cols), dp(out))
    return float(out[0]), float(out[1])


def neighbour_order(prev, keys):
    """Iteration order of a real std::unordered_map after clear/reserve(prev)/ascending inserts."""
    k = np.ascontiguousarray(keys, dtype=np.int64)
    out = np.zeros_like(k)
    n = lib().ora_neighbour_order(int(prev), ip(k), len(k), ip(out))
    assert n == len(k)
    return out


class OracleRng:
    def __init__(self, seed: int):
        self.h = lib().ora_rng_create(seed)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_rng_destroy(self.h)
            self.h = None

    def uniform(self) -> float:
        return lib().ora_rng_uniform(self.h)

    def direction(self):
        d = np.zeros(3)
        lib().ora_rng_direction(self.h, dp(d))
        return d


def inverf(p):
    return lib().ora_inverf(p)


def inverfc(p):
    return lib().ora_inverfc(p)


def verlet_query(n_div, width, cells, src, vec, dist):
    cells = np.ascontiguousarray(np.asarray(cells, np.int64).reshape(-1))
    out = np.zeros(max(1, cells.size // 3 + 1), np.int64)
    n = lib().ora_verlet_query(n_div, width, ip(cells), cells.size // 3, dp(v3(src)), dp(v3(vec)), dist,
                               ip(out), out.size)
    return out[:n]


PHYSICS_KEYS = ("box_length", "lambda", "viscosity", "rpeqmass", "gamma", "u_sg", "box_volume",
                "aggregate_concentration")


def physics(ini_text):
    out = np.zeros(8)
    rc = lib().ora_physics(ini_text.encode(), dp(out))
    assert rc == 0
    return dict(zip(PHYSICS_KEYS, out.tolist()))
