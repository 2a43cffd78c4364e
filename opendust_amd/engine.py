"""ctypes binding of the C ABI declared in include/opendust_amd.h.

The structures below mirror ``od_table``, ``od_config`` and
``od_dust_record`` field by field; ``check_abi()`` compares their sizes with
the library's own ``sizeof`` (no GPU needed).  Every engine call raises
``RuntimeError`` with the library's message on failure; there is no fallback
path when the library is missing.
"""

import ctypes
import os

import numpy as np

from . import host

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libopendust_amd.so")

MAX_DUST = host.MAX_DUST
MAX_TABLE = host.MAX_TABLE
ABI_VERSION = 3

RUN_GRAPH = 1
RUN_PROFILE_PAIR = 2
RUN_PROFILE_STAGES = 4
# stages of a step as od_stage_timing reports them (include/opendust_amd.h OD_STAGE_*)
STAGES = ("kick_drift", "list", "cells", "pair", "rows_absorbed", "corr", "kick2", "rows_reinjected", "inflow",
          "dust")

PAIR_AUTO = 0
PAIR_ROWS = 1
PAIR_SYM = 2
PAIR_CELL = 3

_dv = ctypes.c_double * MAX_DUST


class Table(ctypes.Structure):
    _fields_ = [
        ("n_seg", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("v0", ctypes.c_double * MAX_TABLE),
        ("xi0", ctypes.c_double * MAX_TABLE),
        ("slope", ctypes.c_double * MAX_TABLE),
        ("center", ctypes.c_double * MAX_TABLE),
        ("half", ctypes.c_double * MAX_TABLE),
    ]

    def load(self, tab):
        if tab is None:
            self.n_seg = 0
            return
        n = tab.n_seg
        self.n_seg = n
        for name in ("v0", "xi0", "slope", "center", "half"):
            arr = getattr(self, name)
            vals = getattr(tab, name)
            for k in range(n):
                arr[k] = vals[k]


class Config(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("struct_size", ctypes.c_int32),
        ("n_ions", ctypes.c_int32),
        ("n_dust", ctypes.c_int32),
        ("distribution", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("use_bottom", ctypes.c_int32),
        ("j_split", ctypes.c_int32),
        ("pair_mode", ctypes.c_int32),
        ("sym_chunk", ctypes.c_int32),
        ("sim_id", ctypes.c_int32),
        ("sym_tile", ctypes.c_int32),
        ("sym_band_mib", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("R", ctypes.c_double), ("H", ctypes.c_double), ("dt", ctypes.c_double),
        ("r_D_e", ctypes.c_double), ("r_0", ctypes.c_double),
        ("m_s", ctypes.c_double),
        ("v_T", ctypes.c_double), ("v_fl", ctypes.c_double),
        ("kqq", ctypes.c_double),
        ("k_dust", ctypes.c_double),
        ("trap_C", ctypes.c_double),
        ("trap_alpha", ctypes.c_double * 15),
        ("field_force", ctypes.c_double),
        ("cutoff", ctypes.c_double),
        ("switch_dist", ctypes.c_double),
        ("thr_top", ctypes.c_uint64),
        ("thr_side", ctypes.c_uint64),
        ("thr_coll", ctypes.c_uint64),
        ("dust_x", _dv), ("dust_y", _dv), ("dust_z", _dv), ("dust_r", _dv),
        ("dust_q0", _dv),
        ("dust_method", ctypes.c_int32 * MAX_DUST),
        ("si_dust_x", _dv), ("si_dust_y", _dv), ("si_dust_z", _dv), ("si_dust_r", _dv),
        ("si_k", ctypes.c_double),
        ("si_e", ctypes.c_double),
        ("si_q_s", ctypes.c_double),
        ("si_r_D_e", ctypes.c_double),
        ("si_dt", ctypes.c_double),
        ("si_kB", ctypes.c_double), ("si_Te", ctypes.c_double), ("si_Tn", ctypes.c_double),
        ("si_E", ctypes.c_double),
        ("gamma_e_pref", ctypes.c_double),
        ("oml_A", ctypes.c_double),
        ("oml_B", ctypes.c_double),
        ("oml_u2", ctypes.c_double),
        ("oml_erf_u", ctypes.c_double),
        ("oml_tail", ctypes.c_double),
        ("collect_scale", ctypes.c_double),
        ("top", Table), ("bottom", Table), ("side", Table), ("vz", Table),
    ]


class DustRecord(ctypes.Structure):
    _fields_ = [
        ("q", ctypes.c_double),
        ("force_orbit", ctypes.c_double * 3),
        ("force_collect", ctypes.c_double * 3),
        ("force_dust", ctypes.c_double * 3),
        ("force_field", ctypes.c_double * 3),
        ("orbit_sum", ctypes.c_double * 3),
        ("momentum_sum", ctypes.c_double * 3),
        ("absorbed", ctypes.c_int64),
    ]


# numpy view of an od_dust_record array
RECORD_DTYPE = np.dtype([
    ("q", "<f8"),
    ("force_orbit", "<f8", (3,)),
    ("force_collect", "<f8", (3,)),
    ("force_dust", "<f8", (3,)),
    ("force_field", "<f8", (3,)),
    ("orbit_sum", "<f8", (3,)),
    ("momentum_sum", "<f8", (3,)),
    ("absorbed", "<i8"),
])

_lib = None

# every symbol the header declares
EXPORTS = [
    "od_create", "od_set_ions", "od_get_ions", "od_set_dust_charges", "od_get_dust_charges",
    "od_set_step", "od_run", "od_pair_timing", "od_pair_info", "od_get_step_info", "od_pair_forces",
    "od_get_cells", "od_cutoff_pair_count",
    "od_trap_potential_grid", "od_trap_samples_host", "od_abi_sizes", "od_destroy",
    "od_last_error", "od_build_info",
    "od_batch_create", "od_batch_set_ions", "od_batch_get_ions", "od_batch_run", "od_batch_get_step_info",
    "od_batch_pair_timing", "od_batch_destroy", "od_tune", "od_batch_tune", "od_sym_variants", "od_sym_variant_info",
    "od_sym_choice",
    "od_density_profile", "od_density_profile_z", "od_write_xyz", "od_write_csv", "od_write_restart",
    "od_update_config", "od_density_profile_noalloc", "od_stage_timing",
]


def load_library(path=LIB_PATH):
    """Load the engine .so (built by opendust_amd.build); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            "opendust_amd engine library not found at %s; build it with "
            "`python -m opendust_amd.build` (there is no CPU fallback)" % path)
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    c_ctx = ctypes.c_void_p
    dp = P(ctypes.c_double)
    lib.od_create.argtypes = [P(Config), P(c_ctx)]
    lib.od_set_ions.argtypes = [c_ctx, dp, dp, ctypes.c_int64]
    lib.od_get_ions.argtypes = [c_ctx, dp, dp]
    lib.od_set_dust_charges.argtypes = [c_ctx, dp]
    lib.od_get_dust_charges.argtypes = [c_ctx, dp]
    lib.od_set_step.argtypes = [c_ctx, ctypes.c_int64]
    lib.od_run.argtypes = [c_ctx, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32]
    lib.od_pair_timing.argtypes = [c_ctx, dp, P(ctypes.c_int64), ctypes.c_int32]
    lib.od_pair_info.argtypes = [c_ctx, P(ctypes.c_int32), P(ctypes.c_int64)]
    lib.od_stage_timing.argtypes = [c_ctx, dp, P(ctypes.c_int64), ctypes.c_int32]
    lib.od_get_cells.argtypes = [c_ctx, P(ctypes.c_int32), P(ctypes.c_int32), P(ctypes.c_int64), P(ctypes.c_int32)]
    lib.od_cutoff_pair_count.argtypes = [c_ctx, P(ctypes.c_int64), P(ctypes.c_int64)]
    lib.od_get_step_info.argtypes = [c_ctx, P(ctypes.c_uint32), P(ctypes.c_uint8), P(ctypes.c_int64)]
    lib.od_pair_forces.argtypes = [c_ctx, dp, dp]
    lib.od_trap_potential_grid.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_int32, dp,
                                           ctypes.c_int32, dp, ctypes.c_int64, ctypes.c_uint64, dp]
    lib.od_trap_samples_host.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_uint64, ctypes.c_int64,
                                         ctypes.c_int64, dp]
    lib.od_abi_sizes.argtypes = [P(ctypes.c_int64), P(ctypes.c_int64)]
    lib.od_destroy.argtypes = [c_ctx]
    lib.od_batch_create.argtypes = [P(Config), ctypes.c_int32, P(c_ctx)]
    lib.od_batch_set_ions.argtypes = [c_ctx, ctypes.c_int32, dp, dp, ctypes.c_int64]
    lib.od_batch_get_ions.argtypes = [c_ctx, ctypes.c_int32, dp, dp]
    lib.od_batch_run.argtypes = [c_ctx, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32]
    lib.od_batch_get_step_info.argtypes = [c_ctx, ctypes.c_int32, P(ctypes.c_uint32), P(ctypes.c_uint8)]
    lib.od_batch_pair_timing.argtypes = [c_ctx, dp, P(ctypes.c_int64), P(ctypes.c_int64), ctypes.c_int32]
    lib.od_batch_destroy.argtypes = [c_ctx]
    lib.od_tune.argtypes = [c_ctx, ctypes.c_int32, P(ctypes.c_int32), P(ctypes.c_int32)]
    lib.od_sym_variants.argtypes = []
    lib.od_sym_choice.argtypes = [c_ctx, P(ctypes.c_int32), P(ctypes.c_int32), P(ctypes.c_int32)]
    lib.od_sym_variant_info.argtypes = [ctypes.c_int32, P(ctypes.c_int32), P(ctypes.c_int32), P(ctypes.c_int32),
                                        P(ctypes.c_int32)]
    lib.od_update_config.argtypes = [c_ctx, P(Config)]
    lib.od_write_xyz.argtypes = [ctypes.c_char_p, dp, ctypes.c_int64, ctypes.c_int64]
    lib.od_write_csv.argtypes = [ctypes.c_char_p, dp, ctypes.c_int64, ctypes.c_double, ctypes.c_double]
    lib.od_write_restart.argtypes = [ctypes.c_char_p, dp, dp, ctypes.c_int64]
    lib.od_density_profile.argtypes = [c_ctx, ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       dp, dp]
    lib.od_density_profile_noalloc.argtypes = [c_ctx, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_double, dp, dp]
    lib.od_density_profile_z.argtypes = [ctypes.c_int32, dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, dp, dp]
    lib.od_batch_tune.argtypes = [c_ctx, ctypes.c_int32, P(ctypes.c_int32), P(ctypes.c_int32)]
    lib.od_last_error.restype = ctypes.c_char_p
    lib.od_build_info.restype = ctypes.c_char_p
    for name in EXPORTS:
        if name not in ("od_last_error", "od_build_info"):
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def _check(rc, what):
    if rc != 0:
        msg = _lib.od_last_error().decode(errors="replace")
        raise RuntimeError("%s failed (status %d): %s" % (what, rc, msg))


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def check_abi():
    """Sizes of the mirrored structures must equal the compiled ones."""
    lib = load_library()
    cs = ctypes.c_int64()
    rs = ctypes.c_int64()
    _check(lib.od_abi_sizes(ctypes.byref(cs), ctypes.byref(rs)), "od_abi_sizes")
    return cs.value == ctypes.sizeof(Config) and rs.value == ctypes.sizeof(DustRecord)


def sym_variants():
    """[(64-target sub-blocks per wave, waves per SIMD, operand form (0 DPP/SGPR
    operands, 1 VGPR operands, 2 sources read from LDS), target blocks per
    work-group)] of the symmetric kernel variants;
    od_config.sym_tile = index + 1 selects one."""
    lib = load_library()
    out = []
    for v in range(1, lib.od_sym_variants() + 1):
        mt, wps, vg, rows = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(lib.od_sym_variant_info(v, ctypes.byref(mt), ctypes.byref(wps), ctypes.byref(vg), ctypes.byref(rows)),
               "od_sym_variant_info")
        out.append((mt.value, wps.value, vg.value, rows.value))
    return out


def trap_potential_grid(R_over_lambda, H_over_lambda, rho, h, n_mc, seed, device=0):
    lib = load_library()
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float64)
    out = np.empty((len(rho), len(h)), dtype=np.float64)
    _check(lib.od_trap_potential_grid(int(device), float(R_over_lambda), float(H_over_lambda), len(rho), _dptr(rho),
                                      len(h), _dptr(h), int(n_mc), int(seed) & (2 ** 64 - 1), _dptr(out)),
           "od_trap_potential_grid")
    return out


def _path(p):
    return os.fsencode(os.fspath(p))


def write_xyz(path, pos, frame):
    """Append one XYZ frame (opendust.py:1080-1088)."""
    lib = load_library()
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    _check(lib.od_write_xyz(_path(path), _dptr(pos), pos.shape[0], int(frame)), "od_write_xyz")


def write_csv(path, pos, density, mass):
    """Write one CSV frame (opendust.py:1089-1099)."""
    lib = load_library()
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    _check(lib.od_write_csv(_path(path), _dptr(pos), pos.shape[0], float(density), float(mass)), "od_write_csv")


def write_restart(path, pos, vel):
    """Write the RESTART file (opendust.py:1100-1115)."""
    lib = load_library()
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
    if pos.shape != vel.shape:
        raise ValueError("positions and velocities differ in shape")
    _check(lib.od_write_restart(_path(path), _dptr(pos), _dptr(vel), pos.shape[0]), "od_write_restart")


def density_profile_z(z_nm, R, L, N_s, n=100, device=0):
    """Axial density profile (tools/concentration.cpp) of host positions
    z_nm (nm) on the GPU: (z_grid[n] m, concentration[n] 1/m^3)."""
    lib = load_library()
    z = np.ascontiguousarray(z_nm, dtype=np.float64).ravel()
    zg = np.empty(n)
    c = np.empty(n)
    _check(lib.od_density_profile_z(int(device), _dptr(z), z.size, int(n), float(R), float(L), float(N_s),
                                    _dptr(zg), _dptr(c)), "od_density_profile_z")
    return zg, c


def trap_samples_host(R_over_lambda, H_over_lambda, seed, k0, n):
    lib = load_library()
    out = np.empty((n, 3), dtype=np.float64)
    _check(lib.od_trap_samples_host(float(R_over_lambda), float(H_over_lambda), int(seed) & (2 ** 64 - 1), int(k0),
                                    int(n), _dptr(out)), "od_trap_samples_host")
    return out


class Engine:
    """One engine context (one GPU, one stream)."""

    def __init__(self, cfg):
        self.lib = load_library()
        cfg.struct_size = ctypes.sizeof(Config)
        self.cfg = cfg
        self.n = cfg.n_ions
        self.n_dust = cfg.n_dust
        self._ctx = ctypes.c_void_p()
        _check(self.lib.od_create(ctypes.byref(cfg), ctypes.byref(self._ctx)), "od_create")

    def close(self):
        if self._ctx:
            self.lib.od_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_ions(self, x, v):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        _check(self.lib.od_set_ions(self._ctx, _dptr(x), _dptr(v), x.shape[0]), "od_set_ions")

    def get_ions(self):
        x = np.empty((self.n, 3))
        v = np.empty((self.n, 3))
        _check(self.lib.od_get_ions(self._ctx, _dptr(x), _dptr(v)), "od_get_ions")
        return x, v

    def update_config(self, cfg):
        """New parameters (dusts, plasma, trap, tables, seed) for the next runs;
        the ions stay on the device (od_update_config)."""
        cfg.struct_size = ctypes.sizeof(Config)
        _check(self.lib.od_update_config(self._ctx, ctypes.byref(cfg)), "od_update_config")
        self.cfg = cfg

    def set_dust_charges(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        _check(self.lib.od_set_dust_charges(self._ctx, _dptr(q)), "od_set_dust_charges")

    def get_dust_charges(self):
        q = np.zeros(max(self.n_dust, 1))
        _check(self.lib.od_get_dust_charges(self._ctx, _dptr(q)), "od_get_dust_charges")
        return q[: self.n_dust]

    def set_step(self, step):
        _check(self.lib.od_set_step(self._ctx, int(step)), "od_set_step")

    def run(self, nsteps, flags=0):
        """Run nsteps steps; returns records [nsteps][n_dust] (RECORD_DTYPE)."""
        rec = np.zeros((nsteps, max(self.n_dust, 1)), dtype=RECORD_DTYPE)
        ptr = rec.ctypes.data_as(ctypes.c_void_p) if self.n_dust > 0 else None
        _check(self.lib.od_run(self._ctx, int(nsteps), ptr, int(flags)), "od_run")
        return rec[:, : self.n_dust]

    def pair_timing(self, reset=False):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        _check(self.lib.od_pair_timing(self._ctx, ctypes.byref(ms), ctypes.byref(n), int(bool(reset))),
               "od_pair_timing")
        return ms.value, n.value

    def stage_timing(self, reset=False):
        """{stage: ms per step} of the steps run with RUN_PROFILE_STAGES since the
        last reset (HIP events around each stage), and their count."""
        ms = np.zeros(len(STAGES))
        steps = ctypes.c_int64()
        _check(self.lib.od_stage_timing(self._ctx, _dptr(ms), ctypes.byref(steps), int(bool(reset))),
               "od_stage_timing")
        n = max(steps.value, 1)
        return {name: float(v) / n for name, v in zip(STAGES, ms)}, steps.value

    def pair_info(self):
        """(mode, ordered pairs covered by one timed main pair launch)"""
        mode = ctypes.c_int32()
        pairs = ctypes.c_int64()
        _check(self.lib.od_pair_info(self._ctx, ctypes.byref(mode), ctypes.byref(pairs)), "od_pair_info")
        return mode.value, pairs.value

    def cells(self):
        """(cell_of_ion[n], sorted_ions[n_binned], dims) of the last cell list (cutoff mode)."""
        cell = np.empty(self.n, dtype=np.int32)
        order = np.empty(self.n, dtype=np.int32)
        nb = ctypes.c_int64()
        dims = (ctypes.c_int32 * 3)()
        i32 = ctypes.POINTER(ctypes.c_int32)
        _check(self.lib.od_get_cells(self._ctx, cell.ctypes.data_as(i32), order.ctypes.data_as(i32),
                                     ctypes.byref(nb), dims), "od_get_cells")
        return cell, order[: nb.value], tuple(dims)

    def cutoff_pair_count(self):
        a = ctypes.c_int64()
        b = ctypes.c_int64()
        _check(self.lib.od_cutoff_pair_count(self._ctx, ctypes.byref(a), ctypes.byref(b)), "od_cutoff_pair_count")
        return a.value, b.value

    def step_info(self):
        flags = np.empty(self.n, dtype=np.uint32)
        face = np.empty(self.n, dtype=np.uint8)
        nr = ctypes.c_int64()
        _check(self.lib.od_get_step_info(self._ctx, flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                         face.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(nr)),
               "od_get_step_info")
        return flags, face, nr.value

    def density_profile(self, R, L, N_s, n=100):
        """Axial density profile of the current ions (tools/concentration.cpp):
        (z_grid[n] m, concentration[n] 1/m^3); R, L in m."""
        zg = np.empty(n)
        c = np.empty(n)
        _check(self.lib.od_density_profile(self._ctx, int(n), float(R), float(L), float(N_s), _dptr(zg), _dptr(c)),
               "od_density_profile")
        return zg, c

    def density_profile_noalloc(self, R, L, N_s, n=100):
        """density_profile through the entry point that allocates nothing on
        the device (its scratch is a fixed per-device array; n <= 65536)."""
        zg = np.empty(n)
        c = np.empty(n)
        _check(self.lib.od_density_profile_noalloc(self._ctx, int(n), float(R), float(L), float(N_s), _dptr(zg),
                                                   _dptr(c)), "od_density_profile_noalloc")
        return zg, c

    def tune(self, reps=3):
        """Autotune the symmetric kernel: source blocks per work-group and kernel
        variant (tile shape, waves per SIMD).  Returns (sym_chunk, sym_tile),
        (0, 0) when the context does not run it; setting both in the
        configuration reproduces the tuned run."""
        ch = ctypes.c_int32()
        tile = ctypes.c_int32()
        _check(self.lib.od_tune(self._ctx, int(reps), ctypes.byref(ch), ctypes.byref(tile)), "od_tune")
        return ch.value, tile.value

    def sym_choice(self):
        """(source blocks per work-group, kernel variant) the symmetric kernel
        runs with now; (0, 0) when the context does not run it."""
        return self.sym_schedule()[:2]

    def sym_schedule(self):
        """(source blocks per work-group, kernel variant, bands of rows) of the
        symmetric kernel; zeros when the context does not run it (cutoff
        mode: (0, cell kernel variant, 0))."""
        ch, tile, bands = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(self.lib.od_sym_choice(self._ctx, ctypes.byref(ch), ctypes.byref(tile), ctypes.byref(bands)),
               "od_sym_choice")
        return ch.value, tile.value, bands.value

    def pair_forces(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        f = np.empty_like(x)
        _check(self.lib.od_pair_forces(self._ctx, _dptr(x), _dptr(f)), "od_pair_forces")
        return f


class Batch:
    """Independent simulations advanced together on one device (od_batch_*).
    cfgs: list of Config; cfgs[s].sim_id should differ between simulations."""

    def __init__(self, cfgs):
        self.lib = load_library()
        arr = (Config * len(cfgs))()
        for s, c in enumerate(cfgs):
            c.struct_size = ctypes.sizeof(Config)
            arr[s] = c
        self.cfgs = list(cfgs)
        self.n_sims = len(cfgs)
        self._b = ctypes.c_void_p()
        _check(self.lib.od_batch_create(arr, self.n_sims, ctypes.byref(self._b)), "od_batch_create")

    def close(self):
        if self._b:
            self.lib.od_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_ions(self, sim, x, v):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        _check(self.lib.od_batch_set_ions(self._b, int(sim), _dptr(x), _dptr(v), x.shape[0]), "od_batch_set_ions")

    def get_ions(self, sim):
        n = self.cfgs[sim].n_ions
        x = np.empty((n, 3))
        v = np.empty((n, 3))
        _check(self.lib.od_batch_get_ions(self._b, int(sim), _dptr(x), _dptr(v)), "od_batch_get_ions")
        return x, v

    def run(self, nsteps, flags=0, records=True):
        """nsteps steps of all simulations; records [nsteps][n_sims][OD_MAX_DUST]."""
        if not records:
            _check(self.lib.od_batch_run(self._b, int(nsteps), None, int(flags)), "od_batch_run")
            return None
        rec = np.zeros((nsteps, self.n_sims, MAX_DUST), dtype=RECORD_DTYPE)
        _check(self.lib.od_batch_run(self._b, int(nsteps), rec.ctypes.data_as(ctypes.c_void_p), int(flags)),
               "od_batch_run")
        return rec

    def step_info(self, sim):
        n = self.cfgs[sim].n_ions
        flags = np.empty(n, dtype=np.uint32)
        face = np.empty(n, dtype=np.uint8)
        _check(self.lib.od_batch_get_step_info(self._b, int(sim), flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                               face.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))),
               "od_batch_get_step_info")
        return flags, face

    def tune(self, reps=3):
        """Autotune the batch's symmetric kernel (see Engine.tune); returns
        (sym_chunk, sym_tile), fixed through the first configuration."""
        ch = ctypes.c_int32()
        tile = ctypes.c_int32()
        _check(self.lib.od_batch_tune(self._b, int(reps), ctypes.byref(ch), ctypes.byref(tile)), "od_batch_tune")
        return ch.value, tile.value

    def pair_timing(self, reset=False):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        pairs = ctypes.c_int64()
        _check(self.lib.od_batch_pair_timing(self._b, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(pairs),
                                             int(bool(reset))), "od_batch_pair_timing")
        return ms.value, n.value, pairs.value
