"""The C-ABI library loads without a GPU, exports every function that
include/opendust_amd.h declares, and its structure sizes match the ctypes
mirrors; argument errors come back as status codes with a message."""

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "opendust_amd.h")


@pytest.fixture(scope="module")
def eng():
    from opendust_amd import engine

    if not os.path.exists(engine.LIB_PATH):
        from opendust_amd import build

        build.build()
    engine.load_library()
    return engine


def declared_functions():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(od_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_the_abi():
    names = declared_functions()
    assert "od_create" in names and "od_run" in names and len(names) >= 15


def test_every_declared_symbol_is_exported(eng):
    lib = ctypes.CDLL(eng.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert set(declared_functions()) == set(eng.EXPORTS)


def test_struct_sizes_match(eng):
    assert eng.check_abi()


def test_invalid_config_is_an_error_not_a_crash(eng):
    cfg = eng.Config()
    cfg.abi_version = eng.ABI_VERSION
    cfg.struct_size = ctypes.sizeof(eng.Config)
    cfg.n_ions = 0
    ctx = ctypes.c_void_p()
    rc = eng.load_library().od_create(ctypes.byref(cfg), ctypes.byref(ctx))
    assert rc == 1 and not ctx
    assert b"n_ions" in eng.load_library().od_last_error()
    cfg.n_ions = 16
    cfg.abi_version = 99
    rc = eng.load_library().od_create(ctypes.byref(cfg), ctypes.byref(ctx))
    assert rc == 1 and b"ABI" in eng.load_library().od_last_error()
    with pytest.raises(RuntimeError):
        eng.Engine(cfg)
    # a caller with another layout of od_config is refused before anything runs
    cfg.abi_version = eng.ABI_VERSION
    cfg.n_ions = 16
    cfg.struct_size = ctypes.sizeof(eng.Config) - 8
    rc = eng.load_library().od_create(ctypes.byref(cfg), ctypes.byref(ctx))
    assert rc == 1 and b"layout" in eng.load_library().od_last_error()


def test_sym_variant_table_and_range(eng):
    v = eng.sym_variants()
    assert v[0] == (8, 3, 0, 1) and (8, 3, 1, 1) in v and (8, 4, 1, 4) in v and len(v) >= 2
    assert all(mt in (4, 8, 16) and 1 <= w <= 8 and vg in (0, 1, 2) and rows in (1, 2, 4) for mt, w, vg, rows in v)
    lib = eng.load_library()
    mt = ctypes.c_int32()
    assert lib.od_sym_variant_info(len(v) + 1, ctypes.byref(mt), None, None, None) == 1
    assert b"variant" in lib.od_last_error()
    cfg = eng.Config()
    cfg.abi_version = eng.ABI_VERSION
    cfg.struct_size = ctypes.sizeof(eng.Config)
    cfg.n_ions = 16
    cfg.sym_tile = len(v) + 1
    ctx = ctypes.c_void_p()
    assert lib.od_create(ctypes.byref(cfg), ctypes.byref(ctx)) == 1 and not ctx
    assert b"sym_tile" in lib.od_last_error()


def test_engine_host_sampler_matches_oracle(eng):
    from oracle import oracle

    s = eng.trap_samples_host(4.0, 9.0, 31, 1000, 300)
    o = np.array([oracle.trap_sample(4.0, 9.0, 31, 1000 + k) for k in range(300)])
    assert np.max(np.abs(s - o)) <= 1e-15 * 9.0


def test_build_info(eng):
    info = eng.load_library().od_build_info()
    assert b"gfx950" in info
    assert info.endswith(b"ABI %d" % eng.ABI_VERSION)


def test_wrappers_are_context_managers(eng):
    """Engine, Batch and ResidentPlasma release their device state on leaving
    a with-block (close() is idempotent)."""
    from opendust_amd import dynamics

    for cls in (eng.Engine, eng.Batch, dynamics.ResidentPlasma):
        assert hasattr(cls, "__enter__") and hasattr(cls, "__exit__") and hasattr(cls, "close")
    e = eng.Engine.__new__(eng.Engine)
    e.lib = eng.load_library()
    e._ctx = ctypes.c_void_p()
    with e as same:
        assert same is e
    e.close()
