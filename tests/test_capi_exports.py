"""The C-ABI libraries load and export every symbol include/*.h declares (CPU:
no compute call is made)."""
import ctypes
import os
import re
import subprocess

from pincosimulator_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"^\s*(?:[A-Za-z_][\w\s\*]*?\b)?(twmc_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_header_declares_what_the_bindings_list():
    names = declared("twmc_hip.h")
    assert set(names) == set(_lib.EXPORTS), set(names) ^ set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared("twmc_hip.h"):
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    for name in declared("twmc_hip.h"):
        assert re.search(r"\bT " + name + r"\b", out), name


def test_abi_version_and_error_paths_without_gpu():
    L = _lib.load()
    assert L.twmc_abi_version() == 1
    # invalid arguments are rejected before any device work
    assert L.twmc_create(None, 0, None) == -1
    assert L.twmc_last_error(None)
    assert L.twmc_batch_advance(None, 1) == -1


def test_warm_start_and_sweep_entry_points_reject_bad_arguments_without_gpu():
    import ctypes as C
    L = _lib.load()
    ctx = C.c_void_p()
    bad = _lib.MlpConfig(n_in=0, n_hidden=8, n_hidden_layers=1, n_out=2, seed=0)
    assert L.twmc_mlp_create(C.byref(bad), 0, C.byref(ctx)) == -1
    assert b"shape" in L.twmc_mlp_last_error(None)
    assert L.twmc_mlp_create(None, 0, C.byref(ctx)) == -1
    assert L.twmc_mlp_forward(None, None, 0, None) == -1
    assert L.twmc_mlp_train(None, None, None, None, 0, None, None) == -1
    assert L.twmc_sweep_create(None, 0, 0.1, 1, 1, 0, C.byref(ctx)) == -1
    assert L.twmc_sweep_run(None, 0.0, 1.0, 0, None, None, None, None, None) == -1
    assert L.twmc_sweep_mode_bins(None, 0, None) == -1


def test_kernel_ids_match_the_header():
    """Every TWMC_KERNEL_* of the header has the same value and a name in the
    Python bindings, and the Engine accepts each name."""
    from pincosimulator_amd import engine
    text = open(os.path.join(ROOT, "include", "twmc_hip.h")).read()
    ids = {m.group(1).lower(): int(m.group(2)) for m in re.finditer(r"TWMC_KERNEL_(\w+)\s*=\s*(\d+)", text)}
    assert ids and set(ids.values()) == set(_lib.KERNEL_NAMES)
    for name, value in ids.items():
        assert _lib.KERNEL_NAMES[value] == name
        assert engine._KERNELS[name] == value


def _compile(snippet):
    src = '#include "twmc_hip.h"\n' + snippet
    return subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-x", "c++", "-I", os.path.join(ROOT, "include"),
                           "-"], input=src, capture_output=True, text=True)


def test_handles_are_distinct_types():
    """An engine, sweep or network handle passed where another kind is expected does not compile."""
    ok = _compile("void f(twmc_ctx* c, twmc_sweep_ctx* s, twmc_mlp_ctx* m) {"
                  " twmc_batch_sync(c); twmc_sweep_destroy(s); twmc_mlp_destroy(m); }")
    assert ok.returncode == 0, ok.stderr
    for bad in ("void f(twmc_sweep_ctx* s) { twmc_run(s, 0, 0); }",
                "void f(twmc_mlp_ctx* m) { twmc_batch_sync(m); }",
                "void f(twmc_ctx* c) { twmc_sweep_destroy(c); }"):
        r = _compile(bad)
        assert r.returncode != 0 and "twmc_" in r.stderr, bad
