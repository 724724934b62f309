"""The C-ABI library loads and exports every entry point include/knc_gpu.h
declares; host-side argument checks answer without touching a GPU."""
import ctypes
import os
import re

import pytest

import knc
from knc import capi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "knc_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(knc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("knc_context_create", "knc_block_insert", "knc_block_remove", "knc_time_step",
                     "knc_advance", "knc_hydro_try_to_primitive", "knc_hydro_intercell_flux"):
        assert required in names


def _c_arity(name):
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    args = re.search(r"\b" + name + r"\s*\(([^;]*?)\)\s*;", text, re.S).group(1).strip()
    return 0 if args in ("", "void") else args.count(",") + 1


def test_integration_notes_bind_every_entry_point():
    """INTEGRATION.md declares every header function as a Rust extern, with
    the C declaration's number of parameters."""
    notes = open(os.path.join(os.path.dirname(HEADER), "..", "INTEGRATION.md")).read()
    rust = {m.group(1): m.group(2) for m in re.finditer(r"pub fn (knc_[a-z0-9_]+)\s*\(([^)]*)\)", notes)}
    missing = [n for n in declared_functions() if n not in rust]
    assert not missing, missing
    for n in declared_functions():
        args = rust[n].strip()
        assert (0 if not args else args.count(",") + 1) == _c_arity(n), n


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_missing_library_fails_loudly(monkeypatch):
    """Without the HIP library the host package raises; it has no CPU
    fallback, and no module of the package imports the oracle."""
    monkeypatch.setattr(capi, "_lib", None)
    monkeypatch.setattr(capi, "LIB_PATH", os.path.join(os.path.dirname(capi.LIB_PATH), "absent.so"))
    with pytest.raises(RuntimeError):
        capi.lib()
    with pytest.raises(RuntimeError):
        knc.Context(knc.relativistic_hydro())
    pkg = os.path.dirname(os.path.abspath(knc.__file__))
    for name in os.listdir(pkg):
        if name.endswith(".py"):
            src = open(os.path.join(pkg, name)).read()
            assert not re.search(r"^\s*(import|from)\s+oracle\b", src, re.M), name


def test_abi_version():
    assert knc.lib().knc_abi_version() == 4


def test_invalid_hydro_is_rejected_before_any_device_call():
    h = knc.relativistic_hydro(plm_theta=2.5)
    out = ctypes.c_void_p()
    rc = knc.lib().knc_context_create(ctypes.byref(h), None, ctypes.byref(out))
    assert rc == 10
    assert b"plm_theta" in knc.lib().knc_last_error_message()
    h = knc.relativistic_hydro(cfl_number=0.9)
    assert knc.lib().knc_context_create(ctypes.byref(h), None, ctypes.byref(out)) == 10
    assert out.value is None


def test_context_errors_are_reported_not_raised_in_c():
    with pytest.raises(RuntimeError):
        knc.Context(knc.relativistic_hydro(runge_kutta_order=4))


def test_invalid_options_are_rejected_before_any_device_call():
    h = knc.relativistic_hydro()
    out = ctypes.c_void_p()
    for bad in (capi.Options(1, 5, 0, 0, 0, 1, 0, 0),       # unknown c2p guess mode
                capi.Options(1, 0, 0, 0, 0, 1, 0, 5)):      # stage_passes must be 0 to 4
        rc = knc.lib().knc_context_create(ctypes.byref(h), ctypes.byref(bad), ctypes.byref(out))
        assert rc == 10 and out.value is None


def test_null_context_arguments_are_errors():
    L = knc.lib()
    assert L.knc_set_excision(None, None) == 10
    assert L.knc_excision_counts(None, None, None) == 10
    assert L.knc_autotune(None, 1, 0, None) == 10
    assert L.knc_advance(None, 1, None, None) == 10
    assert L.knc_block_remove(None, 0) == 10
    rate = ctypes.c_double()
    for chains, waves, iters in ((3, 2, 16), (8, 0, 16), (8, 9, 16), (8, 2, 0)):   # checked before any device call
        assert L.knc_fp64_fma_rate(chains, waves, iters, ctypes.byref(rate), None) == 10
    assert L.knc_fp64_fma_rate(8, 2, 16, None, None) == 10


DRIVER = os.path.join(os.path.dirname(capi.LIB_PATH), "knc_drive")


def test_compiled_driver_is_built_and_checks_its_arguments():
    """knc_drive is a plain C++ caller of the C ABI (no Python, no HIP
    headers); bad arguments are rejected before any device call."""
    import subprocess
    assert os.path.exists(DRIVER)
    r = subprocess.run([DRIVER, "100", "16", "3"], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_compiled_driver_advances_a_mesh():
    import subprocess
    r = subprocess.run([DRIVER, "512", "128", "4", "40", "10", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "G zone-updates/s" in r.stdout and "iteration 50" in r.stdout


@pytest.mark.gpu
def test_compiled_driver_writes_and_reads_a_checkpoint(tmp_path):
    """The compiled caller writes the reference's checkpoint file through
    the C ABI, reads it back and finds every block equal to the device
    state; the independent Python decoder reads the same file."""
    import subprocess

    from knc import cbor
    path = tmp_path / "chkpt.0000.cbor"
    r = subprocess.run([DRIVER, "256", "64", "4", "20", "10", "1", str(path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "4 blocks identical" in r.stdout
    doc = cbor.decode(path.read_bytes())
    assert list(doc) == ["state", "tasks", "config", "version"] and doc["version"] == "knc_drive"
    st = doc["state"]["Relativistic"]
    assert st["iteration"] == [30, 1] and sorted(st["solution"]) == [(k, 0) for k in range(4)]


def _rust_impl(notes, name):
    m = re.search(r"impl GpuHydro for " + name + r"\b.*?\n\}\n", notes, re.S)
    assert m, name
    return m.group(0)


@pytest.mark.parametrize("name", ["RelativisticHydro", "NewtonianHydro"])
def test_integration_hydro_mapping_matches_the_reference_fields(name):
    """The KncHydro built in INTEGRATION.md reads every field of the
    reference's struct with an expression its type allows: f64 fields
    directly, bool fields cast, enum fields through a match on their
    variants (field types from tests/golden/serde_layout.json, read from the
    reference's declarations)."""
    import json
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "serde_layout.json")))
    fields = {f["name"]: f["type"] for f in golden["structs"][name]["fields"]}
    notes = open(os.path.join(os.path.dirname(HEADER), "..", "INTEGRATION.md")).read()
    body = _rust_impl(notes, name)
    used = set(re.findall(r"self\.([a-z_]+)", body))
    assert used == set(fields), (used, fields)
    for f, t in fields.items():
        uses = [(body[max(0, m.start() - 12):m.start()], m.group(1))
                for m in re.finditer(r"self\." + f + r"\b([^,\n]*?)(?=,|\n| \{)", body)]
        assert uses, f
        for pre, post in uses:
            if t == "f64":
                assert post.strip() == "" and pre.endswith(": "), (f, pre, post)
            elif t == "bool":
                assert post.strip() == "as i32", (f, post)
            else:   # an enum: matched on its variants, never cast
                assert "as " not in post and (pre.endswith("match ") or pre.endswith("rk_order(&")), (f, pre, post)
    if "runge_kutta_order" in fields:
        fn = re.search(r"fn rk_order\(.*?\n\}", notes, re.S).group(0)
        assert all(f"RungeKuttaOrder::RK{k} => {k}" in fn for k in (1, 2, 3))
