"""CPU checks of the reference's file formats (src/io.rs): the CBOR codecs
against the examples of RFC 8949 Appendix A, the serialised field order of
the configuration and tasks against the reference's struct declarations
(tests/golden/serde_layout.json, read from its sources), and the library's
checkpoint reader (knc_checkpoint_open, host code only) on files written by
the independent Python encoder, with every float width, indefinite
containers and shuffled map keys."""
import json
import math
import os
import struct

import numpy as np
import pytest

from knc import app as kapp
from knc import capi, cbor, presets

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def _value(kind, v):
    if kind == "float":
        return float(v)
    if kind == "bytes":
        return bytes.fromhex(v)
    if kind == "pairs":
        return {k: x for k, x in v}
    return v


def _same(a, b):
    if isinstance(a, float) and isinstance(b, float):
        return (math.isnan(a) and math.isnan(b)) or (a == b and math.copysign(1, a) == math.copysign(1, b))
    return a == b


def test_cbor_rfc8949_examples():
    g = _golden("cbor_rfc8949.json")
    for kind, v, hexs in g["preferred"]:
        x = _value(kind, v)
        assert cbor.encode(x).hex() == hexs, (kind, v)
        assert _same(cbor.decode(bytes.fromhex(hexs)), x), (kind, v)
    for kind, v, hexs in g["decode_only"]:
        assert _same(cbor.decode(bytes.fromhex(hexs)), _value(kind, v)), hexs


def test_cbor_float_width_boundaries():
    """Shortest exact width: the binary16 normal/subnormal edges, a value one
    ulp off each, binary32-only values and binary64-only values."""
    cases = {2.0 ** -24: 3, 2.0 ** -25: 5, 3 * 2.0 ** -24: 3, 2.0 ** -14 * (1 + 2.0 ** -10): 3,
             2.0 ** -14 * (1 + 2.0 ** -11): 5, 65504.0: 3, 65505.0: 5, 65536.0: 5, 2.0 ** 127: 5,
             2.0 ** 128: 9, 1.0 + 2.0 ** -23: 5, 1.0 + 2.0 ** -24: 9, 2.0 ** -149: 5, 2.0 ** -150: 9,
             1e-320: 9, 0.1: 9, 4096.0: 3, 4100.0: 3, 4098.0: 5, 1023.5: 3, 1024.5: 5}
    for x, n in cases.items():
        for s in (x, -x):
            b = cbor.encode(s)
            assert len(b) == n, (s, b.hex())
            assert cbor.decode(b) == s


def test_cbor_rejects_malformed():
    for bad in ("", "18", "1b00", "62c3", "82 01".replace(" ", ""), "a1 01".replace(" ", ""), "ff", "1c", "9f01"):
        with pytest.raises(cbor.CBORError):
            cbor.decode(bytes.fromhex(bad))
    with pytest.raises(cbor.CBORError):
        cbor.decode(bytes.fromhex("0102"))            # trailing bytes


# -- the configuration and tasks items --------------------------------------
def _serialised(struct_name):
    fields = _golden("serde_layout.json")["structs"][struct_name]["fields"]
    return [f["name"] for f in fields if not any(a.startswith("skip") or ", skip" in a for a in f["serde"])]


MODEL_STRUCTS = {"kinetic_bomb": "KineticBomb", "jet_in_cloud": "JetInCloud", "jet_in_star": "JetInStar",
                 "halo_kilonova": "HaloKilonova", "wind_shock": "WindShock"}


def _preset(name, overrides=()):
    return presets.parse(presets.apply_overrides(presets.SETUPS[name], overrides))


def test_configuration_serialises_in_declaration_order():
    layout = _golden("serde_layout.json")
    assert list(_preset("kinetic_bomb").to_serde()) == _serialised("Configuration")
    snake = lambda s: "".join("_" + c.lower() if c.isupper() else c for c in s).lstrip("_")
    assert layout["enums"]["AnyHydro"]["rename_all"] == "snake_case"
    assert {snake(v) for v in layout["enums"]["AnyModel"]["variants"]} == set(MODEL_STRUCTS)
    for name in ("kinetic_bomb", "jet_in_cloud", "halo_kilonova", "wind_shock"):
        s = _preset(name).to_serde()
        assert list(s["hydro"]["relativistic"]) == _serialised("RelativisticHydro")
        assert list(s["model"][name]) == _serialised(MODEL_STRUCTS[name]), name
        assert list(s["mesh"]) == _serialised("Mesh")
        assert list(s["control"]) == _serialised("Control")
    newton = presets.apply_overrides(presets.SETUPS["kinetic_bomb"], [])
    newton["hydro"] = {"newtonian": {"gamma_law_index": 1.4, "plm_theta": 1.5, "cfl_number": 0.3,
                                     "runge_kutta_order": "RK1"}}
    assert list(presets.parse(newton).to_serde()["hydro"]["newtonian"]) == _serialised("NewtonianHydro")
    t = kapp.Tasks(0.0).to_serde()
    assert list(t) == _serialised("Tasks")
    assert all(list(v) == _serialised("RecurringTask") for v in t.values())


def test_configuration_types_and_round_trip():
    """f64 fields are floats and usize fields ints even where the setup gives
    1 or 1e20; absent options are None; the encoded item parses back to the
    same configuration."""
    p = _preset("jet_in_cloud", ["control.products_interval=1", "mesh.excision_delay=2"])
    s = p.to_serde()
    assert s["control"]["products_interval"] == 1.0 and isinstance(s["control"]["products_interval"], float)
    assert isinstance(s["mesh"]["excision_delay"], float) and s["mesh"]["num_radial_zones"] is None
    assert isinstance(s["control"]["fold"], int) and s["control"]["num_threads"] is None
    assert s["hydro"]["relativistic"]["adaptive_time_step"] is False
    item = cbor.encode(s)
    back = presets.parse(cbor.decode(item))
    assert back.to_serde() == s and cbor.encode(back.to_serde()) == item
    # floats in their shortest exact width: 0.1 needs binary64, 4.0 binary16
    assert cbor.encode({"cfl_number": 0.4})[-9:] == b"\xfb" + struct.pack(">d", 0.4)
    assert cbor.encode(s["control"]["start_time"]) == b"\xf9\x3c\x00"


def test_named_setups_load_or_fail_like_the_reference():
    for name in presets.SETUPS:
        if name == "jet_in_star":
            with pytest.raises(presets.ConfigError):
                _preset(name)
        else:
            presets.validate(_preset(name))
    with pytest.raises(presets.ConfigError):
        presets.validate(_preset("kinetic_bomb", ["model.kinetic_bomb.shell_mass=1e20"]))   # v > c/4
    with pytest.raises(presets.ConfigError):
        presets.validate(_preset("jet_in_cloud", ["mesh.num_polar_zones=8"]))


# -- the library's checkpoint reader (host code, no GPU) ---------------------
def _ndarray(shape, data):
    return {"v": 1, "dim": list(shape), "data": data}


def _f(x, width):
    if width == 16:
        return cbor.Raw(b"\xf9" + struct.pack(">e", x))
    if width == 32:
        return cbor.Raw(b"\xfa" + struct.pack(">f", x))
    return cbor.Raw(b"\xfb" + struct.pack(">d", x))


def _app_item(blocks, system="Relativistic", time=2.5, iteration=(7, 1), widths=(64,), key_order=None,
              config=None, tasks=None):
    sol = {}
    for index, (cons, sm) in blocks.items():
        nr, nq = sm.shape
        w = lambda k: widths[k % len(widths)]
        cdata = [[_f(float(v), w(4 * k + q)) for q, v in enumerate(cons.reshape(-1, 4)[k])] for k in range(nr * nq)]
        sdata = [_f(float(v), w(k)) for k, v in enumerate(sm.reshape(-1))]
        sol[(index, 0)] = {"conserved": _ndarray((nr, nq), cdata), "scalar_mass": _ndarray((nr, nq), sdata)}
    state = {"time": time, "iteration": list(iteration), "solution": sol}
    app = {"state": {system: state},
           "tasks": tasks if tasks is not None else kapp.Tasks(0.0).to_serde(),
           "config": config if config is not None else _preset("kinetic_bomb").to_serde(),
           "version": "v0.1.2 test"}
    if key_order:
        app = {k: app[k] for k in key_order}
    return cbor.encode(app)


def _blocks(seed, shapes):
    rng = np.random.default_rng(seed)
    out = {}
    for index, (nr, nq) in shapes.items():
        # values exact in binary16 so every width holds them
        cons = np.round(rng.uniform(-8, 8, (nr, nq, 4)) * 64) / 64
        sm = np.round(rng.uniform(0, 4, (nr, nq)) * 64) / 64
        out[index] = (cons, sm)
    return out


def test_checkpoint_reader_on_python_written_files(tmp_path):
    blocks = _blocks(1, {-2: (3, 5), 0: (4, 5), 7: (2, 5)})
    tasks = {"write_checkpoint": {"count": 3, "next_time": 0.75}, "write_products": {"count": 0, "next_time": 0.0},
             "iteration_message": {"count": 11, "next_time": 0.0}, "report_progress": {"count": 0, "next_time": 0.0}}
    for n, widths in enumerate([(64,), (16,), (32,), (16, 32, 64)]):
        path = tmp_path / f"chkpt.{n:04d}.cbor"
        order = ["version", "config", "state", "tasks"] if n % 2 else None
        path.write_bytes(_app_item(blocks, widths=widths, key_order=order, tasks=tasks))
        with capi.Checkpoint(path) as ck:
            assert ck.system == capi.SRHD and ck.time == 2.5 and ck.iteration == (7, 1)
            assert ck.version == "v0.1.2 test"
            assert [b[0] for b in ck.blocks] == [-2, 0, 7]
            assert cbor.decode(ck.tasks) == tasks
            assert cbor.decode(ck.config) == _preset("kinetic_bomb").to_serde()
            for k, (index, nr, nq) in enumerate(ck.blocks):
                cons, sm = ck.block_data(k)
                assert (nr, nq) == blocks[index][1].shape
                assert np.array_equal(cons, blocks[index][0]) and np.array_equal(sm, blocks[index][1])


def test_checkpoint_reader_indefinite_containers_and_newtonian(tmp_path):
    """A map and arrays of indefinite length, a tag and an unknown field are
    accepted as serde accepts them; the Newtonian state variant is Euler."""
    blocks = _blocks(2, {3: (2, 3)})
    item = bytearray(_app_item(blocks, system="Newtonian", iteration=(40, 1)))
    body = cbor.decode(bytes(item))
    st = body["state"]["Newtonian"]
    cons, sm = blocks[3]
    head = b"\xbf"           # indefinite map of the App
    parts = [cbor.encode("extra"), cbor.encode([1, 2]),
             cbor.encode("state"), cbor.encode({"Newtonian": {"time": st["time"], "iteration": [40, 1]}})[:-0]]
    # rebuild the state with an indefinite solution map and a tagged time
    sol = b"\xbf" + cbor.encode([3, 0]) + cbor.encode(
        {"scalar_mass": _ndarray((2, 3), [float(v) for v in sm.reshape(-1)]),
         "conserved": {"data": cbor.Raw(b"\x9f" + b"".join(cbor.encode([float(v) for v in r]) for r in
                                                           cons.reshape(-1, 4)) + b"\xff"),
                       "dim": [2, 3], "v": 1}}) + b"\xff"
    state = (b"\xbf" + cbor.encode("Newtonian") + b"\xa3" + cbor.encode("iteration") + cbor.encode([40, 1]) +
             cbor.encode("solution") + sol + cbor.encode("time") + b"\xc1" + cbor.encode(st["time"]) + b"\xff")
    doc = (head + parts[0] + parts[1] + cbor.encode("state") + state + cbor.encode("tasks") +
           cbor.encode(body["tasks"]) + cbor.encode("config") + cbor.encode(body["config"]) +
           cbor.encode("version") + b"\x7f" + cbor.encode("v0.1")[0:0] + b"\x62v0\x62.1" + b"\xff" + b"\xff")
    path = tmp_path / "chkpt.0000.cbor"
    path.write_bytes(doc)
    with capi.Checkpoint(path) as ck:
        assert ck.system == capi.EULER and ck.iteration == (40, 1) and ck.version == "v0.1"
        c2, s2 = ck.block_data(0)
        assert np.array_equal(c2, cons) and np.array_equal(s2, sm)


@pytest.mark.parametrize("mutate", ["truncate", "variant", "rank", "short", "key", "dup", "trailing", "version"])
def test_checkpoint_reader_rejects_malformed_files(tmp_path, mutate):
    blocks = _blocks(3, {0: (2, 2), 1: (2, 2)})
    good = _app_item(blocks)
    body = cbor.decode(good)
    st = body["state"]["Relativistic"]
    if mutate == "truncate":
        doc = good[:-7]
    elif mutate == "variant":
        body["state"] = {"Quantum": st}
        doc = cbor.encode(body)
    elif mutate == "rank":
        st["solution"][(0, 0)]["conserved"]["dim"] = [4]
        doc = cbor.encode(body)
    elif mutate == "short":
        st["solution"][(1, 0)]["scalar_mass"]["data"].pop()
        doc = cbor.encode(body)
    elif mutate == "key":
        st["solution"][(0, 5)] = st["solution"].pop((0, 0))
        doc = cbor.encode(body)
    elif mutate == "dup":
        doc = good.replace(cbor.encode([1, 0]), cbor.encode([0, 0]))
    elif mutate == "trailing":
        doc = good + b"\x00"
    else:
        st["solution"][(0, 0)]["conserved"]["v"] = 2
        doc = cbor.encode(body)
    path = tmp_path / "bad.cbor"
    path.write_bytes(doc)
    with pytest.raises(RuntimeError, match="knc error 16"):
        capi.Checkpoint(path)


def test_checkpoint_reader_missing_file(tmp_path):
    with pytest.raises(RuntimeError, match="knc error 15"):
        capi.Checkpoint(tmp_path / "none.cbor")


def test_array_fast_path_and_records_reader(tmp_path):
    """decode(arrays=True) hands long number arrays (and arrays of number
    tuples) to the library's decoder, in any float width; the result equals
    the pure-Python decode.  knc.records reads Products and App files into
    numpy with the reference's layout."""
    from knc import records
    rng = np.random.default_rng(4)
    nr, nq = 40, 8
    prim = np.round(rng.uniform(-4, 4, (nr, nq, 4)) * 256) / 256
    prim[:, :, 1] = 0.0                                      # u_theta = 0 -> binary16 zeros among binary64s
    prim[::3, :, 3] = rng.uniform(0, 1, (len(range(0, nr, 3)), nq))
    scal = rng.uniform(0, 1, (nr, nq))
    rv = np.geomspace(1.0, 2.0, nr + 1)
    qv = np.linspace(0.0, np.pi, nq + 1)
    cfg = _preset("kinetic_bomb").to_serde()
    doc = {"time": 3.25, "blocks": {(2, 0): {"radial_vertices": _ndarray((nr + 1,), rv.tolist()),
                                             "polar_vertices": _ndarray((nq + 1,), qv.tolist()),
                                             "primitive": _ndarray((nr, nq), prim.reshape(-1, 4).tolist()),
                                             "scalar": _ndarray((nr, nq), scal.reshape(-1).tolist())}},
           "config": cfg, "version": "v0.1.2 test"}
    raw = cbor.encode(doc)
    fast = cbor.decode(raw, arrays=True)
    slow = cbor.decode(raw)
    blk = fast["blocks"][(2, 0)]
    assert isinstance(blk["primitive"]["data"], np.ndarray) and blk["primitive"]["data"].shape == (nr * nq, 4)
    assert np.array_equal(blk["primitive"]["data"], np.asarray(slow["blocks"][(2, 0)]["primitive"]["data"]))
    assert np.array_equal(blk["scalar"]["data"], np.asarray(slow["blocks"][(2, 0)]["scalar"]["data"]))
    path = tmp_path / "prods.0000.cbor"
    path.write_bytes(raw)
    p = records.read_products(path)
    b = p["blocks"][(2, 0)]
    assert p["time"] == 3.25 and p["config"] == cfg
    assert np.array_equal(b["primitive"], prim) and np.array_equal(b["scalar"], scal)
    assert np.array_equal(b["radial_vertices"], rv) and np.array_equal(b["polar_vertices"], qv)
    assert np.array_equal(records.radial_profile(p, 3, "gas_pressure"), prim[:, 3, 3])
    assert np.array_equal(records.radial_profile(p, 0, "vertices"), rv[:-1])
    # a checkpoint, in every float width
    blocks = _blocks(5, {0: (16, 8), 1: (16, 8)})
    ck = tmp_path / "chkpt.0000.cbor"
    ck.write_bytes(_app_item(blocks, widths=(16, 32, 64)))
    a = records.read_app(ck)
    assert a["system"] == "Relativistic" and a["iteration"] == (7, 1) and a["time"] == 2.5
    for k, (c, m) in blocks.items():
        assert np.array_equal(a["blocks"][(k, 0)][0], c) and np.array_equal(a["blocks"][(k, 0)][1], m)
