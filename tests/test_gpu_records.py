"""GPU checks of the reference's output files written from HBM
(knc_write_checkpoint, knc_write_products) and of restarts from them
(knc_checkpoint_load), and of the driver (knc.app: main.rs::run and
side_effects).

- A tiny state's checkpoint equals, byte for byte, the App assembled by hand
  from the serde data model (struct = map of fields in declaration order,
  enum = {variant: value}, ndarray = {v, dim, data}, shortest floats).
- A setup's checkpoint decodes (independent Python decoder) to the
  reference's field order and to the resident state bit for bit, and
  re-encodes to the same bytes.
- A run restarted from a checkpoint continues bit for bit like the run that
  wrote it, also across excision births and removals.
- Products equal the oracle's BlockProducts of the same state.
"""
import json
import math
import os

import numpy as np
import pytest

import knc
import oracle
from knc import app as kapp
from knc import capi, cbor, presets
from util import oracle_hydro_from_capi

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLDEN, "serde_layout.json")) as f:
    LAYOUT = json.load(f)


def _fields(name):
    return [f["name"] for f in LAYOUT["structs"][name]["fields"]
            if not any(a.startswith("skip") or ", skip" in a for a in f["serde"])]


JIC_EXCISION = ["mesh.num_polar_zones=32", "mesh.inner_radius=3e8", "mesh.outer_radius=3e10",
                "mesh.excision_delay=1.0", "mesh.inner_excision_speed=3e10", "mesh.outer_excision_speed=3e12"]


def test_checkpoint_bytes_known_answer(tmp_path):
    print("checkpoint known answer: 1 block 2 x 1", flush=True)
    ctx = capi.Context(knc.relativistic_hydro())
    cons = np.array([[[1.0, 0.5, 0.0, 2.0]], [[0.25, -0.5, 0.0, 1.5]]])
    sm = np.array([[1.0], [0.125]])
    ctx.insert_block(0, [1.0, 2.0, 3.0], [math.pi / 2 - 0.5, math.pi / 2 + 0.5], cons, sm)
    ctx.set_time(1.5, 3)
    path = tmp_path / "chkpt.0000.cbor"
    ctx.write_checkpoint(path, cbor.encode({"c": 1}), cbor.encode({"t": 2}), "v")
    ndarray_head = "a3" "6176" "01" "63" "64696d" "820201" "64" "64617461" "82"
    expected = ("a4"
                "65" "7374617465" "a1" "6c" "52656c617469766973746963" "a3"     # state: {Relativistic: State
                "64" "74696d65" "f93e00"                                       # time: 1.5
                "69" "697465726174696f6e" "820301"                             # iteration: [3, 1]
                "68" "736f6c7574696f6e" "a1" "820000" "a2"                     # solution: {[0, 0]: BlockState
                "69" "636f6e736572766564" + ndarray_head +                     # conserved
                "84" "f93c00" "f93800" "f90000" "f94000"
                "84" "f93400" "f9b800" "f90000" "f93e00"
                "6b" "7363616c61725f6d617373" + ndarray_head + "f93c00" "f93000"   # scalar_mass
                "65" "7461736b73" "a1" "6174" "02"                             # tasks (the caller's item)
                "66" "636f6e666967" "a1" "6163" "01"                           # config
                "67" "76657273696f6e" "61" "76")                               # version
    assert path.read_bytes().hex() == expected
    ctx.close()


def _jic_app(**options):
    return kapp.App.from_config(presets.SETUPS["jet_in_cloud"], JIC_EXCISION, launch_mode=capi.BATCHED, **options)


def test_checkpoint_layout_state_and_reencoding(tmp_path):
    print("checkpoint of jet_in_cloud 32 polar zones after 4 folds of 10", flush=True)
    a = _jic_app()
    for _ in range(4):
        a.ctx.advance(10)
    path = tmp_path / "chkpt.0003.cbor"
    a.write_checkpoint(path)
    raw = path.read_bytes()
    doc = cbor.decode(raw)
    assert list(doc) == _fields("App")
    assert list(doc["state"]) == ["Relativistic"]
    st = doc["state"]["Relativistic"]
    assert list(st) == _fields("State")
    assert st["time"] == a.ctx.time and st["iteration"] == [a.ctx.iteration, 1]
    assert doc["version"] == kapp.VERSION
    assert doc["config"] == a.preset.to_serde() and doc["tasks"] == a.tasks.to_serde()
    assert sorted(k[0] for k in st["solution"]) == a.ctx.block_indexes()
    for (index, j), bs in st["solution"].items():
        assert j == 0 and list(bs) == _fields("BlockState")
        cons, sm = a.ctx.download_block(index)
        for name, want in (("conserved", cons), ("scalar_mass", sm)):
            arr = bs[name]
            assert list(arr) == ["v", "dim", "data"] and arr["v"] == 1 and arr["dim"] == list(sm.shape)
            assert np.array_equal(np.asarray(arr["data"], dtype=np.float64).reshape(want.shape), want)
    # both encoders write the shortest form, so the Python encoder reproduces the file
    assert cbor.encode(doc) == raw
    a.ctx.close()


def test_restart_from_checkpoint_continues_bit_for_bit(tmp_path):
    """Run A advances, writes a checkpoint, advances on; run B starts from the
    checkpoint and advances as far: the states, times, iterations and block
    sets agree exactly, across excision removals and births."""
    print("restart: jet_in_cloud 32 polar zones, 3 folds, checkpoint, 3 more", flush=True)
    a = _jic_app()
    for _ in range(3):
        a.ctx.advance(10)
    path = tmp_path / "chkpt.0000.cbor"
    a.write_checkpoint(path)
    b = kapp.App.from_checkpoint(path, launch_mode=capi.BATCHED)
    assert b.ctx.block_indexes() == a.ctx.block_indexes()
    assert (b.ctx.time, b.ctx.iteration) == (a.ctx.time, a.ctx.iteration)
    for _ in range(3):
        a.ctx.advance(10)
        b.ctx.advance(10)
    removed, added = a.ctx.excision_counts()
    assert removed > 0 and added > 0
    assert b.ctx.block_indexes() == a.ctx.block_indexes()
    assert (b.ctx.time, b.ctx.iteration) == (a.ctx.time, a.ctx.iteration)
    for i in a.ctx.block_indexes():
        ca, sa = a.ctx.download_block(i)
        cb, sb = b.ctx.download_block(i)
        assert np.array_equal(ca, cb) and np.array_equal(sa, sb), i
    a.ctx.close()
    b.ctx.close()


def test_products_match_oracle(tmp_path):
    print("products of jet_in_cloud 32 polar zones after 2 folds", flush=True)
    a = _jic_app()
    a.ctx.advance(10)
    a.ctx.advance(10)
    path = tmp_path / "prods.0000.cbor"
    a.write_products(path)
    doc = cbor.decode(path.read_bytes())
    assert list(doc) == _fields("Products")
    assert doc["time"] == a.ctx.time and doc["config"] == a.preset.to_serde()
    h = oracle_hydro_from_capi(a.preset.hydro)
    worst = 0.0
    for (index, _), bp in doc["blocks"].items():
        assert list(bp) == _fields("BlockProducts")
        rv, qv = a.ctx.vertices[index]
        assert np.array_equal(bp["radial_vertices"]["data"], rv) and bp["radial_vertices"]["dim"] == [rv.size]
        assert np.array_equal(bp["polar_vertices"]["data"], qv)
        cons, sm = a.ctx.download_block(index)
        prim, scal = oracle.products(h, rv, qv, cons, sm)
        got = np.asarray(bp["primitive"]["data"], dtype=np.float64).reshape(prim.shape)
        gs = np.asarray(bp["scalar"]["data"], dtype=np.float64).reshape(scal.shape)
        assert np.array_equal(gs, scal)
        worst = max(worst, float(np.max(np.abs(got - prim) / np.maximum(np.abs(prim), 1e-300))))
    # the same c2p and operation order on both sides: equal to the last bit
    # up to the device's own rounding of sqrt/division (checked bitwise elsewhere)
    assert worst <= 1e-14, worst
    a.ctx.close()


def test_products_report_a_failed_zone_and_write_nothing(tmp_path):
    print("products with a negative density zone", flush=True)
    ctx = capi.Context(knc.relativistic_hydro())
    cons = np.zeros((4, 1, 4))
    cons[..., 0] = 1.0
    cons[..., 3] = 1.0
    cons[2, 0, 0] = -1.0
    ctx.insert_block(5, [1.0, 2.0, 3.0, 4.0, 5.0], [1.0, 2.0], cons, np.ones((4, 1)))
    path = tmp_path / "prods.0000.cbor"
    with pytest.raises(capi.HydroError) as e:
        ctx.write_products(path, cbor.encode({}), "v")
    assert e.value.kind == "NegativeDensity" and (e.value.block, e.value.i, e.value.j) == (5, 2, 0)
    assert not path.exists()
    ctx.close()


def test_driver_runs_writes_and_restarts(tmp_path):
    """knc.app.main on kinetic_bomb cut to 64-zone blocks: the checkpoint
    cadence of side_effects, the iteration message, and a restart from the
    middle checkpoint that ends in the same final checkpoint."""
    print("driver: kinetic_bomb 64-zone blocks to t = 3e7", flush=True)
    out1 = tmp_path / "run1"
    common = ["mesh.num_radial_zones=64", "mesh.block_size=64", "control.final_time=3e7",
              "control.checkpoint_interval=1e7", "control.fold=5"]
    lines = []
    assert kapp.main(["kinetic_bomb"] + common + [f"control.output_directory={out1}"], out=lines.append) == 0
    names = sorted(p.name for p in out1.iterdir())
    assert names[0] == "chkpt.0000.cbor" and len(names) >= 4
    assert any(l.startswith("[") and "Mzps=" in l for l in lines)
    assert f"write {out1}/chkpt.0000.cbor" in lines
    last = out1 / names[-1]
    with capi.Checkpoint(last) as ck:
        t_end = ck.time
        final = [ck.block_data(k) for k in range(len(ck.blocks))]
    assert t_end >= 3e7
    # restart from chkpt.0001 into a second directory
    out2 = tmp_path / "run2"
    lines2 = []
    assert kapp.main([str(out1 / "chkpt.0001.cbor"), f"control.output_directory={out2}"], out=lines2.append) == 0
    with capi.Checkpoint(out2 / names[-1]) as ck:
        assert ck.time == t_end
        for k in range(len(ck.blocks)):
            c, s = ck.block_data(k)
            assert np.array_equal(c, final[k][0]) and np.array_equal(s, final[k][1])
