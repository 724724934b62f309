"""The ctypes mirror of the C ABI (knc.capi) has the same struct layouts as
include/knc_gpu.h: a small C program built with the host compiler reports
sizeof and offsetof of every struct field, which must match ctypes."""
import ctypes
import os
import shutil
import subprocess

import pytest

from knc import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "knc_hydro": (capi.Hydro, ["system", "riemann_solver", "runge_kutta_order", "adaptive_time_step",
                               "gamma_law_index", "plm_theta", "cfl_number", "light_speed"]),
    "knc_options": (capi.Options, ["launch_mode", "c2p_guess", "collect_stats", "rows_per_item", "device",
                                   "use_graph", "stage_threads", "stage_passes"]),
    "knc_error": (capi.Error, ["code", "block", "i", "j", "r", "theta", "value", "conserved"]),
    "knc_c2p_stats": (capi.C2PStats, ["hist", "zones", "fallbacks", "floors", "learned_iters"]),
    "knc_block_data": (capi.BlockData, ["nr", "radial_vertices", "conserved", "scalar_mass"]),
    "knc_excision": (capi.Excision, ["inner_radius", "outer_radius", "inner_speed", "outer_speed", "delay",
                                     "check_every", "source", "user"]),
    "knc_tune_result": (capi.TuneResult, ["stage_threads", "rows_per_item", "stage_passes", "candidates",
                                          "from_cache", "ms_per_stage", "default_ms_per_stage"]),
    "knc_model": (capi.Model, ["kind", "table_rows", "params", "table"]),
    "knc_mesh": (capi.Mesh, ["reference_radius", "num_radial_zones", "num_polar_zones", "block_size"]),
    "knc_record_meta": (capi.RecordMeta, ["config", "config_len", "tasks", "tasks_len", "version"]),
    "knc_checkpoint_info": (capi.CheckpointInfo, ["system", "num_blocks", "time", "iteration_numer",
                                                  "iteration_denom", "config", "config_len", "tasks", "tasks_len",
                                                  "version"]),
}

ENUMS = {"KNC_EULER": capi.EULER, "KNC_SRHD": capi.SRHD, "KNC_HLLE": capi.HLLE, "KNC_HLLC": capi.HLLC,
         "KNC_BOUNDARY_FIXED": capi.FIXED, "KNC_BOUNDARY_OUTFLOW": capi.OUTFLOW, "KNC_SIDE_INNER": capi.INNER,
         "KNC_SIDE_OUTER": capi.OUTER, "KNC_LAUNCH_PER_BLOCK": capi.PER_BLOCK, "KNC_LAUNCH_BATCHED": capi.BATCHED,
         "KNC_GUESS_PLAIN": capi.GUESS_PLAIN, "KNC_GUESS_LEARNED": capi.GUESS_LEARNED,
         "KNC_GUESS_LEARNED_STEP": capi.GUESS_LEARNED_STEP, "KNC_GUESS_PREVIOUS": capi.GUESS_PREVIOUS,
         "KNC_GUESS_LEARNED_FOLD": capi.GUESS_LEARNED_FOLD, "KNC_ERR_MODEL": capi.ERR_MODEL,
         "KNC_MODEL_KINETIC_BOMB": capi.KINETIC_BOMB, "KNC_MODEL_JET_IN_CLOUD": capi.JET_IN_CLOUD,
         "KNC_MODEL_JET_IN_STAR": capi.JET_IN_STAR, "KNC_MODEL_HALO_KILONOVA": capi.HALO_KILONOVA,
         "KNC_MODEL_WIND_SHOCK": capi.WIND_SHOCK, "KNC_MODEL_NPARAM": capi.MODEL_NPARAM,
         "KNC_ERR_IO": capi.ERR_IO, "KNC_ERR_FORMAT": capi.ERR_FORMAT}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    d = tmp_path_factory.mktemp("abi")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "knc_gpu.h"', "int main(void) {"]
    for name, (_, fields) in STRUCTS.items():
        lines.append(f'  printf("{name} size %zu\\n", sizeof({name}));')
        for f in fields:
            lines.append(f'  printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    for name in ENUMS:
        lines.append(f'  printf("{name} %d\\n", (int){name});')
    lines.append("  return 0;\n}")
    src = d / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "layout"
    subprocess.run([cc, "-std=c99", "-Wall", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    vals = {}
    for line in out.splitlines():
        k, v = line.rsplit(" ", 1)
        vals[k] = int(v)
    return vals


@pytest.mark.parametrize("name", sorted(STRUCTS))
def test_struct_layout_matches_header(c_layout, name):
    cls, fields = STRUCTS[name]
    assert ctypes.sizeof(cls) == c_layout[f"{name} size"]
    for f in fields:
        assert getattr(cls, f).offset == c_layout[f"{name}.{f}"], f


def test_enum_values_match_header(c_layout):
    for name, value in ENUMS.items():
        assert c_layout[name] == value, name
