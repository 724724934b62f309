"""Host logic of knc.presets: the reference's Configuration checks (serde's
deny_unknown_fields and required fields), the key=value overrides of
Configuration::patch_from, the mesh's block extents and grid_blocks.  No GPU
is used."""
import math

import numpy as np
import pytest

from knc import capi, presets
from test_gpu_models import SETUP_HALO, SETUP_JIC, SETUP_KB, SETUP_STAR, SETUP_WIND



def test_setups_parse_to_model_parameters_in_field_order():
    p = presets.parse(SETUP_KB)
    assert p.model_kind == capi.KINETIC_BOMB and p.model_params == [1e-23, 1e18, 1e18, 1e51, 2e32]
    assert p.hydro.system == capi.SRHD and p.hydro.riemann_solver == capi.HLLC and p.hydro.adaptive_time_step == 1
    h = presets.parse(SETUP_HALO)
    assert h.model_params == [1e20, 1e18, 1e18, 1e51, 2e32, 1e22]   # external_medium_density is skipped
    w = presets.parse(SETUP_WIND)
    assert w.model_table is None and w.model_params[8:] == [1.0, 0.01]


def test_yaml_numbers_without_a_dot_are_numbers():
    """PyYAML's safe loader reads 1e20 as a string; the reference's loader as
    a number.  The parser converts, and rejects text that is not a number."""
    cfg = presets.apply_overrides(SETUP_KB, ["model.kinetic_bomb.shell_mass=2e32", "mesh.outer_radius=2e19"])
    assert cfg["model"]["kinetic_bomb"]["shell_mass"] == "2e32"
    p = presets.parse(cfg)
    assert p.model_params[4] == 2e32 and p.mesh["outer_radius"] == 2e19
    bad = presets.apply_overrides(SETUP_KB, ["model.kinetic_bomb.shell_mass=heavy"])
    with pytest.raises(presets.ConfigError, match="shell_mass"):
        presets.parse(bad)


def test_serde_checks_are_restated():
    """deny_unknown_fields and missing fields are errors (the reference's
    setups/jet_in_star.yaml names eta_0 and lacks envelope_mass and
    volume_factor, so it does not load); num_radial_zones is required in 1-D."""
    cfg = presets.apply_overrides(SETUP_STAR, ["model.jet_in_star.eta_0=1.0"])
    with pytest.raises(presets.ConfigError, match="eta_0"):
        presets.parse(cfg)
    cfg = presets.apply_overrides(SETUP_STAR, [])
    del cfg["model"]["jet_in_star"]["volume_factor"]
    with pytest.raises(presets.ConfigError, match="volume_factor"):
        presets.parse(cfg)
    cfg = presets.apply_overrides(SETUP_KB, ["mesh.num_radial_zones=null"])
    with pytest.raises(presets.ConfigError, match="num_radial_zones"):
        presets.parse(cfg)
    cfg = presets.apply_overrides(SETUP_KB, ["hydro.relativistic.runge_kutta_order=RK4"])
    with pytest.raises(presets.ConfigError, match="runge_kutta_order"):
        presets.parse(cfg)


def test_overrides_merge_yaml_files(tmp_path):
    f = tmp_path / "more.yaml"
    f.write_text("mesh:\n  block_size: 8\ncontrol:\n  fold: 3\n")
    cfg = presets.apply_overrides(SETUP_JIC, [str(f), "mesh.num_polar_zones=64"])
    p = presets.parse(cfg)
    assert p.mesh["block_size"] == 8 and p.mesh["num_polar_zones"] == 64 and p.control["fold"] == 3
    assert p.mesh["reference_radius"] == 3e8   # untouched fields kept
    with pytest.raises(presets.ConfigError):
        presets.apply_overrides(SETUP_JIC, ["no_equals_sign"])


def test_grid_blocks_cover_the_outer_excision_surface():
    """Mesh::grid_blocks(t): blocks 0.. while the inner radius is inside
    OES(t); block i starts at R (1 + b)^i."""
    p = presets.parse(SETUP_JIC)
    b = 4 * math.pi / 512
    idx = p.grid_block_indexes(1.0)
    oes = 0.0 + 1.0 * 3e10
    assert idx == list(range(len(idx)))
    assert 3e8 * (1 + b) ** idx[-1] < oes <= 3e8 * (1 + b) ** (idx[-1] + 1)
    k = presets.parse(SETUP_KB)
    assert k.grid_block_indexes(0.0) == [0, 1, 2, 3, 4]          # 1e18 * 2^i < 2e19
    assert k.excision_surface(0, 5.0) == 1e18 and not k.moving_excision
    s = presets.parse(SETUP_STAR)
    assert s.excision_surface(1, 5.0) == 1e12 and s.excision_surface(1, 11.0) == 1e12 + 3e10


def test_ascii_table_groups_of_four(tmp_path):
    f = tmp_path / "t.dat"
    f.write_text("1 2 3 4\n5 6\n7 8\n")
    assert np.array_equal(presets.ascii_table(f), [[1, 2, 3, 4], [5, 6, 7, 8]])
