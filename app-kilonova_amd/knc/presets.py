"""The reference's setups (``setups/*.yaml``) run on the GPU.

A setup is the reference's ``Configuration`` (``src/app.rs:156-172``): a
``hydro`` section (``relativistic`` or ``newtonian``, ``src/app.rs:79-84``),
a ``model`` section naming one of the five ``InitialModel``s
(``src/app.rs:66-74``), a ``mesh`` (``src/mesh.rs:84-113``) and a
``control`` section (``src/app.rs:105-145``).  :func:`load` parses and
checks it the way serde would (unknown or missing fields are errors),
:func:`apply_overrides` applies the command-line overrides of
``Configuration::patch_from`` (``src/app.rs:285-294``: ``key.path=value``
with the value parsed as YAML, or a ``.yaml`` file merged in), and
:func:`context_from_preset` builds the state of ``App::from_config``
(``src/app.rs:316-327``) on the device: every block of
``Mesh::grid_blocks(start_time)`` from the model, model ghost rows on both
sides and, with moving excision surfaces, blocks born from the model --
all evaluated on the GPU (``knc_block_insert_model``,
``knc_set_boundary_model``, ``knc_set_birth_model``).
"""
import copy
import math
from dataclasses import dataclass, field

import numpy as np
import yaml

from . import capi

# Field order of each model struct (the knc_model params order) and the
# fields serde fills with a default when omitted.
MODELS = {
    "kinetic_bomb": (capi.KINETIC_BOMB, ["external_medium_density", "launch_radius", "shell_thickness",
                                         "kinetic_energy", "shell_mass"], {}),
    "jet_in_cloud": (capi.JET_IN_CLOUD, ["cloud_mass", "engine_delay", "engine_duration", "engine_strength",
                                         "engine_theta", "engine_u", "envelop_m1", "envelop_slowest_beta",
                                         "envelop_fastest_beta", "envelop_psi"], {}),
    "jet_in_star": (capi.JET_IN_STAR, ["star_mass", "engine_duration", "engine_energy", "engine_theta",
                                       "engine_u", "envelope_radius", "envelope_mass", "volume_factor"], {}),
    "halo_kilonova": (capi.HALO_KILONOVA, ["altitude", "launch_radius", "shell_thickness", "kinetic_energy",
                                           "shell_mass", "radial_distance"], {"external_medium_density": 0.0}),
    "wind_shock": (capi.WIND_SHOCK, ["wind_mass_outflow_rate", "wind_gamma_beta", "wind_pressure",
                                     "post_shock_pressure", "shock_location", "post_shock_gamma_beta",
                                     "flare_outflow_rate", "flare_gamma_beta", "flare_time", "flare_duration"],
                   {"flare_outflow_rate": 0.0, "flare_gamma_beta": 0.0, "flare_time": 0.0, "flare_duration": 0.0,
                    "initial_data_table": None}),
}

RK_ORDERS = {"RK1": 1, "RK2": 2, "RK3": 3}
SOLVERS = {"HLLE": capi.HLLE, "HLLC": capi.HLLC}


class ConfigError(ValueError):
    """A setup the reference would reject when deserialising it."""


def _f(v, where):
    """An f64 field: YAML 1.1 (the safe loader) reads 1e20 as a string, the
    reference's YAML loader as a number."""
    try:
        return float(v)
    except (TypeError, ValueError):
        raise ConfigError(f"{where}: expected a number, got {v!r}") from None


def _i(v, where):
    if isinstance(v, bool) or not isinstance(v, int):
        raise ConfigError(f"{where}: expected an integer, got {v!r}")
    return v


def _fields(section, where, required, optional):
    if not isinstance(section, dict):
        raise ConfigError(f"{where}: expected a mapping")
    unknown = set(section) - set(required) - set(optional)
    if unknown:
        raise ConfigError(f"{where}: unknown field(s) {sorted(unknown)}")
    missing = [k for k in required if k not in section]
    if missing:
        raise ConfigError(f"{where}: missing field(s) {missing}")
    out = dict(optional)
    out.update(section)
    return out


@dataclass
class Preset:
    hydro: capi.Hydro
    model_name: str
    model_kind: int
    model_params: list
    model_table: np.ndarray = None
    mesh: dict = field(default_factory=dict)
    control: dict = field(default_factory=dict)

    @property
    def knc_mesh(self):
        m = self.mesh
        return capi.Mesh(float(m["reference_radius"]), int(m["num_radial_zones"] or 0), int(m["num_polar_zones"]),
                         int(m["block_size"]), 0)

    def block_dlogr(self):
        """Mesh::block_dlogr (src/mesh.rs:350-352)."""
        m = self.mesh
        zone = 1.0 / m["num_radial_zones"] if m["num_radial_zones"] else math.pi / m["num_polar_zones"]
        return m["block_size"] * zone

    def block_inner_radius(self, index):
        """Mesh::subgrid_extent(index).inner_radius (src/mesh.rs:358-375)."""
        return self.mesh["reference_radius"] * (1.0 + self.block_dlogr()) ** float(index)

    def excision_surface(self, which, time):
        """Mesh::inner/outer_excision_surface (src/mesh.rs:318-336)."""
        m = self.mesh
        r0 = m["inner_radius"] if which == 0 else m["outer_radius"]
        v = m["inner_excision_speed"] if which == 0 else m["outer_excision_speed"]
        return r0 + max(time - (m["excision_delay"] or 0.0), 0.0) * v

    def grid_block_indexes(self, time):
        """Mesh::grid_blocks(time) (src/mesh.rs:384-397): blocks 0, 1, ...
        while the block's inner radius is inside the OES."""
        oes = self.excision_surface(1, time)
        out = []
        i = 0
        while self.block_inner_radius(i) < oes:
            out.append(i)
            i += 1
        return out

    @property
    def moving_excision(self):
        return self.mesh["inner_excision_speed"] > 0.0 or self.mesh["outer_excision_speed"] > 0.0


def parse(config, table_loader=None):
    """A Preset from a parsed setup mapping (serde's checks restated)."""
    cfg = _fields(config, "configuration", ["hydro", "model", "mesh", "control"], {})
    hy = cfg["hydro"]
    if not isinstance(hy, dict) or len(hy) != 1 or next(iter(hy)) not in ("relativistic", "newtonian"):
        raise ConfigError("hydro: expected exactly one of relativistic, newtonian")
    (hname, hsec), = hy.items()
    if hname == "relativistic":
        f = _fields(hsec, "hydro.relativistic", ["gamma_law_index", "plm_theta", "cfl_number", "runge_kutta_order",
                                                 "riemann_solver"], {"adaptive_time_step": False})
        if f["riemann_solver"] not in SOLVERS:
            raise ConfigError(f"unknown riemann_solver {f['riemann_solver']!r}")
        if f["runge_kutta_order"] not in RK_ORDERS:
            raise ConfigError(f"unknown runge_kutta_order {f['runge_kutta_order']!r}")
        hydro = capi.relativistic_hydro(_f(f["gamma_law_index"], "gamma_law_index"), _f(f["plm_theta"], "plm_theta"),
                                        _f(f["cfl_number"], "cfl_number"), RK_ORDERS[f["runge_kutta_order"]],
                                        SOLVERS[f["riemann_solver"]], bool(f["adaptive_time_step"]))
    else:
        f = _fields(hsec, "hydro.newtonian", ["gamma_law_index", "plm_theta", "cfl_number", "runge_kutta_order"], {})
        if f["runge_kutta_order"] not in RK_ORDERS:
            raise ConfigError(f"unknown runge_kutta_order {f['runge_kutta_order']!r}")
        hydro = capi.newtonian_hydro(_f(f["gamma_law_index"], "gamma_law_index"), _f(f["plm_theta"], "plm_theta"),
                                     _f(f["cfl_number"], "cfl_number"), RK_ORDERS[f["runge_kutta_order"]])
    mo = cfg["model"]
    if not isinstance(mo, dict) or len(mo) != 1 or next(iter(mo)) not in MODELS:
        raise ConfigError(f"model: expected exactly one of {sorted(MODELS)}")
    (mname, msec), = mo.items()
    kind, names, optional = MODELS[mname]
    mf = _fields(msec, f"model.{mname}", names if not optional else [n for n in names if n not in optional],
                 optional)
    params = [_f(mf[n], f"model.{mname}.{n}") for n in names]
    table = None
    if mname == "wind_shock" and mf.get("initial_data_table"):
        if table_loader is None:
            raise ConfigError("wind_shock: initial_data_table given but no table loader")
        table = np.asarray(table_loader(mf["initial_data_table"]), dtype=np.float64).reshape(-1, 4)
    me = _fields(cfg["mesh"], "mesh", ["reference_radius", "inner_radius", "outer_radius", "inner_excision_speed",
                                        "outer_excision_speed", "num_polar_zones", "block_size"],
                 {"num_radial_zones": None, "excision_delay": None})
    co = _fields(cfg["control"], "control", ["start_time", "final_time", "checkpoint_interval", "fold"],
                 {"products_interval": None, "num_threads": None, "snappy_compression": False,
                  "output_directory": "."})
    ints = ("num_polar_zones", "block_size", "num_radial_zones")
    mesh = {k: (_i(v, f"mesh.{k}") if k in ints else _f(v, f"mesh.{k}")) if v is not None else None
            for k, v in me.items()}
    for k in ("start_time", "final_time", "checkpoint_interval"):
        co[k] = _f(co[k], f"control.{k}")
    co["fold"] = _i(co["fold"], "control.fold")
    if mesh["block_size"] < 2:
        raise ConfigError("must have at least 2 radial zones per block")
    if mesh["num_polar_zones"] == 1 and not mesh["num_radial_zones"]:
        raise ConfigError("num_radial_zones is not optional when num_polar_zones=1")
    return Preset(hydro, mname, kind, params, table, mesh, co)


def ascii_table(path):
    """LookupTable::from_ascii_file (src/lookup_table_v2.rs:52-83): whitespace
    separated floats in groups of four."""
    with open(path) as f:
        vals = [float(x) for x in f.read().split()]
    return np.asarray(vals).reshape(-1, 4)


def apply_overrides(config, overrides, loader=yaml.safe_load):
    """Configuration::patch_from (src/app.rs:285-294): each override is a
    ``.yaml`` file merged into the mapping, or ``a.b.c=value`` with the value
    parsed as YAML."""
    config = copy.deepcopy(config)

    def merge(dst, src):
        for k, v in src.items():
            if isinstance(v, dict) and isinstance(dst.get(k), dict):
                merge(dst[k], v)
            else:
                dst[k] = v

    for item in overrides:
        if item.endswith(".yaml"):
            with open(item) as f:
                merge(config, loader(f) or {})
            continue
        if "=" not in item:
            raise ConfigError(f"override {item!r} is neither key=value nor a .yaml file")
        key, value = item.split("=", 1)
        node = config
        parts = key.split(".")
        for p in parts[:-1]:
            node = node.setdefault(p, {})
        node[parts[-1]] = loader(value)
    return config


def load(path, overrides=(), table_loader=ascii_table):
    """Reads a setup file (YAML, loaded with the safe loader), applies the
    overrides and returns the Preset."""
    with open(path) as f:
        config = yaml.safe_load(f)
    return parse(apply_overrides(config, overrides), table_loader=table_loader)


def context_from_preset(preset, **options):
    """App::from_config on the GPU: the blocks of Mesh::grid_blocks at
    control.start_time built from the model on the device, model ghost rows
    on both sides and, with moving surfaces, excision with births from the
    model.  Returns the Context (time = start_time)."""
    ctx = capi.Context(preset.hydro, **options)
    mesh = preset.knc_mesh
    t0 = float(preset.control["start_time"])
    args = (preset.model_kind, preset.model_params)
    for i in preset.grid_block_indexes(t0):
        ctx.insert_block_model(i, mesh, *args, t0, table=preset.model_table)
    ctx.set_time(t0, 0)
    for side in (capi.INNER, capi.OUTER):
        ctx.set_boundary_model(side, mesh, *args, table=preset.model_table)
    if preset.moving_excision:
        m = preset.mesh
        ctx.set_excision(m["inner_radius"], m["outer_radius"], m["inner_excision_speed"],
                         m["outer_excision_speed"], m["excision_delay"] or 0.0, 1, None)
        ctx.set_birth_model(mesh, *args, table=preset.model_table)
    return ctx
