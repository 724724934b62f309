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
import os
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
    # the model struct's serialised fields in declaration order (serde skips
    # halo_kilonova's external_medium_density; wind_shock ends with
    # initial_data_table)
    model_fields: dict = field(default_factory=dict)

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

    def to_serde(self):
        """The Configuration as serde serialises it (src/app.rs:163-172): the
        hydro and model enums externally tagged in snake_case, every struct's
        fields in declaration order, f64 fields as floats, usize fields as
        ints, absent options as None."""
        h = self.hydro
        rk = f"RK{h.runge_kutta_order}"
        common = {"gamma_law_index": float(h.gamma_law_index), "plm_theta": float(h.plm_theta),
                  "cfl_number": float(h.cfl_number), "runge_kutta_order": rk}
        if h.system == capi.SRHD:
            hydro = {"relativistic": dict(common, riemann_solver="HLLC" if h.riemann_solver == capi.HLLC else "HLLE",
                                          adaptive_time_step=bool(h.adaptive_time_step))}
        else:
            hydro = {"newtonian": common}
        m = self.mesh
        opt_f = lambda v: None if v is None else float(v)
        opt_i = lambda v: None if v is None else int(v)
        mesh = {"reference_radius": float(m["reference_radius"]), "inner_radius": float(m["inner_radius"]),
                "outer_radius": float(m["outer_radius"]), "inner_excision_speed": float(m["inner_excision_speed"]),
                "outer_excision_speed": float(m["outer_excision_speed"]),
                "num_radial_zones": opt_i(m["num_radial_zones"]), "num_polar_zones": int(m["num_polar_zones"]),
                "block_size": int(m["block_size"]), "excision_delay": opt_f(m["excision_delay"])}
        c = self.control
        control = {"start_time": float(c["start_time"]), "final_time": float(c["final_time"]),
                   "checkpoint_interval": float(c["checkpoint_interval"]),
                   "products_interval": opt_f(c["products_interval"]), "fold": int(c["fold"]),
                   "num_threads": opt_i(c["num_threads"]), "snappy_compression": bool(c["snappy_compression"]),
                   "output_directory": str(c["output_directory"])}
        return {"hydro": hydro, "model": {self.model_name: dict(self.model_fields)}, "mesh": mesh,
                "control": control}


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
    model_fields = dict(zip(names, params))
    if mname == "wind_shock":
        path = mf.get("initial_data_table")
        if path is not None and not isinstance(path, str):
            raise ConfigError("model.wind_shock.initial_data_table: expected a string")
        model_fields["initial_data_table"] = path
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
    for k in ("products_interval",):
        if co[k] is not None:
            co[k] = _f(co[k], f"control.{k}")
    for k in ("num_threads",):
        if co[k] is not None:
            co[k] = _i(co[k], f"control.{k}")
    if not isinstance(co["snappy_compression"], bool):
        raise ConfigError("control.snappy_compression: expected a bool")
    if not isinstance(co["output_directory"], str):
        raise ConfigError("control.output_directory: expected a string")
    return Preset(hydro, mname, kind, params, table, mesh, co, model_fields)


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


def configure_context(ctx, preset):
    """The parts of App::from_config that are not the state: model ghost rows
    on both sides and, with moving surfaces, excision with births from the
    model -- all evaluated on the device."""
    mesh = preset.knc_mesh
    args = (preset.model_kind, preset.model_params)
    for side in (capi.INNER, capi.OUTER):
        ctx.set_boundary_model(side, mesh, *args, table=preset.model_table)
    if preset.moving_excision:
        m = preset.mesh
        ctx.set_excision(m["inner_radius"], m["outer_radius"], m["inner_excision_speed"],
                         m["outer_excision_speed"], m["excision_delay"] or 0.0, 1, None)
        ctx.set_birth_model(mesh, *args, table=preset.model_table)
    return ctx


def context_from_preset(preset, **options):
    """App::from_config on the GPU: the blocks of Mesh::grid_blocks at
    control.start_time built from the model on the device, then
    configure_context.  Returns the Context (time = start_time)."""
    ctx = capi.Context(preset.hydro, **options)
    mesh = preset.knc_mesh
    t0 = float(preset.control["start_time"])
    for i in preset.grid_block_indexes(t0):
        ctx.insert_block_model(i, mesh, preset.model_kind, preset.model_params, t0, table=preset.model_table)
    ctx.set_time(t0, 0)
    return configure_context(ctx, preset)


def validate(preset):
    """Configuration::validate (src/app.rs:273-279): the hydro's ranges
    (src/physics/relativistic_hydro.rs:46-54), the models' own checks
    (src/models/kinetic_bomb.rs:66-76, halo_kilonova.rs:47-59,
    wind_shock.rs:75-82), Mesh::validate (src/mesh.rs:273-298) and
    Control::validate (src/app.rs:202-213).  Raises ConfigError."""
    h = preset.hydro
    if not 1.0 <= h.plm_theta <= 2.0:
        raise ConfigError("plm_theta must be in the range [1, 2]")
    if not 0.0 <= h.cfl_number <= 0.7:
        raise ConfigError("cfl_number must be in the range [0.0, 0.7]")
    f = preset.model_fields
    if preset.model_name in ("kinetic_bomb", "halo_kilonova"):
        v = math.sqrt(2.0 * f["kinetic_energy"] / f["shell_mass"])
        if v > 0.25 * capi.LIGHT_SPEED:
            raise ConfigError(f"The shell is moving faster (v/c = {v / capi.LIGHT_SPEED}) than 0.25 c, but this "
                              "problem assumes Newtonian expressions for the kinetic energy. Consider reducing the "
                              "kinetic energy or increasing the shell mass.")
    if preset.model_name == "wind_shock" and f["wind_gamma_beta"] < 0.0:
        raise ConfigError("the wind four-velocity must be positive")
    m = preset.mesh
    t0 = float(preset.control["start_time"])
    if m["reference_radius"] <= 0.0 or m["inner_radius"] < 0.0 or m["outer_radius"] < 0.0:
        raise ConfigError("all radii must be positive")
    if m["reference_radius"] > preset.excision_surface(1, t0):
        raise ConfigError("the reference radius is outside the outer excision surface")
    if (m["excision_delay"] or 0.0) < 0.0:
        raise ConfigError("the excision delay time must be non-negative")
    if m["inner_excision_speed"] < 0.0 or m["outer_excision_speed"] < 0.0:
        raise ConfigError("the excision surface speeds must be non-negative")
    if m["outer_excision_speed"] < m["inner_excision_speed"]:
        raise ConfigError("outer_excision_speed < inner_excision_speed (the IES would eventually overtake the OES)")
    if m["num_polar_zones"] != 1 and m["num_polar_zones"] < 16:
        raise ConfigError("must have at least 2 radial zones per block")   # the reference's message
    c = preset.control
    threads = c["num_threads"] if c["num_threads"] is not None else 2 * (os.cpu_count() or 1)
    if threads == 0 or threads >= 1024:
        raise ConfigError("num_threads must be > 0 and < 1024")
    if c["checkpoint_interval"] < 0.0:
        raise ConfigError("checkpoint_interval <= 0.0")
    if (c["products_interval"] or 0.0) < 0.0:
        raise ConfigError("products_interval <= 0.0")
    return preset


# The reference's preset setups (setups/*.yaml, embedded by App::presets,
# src/app.rs:385-393), restated as mappings.  jet_in_star carries the field
# eta_0 and lacks envelope_mass and volume_factor, so that setup is rejected
# here as serde rejects it.
SETUPS = {
    "jet_in_cloud": {
        "hydro": {"relativistic": {"gamma_law_index": 1.333, "plm_theta": 1.5, "cfl_number": 0.4,
                                   "runge_kutta_order": "RK2", "riemann_solver": "HLLC"}},
        "model": {"jet_in_cloud": {"cloud_mass": 1e32, "engine_delay": 1.0, "engine_duration": 1.0,
                                   "engine_strength": 1.0, "engine_theta": 0.1, "engine_u": 30.0,
                                   "envelop_m1": 2e27, "envelop_slowest_beta": 0.1, "envelop_fastest_beta": 0.995,
                                   "envelop_psi": 0.25}},
        "mesh": {"inner_radius": 0.0, "outer_radius": 0.0, "inner_excision_speed": 3e8, "outer_excision_speed": 3e10,
                 "reference_radius": 3e8, "num_polar_zones": 512, "block_size": 4},
        "control": {"final_time": 5.0, "start_time": 1.0, "checkpoint_interval": 0.1, "fold": 100,
                    "num_threads": None, "output_directory": "data"},
    },
    "jet_in_star": {
        "hydro": {"relativistic": {"gamma_law_index": 1.33, "plm_theta": 1.5, "cfl_number": 0.3,
                                   "runge_kutta_order": "RK2", "riemann_solver": "HLLC"}},
        "model": {"jet_in_star": {"star_mass": 1e34, "engine_duration": 10.0, "engine_energy": 1e51,
                                  "engine_theta": 0.1, "engine_u": 50.0, "envelope_radius": 1e11, "eta_0": 1.0}},
        "mesh": {"inner_radius": 1e9, "outer_radius": 1e12, "excision_delay": 10.0, "inner_excision_speed": 1e9,
                 "outer_excision_speed": 3e10, "reference_radius": 1e9, "num_polar_zones": 64, "block_size": 4},
        "control": {"final_time": 3.0, "start_time": 0.0, "checkpoint_interval": 0.1, "fold": 100,
                    "num_threads": None, "output_directory": "data"},
    },
    "halo_kilonova": {
        "hydro": {"relativistic": {"gamma_law_index": 1.333, "plm_theta": 1.1, "cfl_number": 0.2,
                                   "runge_kutta_order": "RK2", "riemann_solver": "HLLC",
                                   "adaptive_time_step": True}},
        "model": {"halo_kilonova": {"altitude": 1e20, "external_medium_density": 1e-23, "launch_radius": 1e18,
                                    "shell_thickness": 1e18, "kinetic_energy": 1e51, "shell_mass": 2e32,
                                    "radial_distance": 1e22}},
        "mesh": {"inner_radius": 1e17, "outer_radius": 1e20, "inner_excision_speed": 0.0,
                 "outer_excision_speed": 0.0, "reference_radius": 1e18, "num_polar_zones": 256, "block_size": 4},
        "control": {"final_time": 1e10, "start_time": 0.0, "checkpoint_interval": 1e8, "output_directory": "data",
                    "num_threads": 16, "fold": 10},
    },
    "wind_shock": {
        "hydro": {"relativistic": {"gamma_law_index": 1.333, "plm_theta": 1.5, "cfl_number": 0.1,
                                   "runge_kutta_order": "RK2", "riemann_solver": "HLLC"}},
        "model": {"wind_shock": {"wind_mass_outflow_rate": 1e20, "flare_outflow_rate": 1e33, "flare_gamma_beta": 1e2,
                                 "wind_gamma_beta": 1e2, "post_shock_gamma_beta": 1e-2, "wind_pressure": 1e-17,
                                 "post_shock_pressure": 1e-9, "shock_location": 5e8, "flare_time": 1.0,
                                 "flare_duration": 0.01}},
        "mesh": {"inner_radius": 5e8, "outer_radius": 50e9, "inner_excision_speed": 0.0,
                 "outer_excision_speed": 0.0, "reference_radius": 2e8, "num_polar_zones": 1,
                 "num_radial_zones": 32, "block_size": 32},
        "control": {"final_time": 8.0, "start_time": 0.0, "checkpoint_interval": 0.05, "fold": 10,
                    "num_threads": 4, "output_directory": "data"},
    },
    "kinetic_bomb": {
        "hydro": {"relativistic": {"gamma_law_index": 1.333, "plm_theta": 1.5, "cfl_number": 0.3,
                                   "runge_kutta_order": "RK2", "riemann_solver": "HLLC",
                                   "adaptive_time_step": True}},
        "model": {"kinetic_bomb": {"external_medium_density": 1e-23, "launch_radius": 1e18,
                                   "shell_thickness": 1e18, "kinetic_energy": 1e51, "shell_mass": 2e32}},
        "mesh": {"inner_radius": 1e18, "outer_radius": 2e19, "inner_excision_speed": 0.0,
                 "outer_excision_speed": 0.0, "reference_radius": 1e18, "num_polar_zones": 1,
                 "num_radial_zones": 512, "block_size": 512},
        "control": {"final_time": 1e9, "start_time": 0.0, "checkpoint_interval": 1e7, "output_directory": "data",
                    "num_threads": None, "fold": 10},
    },
}
