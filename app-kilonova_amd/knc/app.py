"""The reference's driver (``src/main.rs``) over the GPU library.

``App`` is ``app::App`` (``src/app.rs:179-186``): the state -- here a
:class:`knc.capi.Context` holding every block in HBM --, the :class:`Tasks`,
the configuration (a :class:`knc.presets.Preset`) and the version string.
It is built the reference's ways (``App::from_preset_or_file``,
``src/app.rs:344-365``): from a preset name, a setup ``.yaml`` or a
``chkpt.NNNN.cbor`` checkpoint, each with ``key=value`` / ``.yaml``
overrides.  :func:`run` is ``main.rs::run`` (``src/main.rs:76-99``):
``side_effects`` then ``scheme::advance`` with the configured ``fold`` until
``final_time``, then ``side_effects`` once more.  :func:`side_effects`
(``src/main.rs:32-71``) prints the iteration message with its Mzps, and
writes products and checkpoints in the reference's CBOR layout (the library
encodes the state from HBM), named ``prods.NNNN.cbor`` / ``chkpt.NNNN.cbor``
in ``control.output_directory``.

Run it as ``python -m knc <input.yaml|chkpt.cbor|preset> [overrides ...]``.
"""
import os
import sys
import time as _time

import yaml

from . import capi, cbor, presets

DESCRIPTION = "Clemson Kilonova Code"          # the crate's description (Cargo.toml)
VERSION = "v0.1.2 knc-mi355x"                   # the crate version this package mirrors, and its name
TASK_NAMES = ("write_checkpoint", "write_products", "iteration_message", "report_progress")


class RecurringTask:
    """tasks::RecurringTask (src/tasks.rs:10-27, 55-82): count and next_time
    are saved in checkpoints; the wall clock of the last run and the count of
    this run are not."""

    def __init__(self, next_time=0.0, count=0):
        self.count = int(count)
        self.next_time = float(next_time)
        self.last_performed = _time.perf_counter()
        self.count_this_run = 0

    def advance(self, interval):
        """Marks the task done, schedules it `interval` later and returns the
        wall seconds since it was last done."""
        now = _time.perf_counter()
        seconds = now - self.last_performed
        self.count += 1
        self.count_this_run += 1
        self.next_time += interval
        self.last_performed = now
        return seconds

    def to_serde(self):
        return {"count": self.count, "next_time": float(self.next_time)}


class Tasks:
    """tasks::Tasks (src/tasks.rs:33-47, 86-95)."""

    def __init__(self, start_time=0.0, tasks=None):
        for name in TASK_NAMES:
            setattr(self, name, RecurringTask(start_time) if tasks is None else tasks[name])

    def to_serde(self):
        return {name: getattr(self, name).to_serde() for name in TASK_NAMES}

    @classmethod
    def from_serde(cls, d):
        try:
            return cls(tasks={n: RecurringTask(float(d[n]["next_time"]), int(d[n]["count"])) for n in TASK_NAMES})
        except (KeyError, TypeError, ValueError) as e:
            raise presets.ConfigError(f"the checkpoint's tasks are malformed: {e}") from None


class App:
    """app::App on the GPU: ``ctx`` holds the state, ``preset`` the
    configuration."""

    def __init__(self, ctx, tasks, preset, version=VERSION):
        self.ctx = ctx
        self.tasks = tasks
        self.preset = preset
        self.version = version

    # -- construction (src/app.rs:307-365) -----------------------------------
    @classmethod
    def from_config(cls, config, overrides=(), **options):
        """App::from_config: the blocks of the mesh at start_time from the
        model, fresh tasks."""
        preset = presets.parse(presets.apply_overrides(config, overrides), table_loader=presets.ascii_table)
        ctx = presets.context_from_preset(preset, **options)
        return cls(ctx, Tasks(float(preset.control["start_time"])), preset)

    @classmethod
    def from_checkpoint(cls, path, overrides=(), **options):
        """io::read_cbor::<App> then with_patched_config: the state, tasks
        and version as saved, the configuration patched by the overrides."""
        with capi.Checkpoint(path) as ck:
            config = cbor.decode(ck.config)
            tasks = Tasks.from_serde(cbor.decode(ck.tasks))
            preset = presets.parse(presets.apply_overrides(config, overrides), table_loader=presets.ascii_table)
            if ck.system != preset.hydro.system:
                raise presets.ConfigError("the checkpoint's state and its hydro section are different systems")
            ctx = capi.Context(preset.hydro, **options)
            ck.load(ctx, preset.knc_mesh)
            version = ck.version
        presets.configure_context(ctx, preset)
        return cls(ctx, tasks, preset, version)

    @classmethod
    def from_file(cls, path, overrides=(), **options):
        """App::from_file: a setup .yaml or a checkpoint .cbor."""
        ext = os.path.splitext(path)[1]
        if ext == ".yaml":
            with open(path) as f:
                return cls.from_config(yaml.safe_load(f), overrides, **options)
        if ext == ".cbor":
            return cls.from_checkpoint(path, overrides, **options)
        raise presets.ConfigError(f"unknown input file type '{path}'")

    @classmethod
    def from_preset_or_file(cls, name, overrides=(), **options):
        if name in presets.SETUPS:
            return cls.from_config(presets.SETUPS[name], overrides, **options)
        return cls.from_file(name, overrides, **options)

    # -- output ---------------------------------------------------------------
    def config_cbor(self):
        return cbor.encode(self.preset.to_serde())

    def write_checkpoint(self, path):
        """App::package + io::write_cbor; like App::package
        (src/app.rs:368-382) it stamps the running code's version, also after
        a restart from another version's checkpoint."""
        self.ctx.write_checkpoint(path, self.config_cbor(), cbor.encode(self.tasks.to_serde()), VERSION)

    def write_products(self, path):
        """Products::try_from_state + io::write_cbor."""
        self.ctx.write_products(path, self.config_cbor(), VERSION)

    def total_zones(self):
        """State::total_zones (src/state.rs:120-122)."""
        return sum(nr * nq for nr, nq in self.ctx.shapes.values())


def _write_message(path, out):
    out(f"write {path}")          # io::write_cbor (src/io.rs:25)


def side_effects(app, out=print):
    """main.rs::side_effects (src/main.rs:32-71)."""
    tasks, control = app.tasks, app.preset.control
    t = app.ctx.time
    if tasks.iteration_message.next_time <= t:
        seconds = tasks.iteration_message.advance(0.0)
        if tasks.iteration_message.count_this_run > 1:
            mzps = 1e-6 * app.total_zones() / seconds * control["fold"] if seconds > 0 else float("inf")
            out(f"[{app.ctx.iteration:05d}] t={t:.5f} blocks={len(app.ctx.shapes)} Mzps={mzps:.2f})")
    interval = control["products_interval"]
    if interval is not None and tasks.write_products.next_time <= t:
        tasks.write_products.advance(interval)
        path = f"{control['output_directory']}/prods.{tasks.write_products.count - 1:04d}.cbor"
        os.makedirs(control["output_directory"], exist_ok=True)
        _write_message(path, out)
        app.write_products(path)
    if tasks.write_checkpoint.next_time <= t:
        tasks.write_checkpoint.advance(control["checkpoint_interval"])
        path = f"{control['output_directory']}/chkpt.{tasks.write_checkpoint.count - 1:04d}.cbor"
        os.makedirs(control["output_directory"], exist_ok=True)
        _write_message(path, out)
        app.write_checkpoint(path)


def run(app, out=print):
    """main.rs::run (src/main.rs:76-99)."""
    final_time = app.preset.control["final_time"]
    fold = app.preset.control["fold"]
    while app.ctx.time < final_time:
        side_effects(app, out)
        app.ctx.advance(fold)
    side_effects(app, out)
    return app


def main(argv=None, out=print, autotune=True, **options):
    """main.rs::main (src/main.rs:104-151).  With autotune the stage shape is
    chosen by the library's autotuner on the initial mesh before the run (the
    shape changes no bits of the result; the small blocks of the reference's
    setups run 6-16 % faster in the two-pass form it picks)."""
    argv = sys.argv[1:] if argv is None else argv
    out("")
    out(DESCRIPTION)
    out(VERSION)
    out("")
    if not argv:
        out("usage: kilonova <input.yaml|chkpt.cbor|preset> [opts.yaml|group.key=value] [...]")
        out("")
        out("These are the preset model setups:")
        out("")
        for key in presets.SETUPS:
            out(f"  {key}")
        out("")
        out("To run any of these presets, run e.g. `kilonova jet_in_star`.")
        return 0
    options.setdefault("launch_mode", capi.BATCHED)
    app = App.from_preset_or_file(argv[0], argv[1:], **options)
    presets.validate(app.preset)
    for line in yaml.safe_dump(app.preset.to_serde(), sort_keys=False).splitlines():
        out(line)
    out("")
    c = app.preset.control
    out(f"worker threads ...... {c['num_threads'] if c['num_threads'] is not None else 2 * (os.cpu_count() or 1)}")
    out(f"compute cores ....... {os.cpu_count()}")
    if autotune:
        t = app.ctx.autotune()
        out(f"stage shape ......... {t['stage_threads']} polar columns x {t['rows_per_item']} rows, "
            f"form {t['stage_passes']} (autotuned)")
    out("")
    run(app, out)
    return 0
