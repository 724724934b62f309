"""ctypes binding of ``libknc_gpu.so`` (the C ABI in ``include/knc_gpu.h``).

The names mirror the reference's Rust interface: ``Hydro`` carries the fields
of ``RelativisticHydro`` / ``NewtonianHydro``
(``src/physics/relativistic_hydro.rs:13-36``, ``src/physics/newtonian_hydro.rs:13-28``),
``Context.time_step`` is ``State::time_step`` (``src/state.rs:143-162``),
``Context.advance`` is ``scheme::advance`` (``src/scheme.rs:212-240``) and the
``try_to_primitive`` ... ``plm_gradient`` functions are the per-zone methods of
``trait Hydrodynamics`` (``src/traits.rs:41-136``).  Hydro failures raise
``HydroError`` with the reference's error kinds (``src/physics/mod.rs:82-117``).

There is no CPU fallback: every call runs on the GPU through the HIP library
and raises ``RuntimeError`` when the library is missing or fails.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "build", "libknc_gpu.so")

LIGHT_SPEED = 3e10  # src/physics/mod.rs:7

EULER, SRHD = 0, 1
HLLE, HLLC = 0, 1
RADIAL, POLAR = 0, 1
FIXED, OUTFLOW = 0, 1
INNER, OUTER = 0, 1
PER_BLOCK, BATCHED = 0, 1
GUESS_PLAIN, GUESS_LEARNED, GUESS_LEARNED_STEP, GUESS_PREVIOUS, GUESS_LEARNED_FOLD = 0, 1, 2, 3, 4

ERROR_KINDS = {1: "NegativeDensity", 2: "NegativePressure", 3: "NegativeEnergyDensity",
               4: "RootFinderFailed"}
ERR_MODEL = 14   # an initial model evaluated where the reference panics
ERR_IO = 15      # a file could not be opened, read or written
ERR_FORMAT = 16  # a checkpoint is not an App of the expected shape

# initial models on the device (src/models/*.rs); params in struct field order
KINETIC_BOMB, JET_IN_CLOUD, JET_IN_STAR, HALO_KILONOVA, WIND_SHOCK = 1, 2, 3, 4, 5
MODEL_NPARAM = 12


class Hydro(ctypes.Structure):
    _fields_ = [("system", ctypes.c_int32), ("riemann_solver", ctypes.c_int32),
                ("runge_kutta_order", ctypes.c_int32), ("adaptive_time_step", ctypes.c_int32),
                ("gamma_law_index", ctypes.c_double), ("plm_theta", ctypes.c_double),
                ("cfl_number", ctypes.c_double), ("light_speed", ctypes.c_double)]


class Options(ctypes.Structure):
    _fields_ = [("launch_mode", ctypes.c_int32), ("c2p_guess", ctypes.c_int32),
                ("collect_stats", ctypes.c_int32), ("rows_per_item", ctypes.c_int32),
                ("device", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("stage_threads", ctypes.c_int32), ("stage_passes", ctypes.c_int32)]


class Error(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int32), ("block", ctypes.c_int32), ("i", ctypes.c_int32),
                ("j", ctypes.c_int32), ("r", ctypes.c_double), ("theta", ctypes.c_double),
                ("value", ctypes.c_double), ("conserved", ctypes.c_double * 4)]


class BlockData(ctypes.Structure):
    _fields_ = [("nr", ctypes.c_int32), ("pad", ctypes.c_int32), ("radial_vertices", ctypes.c_void_p),
                ("conserved", ctypes.c_void_p), ("scalar_mass", ctypes.c_void_p)]


BOUNDARY_FN = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                               ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_double))

BLOCK_SOURCE_FN = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_double,
                                   ctypes.POINTER(BlockData))


class Excision(ctypes.Structure):
    _fields_ = [("inner_radius", ctypes.c_double), ("outer_radius", ctypes.c_double),
                ("inner_speed", ctypes.c_double), ("outer_speed", ctypes.c_double), ("delay", ctypes.c_double),
                ("check_every", ctypes.c_int32), ("pad", ctypes.c_int32), ("source", BLOCK_SOURCE_FN),
                ("user", ctypes.c_void_p)]


class TuneResult(ctypes.Structure):
    _fields_ = [("stage_threads", ctypes.c_int32), ("rows_per_item", ctypes.c_int32),
                ("stage_passes", ctypes.c_int32), ("candidates", ctypes.c_int32), ("from_cache", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("ms_per_stage", ctypes.c_double), ("default_ms_per_stage", ctypes.c_double)]


class Model(ctypes.Structure):
    """knc_model: kind, wind_shock table rows, parameters, table pointer."""
    _fields_ = [("kind", ctypes.c_int32), ("table_rows", ctypes.c_int32),
                ("params", ctypes.c_double * MODEL_NPARAM), ("table", ctypes.c_void_p)]


class Mesh(ctypes.Structure):
    """knc_mesh: the reference's Mesh fields that fix block geometry
    (src/mesh.rs:84-113); num_radial_zones = 0 stands for None."""
    _fields_ = [("reference_radius", ctypes.c_double), ("num_radial_zones", ctypes.c_int32),
                ("num_polar_zones", ctypes.c_int32), ("block_size", ctypes.c_int32), ("pad", ctypes.c_int32)]


def make_model(kind, params, table=None):
    """A Model and the array that keeps its table alive."""
    m = Model()
    m.kind = int(kind)
    vals = list(params)
    if len(vals) > MODEL_NPARAM:
        raise ValueError("too many model parameters")
    for k, v in enumerate(vals):
        m.params[k] = float(v)
    tab = None
    if table is not None:
        tab = _f64(table).reshape(-1, 4)
        m.table_rows = tab.shape[0]
        m.table = tab.ctypes.data
    return m, tab


def mesh_block_vertices(mesh, index):
    """Radial and polar vertices of block `index` of a Mesh."""
    rv = np.empty(mesh.block_size + 1)
    qv = np.empty(mesh.num_polar_zones + 1)
    _check(lib().knc_mesh_block_vertices(ctypes.byref(mesh), int(index), _p(rv), _p(qv)))
    return rv, qv


class RecordMeta(ctypes.Structure):
    _fields_ = [("config", ctypes.c_void_p), ("config_len", ctypes.c_uint64), ("tasks", ctypes.c_void_p),
                ("tasks_len", ctypes.c_uint64), ("version", ctypes.c_char_p)]


class CheckpointInfo(ctypes.Structure):
    _fields_ = [("system", ctypes.c_int32), ("num_blocks", ctypes.c_int32), ("time", ctypes.c_double),
                ("iteration_numer", ctypes.c_int64), ("iteration_denom", ctypes.c_int64),
                ("config", ctypes.c_void_p), ("config_len", ctypes.c_uint64), ("tasks", ctypes.c_void_p),
                ("tasks_len", ctypes.c_uint64), ("version", ctypes.c_char_p)]


def _record_meta(config, tasks, version):
    """A RecordMeta over copies of the CBOR items; returns (meta, buffers to keep)."""
    cbuf = ctypes.create_string_buffer(bytes(config), len(config))
    tbuf = ctypes.create_string_buffer(bytes(tasks), len(tasks)) if tasks is not None else None
    meta = RecordMeta(ctypes.cast(cbuf, ctypes.c_void_p), len(config),
                      ctypes.cast(tbuf, ctypes.c_void_p) if tbuf is not None else None,
                      len(tasks) if tasks is not None else 0, version.encode())
    return meta, (cbuf, tbuf)


class C2PStats(ctypes.Structure):
    _fields_ = [("hist", ctypes.c_uint64 * 64), ("zones", ctypes.c_uint64), ("fallbacks", ctypes.c_uint64),
                ("floors", ctypes.c_uint64), ("learned_iters", ctypes.c_uint64)]


def relativistic_hydro(gamma_law_index=4.0 / 3.0, plm_theta=1.5, cfl_number=0.4, runge_kutta_order=2,
                       riemann_solver=HLLC, adaptive_time_step=False):
    """``RelativisticHydro`` (relativistic_hydro.rs:13-36)."""
    return Hydro(SRHD, riemann_solver, runge_kutta_order, int(adaptive_time_step), gamma_law_index,
                 plm_theta, cfl_number, LIGHT_SPEED)


def newtonian_hydro(gamma_law_index=5.0 / 3.0, plm_theta=1.5, cfl_number=0.4, runge_kutta_order=2,
                    riemann_solver=HLLE):
    """``NewtonianHydro`` (newtonian_hydro.rs:13-28); HLLC is an extension."""
    return Hydro(EULER, riemann_solver, runge_kutta_order, 1, gamma_law_index, plm_theta, cfl_number, 1.0)


class HydroError(RuntimeError):
    """``HydroError``: kind, block, zone (i, j) and (r, theta) of the failure."""

    def __init__(self, code, err=None, message=""):
        self.code = code
        self.kind = ERROR_KINDS.get(code, f"error {code}")
        if err is not None:
            self.block, self.i, self.j = err.block, err.i, err.j
            self.r, self.theta, self.value = err.r, err.theta, err.value
            self.conserved = tuple(err.conserved)
            message = (f"{self.kind}: {err.value:.4e} at position (r, theta) = ({err.r:.4e}, {err.theta:.4f}) "
                       f"block {err.block} zone ({err.i}, {err.j})")
        super().__init__(message)


_lib = None


def lib():
    """Loads the HIP library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libknc_gpu.so not found at {LIB_PATH}: run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    P, I32, I64, D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    L.knc_last_error_message.restype = ctypes.c_char_p
    L.knc_context_create.argtypes = [ctypes.POINTER(Hydro), ctypes.POINTER(Options), ctypes.POINTER(P)]
    L.knc_context_destroy.argtypes = [P]
    L.knc_context_destroy.restype = None
    L.knc_block_insert.argtypes = [P, I32, I32, I32, P, P, P, P]
    L.knc_block_remove.argtypes = [P, I32]
    L.knc_block_download.argtypes = [P, I32, P, P]
    L.knc_block_count.argtypes = [P, P, P, P]
    L.knc_set_boundary.argtypes = [P, I32, I32, P, P]
    L.knc_set_time.argtypes = [P, D, I64]
    L.knc_get_time.argtypes = [P, P, P]
    L.knc_time_step.argtypes = [P, P, ctypes.POINTER(Error)]
    L.knc_block_time_step_noalloc.argtypes = [ctypes.POINTER(Hydro), I32, I32, I32, P, P, P, P, P,
                                              ctypes.POINTER(Error)]
    L.knc_advance_steps.argtypes = [P, D, I32, ctypes.POINTER(Error)]
    L.knc_advance.argtypes = [P, I32, P, ctypes.POINTER(Error)]
    L.knc_get_c2p_stats.argtypes = [P, ctypes.POINTER(C2PStats)]
    L.knc_reset_c2p_stats.argtypes = [P]
    L.knc_synchronize.argtypes = [P]
    L.knc_set_timing.argtypes = [P, I32]
    L.knc_last_stage_time.argtypes = [P, P, P]
    L.knc_set_excision.argtypes = [P, ctypes.POINTER(Excision)]
    L.knc_set_boundary_source.argtypes = [P, I32, BOUNDARY_FN, P]
    L.knc_excision_counts.argtypes = [P, P, P]
    L.knc_autotune.argtypes = [P, I32, I32, ctypes.POINTER(TuneResult)]
    L.knc_hydro_try_to_primitive.argtypes = [ctypes.POINTER(Hydro), I64, P, P, P, P, P]
    L.knc_hydro_to_conserved.argtypes = [ctypes.POINTER(Hydro), I64, P, P]
    L.knc_hydro_intercell_flux.argtypes = [ctypes.POINTER(Hydro), I64, P, P, P, P, I32, P, P]
    L.knc_hydro_geometrical_source_terms.argtypes = [ctypes.POINTER(Hydro), I64, P, P, P, P]
    L.knc_hydro_max_signal_speed.argtypes = [ctypes.POINTER(Hydro), I64, P, P]
    L.knc_hydro_plm_gradient.argtypes = [D, I64, P, P, P, P]
    L.knc_fp64_probe.argtypes = [I64, P, P, P, P, P, P, P]
    L.knc_model_rows_check.argtypes = [P, I32, D, P, P]
    L.knc_fp64_fma_rate.argtypes = [I32, I32, I32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)]
    L.knc_mesh_block_vertices.argtypes = [ctypes.POINTER(Mesh), I32, P, P]
    L.knc_model_eval.argtypes = [P, ctypes.POINTER(Model), I64, P, P, D, P, P]
    L.knc_block_insert_model.argtypes = [P, I32, ctypes.POINTER(Mesh), ctypes.POINTER(Model), D]
    L.knc_set_boundary_model.argtypes = [P, I32, ctypes.POINTER(Mesh), ctypes.POINTER(Model)]
    L.knc_set_birth_model.argtypes = [P, ctypes.POINTER(Mesh), ctypes.POINTER(Model)]
    L.knc_write_checkpoint.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(RecordMeta)]
    L.knc_write_products.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(RecordMeta), ctypes.POINTER(Error)]
    L.knc_checkpoint_open.argtypes = [ctypes.c_char_p, ctypes.POINTER(P)]
    L.knc_checkpoint_get_info.argtypes = [P, ctypes.POINTER(CheckpointInfo)]
    L.knc_checkpoint_block_shape.argtypes = [P, I32, P, P, P]
    L.knc_checkpoint_block_data.argtypes = [P, I32, P, P]
    L.knc_checkpoint_load.argtypes = [P, P, ctypes.POINTER(Mesh)]
    L.knc_checkpoint_close.argtypes = [P]
    L.knc_checkpoint_close.restype = None
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"knc error {rc}: {lib().knc_last_error_message().decode()}")


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """A KNC ``State`` resident on one GPU plus the scheme that advances it."""

    def __init__(self, hydro, launch_mode=PER_BLOCK, c2p_guess=GUESS_PLAIN, collect_stats=False,
                 rows_per_item=0, device=0, use_graph=True, stage_threads=0, stage_passes=0):
        self.hydro = hydro
        self.options = Options(launch_mode, c2p_guess, int(collect_stats), rows_per_item, device, int(use_graph),
                               stage_threads, stage_passes)
        h = ctypes.c_void_p()
        _check(lib().knc_context_create(ctypes.byref(hydro), ctypes.byref(self.options), ctypes.byref(h)))
        self._h = h
        self.shapes = {}
        self.vertices = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().knc_context_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- block set -----------------------------------------------------------
    def insert_block(self, index, radial_vertices, polar_vertices, conserved, scalar_mass):
        rv, qv = _f64(radial_vertices), _f64(polar_vertices)
        nr, nq = rv.size - 1, qv.size - 1
        cons = _f64(conserved).reshape(nr, nq, 4)
        sm = _f64(scalar_mass).reshape(nr, nq)
        _check(lib().knc_block_insert(self._h, int(index), nr, nq, _p(rv), _p(qv), _p(cons), _p(sm)))
        self.shapes[int(index)] = (nr, nq)
        self.vertices[int(index)] = (rv.copy(), qv.copy())

    def insert_block_model(self, index, mesh, kind, params, time, table=None):
        """Block `index` of `mesh` from the initial model on the device
        (BlockState::from_model, src/state.rs:55-69)."""
        m, _keep = make_model(kind, params, table)
        _check(lib().knc_block_insert_model(self._h, int(index), ctypes.byref(mesh), ctypes.byref(m), float(time)))
        rv, qv = mesh_block_vertices(mesh, index)
        self.shapes[int(index)] = (rv.size - 1, qv.size - 1)
        self.vertices[int(index)] = (rv, qv)

    def set_boundary_model(self, side, mesh, kind=None, params=(), table=None):
        """Ghost rows of `side` from the initial model on the device at every
        stage time (src/scheme.rs:30-52); kind None returns the side to fixed
        or outflow rows."""
        if kind is None:
            _check(lib().knc_set_boundary_model(self._h, int(side), None, None))
            return
        m, _keep = make_model(kind, params, table)
        _check(lib().knc_set_boundary_model(self._h, int(side), ctypes.byref(mesh), ctypes.byref(m)))

    def model_rows_check(self, side, time):
        """The ghost rows of `side`'s device boundary model at `time`, as the
        stage builds them and with one thread per zone: two (5, 2, nq)
        arrays (rho, u_r, u_theta, p, s by the two rows and the columns)."""
        nq = next(iter(self.shapes.values()))[1]
        a = np.zeros((5, 2, nq))
        b = np.zeros((5, 2, nq))
        _check(lib().knc_model_rows_check(self._h, int(side), float(time), _p(a), _p(b)))
        return a, b

    def set_birth_model(self, mesh, kind=None, params=(), table=None):
        """Blocks born by excision built from the initial model on the device
        (src/scheme.rs:197-205); kind None clears it."""
        if kind is None:
            _check(lib().knc_set_birth_model(self._h, None, None))
            self._birth_mesh = None
            return
        m, _keep = make_model(kind, params, table)
        _check(lib().knc_set_birth_model(self._h, ctypes.byref(mesh), ctypes.byref(m)))
        self._birth_mesh = Mesh(mesh.reference_radius, mesh.num_radial_zones, mesh.num_polar_zones,
                                mesh.block_size, 0)

    def remove_block(self, index):
        _check(lib().knc_block_remove(self._h, int(index)))
        self.shapes.pop(int(index), None)
        self.vertices.pop(int(index), None)

    def download_block(self, index):
        nr, nq = self.shapes[int(index)]
        cons = np.zeros((nr, nq, 4))
        sm = np.zeros((nr, nq))
        _check(lib().knc_block_download(self._h, int(index), _p(cons), _p(sm)))
        return cons, sm

    def block_indexes(self):
        return sorted(self.shapes)

    def set_boundary(self, side, mode, primitive=None, scalar=None):
        if mode == FIXED:
            primitive = _f64(primitive)
            scalar = _f64(scalar)
        _check(lib().knc_set_boundary(self._h, side, mode, _p(primitive), _p(scalar)))

    def set_boundary_source(self, side, fn):
        """Model boundary: ``fn(side, time, inner_index, outer_index)`` returns
        the ghost rows ``(primitive (2, nq, 4), scalar (2, nq))`` at a stage's
        input time (BlockState::from_model on the boundary block, then c2p;
        src/scheme.rs:46-52); None returns the side to its fixed rows."""
        if not hasattr(self, "_bsrc"):
            self._bsrc = [None, None]
        if fn is None:
            self._bsrc[side] = None
            _check(lib().knc_set_boundary_source(self._h, side, BOUNDARY_FN(), None))
            return

        def call(user, sd, t, inner, outer, nq, prim, scalar):
            try:
                P, S = fn(int(sd), float(t), int(inner), int(outer))
                np.ctypeslib.as_array(prim, shape=(8 * nq,))[:] = np.asarray(P, np.float64).reshape(-1)
                np.ctypeslib.as_array(scalar, shape=(2 * nq,))[:] = np.asarray(S, np.float64).reshape(-1)
                return 0
            except Exception as e:  # reported through the advance call's error
                self._birth_error = e
                return 1

        cb = BOUNDARY_FN(call)
        self._bsrc[side] = cb
        _check(lib().knc_set_boundary_source(self._h, side, cb, None))

    # -- time ----------------------------------------------------------------
    def set_time(self, time, iteration=0):
        _check(lib().knc_set_time(self._h, float(time), int(iteration)))

    @property
    def time(self):
        t = ctypes.c_double()
        it = ctypes.c_int64()
        _check(lib().knc_get_time(self._h, ctypes.byref(t), ctypes.byref(it)))
        return t.value

    @property
    def iteration(self):
        t = ctypes.c_double()
        it = ctypes.c_int64()
        _check(lib().knc_get_time(self._h, ctypes.byref(t), ctypes.byref(it)))
        return it.value

    def _raise(self, rc, err):
        if rc in ERROR_KINDS:
            raise HydroError(rc, err)
        if rc == ERR_MODEL and err.code == ERR_MODEL:
            raise ModelError(err=err)
        _check(rc)

    def time_step(self):
        dt = ctypes.c_double()
        err = Error()
        rc = lib().knc_time_step(self._h, ctypes.byref(dt), ctypes.byref(err))
        if rc:
            self._raise(rc, err)
        return dt.value

    def advance_steps(self, dt, nsteps):
        err = Error()
        rc = lib().knc_advance_steps(self._h, float(dt), int(nsteps), ctypes.byref(err))
        if rc:
            self._raise(rc, err)

    def advance(self, fold):
        dt = ctypes.c_double()
        err = Error()
        rc = lib().knc_advance(self._h, int(fold), ctypes.byref(dt), ctypes.byref(err))
        self._sync_block_set()
        self._birth_keep = None
        birth_error = self.__dict__.pop("_birth_error", None)
        if birth_error is not None:
            raise birth_error
        if rc:
            self._raise(rc, err)
        return dt.value

    # -- excision (add_remove_blocks, src/scheme.rs:179-206) -----------------
    def set_excision(self, inner_radius, outer_radius, inner_speed, outer_speed, delay=0.0, check_every=1,
                     source=None):
        """Moving excision surfaces.  ``source(index, time)`` returns
        ``(radial_vertices, conserved (nr, nq, 4), scalar_mass (nr, nq))`` of a
        new outer block (BlockState::from_model), or None for no births."""
        self._source_py = source
        self._birth_keep = None
        cb = BLOCK_SOURCE_FN(self._birth) if source is not None else BLOCK_SOURCE_FN()
        self._source_cb = cb
        ex = Excision(float(inner_radius), float(outer_radius), float(inner_speed), float(outer_speed),
                      float(delay), int(check_every), 0, cb, None)
        _check(lib().knc_set_excision(self._h, ctypes.byref(ex)))

    def clear_excision(self):
        _check(lib().knc_set_excision(self._h, None))
        self._source_py = None
        self._source_cb = None

    def _birth(self, user, index, time, out):
        try:
            rv, cons, sm = self._source_py(int(index), float(time))
            rv = _f64(rv)
            nr = rv.size - 1
            nq = next(iter(self.shapes.values()))[1]
            cons = _f64(cons).reshape(nr, nq, 4)
            sm = _f64(sm).reshape(nr, nq)
            self._birth_keep = (rv, cons, sm)
            out[0].nr = nr
            out[0].radial_vertices = rv.ctypes.data
            out[0].conserved = cons.ctypes.data
            out[0].scalar_mass = sm.ctypes.data
            self.shapes[int(index)] = (nr, nq)
            self.vertices[int(index)] = (rv.copy(), self.vertices[min(self.shapes)][1].copy())
            return 0
        except Exception as e:  # reported through the advance call's error
            self._birth_error = e
            return 1

    def _sync_block_set(self):
        n, lo, hi = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(lib().knc_block_count(self._h, ctypes.byref(n), ctypes.byref(lo), ctypes.byref(hi)))
        for k in [k for k in self.shapes if n.value == 0 or k < lo.value or k > hi.value]:
            self.shapes.pop(k)
            self.vertices.pop(k, None)
        mesh = getattr(self, "_birth_mesh", None)
        if n.value and mesh is not None:
            for k in range(lo.value, hi.value + 1):
                if k not in self.shapes:
                    rv, qv = mesh_block_vertices(mesh, k)
                    self.shapes[k] = (rv.size - 1, qv.size - 1)
                    self.vertices[k] = (rv, qv)

    # -- output files (src/io.rs, src/main.rs:50-68) -------------------------
    def write_checkpoint(self, path, config, tasks, version):
        """chkpt.NNNN.cbor: App {state, tasks, config, version} with the state
        encoded from HBM; config and tasks are CBOR items (bytes)."""
        meta, _keep = _record_meta(config, tasks, version)
        _check(lib().knc_write_checkpoint(self._h, os.fsencode(path), ctypes.byref(meta)))

    def write_products(self, path, config, version):
        """prods.NNNN.cbor: Products {time, blocks, config, version} with
        try_to_primitive on the device; raises HydroError if a zone fails."""
        meta, _keep = _record_meta(config, None, version)
        err = Error()
        rc = lib().knc_write_products(self._h, os.fsencode(path), ctypes.byref(meta), ctypes.byref(err))
        if rc:
            self._raise(rc, err)

    def excision_counts(self):
        r, a = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().knc_excision_counts(self._h, ctypes.byref(r), ctypes.byref(a)))
        return r.value, a.value

    # -- autotuner ----------------------------------------------------------------
    def autotune(self, trials=3, force=False):
        """Times every supported stage shape on the current mesh and keeps the
        fastest (knc_autotune); returns the chosen shape and timings."""
        r = TuneResult()
        _check(lib().knc_autotune(self._h, int(trials), int(force), ctypes.byref(r)))
        return dict(stage_threads=r.stage_threads, rows_per_item=r.rows_per_item, stage_passes=r.stage_passes,
                    candidates=r.candidates,
                    from_cache=bool(r.from_cache), ms_per_stage=r.ms_per_stage,
                    default_ms_per_stage=r.default_ms_per_stage)

    def synchronize(self):
        _check(lib().knc_synchronize(self._h))

    def set_timing(self, enable=True):
        _check(lib().knc_set_timing(self._h, int(enable)))

    def last_stage_time(self):
        ms = ctypes.c_double()
        n = ctypes.c_int32()
        _check(lib().knc_last_stage_time(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def c2p_stats(self):
        s = C2PStats()
        _check(lib().knc_get_c2p_stats(self._h, ctypes.byref(s)))
        return dict(hist=np.array(s.hist[:], dtype=np.int64), zones=s.zones, fallbacks=s.fallbacks,
                    floors=s.floors, learned_iters=s.learned_iters)

    def reset_c2p_stats(self):
        _check(lib().knc_reset_c2p_stats(self._h))


# ---- per-zone Hydrodynamics methods ------------------------------------------
class Checkpoint:
    """A checkpoint file (App, src/app.rs:179-186) parsed by the library:
    ``system``, ``time``, ``iteration`` (numer, denom), ``config`` and
    ``tasks`` (the CBOR items, bytes), ``version`` and the blocks."""

    def __init__(self, path):
        h = ctypes.c_void_p()
        _check(lib().knc_checkpoint_open(os.fsencode(path), ctypes.byref(h)))
        self._h = h
        info = CheckpointInfo()
        _check(lib().knc_checkpoint_get_info(self._h, ctypes.byref(info)))
        self.system = info.system
        self.time = info.time
        self.iteration = (info.iteration_numer, info.iteration_denom)
        self.config = ctypes.string_at(info.config, info.config_len)
        self.tasks = ctypes.string_at(info.tasks, info.tasks_len)
        self.version = info.version.decode()
        self.blocks = []
        for k in range(info.num_blocks):
            i, nr, nq = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
            _check(lib().knc_checkpoint_block_shape(self._h, k, ctypes.byref(i), ctypes.byref(nr), ctypes.byref(nq)))
            self.blocks.append((i.value, nr.value, nq.value))

    def block_data(self, k):
        """(conserved (nr, nq, 4), scalar_mass (nr, nq)) of block k."""
        _, nr, nq = self.blocks[k]
        cons = np.zeros((nr, nq, 4))
        sm = np.zeros((nr, nq))
        _check(lib().knc_checkpoint_block_data(self._h, int(k), _p(cons), _p(sm)))
        return cons, sm

    def load(self, ctx, mesh):
        """Every block into an empty Context with the mesh's vertices; sets
        the context's time and iteration."""
        _check(lib().knc_checkpoint_load(self._h, ctx._h, ctypes.byref(mesh)))
        for index, nr, nq in self.blocks:
            rv, qv = mesh_block_vertices(mesh, index)
            ctx.shapes[index] = (nr, nq)
            ctx.vertices[index] = (rv, qv)

    def close(self):
        if getattr(self, "_h", None):
            lib().knc_checkpoint_close(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


NOALLOC_MAX_NR, NOALLOC_MAX_NQ = 8192, 1024


def block_time_step_noalloc(h, index, radial_vertices, polar_vertices, conserved_device, stream=None):
    """State::time_step of one block whose conserved planes D | S_r | S_theta
    | tau (each nr x nq, theta fastest) are at the device address
    `conserved_device`, through the entry point that allocates nothing."""
    rv, qv = _f64(radial_vertices), _f64(polar_vertices)
    dt = ctypes.c_double()
    err = Error()
    rc = lib().knc_block_time_step_noalloc(ctypes.byref(h), int(index), rv.size - 1, qv.size - 1, _p(rv), _p(qv),
                                           ctypes.c_void_p(conserved_device), stream, ctypes.byref(dt),
                                           ctypes.byref(err))
    if rc in ERROR_KINDS:
        raise HydroError(rc, err)
    _check(rc)
    return dt.value


def try_to_primitive(h, conserved, guess=None):
    cons = _f64(conserved).reshape(-1, 4)
    n = cons.shape[0]
    prim = np.zeros((n, 4))
    code = np.zeros(n, np.int32)
    iters = np.zeros(n, np.int32)
    g = None if guess is None else _f64(guess).reshape(-1)
    _check(lib().knc_hydro_try_to_primitive(ctypes.byref(h), n, _p(cons), _p(g), _p(prim), _p(code), _p(iters)))
    return prim, code, iters


def to_conserved(h, primitive):
    prim = _f64(primitive).reshape(-1, 4)
    cons = np.zeros_like(prim)
    _check(lib().knc_hydro_to_conserved(ctypes.byref(h), prim.shape[0], _p(prim), _p(cons)))
    return cons


def intercell_flux(h, pl, pr, sl, sr, direction):
    pl = _f64(pl).reshape(-1, 4)
    pr = _f64(pr).reshape(-1, 4)
    n = pl.shape[0]
    sl = _f64(np.broadcast_to(sl, (n,)))
    sr = _f64(np.broadcast_to(sr, (n,)))
    f = np.zeros((n, 4))
    g = np.zeros(n)
    _check(lib().knc_hydro_intercell_flux(ctypes.byref(h), n, _p(pl), _p(pr), _p(sl), _p(sr), int(direction),
                                          _p(f), _p(g)))
    return f, g


def geometrical_source_terms(h, primitive, r, theta):
    prim = _f64(primitive).reshape(-1, 4)
    n = prim.shape[0]
    r = _f64(np.broadcast_to(r, (n,)))
    theta = _f64(np.broadcast_to(theta, (n,)))
    out = np.zeros((n, 4))
    _check(lib().knc_hydro_geometrical_source_terms(ctypes.byref(h), n, _p(prim), _p(r), _p(theta), _p(out)))
    return out


def max_signal_speed(h, primitive):
    prim = _f64(primitive).reshape(-1, 4)
    out = np.zeros(prim.shape[0])
    _check(lib().knc_hydro_max_signal_speed(ctypes.byref(h), prim.shape[0], _p(prim), _p(out)))
    return out


def model_eval(kind, params, r, theta, time, table=None, ctx=None):
    """primitive_at (n, 4) and scalar_at (n) of an initial model at points
    (r, theta) and `time`, evaluated on the device."""
    r, q = _f64(r).ravel(), _f64(theta).ravel()
    prim = np.empty((r.size, 4))
    scal = np.empty(r.size)
    m, _keep = make_model(kind, params, table)
    own = ctx is None
    if own:
        ctx = Context(relativistic_hydro())
    try:
        rc = lib().knc_model_eval(ctx._h, ctypes.byref(m), r.size, _p(r), _p(q), float(time), _p(prim), _p(scal))
        if rc == ERR_MODEL:
            raise ModelError(lib().knc_last_error_message().decode())
        _check(rc)
    finally:
        if own:
            ctx.close()
    return prim, scal


class ModelError(RuntimeError):
    """An initial model evaluated where the reference panics; raised by a
    run whose model ghost rows fail, with the zone of the mesh block they
    stand for (block, i, j) and its (r, theta) centre."""

    def __init__(self, message="", err=None):
        self.code = ERR_MODEL
        self.block = self.i = self.j = None
        self.r = self.theta = None
        if err is not None:
            self.block, self.i, self.j = err.block, err.i, err.j
            self.r, self.theta = err.r, err.theta
            message = (f"ModelPanics at position (r, theta) = ({err.r:.4e}, {err.theta:.4f}) "
                       f"block {err.block} zone ({err.i}, {err.j})")
        super().__init__(message)


def plm_gradient(plm_theta, a, b, c):
    a, b, c = (_f64(x).reshape(-1) for x in (a, b, c))
    out = np.zeros_like(a)
    _check(lib().knc_hydro_plm_gradient(float(plm_theta), a.size, _p(a), _p(b), _p(c), _p(out)))
    return out


def fp64_fma_rate(chains=8, waves_per_simd=8, iters=4096):
    """FP64 FMAs per second of the device and the timed launch's ms:
    waves_per_simd 256-thread workgroups per compute unit, each lane running
    `chains` independent chains of `iters` dependent FMAs."""
    rate, ms = ctypes.c_double(), ctypes.c_float()
    _check(lib().knc_fp64_fma_rate(int(chains), int(waves_per_simd), int(iters), ctypes.byref(rate), ctypes.byref(ms)))
    return rate.value, ms.value


def fp64_probe(a, b, c):
    """Device a/b, sqrt|a|, a*b+c (two roundings) and a/|b| as the stage
    divides U by dV."""
    a, b, c = (_f64(x).reshape(-1) for x in (a, b, c))
    div, sq, mad, quot = np.zeros_like(a), np.zeros_like(a), np.zeros_like(a), np.zeros_like(a)
    _check(lib().knc_fp64_probe(a.size, _p(a), _p(b), _p(c), _p(div), _p(sq), _p(mad), _p(quot)))
    return div, sq, mad, quot
