"""The reference's output files read into numpy: what its Python extension
knc_loader offers (``knc_loader.app``, ``knc_loader.products``,
loader/src/lib.rs:337-355) for files written by either code.

``read_products(path)`` gives the Products struct (src/products.rs:14-37):
``time``, ``blocks`` {(i, j): {radial_vertices, polar_vertices, primitive
(nr, nq, 4) in AnyPrimitive order (u_r, u_theta, rho, p), scalar (nr,
nq)}}, ``config`` and ``version``.  ``read_app(path)`` gives the App of a
checkpoint (src/app.rs:179-186) with the state's arrays as numpy: ``system``
("Relativistic" or "Newtonian"), ``time``, ``iteration`` (numer, denom),
``blocks`` {(i, j): (conserved (nr, nq, 4), scalar_mass (nr, nq))},
``tasks``, ``config`` and ``version``.  Long arrays are decoded by the
library (knc_cbor_read_f64) when it is built, by Python otherwise.
"""
import numpy as np

from . import cbor


def _array(a, tail=()):
    """An ndarray struct {v, dim, data} as a numpy array of shape dim + tail."""
    if a.get("v") != 1:
        raise cbor.CBORError("unknown ndarray format version")
    return np.asarray(a["data"], dtype=np.float64).reshape(tuple(a["dim"]) + tuple(tail))


def _load(path):
    with open(path, "rb") as f:
        return cbor.decode(f.read(), arrays=True)


def read_products(path):
    d = _load(path)
    blocks = {}
    for key, b in d["blocks"].items():
        blocks[tuple(key)] = {"radial_vertices": _array(b["radial_vertices"]),
                              "polar_vertices": _array(b["polar_vertices"]),
                              "primitive": _array(b["primitive"], (4,)), "scalar": _array(b["scalar"])}
    return {"time": float(d["time"]), "blocks": blocks, "config": d["config"], "version": d["version"]}


def read_app(path):
    d = _load(path)
    (system, st), = d["state"].items()
    blocks = {tuple(key): (_array(b["conserved"], (4,)), _array(b["scalar_mass"]))
              for key, b in st["solution"].items()}
    numer, denom = st["iteration"]
    return {"system": system, "time": float(st["time"]), "iteration": (int(numer), int(denom)), "blocks": blocks,
            "tasks": d["tasks"], "config": d["config"], "version": d["version"]}


def radial_profile(products, polar_index, field):
    """knc_loader's RadialProfile (loader/src/lib.rs:113-206): `field` of
    every block at one polar index, concatenated in radial order; field is
    "scalar", "radial_four_velocity", "polar_four_velocity",
    "comoving_mass_density", "gas_pressure" or "vertices"."""
    keys = sorted(products["blocks"])
    if field == "vertices":
        return np.concatenate([products["blocks"][k]["radial_vertices"][:-1] for k in keys])
    col = {"radial_four_velocity": 0, "polar_four_velocity": 1, "comoving_mass_density": 2, "gas_pressure": 3}
    if field == "scalar":
        return np.concatenate([products["blocks"][k]["scalar"][:, polar_index] for k in keys])
    return np.concatenate([products["blocks"][k]["primitive"][:, polar_index, col[field]] for k in keys])
