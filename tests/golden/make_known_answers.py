"""Writes tests/golden/known_answers.json: published known-answer values of
two Riemann problems used to pin the oracle's flux functions.

Sod (Toro 2009, Table 4.2 test 1; star state from Table 4.3) and the SRHD
blast wave 1 of Marti & Mueller (2003), Problem 1.  The numbers are copied
from those publications, not computed by this repository's code.
"""
import json
import os

DATA = {
    "_source": "Published known-answer values; tests/golden/make_known_answers.py writes this file.",
    "sod_toro_test1": {
        "reference": "E. F. Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics, 3rd ed. (2009), "
                     "Table 4.2 test 1 / Table 4.3",
        "gamma": 1.4, "left": [1.0, 0.0, 1.0], "right": [0.125, 0.0, 0.1], "t": 0.2,
        "p_star": 0.30313, "u_star": 0.92745,
    },
    "srhd_blast_wave_1": {
        "reference": "J. M. Marti and E. Mueller, Living Rev. Relativ. 6 (2003) 7, sec. 6.2, Problem 1 (Table 5)",
        "gamma": 5.0 / 3.0, "left": [10.0, 0.0, 13.33], "right": [1.0, 0.0, 1e-6], "t": 0.4,
        "p_star": 1.445, "v_star": 0.714,
    },
}

if __name__ == "__main__":
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.json"), "w") as f:
        json.dump(DATA, f, indent=2)
        f.write("\n")
