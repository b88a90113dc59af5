"""Writes tests/golden/exact_dense_fixture.json: the registered dense KSS
fixture of fevc/tests/python/test_dense_oracle.py:15-76 (6 workers x 4
periods, 4 firms, two time controls, match identifiers) with its expected
fixed-offset plug-in and correction (test_dense_oracle.py:54-76), and the
seven-edge design of test_dense_oracle.py:186-200 with its closed-form
leverages.  Run from the repository root: python tests/golden/make_exact_fixture.py
"""
import json
import os

import numpy as np

worker = np.repeat(np.arange(6), 4)
firm = np.array([0, 0, 1, 1, 0, 2, 2, 1, 1, 2, 3, 3, 2, 3, 0, 0, 3, 1, 1, 2, 3, 3, 2, 0])
period = np.tile(np.arange(4), 6)
controls = np.column_stack((period - period.mean(), (period == 2).astype(float)))
noise = np.array([0.2, -0.1, 0.1, -0.2, -0.2, 0.3, -0.1, 0.1, 0.1, -0.2, 0.2, -0.1,
                  -0.1, 0.2, -0.2, 0.1, 0.3, -0.2, 0.1, -0.2, -0.2, 0.1, 0.2, -0.1])
y = 1.5 + 0.3 * worker - 0.2 * firm + controls @ np.array([0.4, -0.15]) + noise
match = np.array([10, 10, 11, 11, 20, 21, 21, 22, 30, 31, 32, 32, 40, 41, 42, 42, 50, 51, 51, 52,
                  60, 60, 61, 62])
out = {
    "dense_fixture": {
        "y": y.tolist(), "worker": worker.tolist(), "firm": firm.tolist(),
        "controls": controls.tolist(), "match": match.tolist(),
        "fixedoffset_match_plugin": [0.23886949630417761, 0.02942901912684345,
                                     -0.02694262626105621, 0.21441326290890872],
        "fixedoffset_match_correction": [-0.00467505214166742, 0.00105611329461117,
                                         -0.00744004087996870, -0.01849902060699364],
        "tolerance": 2e-13,
    },
    "seven_edges": {
        "edges": [[2, 2], [1, 1], [0, 2], [2, 1], [1, 0], [0, 1], [2, 0]],
        "leverage": [11 / 15, 11 / 15, 11 / 15, 3 / 5, 11 / 15, 11 / 15, 11 / 15],
    },
    "known_effects": {
        "worker": [1, 1, 2, 2], "firm": [1, 2, 1, 2], "outcome": [1.5, 0.5, -0.5, -1.5],
        "worker_effect": [1.0, -1.0], "firm_effect": [0.5, -0.5],
    },
}
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "exact_dense_fixture.json")
with open(path, "w") as fh:
    json.dump(out, fh, indent=1)
print(path)
