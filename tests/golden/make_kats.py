"""Writes reference_kats.json: the known-answer values asserted by the
reference's own unit tests (unit_tests/kernel_test.cpp,
unit_tests/smoothing_test.cpp:33-147, unit_tests/grad_smoothing_test.cpp:26-70),
as plain inputs / expected outputs.  The expected values are the closed
forms those tests assert (e.g. 20/(7 pi)), evaluated here in FP64."""
import json
import math
import os

pi = math.pi
kats = {
    "tolerance": 1e-14,
    "kernel": [  # [D, distance, h, expected W]
        [1, 0.5, 1.0, 23 / 48], [1, 1.0, 1.0, 1 / 6], [1, 1.5, 1.0, 1 / 48], [1, 2.0, 1.0, 0.0],
        [2, 0.5, 1.0, 115 / 112 / pi], [2, 1.0, 1.0, 5 / 14 / pi], [2, 1.5, 1.0, 5 / 112 / pi], [2, 2.0, 1.0, 0.0],
        [3, 0.5, 1.0, 23 / 32 / pi], [3, 1.0, 1.0, 1 / 4 / pi], [3, 1.5, 1.0, 1 / 32 / pi], [3, 2.0, 1.0, 0.0],
    ],
    "grad_kernel_2d": [  # [pos_a, pos_b, h, expected grad]
        [[0.25, 0.0], [-0.25, 0.0], 1.0, [-75 / 56 / pi, 0.0]],
        [[-0.25, 0.0], [0.25, 0.0], 1.0, [75 / 56 / pi, 0.0]],
        [[0.0, 0.25], [0.0, -0.25], 1.0, [0.0, -75 / 56 / pi]],
        [[0.5, 0.0], [-0.5, 0.0], 1.0, [-15 / 14 / pi, 0.0]],
        [[0.75, 0.0], [-0.75, 0.0], 1.0, [-15 / 56 / pi, 0.0]],
        [[-0.75, 0.0], [0.75, 0.0], 1.0, [15 / 56 / pi, 0.0]],
        [[0.75, 0.0], [0.75, 0.0], 1.0, [0.0, 0.0]],
        [[0.375, 0.0], [-0.375, 0.0], 0.5, [-15 / 56 / pi / 0.125, 0.0]],
        [[0.0, 0.375], [0.0, -0.375], 0.5, [0.0, -15 / 56 / pi / 0.125]],
        [[0.125, 0.0], [-0.125, 0.0], 0.5, [-75 / 56 / pi / 0.125, 0.0]],
    ],
    "smoothing": {  # hT = 1; particle 0 expectations
        "x": [0.0, 0.0, 0.8, 1.5], "y": [0.0, 0.5, 0.6, 0.0],
        "mass": {"s": 1, "rhoB": 2, "rhoQ": 3, "rhoS": 4},
        "extensive": {"s": 1, "rhoB": 2, "rhoQ": 3, "rhoS": 4},
        "expected_p0": {"s": 20 / 7 / pi, "rhoB": 4 * 20 / 7 / pi, "rhoQ": 9 * 20 / 7 / pi,
                        "rhoS": 16 * 20 / 7 / pi, "sigma_lab": 20 / 7 / pi},
        "hT": 1.0,
    },
    "grad_smoothing": {  # hT = 1, sigma_lab = 1, sph_mass_s = 1
        "x": [0.0, 0.5, 1.0, 1.5, -1.5], "y": [0.0, 0.0, 0.0, 0.0, 1.5],
        "p": [1.0, 0.5, 1.5, 2.5, 1.0],
        "expected_gradP_p0": [45 / 8 / pi, 0.0],
        "hT": 1.0,
    },
}
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_kats.json")
with open(out, "w") as f:
    json.dump(kats, f, indent=1)
print(out)
