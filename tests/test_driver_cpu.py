"""Host-side parts of the main.jl driver: file names and Julia's float text."""
import math

import numpy as np

from hyperelasticsolver_amd.driver import HEADER, get_filename, julia_float


def test_filename_and_header():
    assert get_filename(0) == "sol_000000.csv"
    assert get_filename(1200) == "sol_001200.csv"
    cols = HEADER.split("\t")
    assert len(cols) == 30 and cols[0] == "a1" and cols[15] == "a2" and cols[-1] == "F332"


def test_julia_float_text():
    # Julia's `string(::Float64)`: decimal for decimal exponents -4..5
    cases = [(0.0, "0.0"), (-0.0, "-0.0"), (1.0, "1.0"), (0.5, "0.5"), (12.0, "12.0"),
             (100000.0, "100000.0"), (1e6, "1.0e6"), (1234567.0, "1.234567e6"),
             (123456789.0, "1.23456789e8"), (0.0001, "0.0001"), (0.00012, "0.00012"),
             (1e-5, "1.0e-5"), (1.2345e-7, "1.2345e-7"), (-2.5e-5, "-2.5e-5"), (8.93, "8.93"),
             (1e16, "1.0e16"), (5e-324, "5.0e-324"), (math.inf, "Inf"), (-math.inf, "-Inf")]
    for x, s in cases:
        assert julia_float(x) == s, (x, julia_float(x), s)
    assert julia_float(math.nan) == "NaN"


def test_julia_float_round_trips():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-12, 12, 2000),
                         rng.uniform(-1, 1, 500)])
    for x in xs:
        assert float(julia_float(x)) == x
