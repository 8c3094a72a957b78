"""The reference's problem files (its input.txt and indoor.txt, kept as
fixtures under tests/golden/problems) read by pump_read_problem
(mainPUMP.cu:152-230), and the Halton sampler (sampler.cu:20-57) of the
library against the oracle and against the radical inverse."""
import os
import re
from fractions import Fraction

import numpy as np
import pytest

import oracle
import pump_amd as P

O = oracle.load()
HERE = os.path.dirname(os.path.abspath(__file__))
FILES = ["input.txt", "indoor.txt"]


def expected_values(path):
    """The file's comma-separated fields, each read as the number at its
    start (after white space), as the reference's stringstream does."""
    text = open(path).read().split(",")
    num = re.compile(r"\s*([-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?)")
    vals = []
    for tok in text:
        m = num.match(tok)
        vals.append(m.group(1) if m else None)
    return vals


@pytest.mark.parametrize("name", FILES)
def test_read_problem_file(name):
    path = os.path.join(HERE, "golden", "problems", name)
    pf = P.read_problem(path)
    v = expected_values(path)
    f = [np.float32(x) for x in v[:25]]
    assert np.array_equal(pf["init"], f[0:6]) and np.array_equal(pf["goal"], f[6:12])
    assert np.array_equal(pf["lo"], f[12:18]) and np.array_equal(pf["hi"], f[18:24])
    assert pf["offset_mult"] == f[24]
    n_obs = int(v[25])
    assert pf["obstacles"].shape == (n_obs, 12)
    obs = np.array([np.float32(x) for x in v[26: 26 + 12 * n_obs]], np.float32).reshape(-1, 12)
    assert np.array_equal(pf["obstacles"], obs)
    # lowest corner first, velocity bounds +-3.1, as the files state
    assert np.all(obs[:, :6] <= obs[:, 6:]) and np.all(obs[:, 3:6] == np.float32(-3.1))


def test_read_problem_errors(tmp_path):
    short = tmp_path / "short.txt"
    short.write_text("0.1,0.1,0.1,0,0,0,\n0.9,0.9,\n")
    with pytest.raises(P.PumpError):
        P.read_problem(str(short))
    many = tmp_path / "many.txt"
    many.write_text("0,0,0,0,0,0,1,1,1,0,0,0,0,0,0,-1,-1,-1,1,1,1,1,1,1,0.5,3,"
                    + ",".join(["0"] * 36))
    with pytest.raises(P.PumpError):
        P.read_problem(str(many), obs_cap=2)
    assert P.read_problem(str(many), obs_cap=3)["obstacles"].shape == (3, 12)


def radical_inverse(n, b):
    r, f = Fraction(0), Fraction(1, b)
    while n > 0:
        n, d = divmod(n, b)
        r += f * d
        f /= b
    return r


def test_halton_samples():
    pf = P.read_problem(os.path.join(HERE, "golden", "problems", "indoor.txt"))
    num = 2000
    got = P.halton_samples(num, pf["init"], pf["goal"], pf["lo"], pf["hi"])
    ref = np.zeros((num, 6), np.float32)
    args = [np.ascontiguousarray(pf[k], np.float32) for k in ("init", "goal", "lo", "hi")]
    assert O.po_halton_samples(num, 0, *[oracle.fp(a) for a in args], oracle.fp(ref)) == 0
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(got[0], pf["init"]) and np.array_equal(got[-1], pf["goal"])
    unit = np.zeros((num, 6), np.float32)
    zeros, ones = np.zeros(6, np.float32), np.ones(6, np.float32)
    for skip in (0, 5):
        u = P.halton_samples(num, zeros, zeros, zeros, ones, skip=skip)
        primes = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31]
        for d in range(6):
            b = primes[d + skip]
            exact = np.array([float(radical_inverse(n, b)) for n in range(1, num - 1)])
            col = u[1:-1, d].astype(np.float64)
            if b == 2:
                assert np.array_equal(col, exact)      # dyadic: exact in float
            else:
                assert np.max(np.abs(col - exact) / np.maximum(exact, 1e-30)) < 4e-7
        unit = u
    assert np.all((unit[1:-1] >= 0) & (unit[1:-1] < 1))
    with pytest.raises(P.PumpError):
        P.halton_samples(10, zeros, zeros, zeros, ones, skip=20)
