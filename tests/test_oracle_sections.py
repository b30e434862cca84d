"""The oracle's sample-section list and step placement against the reference's
own known-answer tests (unittests/marching_steps.cpp), run in float and double
as the reference's TEMPLATE_TEST_CASE does."""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle as O

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "density_kats.json")))


def val(lit, dbl):
    v = float(np.float32(lit["value"])) if lit["f32"] else lit["value"]
    return v if dbl else float(np.float32(v))


def run_list(dbl, cap, start, puts):
    arr = np.ascontiguousarray(puts, np.float64).reshape(-1, 3) if puts else np.zeros((1, 3))
    out = np.zeros((cap, 3))
    end = np.zeros(1)
    size = O.lib().orc_sections_run(int(dbl), cap, start, len(puts), O.P(arr), O.P(out), O.P(end))
    assert size >= 0
    return out[:size], end[0]


@pytest.mark.parametrize("dbl", [0, 1], ids=["float", "double"])
@pytest.mark.parametrize("case", [c for c in KATS["cases"] if c["kind"] == "list"], ids=lambda c: c["section"])
def test_section_list_kat(case, dbl):
    start = val(case["start"], dbl)
    puts = []
    state = None
    for op in case["ops"]:
        if op["op"] == "put":
            puts.append([val(a, dbl) for a in op["args"]])
            state = None
            continue
        if state is None:
            state = run_list(dbl, case["cap"], start, puts)
        sections, _end = state
        if op["op"] == "size":
            assert len(sections) == op["expect"], (case["section"], len(puts))
        else:
            col = {"start": 0, "end": 1, "delta_t": 2}[op["field"]]
            assert sections[op["index"], col] == val(op["expect"], dbl), (case["section"], op)


@pytest.mark.parametrize("dbl", [0, 1], ids=["float", "double"])
@pytest.mark.parametrize("case", [c for c in KATS["cases"] if c["kind"] == "combine"], ids=lambda c: c["section"])
def test_combine_kat(case, dbl):
    a = np.array([val(x, dbl) for x in case["a"]])
    b = np.array([val(x, dbl) for x in case["b"]])
    out = np.zeros(9)
    O.lib().orc_sections_combine(int(dbl), O.P(a), O.P(b), O.P(out))
    out = out.reshape(3, 3)
    for chk in case["checks"]:
        col = {"start": 0, "end": 1, "delta_t": 2}[chk["field"]]
        assert out[chk["index"], col] == val(chk["expect"], dbl), chk


@pytest.mark.parametrize("dbl", [0, 1], ids=["float", "double"])
def test_randomised_brute_force(dbl):
    """marching_steps.cpp:939-1003: std::srand(0), 1000 lists x 100 insertions,
    compared with the brute-force minimum step at random positions."""
    assert O.lib().orc_sections_randomised_check(int(dbl)) == 0


def sample(dbl, cap, start, puts, n):
    arr = np.ascontiguousarray(puts, np.float64).reshape(-1, 3) if puts else np.zeros((1, 3))
    out = np.zeros(n)
    assert O.lib().orc_sections_sample(int(dbl), cap, start, len(puts), O.P(arr), n, O.P(out)) == 0
    return out


def test_sampler_empty():
    """marching_steps.cpp:40-47: an empty list yields all-zero borders."""
    assert np.all(sample(0, 4, np.float32(0.1), [], 16) == 0)


def test_sampler_filled():
    """marching_steps.cpp:49-65: steps never exceed the section's step size."""
    f = lambda v: float(np.float32(v))
    puts = [[f(0.2), f(1.0), f(0.1)], [f(1.0), f(2.0), f(0.2)], [f(2.0), f(2.5), f(0.5)], [f(3.0), f(4.0), f(1.0)]]
    s = sample(0, 4, f(0.1), puts, 16).astype(np.float32)
    assert s[0] == np.float32(0.2)
    sections, _ = run_list(0, 4, f(0.1), puts)

    def target(t):
        for st, en, dt in sections:
            if st <= t < en:
                return dt
        return np.inf

    for i in range(1, 16):
        d = s[i] - s[i - 1]
        assert d > 0
        t = s[i - 1]
        while t < s[i]:
            assert d <= target(t) + 0.001
            t = np.float32(t + np.float32(0.01))


def test_sampler_walks_gaps_with_next_step_and_pads_with_last_end():
    """sample<n> (marching_steps.h:202-220): a gap is walked with the next
    section's step; past the last section every border equals its end."""
    puts = [[0.0, 1.0, 0.25], [2.0, 3.0, 0.5]]
    s = sample(1, 16, 0.0, puts, 16)
    np.testing.assert_allclose(s[:5], [0, 0.25, 0.5, 0.75, 1.0])
    np.testing.assert_allclose(s[5:9], [1.5, 2.0, 2.5, 3.0])
    assert np.all(s[9:] == 3.0)
