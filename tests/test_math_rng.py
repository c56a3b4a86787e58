"""The shared Philox stream and the trace's own transcendental functions."""
import math

import numpy as np

from oracle import oracle as O


def test_philox_known_answers():
    """Random123's published philox4x32-10 known-answer vectors, through the first
    draw of a block: u = ((w0>>5)*2^26 + (w1>>6)) * 2^-53."""
    def u01(a, b):
        return ((a >> 5) * 67108864.0 + (b >> 6)) / 9007199254740992.0
    # ctr = 0, key = 0  ->  6627e8d5 e169c58d bc57ac4c 9b00dbd8
    assert O.philox_draw(0, 0, 0) == u01(0x6627e8d5, 0xe169c58d)
    assert O.philox_draw(0, 0, 1) == u01(0xbc57ac4c, 0x9b00dbd8)


def test_philox_stream_layout():
    """Draws 2b and 2b+1 share a block; different paths and seeds differ."""
    a = [O.philox_draw(7, 3, d) for d in range(8)]
    assert len(set(a)) == 8
    assert O.philox_draw(7, 4, 0) != a[0]
    assert O.philox_draw(8, 3, 0) != a[0]
    assert all(0.0 <= x < 1.0 for x in a)


def test_uniform_moments():
    u = np.array([O.philox_draw(11, p, d) for p in range(2000) for d in range(5)])
    assert abs(u.mean() - 0.5) < 5 * math.sqrt(1 / 12 / len(u))
    assert abs((u ** 2).mean() - 1 / 3) < 0.01


def _ulps(a, b):
    return abs(a - b) / math.ulp(abs(b)) if b != 0 else abs(a)


def test_own_log_accuracy_and_special_values():
    L = O.lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.random(20000), np.exp(rng.uniform(-700, 700, 2000)),
                         [1.0, 2.0, 0.5, 2 ** -53, 1 - 2 ** -53, 5e-324]])
    worst = max(_ulps(L.oracle_pt_log(float(x)), math.log(x)) for x in xs if x > 0)
    assert worst <= 4
    assert L.oracle_pt_log(1.0) == 0.0
    assert L.oracle_pt_log(0.0) == -math.inf
    assert math.isnan(L.oracle_pt_log(-1.0))


def test_own_sincos_accuracy():
    import ctypes
    L = O.lib()
    s, c = ctypes.c_double(), ctypes.c_double()
    rng = np.random.default_rng(2)
    worst = 0.0
    for u in rng.random(20000):
        x = float(u) * (math.pi * 2)
        L.oracle_pt_sincos(x, ctypes.byref(s), ctypes.byref(c))
        worst = max(worst, abs(s.value - math.sin(x)), abs(c.value - math.cos(x)))
        assert abs(s.value ** 2 + c.value ** 2 - 1) < 1e-15
    assert worst < 4e-16
    L.oracle_pt_sincos(0.0, ctypes.byref(s), ctypes.byref(c))
    assert (s.value, c.value) == (0.0, 1.0)
