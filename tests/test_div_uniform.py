"""The wave kernel divides by its uniform divisors (sigm in the action
staging, dtheta in the theta flux) as (float)((double)a * RN53(1/b)) when
0 < b < 2 (div_uniform, xbeach_gpu_amd/csrc/xbg_device.h).  That equals the
IEEE float quotient a / b for every float a: the double product is within
2^-52 of a/b, and a float quotient that is not a float lies more than 2^-49
from every rounding midpoint (an exact tie needs a subnormal quotient and
b >= 2).  Checked here on the host, whose double multiply and conversions
are the same IEEE operations: exhaustively over all 2^32 numerators for the
divisors of the test and benchmark configurations and for divisors at the
edges of the range, on random divisors, and with the counter-example that
makes b >= 2 unsafe."""
import math
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("divu") / "div_uniform_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-pthread",
                    os.path.join(HERE, "native", "div_uniform_check.cpp"), "-o", exe], check=True)
    return exe


def run(exe, *args):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=600)
    return r.returncode, r.stdout


def f32(x):
    return float(np.float32(x))


# sigm = 2 pi / Trep for the configurations' Trep, dtheta = 10 and 5 degrees
# in radians, and divisors at the ends of (0, 2)
@pytest.mark.parametrize("b", [f32(2 * math.pi / 10.0), f32(2 * math.pi / 12.0), f32(math.radians(10.0)),
                               f32(math.radians(5.0)), float(np.nextafter(np.float32(2.0), np.float32(0))),
                               f32(2 * math.pi / 3.0), 1.0])
def test_exhaustive_over_numerators(checker, b):
    rc, out = run(checker, "exhaustive", repr(b))
    assert rc == 0, out


def test_random_divisors(checker):
    rc, out = run(checker, "random", 2000, 200000)
    assert rc == 0, out


@pytest.mark.parametrize("dx", [0.5, 1.0, 2.0, 4.0, 8.0, 1024.0])
def test_power_of_two_spacing_exhaustive(checker, dx):
    """The flow kernels' instantiation for a power-of-two dx takes v * (1/dx)
    for v / dx: the same real number, so the same rounding for all 2^32
    numerators (subnormal results, zeros, infinities and NaN included)."""
    rc, out = run(checker, "pow2", repr(dx))
    assert rc == 0, out


def test_divisor_two_or_more_can_tie(checker):
    """Why the kernel keeps the IEEE division for b >= 2."""
    rc, out = run(checker, "tie")
    assert rc == 0, out
