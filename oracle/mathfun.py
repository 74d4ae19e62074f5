"""Scalar helpers of the HS transform (test infrastructure only).

Follows common/mathHao/source/simple_fun.cpp.  Pinned by the literal
known-answer values of common/mathHao/test/simple_fun_test.cpp:8-64
(tests/test_oracle_math.py).
"""
import cmath
import math


def solveCoshxEqExpy(y: float) -> complex:
    """x with cosh(x) = exp(y); pure real or pure imaginary (simple_fun.cpp:7-17)."""
    ey = complex(math.exp(y), 0.0)
    gamma = cmath.log(ey - cmath.sqrt(ey * ey - 1.0))
    if abs(gamma.real) < abs(gamma.imag):
        return complex(0.0, gamma.imag)
    return complex(gamma.real, 0.0)


def solveCosxEqExpy(y: float) -> complex:
    """x with cos(x) = exp(y) (simple_fun.cpp:23-33)."""
    ey = complex(math.exp(y), 0.0)
    gamma = cmath.acos(ey)
    if abs(gamma.real) < abs(gamma.imag):
        return complex(0.0, gamma.imag)
    return complex(gamma.real, 0.0)


def exp2by2Matrix(a: float, b: float, c: complex):
    """exp of [[a, conj(c)], [c, b]] returned as (a', b', c') (simple_fun.cpp:43-91)."""
    if abs(c) < 1e-60:
        return math.exp(a), math.exp(b), 0j
    abs_c = abs(c)
    arg_c = cmath.phase(c)
    im = 1j
    if abs(a - b) < 1e-60:
        norm = 1.0 / math.sqrt(2.0)
        d0 = math.exp(-abs_c + a)
        d1 = math.exp(abs_c + b)
        v00 = complex(-norm)
        v10 = norm * cmath.exp(im * arg_c)
        v01 = complex(norm)
        v11 = norm * cmath.exp(im * arg_c)
    else:
        xi = 0.5 * math.atan(2.0 * abs_c / (a - b))
        d0 = math.exp(0.5 * math.cos(2.0 * xi) * (a - b) + math.sin(2.0 * xi) * abs_c + (a + b) * 0.5)
        d1 = math.exp(-0.5 * math.cos(2.0 * xi) * (a - b) - math.sin(2.0 * xi) * abs_c + (a + b) * 0.5)
        v00 = cmath.exp(-im * arg_c * 0.5) * math.cos(xi)
        v10 = cmath.exp(im * arg_c * 0.5) * math.sin(xi)
        v01 = -cmath.exp(-im * arg_c * 0.5) * math.sin(xi)
        v11 = cmath.exp(im * arg_c * 0.5) * math.cos(xi)
    a2 = (d0 * v00 * v00.conjugate() + d1 * v01 * v01.conjugate()).real
    b2 = (d0 * v10 * v10.conjugate() + d1 * v11 * v11.conjugate()).real
    c2 = d0 * v10 * v00.conjugate() + d1 * v11 * v01.conjugate()
    return a2, b2, c2
