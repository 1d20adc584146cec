"""Arithmetic functions on the host CPU (same functions and errors as
bindings/python/voss/numbers.py)."""

import ctypes

from . import _capi
from .exceptions import VossError

__all__ = ["phi", "tau", "sigma", "mu", "gcd", "lcm", "fibonacci", "factorial"]


def _run(name, fn, args, out):
    rc = fn(*args, ctypes.byref(out))
    if rc != 0:
        msg = _capi._lib.voss_get_last_error()
        raise VossError(msg.decode() if msg else f"{name} failed rc={rc}")
    return int(out.value)


def _unary(name):
    fn = getattr(_capi._lib, name)

    def f(n: int) -> int:
        if not isinstance(n, int):
            raise TypeError(f"n must be int, got {type(n).__name__}")
        return _run(name, fn, (ctypes.c_uint64(n),), ctypes.c_uint64(0))
    return f


phi = _unary("voss_numbers_phi")
phi.__doc__ = "Euler's totient (phi(1) = 1; n >= 1)."
tau = _unary("voss_numbers_tau")
tau.__doc__ = "Number of positive divisors (n >= 1)."
sigma = _unary("voss_numbers_sigma")
sigma.__doc__ = "Sum of positive divisors (n >= 1)."
fibonacci = _unary("voss_numbers_fibonacci")
fibonacci.__doc__ = "F(n), F(0) = 0 (n <= 93)."
factorial = _unary("voss_numbers_factorial")
factorial.__doc__ = "n! (n <= 20)."


def mu(n: int) -> int:
    """Moebius function (n >= 1)."""
    if not isinstance(n, int):
        raise TypeError(f"n must be int, got {type(n).__name__}")
    return _run("voss_numbers_mu", _capi._lib.voss_numbers_mu, (ctypes.c_uint64(n),), ctypes.c_int(0))


def _binary(name):
    fn = getattr(_capi._lib, name)

    def f(a: int, b: int) -> int:
        if not isinstance(a, int) or not isinstance(b, int):
            raise TypeError("a and b must be int")
        return _run(name, fn, (ctypes.c_uint64(a), ctypes.c_uint64(b)), ctypes.c_uint64(0))
    return f


gcd = _binary("voss_numbers_gcd")
gcd.__doc__ = "Greatest common divisor (gcd(0, 0) = 0)."
lcm = _binary("voss_numbers_lcm")
lcm.__doc__ = "Least common multiple (0 if either is 0; error on overflow)."
