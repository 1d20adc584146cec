"""voss_amd: the voss prime library (wheel-30 segmented sieve, gap
statistics, range queries) on AMD Instinct MI355X.

Drop-in for the reference Python package `voss`: `import voss_amd as voss`
gives the same modules (primes, conjectures, numbers) and names.  The
extension modules `stats` (one-pass full statistics, batched ranges, GPU
Miller-Rabin) and `serve` (continuous-batching request loop) have no
reference counterpart.
"""

from . import primes, conjectures, numbers, stats, serve
from .primes import Context, Statistics, ChebyshevBias, LargeGap
from .exceptions import VossError

__version__ = "0.7.0"

__all__ = ["primes", "conjectures", "numbers", "stats", "Context", "Statistics",
           "ChebyshevBias", "LargeGap", "VossError", "serve"]
