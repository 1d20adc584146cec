"""Exception types (same hierarchy as bindings/python/voss/exceptions.py)."""


class VossError(RuntimeError):
    """Base class of every error raised by the library."""
