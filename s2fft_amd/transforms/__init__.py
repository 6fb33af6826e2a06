from . import spherical, wigner  # noqa: F401
