from . import spherical  # noqa: F401
