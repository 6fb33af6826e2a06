from . import price_mcewen  # noqa: F401
