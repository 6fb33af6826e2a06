from . import s2_samples, so3_samples  # noqa: F401
