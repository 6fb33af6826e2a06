"""Learned preconditioner for the iterative forward transform."""

from .precond import MLPPreconditioner, make_training_rows, train  # noqa: F401
