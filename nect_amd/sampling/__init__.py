"""Sampling: the ``ct_sampling`` drop-in, ``Geometry`` and ``Projector``."""
