"""Mesh databases for the single-partition solver loop.

``database.Database`` holds exactly the arrays ``prepare_constants_device`` /
``prepare_fields_*_device`` receive from the Fortran host, built either by the box
generator (``box.build_box``) or from a CUBIT-format external mesh
(``external.read_external_mesh``).
"""
from .database import Database  # noqa: F401
