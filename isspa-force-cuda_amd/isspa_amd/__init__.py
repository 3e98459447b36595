"""MI355X-native IS-SPA MD engine (Python side).

* ``system`` / ``io`` / ``synth``: system container, reference file formats,
  seeded benchmark workloads (numpy only, no GPU needed);
* ``capi`` / ``engine``: the C ABI of libisspa_amd.so and a torch-backed
  device front end (import them explicitly; they require the built library).
"""
from .system import IsspaTables, System, UmbrellaRestraint  # noqa: F401

__all__ = ["IsspaTables", "System", "UmbrellaRestraint"]
