"""Same-GPU continuous-batching ray-trace serving (requests, admission, step batches)."""
from .engine import Admission, Response, ServingEngine, run_offline  # noqa: F401
from .requests import BeamSpec, FieldSpec, RequestSpec, materialize  # noqa: F401
