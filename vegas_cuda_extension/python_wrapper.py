"""Drop-in for vegas_cuda_extension.python_wrapper (python_wrapper.py:9-77)."""
from cuvegas_amd.api import integrate  # noqa: F401
