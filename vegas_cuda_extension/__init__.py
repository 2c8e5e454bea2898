"""Import path of the reference package (vegas_cuda/src/vegas_cuda_extension/__init__.py),
backed by the MI355X implementation in cuvegas_amd."""
from .python_wrapper import integrate  # noqa: F401
