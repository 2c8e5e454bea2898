"""Drop-in for the native module vegas_cuda_extension.cuda_extension (cuda_wrapper.cu:874-876)."""
from cuvegas_amd.cuda_extension import integrate_cuda  # noqa: F401
