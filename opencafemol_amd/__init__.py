"""MI355X-native MD step for OpenCafeMol force fields (HIP kernels behind a C ABI)."""
from . import abi  # noqa: F401

__version__ = "0.1.0"
