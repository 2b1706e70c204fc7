"""Build and load the native parts: ``libtensorforge_rt.so`` and generated kernel objects."""
from .build import (LIB_DIR, build_runtime, compile_generators, hipcc_flags,  # noqa: F401
                    runtime_path)
from .loader import Kernel, load_kernels, runtime  # noqa: F401
