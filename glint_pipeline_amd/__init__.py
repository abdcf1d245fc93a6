"""MI355X-native contact tracing: GLINT's partition -> refine -> evaluate path
as HIP kernels for gfx950 behind the C ABI of include/glint.h."""
from .gen import Generator, Workload, plain_cpu_reference, plain_gpu_form, skewed_density, uniform_city
from .runtime import (CANDIDATES_COMPLETE, CANDIDATES_REFERENCE, PRECISION_FP32_GUARD, PRECISION_FP64,
                      DeviceBuffer, DeviceWalkers, GlintError, Tracer, Workbench, device_count, lib,
                      make_config)

__all__ = [
    "Generator", "Workload", "plain_cpu_reference", "plain_gpu_form", "skewed_density", "uniform_city",
    "CANDIDATES_COMPLETE", "CANDIDATES_REFERENCE", "PRECISION_FP32_GUARD", "PRECISION_FP64", "DeviceBuffer",
    "DeviceWalkers", "GlintError", "Tracer", "Workbench", "device_count", "lib", "make_config",
]
