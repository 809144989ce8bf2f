"""MI355X-native (gfx950) Lyapunov-spectrum parameter sweep.

The product is liblyap.so (hand-written HIP kernels behind the C-ABI in
include/chaos_lyap.h) plus the C++ host layer (TaskLyapunov, config parsing,
output) built into the same library and the chaos_lyap driver.  This Python
package holds the build script, the ctypes bindings and the workload
definitions used by the tests and the benchmark.
"""
__all__ = ["build", "lyap", "workloads"]
