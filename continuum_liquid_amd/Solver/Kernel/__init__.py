"""Host step loop and stage wrappers over the HIP C-ABI (reference Solver/Kernel)."""
