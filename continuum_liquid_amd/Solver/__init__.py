"""Host-side mirror of the reference's Solver package (Solver/General, Solver/Kernel)."""
