"""CPU oracle for the HIP dynamical-core backend (test infrastructure only)."""
