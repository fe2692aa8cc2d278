"""FP64 CPU oracle of the BLD solver (test infrastructure only)."""
