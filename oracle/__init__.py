"""CPU FP64 oracle of the substep (test infrastructure only)."""
