"""CPU FP64 oracle of the Lya M1 step (test infrastructure only; see oracle.py)."""
