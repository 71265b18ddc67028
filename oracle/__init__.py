"""FP64 oracle of the aclfft execution path (test infrastructure only; see oracle.py)."""
