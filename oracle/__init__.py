"""FP64 CPU oracle of the walker (test infrastructure only)."""
