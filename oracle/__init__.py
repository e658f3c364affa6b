"""FP64 CPU oracle package (test infrastructure only)."""
