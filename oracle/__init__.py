"""FP64 CPU oracle (test infrastructure): the checker for the HIP path, never a fallback."""
