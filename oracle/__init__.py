"""FP64 CPU oracle (test infrastructure): the reference the HIP path is compared against,
never a fallback."""
