"""FP64 CPU oracle of the LPMD force and integration path (test infrastructure only)."""
