"""Host side of the MI355X shallow-water step: the ctypes mirror of the C-ABI in include/swe_amd.h,
the reference mesh format and set-up, and the benchmark problems."""
