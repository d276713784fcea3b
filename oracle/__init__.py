"""Test infrastructure only: the FP64 CPU oracle (see dmd_oracle.h)."""
