"""FP64 CPU oracle of the BA LM path — test infrastructure only (see ba_oracle.h)."""
