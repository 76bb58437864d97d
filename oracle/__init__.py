"""CPU oracle (test infrastructure only): see oracle/cvt_oracle.c."""
