"""CPU oracle for the IVF-flat search path (test infrastructure only; see harmony_oracle.cpp)."""
