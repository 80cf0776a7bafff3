"""TEST INFRASTRUCTURE: the CPU oracle of the FLIP substep (see flip_oracle.cpp)."""
