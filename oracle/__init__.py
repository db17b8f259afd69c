"""CPU oracle of the ParHDE layout path (test infrastructure only)."""
