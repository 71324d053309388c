"""CPU oracle of the reference fill (test infrastructure only)."""
