"""CPU oracle of the explicit MPM step (test infrastructure only)."""
