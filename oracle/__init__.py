"""CPU oracle of the OSEM hot path (test infrastructure only)."""
