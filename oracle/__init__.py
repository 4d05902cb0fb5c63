"""CPU oracle of the transport path (test infrastructure only)."""
