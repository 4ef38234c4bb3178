"""CPU oracle of the KMC path (test infrastructure only)."""
