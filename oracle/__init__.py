"""CPU oracle of the cloud compute stage (test infrastructure only)."""
