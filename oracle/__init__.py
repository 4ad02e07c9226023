"""CPU oracle of the SIFT putative-matching rule (test infrastructure only)."""
