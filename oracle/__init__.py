"""CPU restatement of the reference step (test infrastructure only)."""
