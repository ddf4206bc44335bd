"""CPU oracle for the GPP/FF hot path (test infrastructure only)."""
