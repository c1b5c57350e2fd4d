"""CPU oracle for the real8 inner loop (test infrastructure only)."""
