"""CPU oracle (test infrastructure only; never imported by bddcml_amd)."""
