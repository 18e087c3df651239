"""CPU oracle of the reference IP hot path (test infrastructure only; see oracle/oracle.py)."""
