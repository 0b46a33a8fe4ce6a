"""CPU oracle of the BraWl hot path -- test infrastructure only (see oracle.py)."""
