"""TEST INFRASTRUCTURE ONLY: the CPU oracle (see oracle/reference.py)."""
