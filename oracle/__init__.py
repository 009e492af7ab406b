"""Checker for the HIP path (test infrastructure only; see oracle/bind.py)."""
