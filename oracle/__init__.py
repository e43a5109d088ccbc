"""Parity oracle and CPU baseline (test infrastructure only)."""
