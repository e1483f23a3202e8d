"""The oracle: CPU restatement of the reference spot-finding path (test infrastructure only)."""
