"""Oracle: CPU restatements used only as checkers (tests, smoke, bench cpu_baseline)."""
