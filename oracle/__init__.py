"""CPU checker for the GP path.  Test infrastructure only: imported by tests/, the smoke check and
the CPU-baseline leg of bench.py, never by the product package."""
