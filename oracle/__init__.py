"""CPU oracle for the NICAM hot path -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, by the package smoke check and by the CPU-baseline
leg of bench.py, always as the checker.  The product library never uses it.
"""
