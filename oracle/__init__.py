"""CPU restatement of PIMDAL's DB-operator hot path — test infrastructure.

Only the tests, the smoke check and the benchmark's CPU-baseline leg use
this package; the product library never imports it.
"""
