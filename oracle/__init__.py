"""CPU oracle of the transport path (test infrastructure).

Only the test suite, the smoke check and the CPU-baseline leg of the
benchmark import this package; the product library never does.
"""
