"""CPU oracle of the stencil path (test infrastructure only).

Only the test suite, the smoke check and the benchmark's CPU-baseline leg
may import this package; the product path never does.
"""
