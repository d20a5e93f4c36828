"""CPU oracle for the HIP time loop: test infrastructure, never part of the product.

``scheme`` restates the reference NumPy scheme; ``ref_c`` drives the reference's own C loop
compiled from its sources into ``oracle/_ref`` by ``oracle/Makefile``.
"""
