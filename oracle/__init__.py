"""CPU checkers for the H-matrix apply -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg.  The product package (qhmat_amd) never imports, links or runs anything
here.
"""
