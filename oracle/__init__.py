"""CPU oracle for the FAC Poisson path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package.  The product (samrai_amd) never does.
"""
