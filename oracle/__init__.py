"""CPU oracle (test infrastructure only): numpy/scipy restatements of the reference's
HSS2D and HSS1D dense-block paths, used by tests/, __graft_entry__.smoke() and the CPU
baseline leg of bench.py.  Nothing in the product library imports this package."""
