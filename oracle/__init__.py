"""CPU oracle for the portUrb dycore step and LES closure -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use this package, as the
checker and baseline; the product (porturb_amd) never imports it.
"""
