"""ORACLE package -- CPU restatement of the reference's Newton path, used ONLY as
the checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
Nothing in kadath_frontera_amd/ imports it."""
