"""Oracle (test infrastructure only): CPU restatement of the reference march.

Importable by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg only; the product package never imports it."""
