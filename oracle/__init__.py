"""CPU oracle for the Nenya prover -- TEST INFRASTRUCTURE ONLY.

A bit-exact C restatement (oracle/*.c) of the reference prover and verifier
(tnyuzg/nenya, crates/nenya).  It is the parity checker for the HIP product
path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may load it.  The product package (nenya_amd) never imports it.
"""
from .oracle import Oracle, load, build  # noqa: F401
