"""snarkrs_amd: an MI355X (gfx950) Groth16 prover back end for snarkrs.

The product is the C-ABI library ``libg16hip.so`` (include/g16hip.h) built from
``snarkrs_amd/csrc``; this package is its Python host side: the ctypes binding
(:mod:`snarkrs_amd.native`), a mirror of the reference's ``Backend``/``PreparedCircuit``
contract (:mod:`snarkrs_amd.backend`), the wire format (:mod:`snarkrs_amd.wire`) and the
synthetic workload generator (:mod:`snarkrs_amd.synth`).
"""
__version__ = "0.1.0"
