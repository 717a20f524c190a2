"""snarkrs_amd: an MI355X (gfx950) Groth16 prover back end for snarkrs.

The product is the C-ABI library ``libg16hip.so`` (include/g16hip.h) built from
``snarkrs_amd/csrc``; this package is its Python host side: the ctypes binding that mirrors
the reference's ``Backend``/``PreparedCircuit``/``MsmBackend`` contract
(:mod:`snarkrs_amd.native`: ``Device``, ``Circuit``, ``CeremonyMsm``, ``prove_batch``,
``Server``), the proving-key container (:mod:`snarkrs_amd.key`), the snarkjs ``.zkey`` /
``.wtns`` codec (:mod:`snarkrs_amd.zkey`), the verifier and snarkjs JSON files
(:mod:`snarkrs_amd.verify`), device buffers (:mod:`snarkrs_amd.devmem`), the wire format
(:mod:`snarkrs_amd.wire`) and the synthetic workload generator (:mod:`snarkrs_amd.synth`).
"""
__version__ = "0.2.0"
