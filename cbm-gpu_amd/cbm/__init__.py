"""CBM (compressed binary matrix) product for AMD Instinct MI355X.

Drop-in for the ``cbm`` package of cbm4scale/CBM-GPU: the modules
``cbm_mkl_cpp``, ``cbm_cusparse``, ``csr_cusparse`` and the classes
``cbm4mm``, ``cbm4ad``, ``cbm4dad``, ``csr4mm``, ``csr4ad``, ``csr4dad`` keep
their names and call signatures.  Import them as in the reference, e.g.
``from cbm.cbm4mm import cbm4mm``.
"""
