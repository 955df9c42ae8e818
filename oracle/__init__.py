"""Oracle package (test infrastructure only): the plain-C FP64 restatement of
the reference hydraulic step (swe), the reference-solver harness (deck) and
the mesh-contract restatement (mesh_oracle).  Imported only by tests/, the
smoke check and the CPU-baseline leg of bench.py."""
