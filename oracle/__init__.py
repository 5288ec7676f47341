"""CPU oracle for the sparse-register-tiling SpMM path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package imports this
directory.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only to compare the
GPU results against (or, for ``cpu_baseline``, as the timed CPU
restatement); the product path never routes through it.

Contents
--------
``sop_oracle``  pure Python / numpy restatement of the reference's packing
                rules, mapping files and pattern-set generator.
``cbind``       ctypes loader for ``srt_oracle.c`` (C restatement of the
                packer, the nano-kernel executor and the CSR executor).
``inputs``      seeded generators for the benchmark configurations.
"""
