"""FP64 CPU oracle for the Fourier-slice gather/scatter path.

TEST INFRASTRUCTURE ONLY. Nothing in the ``teamtomo_amd`` package imports this
package: it is the checker that the HIP kernels are compared against, used by
``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py``.  It is a restatement of the reference algorithm in plain NumPy
(``oracle.fourier_slice``) and plain C (``oracle/c/tt_oracle.c``), written from
the behaviour of the reference sources cited in each function docstring.
"""
