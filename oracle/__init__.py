"""CPU oracle for the VESPER_Python map-alignment hot path.

TEST INFRASTRUCTURE ONLY.  This package is a plain NumPy / C restatement of
the reference algorithm (KawaiR/VESPER_Python, ``VESPER_1_prob_new.py`` and
``main.py``).  It is imported only by ``tests/``, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``,
and only as the checker or as the timed CPU baseline -- never by the product
package ``vesper_python_amd``.

Modules:
  vesper_ref   -- NumPy restatement, two dtype paths: "ref32" (the reference's
                  own float32/float64 mix under numpy-1.21 promotion rules) and
                  "ref64" (float64 throughout, the FP64 oracle).
  fastvec_ref.c -- C restatement of ``doFastVEC``/``calc`` in the reference's
                  exact loop and summation order (built into
                  ``oracle/_build/libfastvec_ref.so``).
"""
