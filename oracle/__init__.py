"""CPU restatement of the reference's wave stepping: the checker for the HIP library.

TEST INFRASTRUCTURE ONLY.  ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg
of ``bench.py`` may import it; the product package ``neuralmech_amd`` never does.
"""
