"""oracle -- CPU restatement of QS3's symmetry-adapted Lanczos hot path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import it.  The
product (ed_amd/) never imports, links or calls anything here.

Contents:
  qs3_oracle.c   literal C restatement of state_lists / ham2vec / observables
                 (built by build_oracle.sh into liboracle_qs3.so)
  ref_harness.f90 + build_ref.sh
                 the reference's own Fortran modules compiled from
                 the reference sources into _ref/libqs3ref.so (when present)
  qs3.py         ctypes bindings for both, plus numpy restatements of the
                 solver loops (plain Lanczos, thick restart, DSF recurrence,
                 DSFLan continued fraction) for small sectors.

Parity status of each function is recorded in the module docstring of qs3.py.
"""
