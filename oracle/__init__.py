"""CPU oracle for the PyQuake3D hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
package; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may use it, and only as the checker.

Modules
-------
td_oracle      long-double restatement of the triangular-dislocation (TD)
               displacement field and its gradient (strain/stress), plus the
               dense traction Green's-function assembly built on it.
qdsim_oracle   float64 NumPy restatement of the dense rate-and-state
               stepping (RHS, Dormand-Prince attempt, accept/reject, update).
refbind        ctypes binding to the reference's own TDstressFS_C.cpp compiled
               from source into ``oracle/_ref`` (present only where it was
               built).
"""
