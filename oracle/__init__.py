"""The oracle: a CPU (numpy FP64) restatement of the reference voltax
algorithms on the Newton hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import, call or link anything under
``oracle/``, and only as the checker (or the timed CPU baseline), never as the
thing measured or shipped.  The product (``voltax_amd``) never imports it.

Modules and the reference code each restates:

* `oracle.circuit`   voltax/circuit.py, voltax/builder.py, voltax/elements/
                     {passive,sources,mosfet}.py, voltax/signals.py
* `oracle.sparse`    voltax/sparse.py (pattern, stamping, plan, factor, solve)
* `oracle.analysis`  voltax/analysis.py (Newton, gmin stepping, dc, transient,
                     ac, implicit-function adjoint)
* `oracle.measure`   voltax/measure.py (crossing, delay, period)
* `oracle.dual`      forward-mode derivatives standing in for jax.jvp

How each part is pinned (tests/test_oracle.py): the reference's golden
RC-ladder state and gradient (tests/test_analysis.py:33-49), EKV closed forms
(tests/test_elements.py:190-219), stamped Jacobians against central finite
differences of the assembled residual, factor/solve against LAPACK (real,
transposed, complex), plan invariants (tests/test_sparse.py:233-242),
gradients against central finite differences, and the published 43.5 ps
delay of the 3-stage inverter chain (README.md:78).
"""
