"""CPU oracle for the DPDEM ice-ship frame step.

TEST INFRASTRUCTURE ONLY.  This package is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The shipped library (``mbd-icenav_amd/``) never imports, links or
executes anything under ``oracle/``.

It is a NumPy restatement of the reference path, written from the reference's
text (``medea_ice_ship/jax_simulation.py``, ``jax_geometry.py``,
``jax_bond_fracture.py``, ``jax_vessel_dynamics.py``, ``jax_ice_state.py``,
``jax_bonds.py``); every function cites the lines it follows.  It runs in two
precisions:

* ``float64`` -- the reference's default (``jax_enable_x64``); numpy's own
  ``sin/cos/power`` are used.
* ``float32`` -- the reference's ``--float32`` mode.  Python scalars meet
  arrays the way JAX weak types do (rounded to float32 at the point of use,
  folded in binary64 first when the whole sub-expression is Python scalars),
  transcendentals come from :mod:`oracle.detmath`, reductions follow the
  pairwise tree of :func:`oracle.step.tree_sum`, and scatter-adds run in
  index order (``np.add.at``).  The HIP kernels implement the same operation
  order, so the two agree bit for bit.

Parity status of each function (what pins it) is stated in its docstring.
"""
