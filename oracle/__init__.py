"""CPU oracle for the photon random walk — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import anything here, and only as the checker (or as the timed CPU
baseline).  The product path (``monte_carlompi_amd``) never imports this
package and fails loudly when its HIP library is missing.

Contents
--------
walk.py          Python/numpy restatement of the reference's ``run()`` photon
                 section and ``monte_carlo3D`` walk (sphere / Henyey-Greenstein
                 path), keeping the reference's numpy operation order and its
                 per-step draw structure.  Pinned by the golden records the
                 reference itself wrote (tests/golden) and by closed forms.
mcoracle.c       Plain-C float64 restatement of the same walk with a numpy-
                 compatible MT19937 stream (for config-scale replay streams and
                 native-mode statistics).  Pinned against walk.py record by
                 record and against numpy's RandomState draw by draw.
cpu_baseline.py  Runs walk.py as P processes with the reference's
                 ``np.array_split`` scatter / ordered gather (parallelize.py:14-38).

Parity status per function (what pins it):
  hg2                 closed form + chi^2 vs HG pdf (mc3D:779-788) + golden records
  initial_triples     golden records (reference draw order, mc3D:1027-1044)
  walk_photon         golden records written by the reference (bit-exact) +
                      van de Hulst R/T, direct transmittance, omega=0/1 limits
  draw_wavelengths    golden records (mc3D:1516-1520)
  C walk (mcoracle.c) == walk.py on identical streams; MT19937 == numpy
"""
