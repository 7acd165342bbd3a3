"""Physical constants, as rapt/__init__.py:4-10 defines them.

c and Re stay Python ints on purpose: expressions such as ``c**2`` and
``Re**3`` are then evaluated exactly as integers and rounded once, as in the
reference.
"""
e = 1.602176565e-19      # elementary charge (C)
m_pr = 1.672621777e-27   # proton mass (kg)
m_el = 9.10938291e-31    # electron mass (kg)
c = 299792458            # speed of light (m/s)
B0 = 3.07e-5             # Earth field strength at the magnetic equator (T)
re = Re = 6378137        # Earth radius (m)
