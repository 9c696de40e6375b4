"""Physical constants with the values PyQuante uses (PyQuante/Constants.py:20-21)."""
bohr2ang = 0.529177249
ang2bohr = 1 / bohr2ang
