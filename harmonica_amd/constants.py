"""Physical constants (reference: src/harmonica/constants.py:12)."""

#: Gravitational constant G in SI units (m^3 kg^-1 s^-2).
GRAVITATIONAL_CONST = 6.6743e-11
