"""Physical constants used by the time-dependent KPM drivers, in the
reference's units (jclsquant/constants.py:1-6): eV, s, fs, K."""

hbar = 6.582e-16           # eV s
kb = 8.617333 * 1e-5       # eV / K
hbar_fs = hbar * 1e15      # eV fs
