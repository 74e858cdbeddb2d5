"""CGS constants with the values holodeck takes from astropy (holodeck/constants.py:19-47).

astropy's CODATA-2018 / IAU-2015 values, written out so that no astropy import
is needed.
"""
NWTG = 6.6743e-08                 #: Newton's G [cm^3/g/s^2]
SPLC = 29979245800.0              #: speed of light [cm/s]
MSOL = 1.988409870698051e+33      #: solar mass [g]
PC = 3.0856775814913673e+18       #: parsec [cm]
YR = 31557600.0                   #: Julian year [s]
AU = 14959787070000.0             #: astronomical unit [cm]

SCHW = 2 * NWTG / (SPLC * SPLC)   #: Schwarzschild constant 2G/c^2 [cm/g]
DAY = 86400.0
MYR = 1.0e6 * YR
GYR = 1.0e9 * YR
KPC = 1.0e3 * PC
MPC = 1.0e6 * PC
GPC = 1.0e9 * PC
