"""Physical constants of Particulator.jl (src/constants.jl:6-166).

The values are the reference's CODATA-2014 literals; the derived constants follow
src/constants.jl:157-164 with the same evaluation order.
"""
import math

pi = 3.141592653589793
c = 299792458.0
electron_mass = m_e = 9.10938356e-31
elementary_charge = e = eV = 1.6021766208e-19
fine_structure = alpha = 0.0072973525664
hbar = 1.0545718001391127e-34
epsilon_0 = 8.854187817620389e-12
N_A = Avogadro = 6.022140857e+23
Boltzmann = k = 1.38064852e-23
kilo = 1000.0
centi = 0.01
mega = 1000000.0
atm = 101325.0
nair = 2.6867811e+25
zero_Celsius = 273.15

c2 = c * c
electron_mc2 = electron_mass * c2
electron_mc = electron_mass * c
r_e = elementary_charge * elementary_charge / (electron_mass * c2) / (4.0 * pi * epsilon_0)
a_0 = hbar / (electron_mass * c * fine_structure)
two_pi = 2.0 * pi

assert math.isclose(r_e, 2.8179403e-15, rel_tol=1e-7)
