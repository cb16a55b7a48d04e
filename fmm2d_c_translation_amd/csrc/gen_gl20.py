"""Writes gl20.h: the 20-point generalised Gauss-Laguerre rule with weight
t^(-1/2) e^(-t) (nodes and weights from scipy.special.roots_genlaguerre),
used for H0/H1 in 4 <= |z| < 20 through the Laplace-type integral
H_nu(z) = sqrt(2/(pi z)) e^{i(z - nu pi/2 - pi/4)} / Gamma(nu + 1/2)
          * int_0^inf e^{-t} t^{nu - 1/2} (1 + i t / (2 z))^{nu - 1/2} dt
(the Laplace-type integral representation of the Hankel function).  Usage:
    python gen_gl20.py > gl20.h"""
import numpy as np
from scipy.special import roots_genlaguerre

x, w = roots_genlaguerre(20, -0.5)
print("// gl20.h -- generated by gen_gl20.py: 20-point Gauss-Laguerre rule,")
print("// weight t^(-1/2) e^(-t).  Do not edit.")
print("#pragma once")
print("#define GL20_N 20")
print("static const double GL20_X[20] = {" + ", ".join(f"{v:.17e}" for v in x) + "};")
print("static const double GL20_W[20] = {" + ", ".join(f"{v:.17e}" for v in w) + "};")
