"""CGS constants of the Lya M1 step (src/ConstantsAndFunctions.jl:45-170,
src/M1Closure.jl:36-41), evaluated in the reference's operation order."""
import math

Lsun = 3.839e33
pc = 3.085677581467192e18
kms = 1e5
c = 2.99792458e10
h = 6.626069573e-27
kB = 1.380648813e-16
mu = 1.660539069e-24
mH = 1.0080 * mu
me = 9.109382917e-28
e = 4.80320451e-10
nu_Lya = 2.466068e15
nu_LyC = 3.289842e15
A_Lya = 6.265e8
f_lu_Lya = 0.41641

OneThird = 1 / 3
Four_pi = 4 * math.pi
Four_pi_over_c = Four_pi / c
inv_c = 1 / c
Sqrt_pi = math.sqrt(math.pi)
Sqrt_pi_over_2 = Sqrt_pi / 2
inv_Sqrt_pi_over_2 = 1 / Sqrt_pi_over_2

Jeps = 1e-60
Heps = 1e-100
fmin = 1e-10
fmax = 1.0 - 1e-10

b_PREFACT = math.sqrt(2 * kB / mH)
nu_Lya_over_c = nu_Lya / c
a_PREFACT = A_Lya / (Four_pi * nu_Lya * 1 / c)
sigma0_PREFACT = Sqrt_pi * (e * e) * f_lu_Lya / (me * nu_Lya)

tau_alpha_floor = 1e-5   # src/SetupLymanAlphaArrays.jl:12
