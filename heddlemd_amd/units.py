"""SI <-> Hartree atomic units (CODATA 2018), the values of the reference's
src/units/atomic_constants.rs.  The engine works in atomic units; these
factors are used only at the input/output boundary."""
import math

LENGTH_BOHR_TO_M = 5.29177210903e-11
MASS_ME_TO_KG = 9.1093837015e-31
CHARGE_E_TO_C = 1.602176634e-19
ENERGY_HARTREE_TO_J = 4.359744722207101e-18
TIME_AU_TO_S = 2.4188843265857195e-17
FORCE_AU_TO_N = 8.238723498254079e-08
TEMPERATURE_AU_TO_K = 315775.0248040668
PRESSURE_AU_TO_PA = 29421015696522.1      # Hartree / Bohr^3 in Pa
AVOGADRO = 6.02214076e23
DALTON_KG = 1.66053906660e-27

NM = 1e-9 / LENGTH_BOHR_TO_M            # Bohr per nm
ANGSTROM = 1e-10 / LENGTH_BOHR_TO_M     # Bohr per Angstrom
FS = 1e-15 / TIME_AU_TO_S               # atomic time units per fs
AMU = DALTON_KG / MASS_ME_TO_KG         # electron masses per dalton
KJ_PER_MOL = 1000.0 / AVOGADRO / ENERGY_HARTREE_TO_J   # Hartree per kJ/mol


def kelvin(t):
    """k_B T in Hartree."""
    return t / TEMPERATURE_AU_TO_K


def ns_per_day(steps_per_s, dt_fs):
    return steps_per_s * dt_fs * 86400.0 / 1e6


def erfc_alpha(r_cut, tolerance):
    """alpha with erfc(alpha * r_cut) = tolerance (Ewald tolerance -> alpha)."""
    from scipy.special import erfcinv
    return float(erfcinv(tolerance)) / r_cut


def deg(x):
    return x * math.pi / 180.0
