"""Physical constants and unit conversions (units.py of the reference; CODATA 2018, SI).

Code units G = c = M_BH = 1 with M_BH = 4.3e6 M_sun: length GM/c^2, time GM/c^3, mass M_BH.
Gas masses are given in the same mass unit (a 0.1 M_sun cloud is 0.1 / 4.3e6 in code units),
which fixes the gas density scale used by the cooling law.
"""

import math
from dataclasses import dataclass, field

G = 6.67430e-11
C = 299_792_458.0
M_SUN = 1.98847e30
K_B = 1.380649e-23
M_P = 1.67262192369e-27
MU = 0.6
SGR_A_MASS_MSUN = 4.3e6
T_FLOOR = 1.0e4
SIGMA_SB = 5.670374419e-8  # W m^-2 K^-4
A_RAD = 4.0 * SIGMA_SB / C  # radiation constant, J m^-3 K^-4
CLOUD_MASS_MSUN = 0.1

LENGTH_M = G * SGR_A_MASS_MSUN * M_SUN / C**2
TIME_S = G * SGR_A_MASS_MSUN * M_SUN / C**3
DENSITY_KG_M3 = SGR_A_MASS_MSUN * M_SUN / LENGTH_M**3

# free-free emissivity of fully ionized hydrogen with n_e = n_i = rho / m_p and Gaunt factor 1
# (Rybicki & Lightman 1979, Eq. 5.15b): 1.4e-27 erg cm^3 s^-1 K^-1/2 = 1.4e-40 W m^3 K^-1/2
LAMBDA_FF = 1.4e-40


def gas_u_from_temperature(t_kelvin):
    """Ideal-gas (gamma = 5/3) code specific energy at T (units.py:131-133)."""
    return 1.5 * K_B * t_kelvin / (MU * M_P) / C**2


def ideal_u_floor(t_floor, gamma):
    """k T / ((gamma - 1) mu m_p) / c^2 (eos.py:38-42 with hydro.py:155-163)."""
    return K_B * t_floor / ((gamma - 1.0) * MU * M_P) / C**2


def msun_to_code_mass(m_msun):
    return m_msun / SGR_A_MASS_MSUN


def powerlaw_cooling_coefficient(gamma):
    """k of du/dt = -k rho sqrt(u) in code units, for the optically thin law
    Lambda = LAMBDA_FF (rho / m_p)^2 sqrt(T) per unit volume, T = (gamma - 1) mu m_p u c^2 / k_B."""
    return (TIME_S * DENSITY_KG_M3 * LAMBDA_FF * math.sqrt((gamma - 1.0) * MU * M_P * C**2 / K_B)
            / (M_P**2 * C**2))  # fmt: skip


@dataclass(frozen=True)
class BlackHoleUnits:
    """The code units of a hole of ``mass_msun`` (units.py:52-86): G = c = M = 1."""

    mass_msun: float = SGR_A_MASS_MSUN

    @property
    def mass_kg(self) -> float:
        return self.mass_msun * M_SUN

    @property
    def length_m(self) -> float:
        return G * self.mass_kg / C**2

    @property
    def time_s(self) -> float:
        return G * self.mass_kg / C**3


@dataclass(frozen=True)
class GasUnits:
    """The gas mass scale of the gas + radiation equation of state (units.py:89-113):
    ``cloud_mass_code`` code units of gas mass are ``cloud_mass_msun`` solar masses. It sets the
    physical density that radiation pressure sees; the ideal gas does not depend on it."""

    bh: BlackHoleUnits = field(default_factory=BlackHoleUnits)
    cloud_mass_msun: float = CLOUD_MASS_MSUN
    cloud_mass_code: float = 1.0

    @property
    def mass_kg_per_code(self) -> float:
        return self.cloud_mass_msun * M_SUN / self.cloud_mass_code

    @property
    def density_kg_m3(self) -> float:
        """One code density unit in kg/m^3."""
        return self.mass_kg_per_code / self.bh.length_m**3

    def density_to_si(self, rho_code):
        return rho_code * self.density_kg_m3

    def density_from_si(self, rho_si):
        return rho_si / self.density_kg_m3


def specific_energy_to_si(u_code):
    return u_code * C**2


def specific_energy_from_si(u_si):
    return u_si / C**2
