"""Pinning the oracle's gas + radiation equation of state and, through the reference's Phantom
fixture, its force pass with conductivity and velocity reconstruction (CPU only).

- EOS: the expected-values §9 points (u, P, beta_gas, Gamma_1 to 1e-8; T recovered from (rho, u)
  to 1e-10), the gas and radiation limits, the T -> u -> T round trip on a 200 x 200 grid to
  1e-10, and the code-unit wrapper (tests/physics/test_eos_gasrad.py of the reference).
- Phantom fixture (tests/golden/phantom_force_sample.npz, the reference's
  tests/data/phantom_force_sample.npz): one Phantom derivative evaluation of a 50k-particle run
  with all of the reference's options on. The oracle's density, tightening, velocity gradients
  and force pass, with Phantom's pressure and the gas + radiation sound speed scaled to it, give
  Phantom's hydro acceleration and du/dt to a median relative difference of 1e-5 and a 99th
  percentile of 1e-3, with reconstruction on and off (tests/sim/test_phantom_patch.py:37-107).
  This is an external pin of the pair loop: Phantom is a different code.
"""

from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

from helpers import check

from oracle import coupled as oc
from oracle import neighbours as on
from oracle import physics as ph
from oracle import sph as osph

FIXTURE = Path(__file__).resolve().parent / "golden" / "phantom_force_sample.npz"
MEDIAN_TOL, P99_TOL = 1e-5, 1e-3

# expected-values §9: (rho kg/m^3, T K, u J/kg, P Pa, beta_gas, Gamma_1)
POINTS = [
    (1e-3, 1e4, 2.0636755965e8, 1.3757585119e5, 9.9998166894e-1, 1.6666361185),
    (1e-6, 1e6, 7.5659396100e14, 2.5220486567e8, 5.4548245497e-5, 1.3333424248),
    (1e-9, 1e5, 7.5659396100e13, 2.5220486567e4, 5.4548245497e-5, 1.3333424248),
    (1e-9, 1e7, 7.5657332502e21, 2.5219110835e12, 5.4551221167e-11, 1.3333333333),
]
TORUS = (1.6586853e-6, 9.9188303029e5, 4.415186e14, 2.4412478743e8, 9.2714321654e-5, 1.3333487859)
EXTREMES = [(1e3, 1e3, 2.0635999392e7, 1.3757332928e10, 1.6666666667),
            (1e-15, 1e8, 7.5657332500e31, 2.5219110833e16, 1.3333333333)]  # fmt: skip


def test_eos_points():
    for rho, t, u, p, beta, g1 in POINTS:
        tag = f"rho={rho:g}, T={t:g}"
        check(f"{tag}: u (J/kg)", u, ph.gasrad_u_si(rho, t), rtol=1e-8)
        check(f"{tag}: P (Pa)", p, ph.gasrad_pressure_si(rho, t), rtol=1e-8)
        check(f"{tag}: beta_gas", beta, ph.gasrad_beta_si(rho, t), rtol=1e-8)
        check(f"{tag}: Gamma_1", g1, ph.gamma1(ph.gasrad_beta_si(rho, t)), rtol=1e-8)
        check(f"{tag}: T from (rho, u)", t, ph.gasrad_temperature_si(rho, np.array([u])), rtol=1e-10)
    rho, t, u, p, beta, g1 = TORUS
    t_rec = ph.gasrad_temperature_si(rho, np.array([u]))
    check("torus point: T recovered from u", t, t_rec, rtol=1e-10)
    check("torus point: P", p, ph.gasrad_pressure_si(rho, t_rec), rtol=1e-8)
    check("torus point: Gamma_1", g1, ph.gamma1(ph.gasrad_beta_si(rho, t_rec)), rtol=1e-8)


def test_eos_limits():
    for rho, t, u, p, g1 in EXTREMES:
        tag = f"rho={rho:g}, T={t:g}"
        check(f"{tag}: u", u, ph.gasrad_u_si(rho, t), rtol=1e-8)
        check(f"{tag}: P", p, ph.gasrad_pressure_si(rho, t), rtol=1e-8)
        check(f"{tag}: Gamma_1", g1, ph.gamma1(ph.gasrad_beta_si(rho, t)), atol=1e-6)
    check("Gamma_1(beta = 1) = 5/3, (beta = 0) = 4/3", [5 / 3, 4 / 3], ph.gamma1(np.array([1.0, 0.0])), rtol=1e-15)


def test_eos_round_trip_and_bracket():
    rho, t = np.meshgrid(np.logspace(-15, 3, 200), np.logspace(3, 10, 200))
    t_back = ph.gasrad_temperature_si(rho, ph.gasrad_u_si(rho, t))
    check("round trip T -> u -> T, max relative error", 0.0, float(np.max(np.abs(t_back / t - 1))), atol=1e-10)
    with pytest.raises(ValueError):
        ph.gasrad_temperature_si(1e-6, np.array([ph.gasrad_u_si(1e-6, 0.5 * ph.T_MIN_SI)]))
    with pytest.raises(ValueError):
        ph.gasrad_temperature_si(1e-6, np.array([ph.gasrad_u_si(1e-6, 2.0 * ph.T_MAX_SI)]))


def test_code_unit_wrapper():
    dens = ph.gas_density_kg_m3()
    rho_si, t = 1e-6, 1e6
    rho_code = rho_si / dens
    u_code = ph.gasrad_u_si(rho_si, t) / ph.C**2
    p, cs, t_out, beta = ph.gasrad_state(np.array([rho_code]), np.array([u_code]), dens)
    check("wrapper: T", t, t_out, rtol=1e-10)
    check("wrapper: P in code units back to Pa", ph.gasrad_pressure_si(rho_si, t), p * dens * ph.C**2, rtol=1e-10)
    check("wrapper: c_s^2 = Gamma_1 P / rho", ph.gamma1(beta) * p / rho_code, cs**2, rtol=1e-12)
    # the code density unit: 0.1 M_sun per code mass over (G 4.3e6 M_sun / c^2)^3
    length = ph.G * 4.3e6 * ph.M_SUN / ph.C**2
    check("code density unit", 0.1 * ph.M_SUN / length**3, dens, rtol=1e-15)


# ----------------------------------------------------------------- the Phantom fixture
def phantom_sample():
    """The fixture with the reference's set-up (test_phantom_patch.py:37-50): masses in the
    reference's convention (one code mass = 0.1 M_sun), the density solve at Phantom's h, the
    tightened lists."""
    from nebula.sim.coupled import CoupledConfig

    with np.load(FIXTURE, allow_pickle=False) as f:
        d = {k: np.array(f[k]) for k in f.files}
    hc = CoupledConfig().hydro
    d["hc"] = hc
    m = float(d["mass_phantom"]) * hc.gas_units.bh.mass_kg / hc.gas_units.mass_kg_per_code
    d["mass"] = np.full(len(d["x"]), m)
    d["mratio"] = float(d["mass_phantom"]) / m
    d["inner"] = np.linalg.norm(d["x"] - d["center"], axis=1) < float(d["r_in"])
    return d


def phantom_rel(ours, theirs):
    """test_phantom_patch.py:71-77: per-particle difference over its own magnitude."""
    if ours.ndim == 2:
        mag, diff = np.linalg.norm(ours, axis=1), np.linalg.norm(ours - theirs, axis=1)
    else:
        mag, diff = np.abs(ours), np.abs(ours - theirs)
    return diff / np.maximum(mag, 1e-6 * np.median(mag))


@pytest.fixture(scope="module")
def oracle_sample():
    d = phantom_sample()
    hc = d["hc"]
    dens = osph.compute_density(d["x"], d["mass"], d["h_phantom"], kernel_name=hc.kernel, eta=hc.eta, h_max=hc.h_max)
    d["dens"] = dens
    d["nb"] = on.tighten(dens.neighbours, d["x"], dens.h)
    return d


def _oracle_forces(d, recon):
    hc = replace(d["hc"], eos="gasrad", reconstruction=recon)
    dens = d["dens"]
    p_ours, cs_ours = oc.eos_state(hc, dens.rho, d["u"])
    p = d["pressure"] / d["mratio"]
    cs = cs_ours * np.sqrt(p / p_ours)
    gv = osph.velocity_gradients_numba(d["x"], d["v"], d["mass"], dens.h, d["nb"], hc.kernel) if recon else None
    f = osph.forces_numba(d["x"], d["v"], d["mass"], dens.h, dens.rho, dens.omega, p, d["nb"], hc.kernel, u=d["u"],
                          c_s=cs, alpha=d["alpha"], beta=hc.beta, alpha_u=hc.alpha_u,
                          cond_signal=hc.conductivity_signal, reconstruction=recon, closeness=hc.recon_closeness,
                          eta_crit=1.0 / hc.eta, recon_gate=hc.recon_gate, grad_v=gv)  # fmt: skip
    return f.acc, f.dudt


def test_phantom_sample_contents(oracle_sample):
    d = oracle_sample
    inner = d["inner"]
    capped = d["dens"].h[inner] >= d["hc"].h_max * (1 - 1e-12)
    check("Phantom sample: particles compared (r < 600)", 2203, int(inner.sum()))
    check("... of which at the h cap (h = 200)", 93, int(capped.sum()))
    rel_h = np.abs(d["dens"].h - d["h_phantom"]) / d["dens"].h
    check("h: oracle vs Phantom, median relative difference < 1e-6", True, bool(np.median(rel_h[inner]) < 1e-6))


@pytest.mark.parametrize("recon", [True, False], ids=["recon_on", "recon_off"])
def test_oracle_forces_match_phantom(oracle_sample, recon):
    d = oracle_sample
    inner = d["inner"]
    acc, dudt = _oracle_forces(d, recon)
    tag = "on" if recon else "off"
    ra = phantom_rel(acc, d[f"acc_{tag}"])[inner]
    ru = phantom_rel(dudt, d[f"dudt_{tag}"])[inner]
    print(f"recon {tag}: |a| median {np.median(ra):.3g} p99 {np.percentile(ra, 99):.3g}; "
          f"du/dt median {np.median(ru):.3g} p99 {np.percentile(ru, 99):.3g}")  # fmt: skip
    check(f"recon {tag}: |a| median rel. diff", 0.0, float(np.median(ra)), atol=MEDIAN_TOL)
    check(f"recon {tag}: |a| 99th pct rel. diff", 0.0, float(np.percentile(ra, 99)), atol=P99_TOL)
    check(f"recon {tag}: du/dt median rel. diff", 0.0, float(np.median(ru)), atol=MEDIAN_TOL)
    check(f"recon {tag}: du/dt 99th pct rel. diff", 0.0, float(np.percentile(ru, 99)), atol=P99_TOL)


def test_phantom_reconstruction_changes_the_result(oracle_sample):
    """test_phantom_patch.py:98-107: the comparison is sensitive (reconstruction on vs off
    differs by far more than the tolerance in this shocked region)."""
    d = oracle_sample
    r = phantom_rel(d["acc_on"], d["acc_off"])[d["inner"]]
    check("Phantom recon on vs off: |a| 90th pct rel. diff > 100 x P99_TOL", True,
          bool(np.percentile(r, 90) > 100 * P99_TOL))  # fmt: skip
