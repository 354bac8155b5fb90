"""The gas + radiation equation of state on the GPU, and the force pass on the reference's
Phantom fixture (GPU).

- nb_eos_gasrad (Newton on log T, the device's exp / log) against the oracle's restatement of
  fastphys.py:45-111 over the densities and energies the runs see: P, c_s and T to 1e-11
  relative (both stop at a log-T step below 1e-12, so they can differ by a few 1e-12); u outside
  the solver bracket raises ValueError as in the reference.
- The Phantom fixture through the product's own density, tightening, velocity-gradient and force
  passes (FP64): Phantom's acceleration and du/dt to the reference's tolerances (median 1e-5,
  99th percentile 1e-3), with reconstruction on and off, and the oracle's to 1e-10; the mixed
  precision force pass on the same lists to the same Phantom tolerances.
- The coupled step with gas + radiation and beta cooling (the floor at each density) against
  the oracle, and the reference's hot-blob ledger with that physics (test_coupled.py:80-93).
"""

from dataclasses import replace

import numpy as np
import pytest

from conftest import gpu_available
from helpers import T_END_BLOB, check, hot_blob_config, hot_blob_ic

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def test_eos_kernel_matches_oracle(rng):
    from nebula.sim.hydro import HydroConfig, eos_state

    from oracle import physics as ph

    hc = HydroConfig(eos="gasrad")
    dens = hc.gas_units.density_kg_m3
    rho = 10.0 ** rng.uniform(-9.0, -3.0, 3000)
    u = 10.0 ** rng.uniform(-9.0, -3.0, 3000)
    print("gas + radiation EOS: 3000 random (rho, u) in code units", flush=True)
    p, cs = eos_state(hc, rho, u)
    po, cso, _, _ = ph.gasrad_state(rho, u, dens)
    check("P: device vs oracle (1e-11 relative)", po, p, rtol=1e-11)
    check("c_s: device vs oracle (1e-11 relative)", cso, cs, rtol=1e-11)
    bad = np.array([ph.gasrad_u_si(1e-6, 0.5) / ph.C**2])
    with pytest.raises(ValueError):
        eos_state(hc, np.array([1e-6 / dens]), bad)


@pytest.fixture(scope="module")
def sample():
    from nebula.sim.neighbours import tighten
    from nebula.sim.sph import compute_density
    from test_oracle_gasrad import phantom_sample

    d = phantom_sample()
    hc = d["hc"]
    dens = compute_density(d["x"], d["mass"], d["h_phantom"], kernel_name=hc.kernel, eta=hc.eta, h_max=hc.h_max)
    d["dens"] = dens
    d["nb"] = tighten(dens.neighbours, d["x"], dens.h)
    return d


def _forces(d, recon, precision="fp64"):
    from nebula.sim.hydro import eos_state
    from nebula.sim.sph import forces_numba, velocity_gradients_numba

    hc = replace(d["hc"], eos="gasrad", reconstruction=recon)
    dens = d["dens"]
    p_ours, cs_ours = eos_state(hc, dens.rho, d["u"])
    p = d["pressure"] / d["mratio"]
    cs = cs_ours * np.sqrt(p / p_ours)
    gv = velocity_gradients_numba(d["x"], d["v"], d["mass"], dens.h, d["nb"], hc.kernel) if recon else None
    f = forces_numba(d["x"], d["v"], d["mass"], dens.h, dens.rho, dens.omega, p, d["nb"], hc.kernel, u=d["u"], c_s=cs,
                     alpha=d["alpha"], shock=hc.shock(), grad_v=gv, precision=precision)  # fmt: skip
    return f.acc, f.dudt


def test_phantom_sample_density(sample):
    d = sample
    inner = d["inner"]
    check("Phantom sample: particles compared (r < 600)", 2203, int(inner.sum()))
    check("... of which at the h cap (h = 200)", 93, int((d["dens"].h[inner] >= d["hc"].h_max * (1 - 1e-12)).sum()))
    rel_h = np.abs(d["dens"].h - d["h_phantom"]) / d["dens"].h
    check("h: GPU vs Phantom, median relative difference < 1e-6", True, bool(np.median(rel_h[inner]) < 1e-6))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("recon", [True, False], ids=["recon_on", "recon_off"])
def test_forces_match_phantom(sample, recon, precision):
    from test_oracle_gasrad import MEDIAN_TOL, P99_TOL, _oracle_forces, phantom_rel

    from oracle import neighbours as on
    from oracle import sph as osph

    d = sample
    inner = d["inner"]
    tag = "on" if recon else "off"
    print(f"Phantom fixture: recon {tag}, {precision}", flush=True)
    acc, dudt = _forces(d, recon, precision)
    ra = phantom_rel(acc, d[f"acc_{tag}"])[inner]
    ru = phantom_rel(dudt, d[f"dudt_{tag}"])[inner]
    print(f"recon {tag} {precision}: |a| median {np.median(ra):.3g} p99 {np.percentile(ra, 99):.3g}; "
          f"du/dt median {np.median(ru):.3g} p99 {np.percentile(ru, 99):.3g}")  # fmt: skip
    check(f"{precision} recon {tag}: |a| median rel. diff", 0.0, float(np.median(ra)), atol=MEDIAN_TOL)
    check(f"{precision} recon {tag}: |a| 99th pct rel. diff", 0.0, float(np.percentile(ra, 99)), atol=P99_TOL)
    check(f"{precision} recon {tag}: du/dt median rel. diff", 0.0, float(np.median(ru)), atol=MEDIAN_TOL)
    check(f"{precision} recon {tag}: du/dt 99th pct rel. diff", 0.0, float(np.percentile(ru, 99)), atol=P99_TOL)
    if precision == "fp64":
        od = dict(d)
        dens = osph.compute_density(d["x"], d["mass"], d["h_phantom"], kernel_name=d["hc"].kernel, eta=d["hc"].eta,
                                    h_max=d["hc"].h_max)  # fmt: skip
        od["dens"] = dens
        od["nb"] = on.tighten(dens.neighbours, d["x"], dens.h)
        oacc, odudt = _oracle_forces(od, recon)
        check("FP64: acceleration vs the oracle (1e-10 per vector)", oacc, acc, rtol=1e-10, per_vector=True)
        scale = np.abs(odudt).max()
        check("FP64: du/dt vs the oracle (1e-10 of the largest)", odudt / scale, dudt / scale, atol=1e-10)


def test_coupled_step_gasrad_beta_cooling():
    from nebula.ics import jittered_sphere_on_parabola
    from nebula.sim import coupled

    from oracle import coupled as oc

    ic = jittered_sphere_on_parabola(2048, r0=3000.0, r_p=500.0, seed=5)
    # the reference's mass convention for gas + radiation: the cloud is one code mass
    ic = replace(ic, mass=ic.mass / ic.mass.sum(), u=ic.u * 100.0)
    base = coupled.baseline_config(eos="gasrad")
    cfg = replace(base, cooling_beta=10.0)
    print("coupled: 2048-particle cloud, gas + radiation EOS, beta = 10 cooling, 4 steps", flush=True)
    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    ost = oc.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    for _ in range(4):
        dt = oc.stable_dt(ost, cfg)
        check("stable_dt (1e-10 relative)", dt, coupled.stable_dt(st, cfg), rtol=1e-10)
        st = coupled.step(st, cfg, dt)
        ost = oc.step(ost, cfg, dt)
    got = st.to_numpy()
    for k in ("x", "p"):
        check(f"gasrad + beta cooling: {k} (1e-10 per vector)", ost.caller_order(k), got[k], rtol=1e-10, per_vector=True)
    for k in ("u", "h", "rho", "c_s"):
        check(f"gasrad + beta cooling: {k} (1e-10 relative)", ost.caller_order(k), got[k], rtol=1e-10)
    check("radiated energy (1e-10 relative)", ost.e_radiated, float(st.e_radiated), rtol=1e-10)
    check("energy was radiated", True, ost.e_radiated > 0.0)


def test_hot_blob_ledger_gasrad_beta_cooling():
    """test_coupled.py:80-93: beta = 10 cooling with the gas + radiation EOS, from apocentre
    through pericentre: max |d(sum m (H + u) + E_rad)| / sum m|H| < 1e-5, falling 3-5x when the
    step factors are halved; angular momentum to 1e-10; energy is radiated."""
    from nebula.sim import coupled

    out = {}
    for label, eta, cour in (("fine", 0.00125, 0.15), ("coarse", 0.0025, 0.3)):
        cfg = hot_blob_config(eta, cour, eos="gasrad", cooling_beta=10.0)
        st = coupled.initialise(*hot_blob_ic(), cfg)
        led0 = coupled.ledger(st, cfg)
        e = lm = 0.0
        steps = 0
        while st.t < T_END_BLOB * (1 - 1e-14):
            st = coupled.step(st, cfg, min(coupled.stable_dt(st, cfg), T_END_BLOB - st.t))
            led = coupled.ledger(st, cfg)
            e = max(e, abs(led["energy"] - led0["energy"]) / led0["abs_h"])
            lm = max(lm, float(np.linalg.norm(led["angmom"] - led0["angmom"]) / np.linalg.norm(led0["angmom"])))
            steps += 1
        out[label] = (e, lm, float(st.e_radiated))
        print(f"hot blob, gas + radiation, beta 10 ({label}): {steps} steps, ledger {e:.3e}, |dL|/|L| {lm:.3e}, "
              f"E_rad {out[label][2]:.3e}", flush=True)  # fmt: skip
    check("max ledger error (with E_rad)", 0.0, out["fine"][0], atol=1e-5)
    check("error ratio coarse / fine", 4.0, out["coarse"][0] / out["fine"][0], atol=1.0)
    check("angular momentum", 0.0, out["fine"][1], atol=1e-10)
    check("energy radiated", True, out["fine"][2] > 0.0)
