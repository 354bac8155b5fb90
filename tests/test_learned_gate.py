"""The learned start's gate (CPU, no GPU): when the network runs and when the step starts from
h_prev, from the fraction of rows the plain start leaves for the slow solve. A decision reads
the counts of every step but the last (that step may still be running when the next one is
launched)."""

from helpers import check


def test_gate_probes_switches_and_reprobes():
    from nebula.sim.learned_h import StartGate

    g = StartGate(on_above=0.05, off_below=0.03, probe_every=3)
    n = 1000
    check("first step is a probe (plain start)", False, g.decide())
    g.observe(400, n, False)  # plain start leaves 40 % of the rows slow
    check("its count is not read before the next step's decision: another probe", False, g.decide())
    g.observe(300, n, False)
    seq = [g.decide()]
    g.observe(10, n, True)  # steps with the network say nothing about the plain start
    seq += [g.decide()]
    g.observe(10, n, True)
    seq += [g.decide()]
    check("network on for the next probe_every steps", [True, True, True], seq)
    g.observe(10, n, True)
    check("then a probe", False, g.decide())
    g.observe(20, n, False)  # the plain start now leaves 2 %: below off_below
    g.decide()  # (reads only the network step's count)
    g.observe(10, n, True)
    check("the network is switched off", [False, False], [g.decide(), g.decide()])
    g.observe(40, n, False)  # 4 %: not above on_above (hysteresis)
    g.decide()
    g.observe(60, n, False)  # 6 % > on_above
    check("stays off between the thresholds", False, g.decide())
    g.observe(60, n, False)
    check("switched on again", True, g.decide())
    st = g.stats()
    check("statistics count every decision", st["steps_network"] + st["steps_plain"], 12)
    check("last plain fraction", 0.06, st["last_plain_slow_fraction"])


def test_gate_stays_on_while_probes_find_slow_rows():
    from nebula.sim.learned_h import StartGate

    g = StartGate(on_above=0.05, off_below=0.03, probe_every=2)
    g.decide()
    g.observe(900, 1000, False)
    g.decide()
    g.observe(900, 1000, False)
    pattern = []
    for _ in range(9):
        use = g.decide()
        pattern.append(use)
        g.observe(5 if use else 900, 1000, use)
    check("probe every third step, network otherwise", [True, True, False] * 3, pattern)


def test_snapshot_restores_the_decisions():
    from nebula.sim.learned_h import StartGate

    g = StartGate(on_above=0.05, off_below=0.03, probe_every=3)
    g.decide()
    g.observe(400, 1000, False)
    snap = g.snapshot()
    first = [g.decide()]
    g.observe(300, 1000, False)
    first.append(g.decide())
    g.restore(snap)
    again = [g.decide()]
    g.observe(300, 1000, False)
    again.append(g.decide())
    check("the same decisions after restore", first, again)
