"""Hodgkin-Huxley 1952 channels: the oracle's rates against the published
forms the reference quotes (``channel/sodium.py:293-318``,
``channel/potassium.py:348-370``) and the classic resting gate values, the
host initialisers against the oracle, and the channel-set plumbing."""

import numpy as np
import pytest

from oracle import channels as ch


def test_rates_equal_the_published_forms():
    V = np.linspace(-100.0, 50.0, 601)
    V = V[(np.abs(V + 40.0) > 1e-6) & (np.abs(V + 55.0) > 1e-6)]
    ap, bp, aq, bq = ch.na_hh1952(V)
    np.testing.assert_allclose(ap, 0.1 * (V + 40.0) / (1.0 - np.exp(-(V + 40.0) / 10.0)), rtol=1e-12)
    np.testing.assert_allclose(bp, 4.0 * np.exp(-(V + 65.0) / 18.0), rtol=1e-13)
    np.testing.assert_allclose(aq, 0.07 * np.exp(-(V + 65.0) / 20.0), rtol=1e-13)
    np.testing.assert_allclose(bq, 1.0 / (1.0 + np.exp(-(V + 35.0) / 10.0)), rtol=1e-13)
    an, bn = ch.k_hh1952(V)
    np.testing.assert_allclose(an, 0.01 * (V + 55.0) / (1.0 - np.exp(-(V + 55.0) / 10.0)), rtol=1e-12)
    np.testing.assert_allclose(bn, 0.125 * np.exp(-(V + 65.0) / 80.0), rtol=1e-13)
    # the removable singularities
    assert ch.na_hh1952(np.array([-40.0]))[0][0] == pytest.approx(1.0, rel=1e-14)
    assert ch.k_hh1952(np.array([-55.0]))[0][0] == pytest.approx(0.1, rel=1e-14)


def test_resting_gates_are_the_classic_values():
    ap, bp, aq, bq = ch.na_hh1952(np.array([-65.0]))
    an, bn = ch.k_hh1952(np.array([-65.0]))
    assert ap[0] / (ap[0] + bp[0]) == pytest.approx(0.0529, abs=1e-4)
    assert aq[0] / (aq[0] + bq[0]) == pytest.approx(0.5961, abs=1e-4)
    assert an[0] / (an[0] + bn[0]) == pytest.approx(0.3177, abs=1e-4)


def test_host_initialisers_match_the_oracle():
    from braincell_amd import rates

    V = np.linspace(-90.0, 40.0, 131)
    ap, bp, aq, bq = ch.na_hh1952(V)
    m, h = rates.nahh_steady(V, -45.0)
    np.testing.assert_allclose(m, ap / (ap + bp), rtol=1e-13)
    np.testing.assert_allclose(h, aq / (aq + bq), rtol=1e-13)
    an, bn = ch.k_hh1952(V)
    np.testing.assert_allclose(rates.khh_steady(V, -45.0), an / (an + bn), rtol=1e-13)


def test_channel_set_and_state_slots():
    from braincell_amd import workloads as W
    from braincell_amd.runtime import StaggerLayout, choose_channel_set

    L = StaggerLayout([W.n144_hh_cell(0.7)], probes="soma")
    assert L.slots == [("Na_HH1952", 0), ("Na_HH1952", 1), ("K_HH1952", 0)]
    assert L.n_cv_total == 868
    V, S, _ = L.initial_state()
    np.testing.assert_allclose(S[2], ch.k_hh1952(np.array([-62.0]))[0][0] / sum(x[0] for x in ch.k_hh1952(np.array([-62.0]))))
    with pytest.raises(ValueError):
        choose_channel_set((1 << 9) | (1 << 0))  # Na_HH1952 with Na_Ba2002: no kernel holds both
    cc = L.cc
    assert cc["na_vsh"] == -45.0 and cc["na_phi"] == 1.0 and cc["E_K"] == -77.0


def test_n144_hh_cell_spikes_during_the_clamp_in_the_oracle():
    from braincell_amd import workloads as W
    from oracle.staggered import StaggeredOracle
    from tests.helpers import soma_cv, spike_times

    m = W.n144_hh_cell(0.7)
    o = StaggeredOracle(m)
    o.reset_state(m.V0)
    v = o.run(0.025, 800, probe_cvs=[soma_cv(m)])[:, 0]
    st = spike_times(v, 0.025)
    assert len(st) == 1 and 3.0 < st[0] < 5.5
