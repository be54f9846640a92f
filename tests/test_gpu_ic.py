"""GPU parity of the transverse-polarised initial conditions with a general
occupation, the specific-sender form and the optional momentum-space
correlators (nrq_ic.hip) against oracle/ic_oracle.py.

Bars: links and E within 1e-13 absolute (hipFFT against the oracle's FFT, as
for the box initial condition); correlators within 1e-12 of their largest
value; the random stream left at the same position."""
import numpy as np
import pytest

from oracle import ic_oracle as ico

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _prepare(nrq, oracle, lattice, ext, sp, seed):
    lattice(ext, sp)
    oracle.lattice_init(ext, sp)
    vol = int(np.prod(ext))
    pn = np.array([oracle.norm2_momentum(k + 1) for k in range(vol)])
    nrq.random_init(seed)
    oracle.ranlxd_init(2, seed)
    return vol, pn


def _check(nrq, oracle, g, got_corr, want, ext, sp):
    U, E, aa, ee = want[:4]
    Ug, Eg = g.download()
    assert np.max(np.abs(Ug - U)) < 1e-13
    assert np.max(np.abs(Eg - E)) < 1e-13
    if got_corr is not None:
        assert np.max(np.abs(got_corr[0] - aa)) <= 1e-12 * np.max(aa)
        assert np.max(np.abs(got_corr[1] - ee)) <= 1e-12 * np.max(ee)
    # both sides consumed the same number of uniforms
    assert np.array_equal(nrq.random_uniform(7), oracle.ranlxd(7))


def test_box_ic_with_correlators(nrq, oracle, lattice):
    ext, sp, seed = (6, 5, 4), (0.05, 1.0, 0.9, 1.2), 12
    _say("box IC with correlators 6x5x4")
    vol, pn = _prepare(nrq, oracle, lattice, ext, sp, seed)
    g = nrq.GaugeConfiguration()
    corr = g.TransversePolarisedOccupiedInit_Box(0.6, 1.4, 0.9, correlators=True)
    want = ico.transverse_init(oracle, ext, sp, ico.box_occupation(pn, 0.6, 1.4, 0.9), False)
    _check(nrq, oracle, g, corr, want, ext, sp)


@pytest.mark.parametrize("sender", [False, True])
def test_general_occupation(nrq, oracle, lattice, sender):
    ext, sp, seed = (5, 6, 4), (0.05, 1.0, 1.0, 1.0), 13 + sender
    _say(f"transverse init, random occupation, specific sender {sender}")
    vol, pn = _prepare(nrq, oracle, lattice, ext, sp, seed)
    occ = np.random.default_rng(5).uniform(0.0, 2.0, size=vol) / (1 + pn ** 2)
    g = nrq.GaugeConfiguration()
    init = g.TransversePolarisedOccupiedInit_specificSender if sender else g.TransversePolarisedOccupiedInit
    corr = init(occ, correlators=True)
    want = ico.transverse_init(oracle, ext, sp, occ, sender)
    _check(nrq, oracle, g, corr, want, ext, sp)
    if sender:
        # real fields: the correlators measured on the finished configuration
        # are the momentum-space ones
        aa, ee = g.transverse_correlators()
        assert np.max(np.abs(aa - corr[0])) <= 1e-10 * np.max(corr[0])
        assert np.max(np.abs(ee - corr[1])) <= 1e-10 * np.max(corr[1])


def test_specific_sender_fp32_storage(nrq, oracle, lattice):
    ext, sp, seed = (8, 8, 8), (0.05, 1.0, 1.0, 1.0), 3
    _say("specific sender 8^3 FP32 storage")
    vol, pn = _prepare(nrq, oracle, lattice, ext, sp, seed)
    occ = ico.box_occupation(pn, 0.5, 1.0, 1.0)
    g = nrq.GaugeConfiguration(nrq.NRQ_PRECISION_FP32)
    g.TransversePolarisedOccupiedInit_specificSender(occ)
    U, E = ico.transverse_init(oracle, ext, sp, occ, True)[:2]
    Ug, Eg = g.download()
    assert np.max(np.abs(Ug - U)) < 1e-6
    assert np.max(np.abs(Eg - E)) < 1e-6 * max(1.0, np.max(np.abs(E)))


def test_rejects_bad_occupation(nrq, lattice):
    lattice((4, 4, 4), (0.05, 1.0, 1.0, 1.0))
    nrq.random_init(1)
    g = nrq.GaugeConfiguration()
    occ = np.ones(64)
    occ[5] = -1.0
    with pytest.raises(nrq.NRQError):
        g.TransversePolarisedOccupiedInit(occ)
    with pytest.raises(ValueError):
        g.TransversePolarisedOccupiedInit(np.ones(10))
