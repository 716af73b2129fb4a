"""Error reporting of the team launches: an exchange time-out is kept on the device until anc_sync_status reports it
(a later clean launch cannot hide it), it raises instead of silently re-running, and the device-scope fallback happens
only when the caller opts in.

The time-out is forced from the host (option "stall_test": every workgroup's bounded wait polls a word that nothing
sets until option "timeout_s" expires, raises the error and exits), so every kernel still ends normally.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SCH = dict(S=30, t0=12.0, t1=5.0, q=8.0, lam=1.05, ramp=True)


@pytest.fixture(scope="module")
def k2000():
    from anechoic_amd import instances as I
    J, _ = I.ising_random(20261203, 2000)
    return J


def _engine(**kw):
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    from anechoic_amd import Engine
    return Engine(0, **kw)


def _stalled_launch(eng, inst, R):
    import torch
    eng.set_option("timeout_s", 0.002)
    eng.set_option("stall_test", 1)
    o = eng.alloc_out(inst, R)
    eng.launch(inst, 1, torch.arange(R, dtype=torch.int32, device="cuda"), o)
    eng.set_option("stall_test", 0)
    eng.set_option("timeout_s", 5.0)


def test_timeout_survives_a_later_clean_launch(k2000):
    import torch
    from anechoic_amd import _native as N
    eng = _engine()
    try:
        tb = O.tables(SCH)
        eng.set_tables(tb)
        inst = eng.add_instance(k2000)
        _stalled_launch(eng, inst, 64)
        good = eng.alloc_out(inst, 64, trace=True)
        trials = torch.arange(64, dtype=torch.int32, device="cuda")
        eng.launch(inst, 0x51, trials, good)    # clean launch queued before the one synchronisation
        with pytest.raises(N.AncError, match="timed out"):
            eng.sync()
        assert int(eng.get_option("xflavor")) == 3   # no fallback without the opt-in
        eng.sync()                                    # reported once, then clear
        ref = O.run_batch(k2000, 2000, tb, 0x51, np.arange(64, dtype=np.uint32), trace=True, nthreads=16)
        assert np.array_equal(good["sum_sh"].cpu().numpy(), ref["sum_sh"])
        assert np.array_equal(good["nlin"].cpu().numpy(), ref["nlin"])
        # the plain launch and run() report it as well, and run() does not repeat the run
        eng.set_option("timeout_s", 0.002)
        eng.set_option("stall_test", 1)
        with pytest.raises(N.AncError, match="timed out"):
            eng.run(inst, 3, [0], kernel="plain")
        with pytest.raises(N.AncError, match="timed out"):
            eng.run(inst, 3, np.arange(40, dtype=np.uint32))
        eng.set_option("stall_test", 0)
        eng.set_option("timeout_s", 5.0)
        eng.sync()
    finally:
        eng.close()


def test_fallback_only_when_opted_in(k2000):
    from anechoic_amd import _native as N
    eng = _engine(fallback=True)
    try:
        eng.set_tables(O.tables(SCH))
        inst = eng.add_instance(k2000)
        _stalled_launch(eng, inst, 32)
        with pytest.raises(N.AncError, match="device-scope exchange"):
            eng.sync()
        assert int(eng.get_option("xflavor")) == 1
    finally:
        eng.close()
