"""What the trained closure (ps3d_amd/data/closure64.npz) gains over the coarse
run without it, on two held-out seeds: the 64^3 run with the net, the plain 64^3
run at ps3d's prediss 30 and at the prediss tuned without the net, against the
256^3 run truncated to the modes a 64^3 step passes (genspec spectrum below the
coarse 2/3 cutoff, enstrophy), at two horizons.  The numbers are printed; the
checks are that every run completes, that the net beats the untuned coarse run
and that it stays within 10 % of the tuned one (the trained net's measured gain
over tuning the coarse hyperdiffusion is nil: spectrum 10.0 % against 9.8 % over
20 seeds x 1000 steps)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(os.path.dirname(HERE), "ps3d_amd", "data", "closure64.npz")


@pytest.mark.parametrize("fine_steps", [200, 600])
def test_trained_closure_against_the_coarse_run_without_it(fine_steps):
    from ps3d_amd import closure_run as CR
    net, gain, pd_c, meta = CR.load_net(WEIGHTS, 64)
    pd_p = meta["coarse_prediss_plain_tuned"]
    cfg = [("no_closure", 30.0, None), ("no_closure_tuned", pd_p, None), ("closure", pd_c, gain)]
    print("trained closure, %d fine steps" % fine_steps, flush=True)
    ev = CR.evaluate(256, 64, net, [9003, 9004], fine_steps, cfg)
    net.close()
    for e in ev:
        print({k: (e[k]["steps"], round(100 * e[k]["spectrum_error"], 2), round(100 * e[k]["enstrophy_error"], 1))
               for k, _, _ in cfg}, "t_end %.3f" % e["t_end"], flush=True)
        for k, _, _ in cfg:
            assert e[k]["completed"] and np.isfinite(e[k]["spectrum_error"]), k
        assert e["closure"]["spectrum_error"] < e["no_closure"]["spectrum_error"]
        assert e["closure"]["spectrum_error"] < 1.1 * e["no_closure_tuned"]["spectrum_error"]
