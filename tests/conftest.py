import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def golden_names():
    return sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))


# fixtures of full-size workloads: inputs regenerated from their seeds by
# the tests that use them, not stored
LARGE_FIXTURES = ("cfg2", "cfg3")


def stored_golden_names():
    """Fixtures that store their inputs (U0, M0)."""
    return [n for n in golden_names() if n not in LARGE_FIXTURES]


class Case:
    """One golden fixture: reference inputs and outputs of the wannierise path."""

    def __init__(self, name):
        from wannier90_amd import capi

        z = np.load(os.path.join(GOLDEN, name + ".npz"))
        self.name = name
        self.z = z
        self.ctl = json.loads(str(z["control"]))
        self.kmesh = capi.KMesh(int(z["nnlist"].shape[1]), np.asfortranarray(z["nnlist"].astype(np.int32)),
                                np.asfortranarray(z["nncell"]), np.asfortranarray(z["bk"]),
                                np.ascontiguousarray(z["wb"]), float(z["wbtot"]))
        # large cases store no inputs; tests regenerate them from the seed
        self.m0 = np.asfortranarray(z["m0"]) if "m0" in z else None
        self.u0 = np.asfortranarray(z["u0"]) if "u0" in z else None
        self.nw = int(z["u_final"].shape[0])
        self.nk = int(z["kpt_latt"].shape[0])
        self.nn = self.kmesh.nntot
        # sensitivity of the trajectory: change of the oracle's final state
        # under a 1e-15 relative perturbation of M0 (absent -> 0)
        self.sens_u = float(z["sens_u"]) if "sens_u" in z else 0.0
        self.sens_omega = float(z["sens_omega"]) if "sens_omega" in z else 0.0
        self.sens_centres = float(z["sens_centres"]) if "sens_centres" in z else 0.0

    def gpu_control(self, **over):
        from wannier90_amd import capi

        c = dict(num_iter=self.ctl["num_iter"], num_cg_steps=self.ctl["num_cg_steps"],
                 conv_window=self.ctl["conv_window"], conv_tol=self.ctl["conv_tol"],
                 trial_step=self.ctl["trial_step"], fixed_step=self.ctl["fixed_step"],
                 lfixstep=self.ctl["lfixstep"])
        c.update(over)
        return capi.control(**c)

    def oracle_control(self, **over):
        c = dict(self.ctl)
        c.update(over)
        return c

    def well_conditioned(self, factor=20.0):
        """The reference trajectory can be reproduced to the stated tolerances."""
        return (factor * self.sens_u <= 1e-7 and factor * self.sens_omega <= 1e-8
                and factor * self.sens_centres <= 1e-6)


@pytest.fixture(scope="session")
def cases():
    return {n: Case(n) for n in stored_golden_names()}


def tol(base, sens, factor=20.0):
    """Tolerance of a parity check: the stated bound, or the trajectory's own
    sensitivity to rounding (measured on the oracle) when that is larger."""
    return max(base, factor * sens)
