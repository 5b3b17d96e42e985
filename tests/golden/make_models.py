"""Freezes the seeded synthetic models (bath draws) as tests/golden/models.json.

    python tests/golden/make_models.py

The draws come from ed_solver_amd.aim.synthetic (numpy default_rng, eps_k then V_k);
tests/test_oracle.py checks that the generator still reproduces this file bit for bit,
so the device path and the C oracle never depend on NumPy's RNG.
"""
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from ed_solver_amd.aim import WORKLOADS, synthetic  # noqa: E402

out = {}
for name in WORKLOADS:
    m, sec = synthetic(name)
    out[name] = {"model": m.to_json(), "sector": list(sec)}
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models.json")
with open(path, "w") as f:
    json.dump(out, f, indent=1)
print(path)
