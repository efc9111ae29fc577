"""Writes the golden fixtures of tests/golden from the FP64 oracle (run from the repo root:
``python tests/golden/make_golden.py``).

seeds_random.json   splitmix64 stream -> seed centres and Shoemake unit quaternions for
                    (seed 20240917, 32 seeds, radius 4, 256^3) -- the multi-dendrite case.
seed_geometry.json  flags / grain ids of three seeds on a 24 x 20 x 16 grid (sha256 of the
                    int8 / int16 arrays plus per-class counts), from the oracle's seed
                    geometry.
"""
import hashlib
import json
import math
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from materkit_anisotropiclbm_trt_rbc_facet_amd.mkg import MkgParams  # noqa: E402
from oracle import oracle as O  # noqa: E402

HERE = Path(__file__).resolve().parent


def random_seeds(seed, n, radius, dims):
    u = O.splitmix_stream(seed, 3 * n + 3 * n)
    centers = np.zeros((n, 3))
    k = 0
    for s in range(n):
        for a in range(3):
            lo, hi = radius, dims[a] - 1 - radius
            centers[s, a] = lo + (hi - lo) * u[k]
            k += 1
    quats = np.zeros((n, 4))
    for s in range(n):
        u1, u2, u3 = u[k], u[k + 1], u[k + 2]
        k += 3
        a, b = math.sqrt(1 - u1), math.sqrt(u1)
        q0, q1 = a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2)
        q2, q3 = b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)
        quats[s] = (q3, q0, q1, q2)
    return centers, quats


def seed_geometry():
    p = MkgParams()
    p.nx, p.ny, p.nz = 24, 20, 16
    p.phase_q = 19
    p.pf_wdx2 = 6.25
    p.phi_s = 0.99
    centers = np.array([[6.0, 5.5, 7.25], [17.5, 13.0, 8.0], [12.0, 2.0, 13.5]])
    radii = np.array([4.0, 3.5, 2.0])
    s = O.State.empty(p)
    s.init_seeds(centers, radii)
    return {
        "dims": [24, 20, 16], "centers": centers.tolist(), "radii": radii.tolist(),
        "pf_wdx2": 6.25, "phi_s": 0.99,
        "flag_sha256": hashlib.sha256(s.flag.tobytes()).hexdigest(),
        "grain_sha256": hashlib.sha256(s.grain.tobytes()).hexdigest(),
        "flag_counts": {str(v): int((s.flag == v).sum()) for v in (0, 1, 2)},
        "grain_counts": {str(v): int((s.grain == v).sum()) for v in (1, 2, 3)},
        "phi_sum": float(s.phi.sum()),
    }


def main():
    c, q = random_seeds(20240917, 32, 4.0, (256, 256, 256))
    (HERE / "seeds_random.json").write_text(json.dumps(
        {"seed": 20240917, "n": 32, "radius": 4.0, "dims": [256, 256, 256],
         "centers": c.tolist(), "quats": q.tolist()}, indent=1))
    (HERE / "seed_geometry.json").write_text(json.dumps(seed_geometry(), indent=1))
    print("golden fixtures written")


if __name__ == "__main__":
    main()
