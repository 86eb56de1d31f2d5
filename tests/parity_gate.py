"""The field gate of the GPU parity tests.

rho, u and p are compared by relative L2. v is compared by the L2 norm of its error over
the L2 norm of the velocity (u, v): v is near zero over most of a nozzle flow, where its
own relative L2 measures rounding noise rather than the kernel (the FP32 oracle itself
sits 2e-4 to 0.6 from the FP64 oracle there). One fixed bar applies to all four; the plain
relative L2 of v is printed beside it.
"""
import sys

import numpy as np

BAR = 1e-4


def prims(A, B):
    A = A.reshape(-1, 4).astype(np.float64)
    B = B.reshape(-1, 4).astype(np.float64)
    rho = np.maximum(A[:, 0], 1e-7)
    u, v = A[:, 1] / rho, A[:, 2] / rho
    G = np.maximum(B[:, 2], 0.5)
    p = np.maximum((A[:, 3] - 0.5 * rho * (u * u + v * v)) / G, 1e-7)
    return {"rho": rho, "u": u, "v": v, "p": p}


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def field_errors(pg, po):
    """rho, u, p: relative L2; v: error over the velocity norm; v_rel: plain relative L2."""
    e = {k: rel_l2(pg[k], po[k]) for k in ("rho", "u", "p")}
    vel = np.sqrt(np.sum(po["u"] ** 2 + po["v"] ** 2))
    e["v"] = float(np.linalg.norm(pg["v"] - po["v"]) / max(vel, 1e-300))
    e["v_rel"] = rel_l2(pg["v"], po["v"])
    return e


def gate(pg, po, label, bar=BAR):
    e = field_errors(pg, po)
    print(f"[parity] {label}: rho={e['rho']:.2e} u={e['u']:.2e} v/|U|={e['v']:.2e} p={e['p']:.2e} "
          f"(v rel-L2 {e['v_rel']:.2e})", flush=True)
    sys.stdout.flush()
    for k in ("rho", "u", "v", "p"):
        assert e[k] <= bar, f"{label}: {k} {e[k]:.3e} > {bar}"
    return e
