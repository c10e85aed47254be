"""The benchmark decks at their stated sizes, set up on the GPU library and
on the oracle with identical data (for the full-size parity tests)."""
import os

import numpy as np

from oracle.oracle import OracleFAC


def oracle_threads():
    return max(1, min(16, os.cpu_count() or 1))


def make_oracle(deck, params, dvar, nthreads=None):
    o = OracleFAC(deck["dim"], deck["levels"], deck["ratio"], deck["dlo"], deck["dhi"],
                  [0.0] * deck["dim"], [1.0] * deck["dim"])
    o.set_poisson(dvar, 1.0, 0, 0.0)
    o.set_bc([1.0] * 6, [0.0] * 6, [0.0] * 6)
    o.set_params(max_cycles=params["max_cycles"], residual_tol=params["residual_tol"],
                 relative_residual_tol=params["relative_residual_tol"],
                 pre=params["num_pre_sweeps"], post=params["num_post_sweeps"],
                 coarse_max_iterations=params["coarse_solver_max_iterations"],
                 coarse_tol=params["coarse_solver_tolerance"], ewing=True,
                 nthreads=nthreads or oracle_threads())
    o.set_coarse_mg(params["coarse_solver_choice"] in ("mg", "hypre"))
    o.set_prolongation(params["prolongation_method"])
    o.set_cf_discretization(params.get("cf_discretization", "Ewing"))
    return o


def copy_to_oracle(o, deck, f, D=None):
    """f (and D) from the GPU patch data into the oracle, bit for bit."""
    for ln, boxes in enumerate(deck["levels"]):
        for p in range(len(boxes)):
            o.put("f", ln, p, f.get(ln, p))
            if D is not None:
                for a in range(deck["dim"]):
                    o.put("D", ln, p, D.get(ln, p, axis=a), axis=a)


def interior(deck, ln, p, flat, g=1):
    lo, hi = deck["levels"][ln][p]
    shp = tuple(h - l + 1 + 2 * g for l, h in zip(lo, hi))
    sl = tuple(slice(g, -g) if g else slice(None) for _ in lo)
    return flat.reshape(shp, order="F")[sl]


def compare_u(deck, u, o, rtol):
    """Per patch: max |u_gpu - u_oracle| over interior cells <= rtol * max
    |u_oracle| of that patch.  Returns the worst relative difference."""
    worst = 0.0
    for ln, boxes in enumerate(deck["levels"]):
        for p in range(len(boxes)):
            g = interior(deck, ln, p, u.get(ln, p))
            r = interior(deck, ln, p, o.get("u", ln, p))
            d = float(np.max(np.abs(g - r)))
            scale = float(np.max(np.abs(r)))
            assert d <= rtol * scale, (ln, p, d, scale)
            if scale > 0:
                worst = max(worst, d / scale)
    return worst


def level_errors(deck, u, o):
    """Per level: RMS and max |u_gpu - u_oracle| over interior cells, and
    max |u_oracle|."""
    out = []
    for ln, boxes in enumerate(deck["levels"]):
        sq, n, mx, um = 0.0, 0, 0.0, 0.0
        for p in range(len(boxes)):
            g = interior(deck, ln, p, u.get(ln, p))
            r = interior(deck, ln, p, o.get("u", ln, p))
            d = g - r
            sq += float(np.sum(d * d))
            n += d.size
            mx = max(mx, float(np.max(np.abs(d))))
            um = max(um, float(np.max(np.abs(r))))
        out.append(dict(level=ln, l2=float(np.sqrt(sq / max(n, 1))), max=mx, umax=um))
    return out
