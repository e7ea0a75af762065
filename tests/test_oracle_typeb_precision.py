"""Why the ghost walls' type-B fits stay FP64 on the mixed path.

The virtual markers of the walls interpolate the fluid fields with a type-B
least-squares fit over their fluid neighbors only (interpolate_LSSPH_B_mod
.f90:24-178, q = 6 for Boussinesq convection with WALL_MODEL 3).  Next to a
wall those supports are truncated.  This measures, on the oracle's state of
a 64^2 convection box (jitter 0.2 dx), the 1-norm condition number of every
marker's moment matrix and the error that the mixed path's recipe -- FP32
pair sums, FP64 solve with one refinement step against them -- would put
into the interpolated wall temperature, against the FP64 fit:

  kappa_1 over the 1088 markers: median 75, 90th percentile 670, max 1.4e3;
  relative error of T at the markers: median 2e-7, 99th percentile 9e-7,
  max 1.7e-6.

The type-A mixed path keeps temperatures within ~1e-10 of FP64 over 20
steps (tests/test_gpu_oracle.py), and the Nusselt number is a wall
gradient of these values, so the marker fits compute in FP64 and are not
part of the mixed path.  The test pins the measurement's order of
magnitude."""
import numpy as np
from scipy.spatial import KDTree

from lssph import workloads
from oracle.pyoracle import OracleSim, kernel_w


def _marker_fits():
    param, st = workloads.boussinesq(n=64, ra=1e6, jitter=0.2, seed=2030)
    c = param.to_c()
    o = OracleSim(c, st.num_int, st.num_ext)
    o.init(st)
    o.ghost()
    s, vm = o.get_state(), o.get_vm()
    o.close()
    n = st.num_int
    xf, yf, tf, rf = s["x"][:n], s["y"][:n], s["tem"][:n], s["rho"][:n]
    tree = KDTree(np.column_stack((xf, yf)))
    kappa, err = [], []
    for k in range(st.num_ext):
        mx, my = vm["x"][k], vm["y"][k]
        nb = np.array(sorted(tree.query_ball_point((mx, my), r=c.h_eff)))
        if nb.size < 6:
            continue
        dx, dy = (xf[nb] - mx) / c.h, (yf[nb] - my) / c.h
        a = np.stack([np.ones_like(dx), dx, dy, 0.5 * dx * dx, dx * dy, 0.5 * dy * dy], axis=1)
        r = np.hypot(xf[nb] - mx, yf[nb] - my)
        w = np.array([kernel_w(c.kernel_type, float(q), c.h) for q in r]) * c.sp_mass / rf[nb]
        m = a.T @ (w[:, None] * a)
        b = a.T @ (w * tf[nb])
        kappa.append(np.linalg.norm(m, 1) * np.linalg.norm(np.linalg.inv(m), 1))
        d64 = np.linalg.solve(m, b)
        # FP32 sums in neighbor order, FP64 solve plus one refinement step
        m32 = np.zeros((6, 6), np.float32)
        b32 = np.zeros(6, np.float32)
        a32, w32 = a.astype(np.float32), w.astype(np.float32)
        for j in range(nb.size):
            va = w32[j] * a32[j]
            m32 += np.outer(va, a32[j]).astype(np.float32)
            b32 += (va * np.float32(tf[nb[j]])).astype(np.float32)
        mm, bm = m32.astype(np.float64), b32.astype(np.float64)
        x = np.linalg.solve(mm, bm)
        x += np.linalg.solve(mm, bm - mm @ x)
        err.append(abs(x[0] - d64[0]) / abs(d64[0]))
    return np.array(kappa), np.array(err)


def test_truncated_type_b_fits_are_too_ill_conditioned_for_fp32_sums():
    kappa, err = _marker_fits()
    print(f"kappa_1 median {np.median(kappa):.3g} max {kappa.max():.3g}; "
          f"T error median {np.median(err):.3g} max {err.max():.3g}")
    assert kappa.size > 1000
    assert 30 < np.median(kappa) < 300 and 5e2 < kappa.max() < 1e4
    assert 1e-7 < np.median(err) < 1e-6 and 5e-7 < err.max() < 1e-5
