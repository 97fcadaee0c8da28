"""Load the golden fixtures (written by tests/golden/make_golden.py from the
reference's Fortran routines) as basin objects for the tests."""
import os

import numpy as np

from pyom2_amd.basins import Basin

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["c0_beta_plane_128", "islands_64", "spherical_96", "cyclic_40x32", "maxiter_48",
         "warmstart_56", "tricky_36x28"]


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    d = {k: z[k] for k in z.files}
    nx, ny = int(d["nx"]), int(d["ny"])
    b = Basin(nx, ny, bool(d["cyclic_x"]), d["dxt"], d["dxu"], d["dyt"], d["dyu"], d["cost"],
              d["cosu"], d["kbot"], None, None, None, d["hur"], d["hvr"], None, d["maskU"],
              d["maskV"], d["forc"], name)
    return b, d


STEP_NAMES = ["step_24x20_nz5", "step_cyclic_28x18_nz4"]


def load_step(name):
    """A whole-routine fixture: (basin, Column, Momentum at taum1=0/tau=1, dict).
    Time level 2 (taup1) of the state is zero on entry."""
    from pyom2_amd.basins import Column, Momentum
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    d = {k: z[k] for k in z.files}
    nx, ny, nz = int(d["nx"]), int(d["ny"]), int(d["nz"])
    b = Basin(nx, ny, bool(d["cyclic_x"]), d["dxt"], d["dxu"], d["dyt"], d["dyu"], d["cost"],
              d["cosu"], d["kbot"], None, None, None, d["hur"], d["hvr"], d["maskT"][-1],
              d["maskU"][-1], d["maskV"][-1], np.zeros((ny + 4, nx + 4)), name)
    col = Column(nz, d["dzt"], d["dzw"], d["kbot_levels"], d["maskT"], d["maskU"], d["maskV"])
    z3 = np.zeros((nz, ny + 4, nx + 4))
    z2 = np.zeros((ny + 4, nx + 4))
    mom = Momentum(np.stack([z3, d["rho"], z3]), np.stack([z3, d["u_tau"], z3]),
                   np.stack([z3, d["v_tau"], z3]), np.stack([d["du_taum1"], d["du_tau"], z3]),
                   np.stack([d["dv_taum1"], d["dv_tau"], z3]), d["du_mix"], d["dv_mix"],
                   np.stack([z2, d["psi_tau"], z2]), np.stack([d["dpsi_taum1"], d["dpsi_tau"], z2]))
    return b, col, mom, d


PRESS_NAMES = ["press_fs_24x20_nz5", "press_rigid_26x22_nz3", "press_rigid_cyclic_30x16_nz4"]


def load_press(name):
    """A surface-pressure fixture: (basin with hu/hv, Column, Momentum at
    taum1=0/tau=1 with psi the surface pressure, dict).  Time level 2
    (taup1) of the state is zero on entry."""
    from pyom2_amd.basins import Column, Momentum
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    d = {k: z[k] for k in z.files}
    nx, ny, nz = int(d["nx"]), int(d["ny"]), int(d["nz"])
    kbot = d["kbot_levels"]
    b = Basin(nx, ny, bool(d["cyclic_x"]), d["dxt"], d["dxu"], d["dyt"], d["dyu"], d["cost"],
              d["cosu"], kbot, None, d["hu"], d["hv"], None, None, d["maskT"][-1],
              d["maskU"][-1], d["maskV"][-1], np.zeros((ny + 4, nx + 4)), name)
    col = Column(nz, d["dzt"], d["dzw"], kbot, d["maskT"], d["maskU"], d["maskV"])
    z3 = np.zeros((nz, ny + 4, nx + 4))
    z2 = np.zeros((ny + 4, nx + 4))
    mom = Momentum(np.stack([z3, d["rho"], z3]), np.stack([z3, d["u_tau"], z3]),
                   np.stack([z3, d["v_tau"], z3]), np.stack([d["du_taum1"], d["du_tau"], z3]),
                   np.stack([d["dv_taum1"], d["dv_tau"], z3]), d["du_mix"], d["dv_mix"],
                   np.stack([d["psi_taum1"], d["psi_tau"], z2]), np.stack([z2, z2, z2]))
    return b, col, mom, d


NH_NAMES = ["nh_box_20x16_nz10", "nh_cyclic_22x14_nz8", "nh_stretched_18x16_nz10"]


def load_nonhydro(name):
    """A non-hydrostatic fixture: (basin, Column, maskW, NonHydro state with
    taum1=0/tau=1/taup1=2, dict)."""
    from pyom2_amd.basins import Column, NonHydro
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    d = {k: z[k] for k in z.files}
    nx, ny, nz = int(d["nx"]), int(d["ny"]), int(d["nz"])
    kbot = d["kbot_levels"]
    b = Basin(nx, ny, bool(d["cyclic_x"]), d["dxt"], d["dxu"], d["dyt"], d["dyu"], d["cost"],
              d["cosu"], kbot, None, None, None, None, None, d["maskT"][-1], d["maskU"][-1],
              d["maskV"][-1], np.zeros((ny + 4, nx + 4)), name)
    col = Column(nz, d["dzt"], d["dzw"], kbot, d["maskT"], d["maskU"], d["maskV"])
    z3 = np.zeros((nz, ny + 4, nx + 4))
    nh = NonHydro(np.stack([z3, z3, d["u_taup1"]]), np.stack([z3, z3, d["v_taup1"]]),
                  np.stack([z3, d["w_tau"], d["w_taup1"]]),
                  np.stack([d["dw_taum1"], d["dw_tau"], z3]), d["dw_mix"],
                  np.stack([d["p_taum1"], d["p_tau"], z3]))
    return b, col, d["maskW"], nh, d
