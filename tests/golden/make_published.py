"""Writes tests/golden/published.json: the accuracy figures the reference
publishes for this path, transcribed from its data files (no code is run).

Sources (read as text):
  examples/make_validation_plots.py:86-89   isentropic vortex, 1st-order HLLC + RK3
  examples/make_validation_plots.py:94-97   periodic advected pulse
  examples/make_validation_plots.py:104-107 cylinder drag (marker Cd)
  examples/make_validation_plots.py:140-142 Sod exact-Riemann gate
  crates/cfd_solver/tests/isentropic_vortex_muscl_convergence.rs:176-191  MUSCL order gate
  crates/cfd_solver/tests/poiseuille_channel_validation.rs:88-91          Poiseuille L2
  examples/cylinder_drag_validation/config.toml:24-26                    Cd grid series
  examples/make_validation_plots.py:109-117, examples/sphere_drag_validation/README.md:5,
  examples/sphere_drag_validation/config.toml:21-27 (D/dx = 12 row), the example's own
  setup (main.rs, config.toml)                                             sphere drag
"""
import json
import os

DATA = {
    "vortex_first_order": {"N": [40, 80, 160], "L1": [5.029291e-3, 2.774181e-3, 1.467073e-3], "fitted_order": 0.889,
                            "setup": {"L": 10.0, "beta": 5.0, "gamma": 1.4, "mean": [0.5, 0.5], "cfl": 0.3, "t_final": 1.0,
                                      "bcs": "outflow x/y, slip z", "scheme": "hllc", "integrator": "rk3"}},
    "periodic_pulse": {"N": [40, 80, 160], "L1": [1.491960e-2, 8.918114e-3, 4.986845e-3], "fitted_order": 0.791,
                       "max_mass_drift": 3.035e-14, "mass_tol": 1.0e-12,
                       "setup": {"domain": 1.0, "center": 0.85, "sigma": 0.08, "amplitude": 0.2, "gamma": 1.4,
                                 "velocity": 1.0, "pressure": 1.0, "t_final": 0.3, "cfl": 0.45}},
    "sod": {"N": [100, 200, 400], "L1": [1.733831e-2, 1.070826e-2, 6.782218e-3],
            "setup": {"rho": [1.0, 0.125], "p": [1.0e5, 1.0e4], "x0": 0.5, "t_final": 6.0e-4, "cfl": 0.4}},
    "muscl_mc_gate": {"N": [40, 80, 160], "min_fitted_order": 1.8, "max_err_finest": 3.0e-4},
    "poiseuille": {"rel_l2_measured": 7.6e-4, "gate": 5.0e-3, "wrong_mu_control": 0.37},
    "cylinder_cd": {"D_over_dx": [10, 15, 20], "cd": [5.126, 5.466, 5.6312], "cd_ref": 5.5795},
    "sphere_drag": {"cd": 2.5402, "cd_schiller_naumann": 2.6095,
                    "source": "examples/make_validation_plots.py:109-117; examples/sphere_drag_validation/README.md:5",
                    "ptp_percent": 2.7,
                    "ptp_source": "examples/sphere_drag_validation/config.toml:21-27 (D/dx = 12 row)",
                    "setup": {"grid": [144, 72, 72], "lo": [-0.3, -0.3, -0.3], "hi": [0.9, 0.3, 0.3], "D": 0.1,
                              "rho": 1.0, "p": 100000.0, "u_inf": 56.1, "reynolds": 20.0, "markers": 2000, "cfl": 0.3,
                              "muscl": False, "sponge": {"x_start": 0.5, "sigma": 20000.0}, "runs": 9000,
                              "average_frac": 0.2, "bcs": "subsonic inflow x-lo, pressure outlet x-hi, symmetry y/z",
                              "scheme": "hllc + viscous, rk3"}},
}

if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "published.json")
    with open(path, "w") as f:
        json.dump(DATA, f, indent=1)
    print(path)
