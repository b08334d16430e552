#!/usr/bin/env python
"""Driver of the MI355X photon walk with the settings layout of the reference
driver (monte_carlo3D-run.py): module-level flags for the bottom boundary,
run parameters in run(), one output file per grain size.

    python monte_carlo3D-run.py [--n_photon N] [--theta_0 DEG] [--wvl UM]
                                [--half_width UM] [--rds_snw UM] [--shape NAME]
                                [--roughness NAME] [--stokes I Q U V] [config flags]

The reference picks the grain shape, roughness and initial Stokes vector by
editing its driver (a default and commented alternatives); here they are
flags with the same names and values ('droxtal', 'hexagonal plate', ...;
'smooth', 'moderately rough', 'severely rough').
"""
import argparse

import numpy as np

from monte_carlompi_amd import monte_carlo3D

DEBUG = False
LAMBERTIAN_SURFACE = False      # True: simulate a Lambertian surface instead of snow
LAMBERTIAN_BOTTOM = True        # Lambertian boundary under the snow ...
LAMBERTIAN_REFLECTANCE = 0.5    # ... with this reflectance


def run():
    p = argparse.ArgumentParser(add_help=False)
    p.add_argument('--n_photon', type=int, default=10000)
    p.add_argument('--theta_0', type=float, default=15.)
    p.add_argument('--wvl', type=float, default=1.3)
    p.add_argument('--half_width', type=float, default=0.085)
    p.add_argument('--rds_snw', type=float, default=100.)
    p.add_argument('--shape', type=str, default='sphere')
    p.add_argument('--roughness', type=str, default='smooth')
    p.add_argument('--stokes', type=float, nargs=4, default=[1, 0, 0, 0])
    a, _ = p.parse_known_args()
    stokes_params = np.array(a.stokes)
    single_grain_size(a.n_photon, a.wvl, a.half_width, a.rds_snw, theta_0=a.theta_0,
                      stokes_params=stokes_params, shape=a.shape, roughness=a.roughness)


def single_grain_size(n_photon, wvl, half_width, rds_snw, theta_0=0.,
                      stokes_params=np.array([1, 0, 0, 0]), shape='sphere', roughness='smooth'):
    mc = monte_carlo3D.MonteCarlo()
    mc.run(n_photon, wvl, half_width, rds_snw, theta_0=theta_0, stokes_params=stokes_params,
           shape=shape, roughness=roughness, debug=DEBUG, Lambertian_surface=LAMBERTIAN_SURFACE,
           Lambertian_bottom=LAMBERTIAN_BOTTOM, Lambertian_reflectance=LAMBERTIAN_REFLECTANCE)
    mc.close()


def multiple_grain_sizes(n_photon, wvl, half_width, rds_snw, theta_0=0.,
                         stokes_params=np.array([1, 0, 0, 0]), shape='sphere', roughness='smooth'):
    mc = monte_carlo3D.MonteCarlo()
    for rds in rds_snw:
        mc.run(n_photon, wvl, half_width, rds, theta_0=theta_0, stokes_params=stokes_params,
               shape=shape, roughness=roughness, debug=DEBUG, Lambertian_surface=LAMBERTIAN_SURFACE,
               Lambertian_bottom=LAMBERTIAN_BOTTOM, Lambertian_reflectance=LAMBERTIAN_REFLECTANCE)
    mc.close()


def main():
    run()


if __name__ == '__main__':
    main()
