"""FFEA's internal ("meso") units and the SI->meso conversions of the loaders.

The constants are the reference's (include/dimensions.h:30-40) and must be the
same values for results to agree. Conversions follow SimulationParams.cpp:252-265
(dt, epsilon), :415-421 (kT), Blob.cpp:2282-2287 (material), World.cpp:2628-2632
(scale) and Blob.cpp:266 (stokes drag).
"""
import math

LENGTH = 1.7e-10
ENERGY = 4.141945559999999e-21
MASS = 1.994307387553024e-26
AREA = 2.89e-20
VOLUME = 4.913e-30
FORCE = 2.4364705882352941e-11
VELOCITY = 4.5572845297447225e2
TIME = 3.7302915560886126e-13
PRESSURE = 8.4305832688784826e8
PI = 3.141592653589793


def dt_to_meso(dt_si):
    return dt_si / TIME


def kT_to_meso(kT_si):
    return kT_si / ENERGY


def epsilon2(epsilon):
    e = float(epsilon)
    return e * e


def scale_to_meso(scale_si):
    return scale_si / LENGTH


def material_to_meso(density, shear_visc, bulk_visc, shear_mod, bulk_mod):
    """(rho, A, B, G, E) in meso units; B keeps the reference's mix of a meso
    bulk viscosity and an SI shear viscosity (Blob.cpp:2284)."""
    rho = density * VOLUME / MASS
    A = shear_visc / (PRESSURE * TIME)
    B = bulk_visc / (PRESSURE * TIME) - (2.0 / 3.0) * shear_visc
    G = shear_mod / PRESSURE
    E = bulk_mod / PRESSURE
    return rho, A, B, G, E


def stokes_visc_to_meso(visc_si):
    return visc_si / (PRESSURE * TIME)


def stokes_drag(radius_meso, visc_meso):
    return 6.0 * PI * visc_meso * radius_meso


def force_to_meso(f_si):
    return f_si / FORCE
