"""Scalar part of the reference's unit handling (py_fof/__init__.py:239-326).

Shared by the host-array and the device-tensor front ends: it validates the
physical-unit arguments in the reference's order, with its messages, and
returns the code-unit constants. Arrays are scaled by the caller, with numpy
or torch, using ``dx`` (positions), ``v_scale`` (velocities), ``m_grid``
(masses) and ``mp_fill`` (the per-particle mass when only ``n_sim`` is given).
"""
import math


class UnitScalars:
    __slots__ = ("physical", "a", "b", "s15", "L", "link", "softening", "dx", "v_scale",
                 "m_grid", "mp_fill", "ang_mom_scale")


def unit_scalars(units, bound_check, softening, L_Box, linking_length, z, box, h0, omega_m,
                 omega_l, n_sim, ng_global, include_hubble_flow, have_mp, total_np):
    """py_fof/__init__.py:239-326 without the array arithmetic."""
    U = UnitScalars()
    z_val = float(z)
    if 1.0 + z_val <= 0.0:
        raise ValueError(f"1 + z must be positive, got {1.0 + z_val}")
    a_val = 1.0 / (1.0 + z_val)
    if include_hubble_flow:
        t0 = 1.0 / (150.0 * float(h0) * math.sqrt(float(omega_m)))
        hubble = 100.0 * float(h0) * math.sqrt(float(omega_m) / (a_val ** 3) + float(omega_l))
        b_val = hubble * (a_val ** 2) * t0
    else:
        b_val = 0.0
    U.a, U.b = a_val, b_val
    U.physical = units != 'cube'
    U.mp_fill = None
    if not U.physical:
        U.softening = 0.1 if softening is None else float(softening)
        if bound_check and U.softening <= 0:
            raise ValueError("softening must be positive")
        ng_eff_global = float(ng_global) if ng_global is not None else float(L_Box)
        U.s15 = 15.0 * float(h0) * ng_eff_global / float(box)
        U.L, U.link = float(L_Box), float(linking_length)
        U.dx = U.v_scale = U.m_grid = U.ang_mom_scale = 1.0
        return U
    if bound_check:
        if softening is None or float(softening) <= 0:
            raise ValueError("softening must be explicitly provided and > 0 when units='physical' and bound_check=True")
        if not have_mp and n_sim is None:
            raise ValueError("Either mp_list or n_sim must be provided when units='physical' and bound_check=True")
    if n_sim is not None and float(n_sim) > 0:
        ng_eff = float(round(float(n_sim) ** (1.0 / 3.0)))
    else:
        ng_eff = float(ng_global) if ng_global is not None else 512.0
    box_mpc_h = float(L_Box)
    dx = box_mpc_h / ng_eff
    U.dx = dx
    U.L = ng_eff
    U.link = float(linking_length) / dx
    U.softening = 0.1 if softening is None else (float(softening) / dx)
    U.s15 = 15.0 * float(h0) * ng_eff / box_mpc_h
    U.v_scale = (a_val * ng_eff) / (150.0 * math.sqrt(float(omega_m)) * box_mpc_h)
    rho_crit_0 = 2.77536627e11
    m_grid = float(omega_m) * rho_crit_0 * (box_mpc_h ** 3) / (ng_eff ** 3)
    if not have_mp:
        if n_sim is not None:
            U.mp_fill = (ng_eff ** 3) / float(n_sim)
        else:
            m_grid = 1.0
    U.m_grid = m_grid
    U.ang_mom_scale = m_grid * dx / U.v_scale
    return U
