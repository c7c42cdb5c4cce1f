"""Shared helpers for the tests: mass budgets and simple closed-form cases."""
import numpy as np


def interior_mass(st, t=0):
    """sum over interior cells of concen * kappa, per bin (kg)."""
    nx, ny, nz = st.nx, st.ny, st.nz
    kap = st.kappa_pd[2:nx + 2, 2:ny + 2, 2:nz + 2]
    c = st.concen_pd[2:nx + 2, 2:ny + 2, 2:nz + 2, :, t]
    return np.einsum("ijkn,ijk->n", c, kap)


def outflow_mass(st):
    """Mass that left through the six faces, per bin: each outflow entry is
    the update of a ghost cell, so it is weighted by that ghost cell's kappa."""
    nx, ny, nz = st.nx, st.ny, st.nz
    kap = st.kappa_pd
    m = np.zeros(st.ns)
    # yz faces: outflow_yz{1,2}(j,k,n) at i = 0 / nx+1
    m += np.einsum("jkn,jk->n", st.outflow_yz1_pd, kap[1, :, :])
    m += np.einsum("jkn,jk->n", st.outflow_yz2_pd, kap[nx + 2, :, :])
    # xz faces at j = 0 / ny+1
    m += np.einsum("ikn,ik->n", st.outflow_xz1_pd, kap[:, 1, :])
    m += np.einsum("ikn,ik->n", st.outflow_xz2_pd, kap[:, ny + 2, :])
    # xy faces at k = 0 (deposit) / nz+1
    m += np.einsum("ijn,ij->n", st.outflow_xy1_pd, kap[:, :, 1])
    m += np.einsum("ijn,ij->n", st.outflow_xy2_pd, kap[:, :, nz + 2])
    return m


def total_mass(st):
    return interior_mass(st) + outflow_mass(st)


def rel_linf(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = np.max(np.abs(b))
    if scale == 0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / scale)


def total_variation(x):
    return float(np.sum(np.abs(np.diff(x))))
