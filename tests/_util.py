"""Test helpers: load a workload member into the FP64 oracle."""
import numpy as np


def oracle_state(mem, nthreads=0):
    from oracle import oracle as o

    d = o.make_dom(mem.nx, mem.ny, mem.nz, mem.lower, mem.extent, mem.glat, mem.lambda_c, mem.q0,
                   mem.f_cor, mem.zeta_zero, nthreads=nthreads)
    st = o.State(d, mem.n)
    st.position[:] = mem.position
    st.vorticity[:] = mem.vorticity
    st.B[:] = mem.B
    st.volume[:] = mem.volume
    st.buoyancy[:] = mem.buoyancy
    st.humidity[:] = mem.humidity
    return st


def maxrel(a, ref):
    """max|a - ref| / max|ref| (the error metric of the parity tests)."""
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.abs(ref).max()
    if den == 0.0:
        return float(np.abs(a).max())
    return float(np.abs(a - ref).max() / den)
