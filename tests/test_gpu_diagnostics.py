"""Parcel and field diagnostics on the GPU against the FP64 oracle, through
the C-ABI (parcel_diagnostics.f90:51-159, field_diagnostics.f90:40-154).

The state is the one at the start of ls_rk_step (par2grid, vor2vel,
grid2par); same FP32-rounded inputs on both sides.  Counts and integer
extremes are exact; sums of FP32 fields and velocities agree within 1e-5
relative, the field tolerance of the core."""
import numpy as np
import pytest

from epic_amd import workloads as W
from epic_amd.core import Member
from tests._util import oracle_state

pytestmark = pytest.mark.gpu

TOL = 1e-5
EXACT = ("n_small", "ntot_par", "nsplits", "nbig_iclo", "nmerges", "avg_npar", "avg_nspar", "max_npar",
         "min_npar", "ape", "apeg")


def _close(got, ref, name):
    for k, r in ref.items():
        g = got[k]
        if k in EXACT:
            assert g == r, (name, k, g, r)
        elif k in ("min_buoy", "max_buoy"):
            # extremes of a field: within the field tolerance of its scale (max-rel)
            scale = max(abs(ref["min_buoy"]), abs(ref["max_buoy"]))
            assert abs(g - r) <= TOL * scale, (name, k, g, r)
        elif k.startswith("std_"):
            # sqrt(|sum x^2 / n - mean^2|) cancels: its resolution is sqrt(eps) * mean
            mean = ref["avg_" + k[4:]]
            assert abs(g - r) <= TOL * abs(r) + 1e-6 * abs(mean), (name, k, g, r)
        else:
            assert abs(g - r) <= TOL * max(abs(r), 1e-30) + 1e-12 * (abs(r) == 0), (name, k, g, r)


@pytest.mark.parametrize("name", ["thermal16", "moist16"])
def test_diagnostics_match_oracle(name):
    mem = (W.rising_thermal(16, seed=91) if name == "thermal16"
           else W.rising_thermal(16, seed=92, humidity_plume=True, n_blobs=4))
    print(f"diagnostics: {name}", flush=True)
    gm = Member.from_workload(mem)
    st = oracle_state(mem)
    for x in (gm, st):
        x.par2grid()
        x.vor2vel()
        x.grid2par()
    pg, pr = gm.parcel_diagnostics(), st.parcel_diagnostics()
    fg, fr = gm.field_diagnostics(), st.field_diagnostics()
    print("  ke", pg["ke"], pr["ke"], " keg", fg["keg"], fr["keg"], flush=True)
    assert pr["ke"] > 0 and fr["keg"] > 0
    _close(pg, pr, name)
    _close(fg, fr, name)


def test_split_and_merge_counters():
    mem = W.rising_thermal(16, seed=93)
    gm = Member.from_workload(mem, max_parcels=2 * mem.n)
    ns = gm.parcel_split(lambda_max=2.0)
    nm = gm.parcel_merge()
    gm.par2grid()
    gm.vor2vel()
    gm.grid2par()
    d = gm.parcel_diagnostics(reset_counters=True)
    assert (d["nsplits"], d["nmerges"]) == (ns, nm)
    assert d["ntot_par"] == gm.n
    d2 = gm.parcel_diagnostics()
    assert (d2["nsplits"], d2["nmerges"]) == (0, 0)
