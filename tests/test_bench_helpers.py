"""bench.py's bookkeeping, on the CPU: the kernel names that the library's
profiling mode reports (launch expressions with the lanes-per-atom macros)
map onto the rocprofv3 names of the committed PMC summary, so that the
`roofline` line finds the measured HBM traffic of its kernel, and the bytes
model covers the kernel families the line names."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_lanes_per_atom_macros_come_from_the_header():
    lpa = bench.lanes_per_atom()
    assert lpa["NB_LPA"] == "16" and lpa["NB_LPA_PCG"] == "8"
    # the longer name is replaced first
    assert list(lpa)[0] == "NB_LPA_PCG"


def test_profile_names_map_to_rocprof_names():
    assert bench.profile_key("(k_dip_field<float, NB_LPA_PCG, false>)") == "k_dip_field<float, 8, false>"
    assert bench.profile_key("(k_vdw<float, NB_LPA, false, false>)") == "k_vdw<float, 16, false, false>"
    assert bench.profile_key("(k_spread_brick<T, IND>)") == "k_spread_brick<float, true>"
    assert bench.profile_key("(k_gather_ap<T>)", "fp64") == "k_gather_ap<double>"


def test_roofline_kernels_have_measured_traffic_and_a_bytes_model():
    w = {"n": 20736, "K": 64, "s": 4, "ewald_rows": 2976786, "vdw_rows": 6301742, "nbricks": 512, "ext": 12}
    for name in ("(k_dip_field<float, NB_LPA_PCG, false>)", "(k_spread_brick<T, IND>)", "(k_gather_ap<T>)",
                 "(k_plane<T, 64, 64, 2, false>)"):
        t = bench.measured_traffic(name)
        assert t is not None and t > 1e6, name
        b, model, _ = bench.work_model(name, w)
        assert b is not None and b > 1e6 and model, name
