"""CPU tests of the profile summaries: kernel classes and the copy-calibrated
per-launch traffic computed from rocprofv3 counter CSVs."""
import csv
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "spectrecon.jl_amd"))

from spectrecon_amd import profiling  # noqa: E402

FIELDS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size",
          "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
          "Accum_VGPR_Count", "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]


def _write(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, FIELDS)
        w.writeheader()
        for i, (name, grid, val) in enumerate(rows):
            r = dict.fromkeys(FIELDS, 0)
            r.update(Dispatch_Id=i + 1, Kernel_Name=name, Grid_Size=grid, Counter_Name=counter, Counter_Value=val)
            w.writerow(r)


def test_kernel_classes():
    c = profiling.kernel_class
    assert c("void sr::k_psf_adj_sep<40, 16>(float const*, sr::Taps)") == "psf_adj"
    assert c("sr::k_rot2_adj_scatter(float const*, float*)") == "rot_adj"
    assert c("sr::k_rot2_fwd4(float const*)") == "rot_fwd"
    assert c("sr::k_copy16(uint4*, uint4 const*, long)") == "copy"
    assert c("__amd_rocclr_copyBuffer") is None


def test_traffic_calibration(tmp_path):
    """A counter that reads 2x the true bytes is scaled back by the copy of
    known size; the most frequent launch shape of a class is reported."""
    mib = 64
    f, w = tmp_path / "f.csv", tmp_path / "w.csv"
    copy = "sr::k_copy16(uint4*, uint4 const*, long)"
    adj = "void sr::k_psf_adj_sep<40, 16>(float const*)"
    _write(f, "FETCH_SIZE", [(copy, 10, 2 * mib * 1024), (adj, 100, 2000.0), (adj, 100, 2200.0), (adj, 100, 2100.0),
                             (adj, 999, 9e9)])
    _write(w, "WRITE_SIZE", [(copy, 10, 4 * mib * 1024), (adj, 100, 400.0), (adj, 100, 400.0), (adj, 100, 9e9),
                             (adj, 999, 9e9)])
    t = profiling.traffic(str(f), str(w), mib << 20)
    assert t["calibration"]["fetch_scale"] == pytest.approx(0.5)
    assert t["calibration"]["write_scale"] == pytest.approx(0.25)
    k = t["kernels"]["psf_adj"]
    assert k["grid_size"] == 100 and k["dispatches"] == 3
    assert k["fetch_bytes"] == round(1024 * 0.5 * 2100) and k["write_bytes"] == round(1024 * 0.25 * 400)
    assert k["traffic_bytes"] == k["fetch_bytes"] + k["write_bytes"]
    with pytest.raises(ValueError):
        _write(f, "FETCH_SIZE", [(adj, 100, 1.0)])
        profiling.traffic(str(f), str(w), mib << 20)


def test_traffic_keeps_the_kernel_that_ran(tmp_path):
    """Autotuner candidates of a class (other kernels at the same launch
    shape, fewer dispatches) stay out of the reported median."""
    mib = 64
    f, w = tmp_path / "f.csv", tmp_path / "w.csv"
    copy = "sr::k_copy16(uint4*, uint4 const*, long)"
    ran = "void sr::k_psf_fwd_ws<5, 8>(float const*)"
    cand = "void sr::k_psf_fwd_col<5>(float const*)"
    _write(f, "FETCH_SIZE", [(copy, 10, mib * 1024)] + [(ran, 100, 1000.0)] * 4 + [(cand, 100, 5000.0)] * 3)
    _write(w, "WRITE_SIZE", [(copy, 10, mib * 1024)] + [(ran, 100, 10.0)] * 4 + [(cand, 100, 50.0)] * 3)
    k = profiling.traffic(str(f), str(w), mib << 20)["kernels"]["psf_fwd"]
    assert k["kernel"] == "sr::k_psf_fwd_ws<5, 8>" and k["dispatches"] == 4
    assert k["fetch_bytes"] == 1024 * 1000 and k["write_bytes"] == 1024 * 10


def test_short_name_keeps_kernels_in_anonymous_namespaces():
    from spectrecon_amd.profiling import kernel_class, short_name
    n = "void sr::(anonymous namespace)::k_psf_fwd_fft<256, 128>(float const*, float const*, int)"
    assert short_name(n) == "sr::k_psf_fwd_fft<256, 128>"
    assert kernel_class(n) == "psf_fwd"
    assert kernel_class("void sr::k_atten_scan<1>(float*, sr::Geo, float, int)") == "atten"
