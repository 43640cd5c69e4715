"""The user-plugin test program (tests/plugin) builds against include/mppi_amd/device_plugin.hpp and loads next to
the engine library on a host without a GPU: every entry point the GPU tests call is exported."""
import ctypes as ct
import subprocess

from mppi_generic_vendor_amd import capi
from test_gpu_plugin import PLUGIN_LIB


def test_plugin_library_exports_its_entry_points(engine):
    assert PLUGIN_LIB.exists(), "run __graft_entry__.build() first"
    out = subprocess.run(["nm", "-D", "--defined-only", str(PLUGIN_LIB)], capture_output=True, text=True,
                         check=True).stdout
    for name in ("unicycle_rollout_gpu", "unicycle_rollout_oracle_f32", "unicycle_rollout_oracle_f64",
                 "unicycle_state_trajectory_f32", "unicycle_controller_create", "unicycle_cpp_closed_loop"):
        assert f" T {name}" in out, name
    capi.load()
    lib = ct.CDLL(str(PLUGIN_LIB))
    assert hasattr(lib, "unicycle_controller_create")
    assert hasattr(capi.load(), "mppi_controller_create_plugin")
