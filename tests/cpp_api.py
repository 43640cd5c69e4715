"""Build helper for the C++ API test programs (g++ against include/ and libmppi_amd.so)."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "mppi_generic_vendor_amd"


def build_program(name: str, out_dir: Path) -> Path:
    src = ROOT / "tests" / "cpp" / f"{name}.cpp"
    exe = out_dir / name
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-pthread", "-I", str(ROOT / "include"), str(src), "-o",
           str(exe),
           "-L", str(PKG), "-lmppi_amd", f"-Wl,-rpath,{PKG}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe
