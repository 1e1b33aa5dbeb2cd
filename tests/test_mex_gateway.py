"""The MATLAB MEX gateway (ustb_amd/mex/das_hip.cpp), compiled against the
test stand-in of the MEX API (tests/mexshim/mex.h): argument errors carry
das_c's identifiers, or none where das_c raises them with mexErrMsgTxt (CPU), and real calls in all four sum modes match the
das_c oracle (GPU). Not exercised inside MATLAB (none in this environment)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("mex")
    obj = str(d / "oracle.o")
    exe = str(d / "driver")
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-ffp-contract=off", "-c", os.path.join(ROOT, "oracle", "das_c_oracle.c"),
                    "-o", obj], check=True)
    lib = os.path.join(ROOT, "ustb_amd")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "tests", "mexshim"), "-I",
                    os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "mexshim", "driver.cpp"),
                    os.path.join(ROOT, "ustb_amd", "mex", "das_hip.cpp"), obj, "-L", lib, "-lustbx",
                    f"-Wl,-rpath,{lib}", "-lpthread", "-lm", "-o", exe], check=True)
    return exe


def test_gateway_argument_errors(driver):
    out = subprocess.run([driver, "errors"], check=True, capture_output=True, text=True).stdout
    got = dict(line.split() for line in out.strip().splitlines())
    # das_c: mexErrMsgIdAndTxt for nrhs / nlhs / Dimensions (das_c.cpp:53-55,
    # 136-199), mexErrMsgTxt with no identifier for the sum mode, the
    # single-precision and scalar checks and real data with fd (das_c.cpp:79,
    # 208-242)
    assert got == {
        "nrhs8": "Toolbox:SRP_SRC:nrhs",
        "nlhs2": "Toolbox:SRP_SRC:nlhs",
        "summode": "(no-identifier)",
        "txdelay_waves": "Toolbox:SRP_SRC:Dimensions",
        "rxdelay_pixels": "Toolbox:SRP_SRC:Dimensions",
        "fs_double": "(no-identifier)",
        "real_with_fd": "(no-identifier)",
        "fd_vector": "(no-identifier)",
    }


@pytest.mark.gpu
def test_gateway_runs_all_modes(driver):
    out = subprocess.run([driver, "run"], check=True, capture_output=True, text=True, timeout=120).stdout
    lines = out.strip().splitlines()
    assert len(lines) == 8, out
    shapes = {0: (500, 12, 5, 2), 1: (500, 1, 5, 2), 2: (500, 12, 1, 2), 3: (500, 1, 1, 2)}
    for line in lines:
        f = line.split()
        dim = int(f[1])
        assert tuple(int(x) for x in f[5:9]) == shapes[dim], line
        assert float(f[10]) <= 1e-5, line
