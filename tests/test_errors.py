"""Error semantics of the C ABI: misuse ends the program with
`Error in <task>: <msg>.` on stderr and exit status 1 (global.f90:51-56),
with the reference's task names and messages.  Each case runs in a child
Python process, against the CPU oracle (CPU tests) and against the HIP
library (GPU tests), so both libraries are held to the same behaviour."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PRELUDE = """
import sys, ctypes
import numpy as np
sys.path[:0] = [{root!r}]
from emdee_amd.emdee import EmDeeLibrary
lib = EmDeeLibrary({so!r})
R = np.array([[0.0, 0.0, 0.0], [1.5, 0.0, 0.0], [0.0, 1.5, 0.0]])
types = np.array([1, 2, 2], np.int32)
def fresh():
    return lib.system(1, 1, 2.5, 0.3, 3, types, None, None)
def initialised():
    md = fresh()
    for t in (1, 2):
        lib.EmDee_set_pair_model(md, t, t, lib.EmDee_pair_lj_cut(1.0, 1.0), 0.0)
    lib.upload(md, "box", np.array([10.0]))
    lib.upload(md, "coordinates", R)
    return md
"""

# (case, code that misuses the API, expected stderr line), reference lines cited
CASES = [
    ("model_after_init",      # EmDeeCode.f90:323-325
     "md = initialised(); lib.EmDee_set_pair_model(md, 1, 1, lib.EmDee_pair_lj_cut(1.0, 1.0), 0.0)",
     "Error in pair model setup: cannot set model after coordinates have been defined."),
    ("type_out_of_range",     # EmDeeCode.f90:327-329
     "md = fresh(); lib.EmDee_set_pair_model(md, 1, 3, lib.EmDee_pair_lj_cut(1.0, 1.0), 0.0)",
     "Error in pair model setup: provided type index is out of range."),
    ("charges_after_init",    # EmDeeCode.f90:878-880
     "md = initialised(); lib.upload(md, 'charges', np.zeros(3))",
     "Error in upload: cannot set charges after box and coordinates initialization."),
    ("bad_upload_option",     # EmDeeCode.f90:889-890
     "md = initialised(); lib.upload(md, 'velocities', np.zeros(9))",
     "Error in upload: invalid option."),
    ("bad_download_option",   # EmDeeCode.f90:719-720
     "md = initialised(); lib.download(md, 'velocities', (3, 3))",
     "Error in download: invalid option."),
    ("kspace_missing",        # EmDeeData.f90:394-400
     "md = fresh()\n"
     "for t in (1, 2): lib.EmDee_set_pair_model(md, t, t, lib.EmDee_pair_lj_cut(1.0, 1.0), 1.0)\n"
     "lib.EmDee_set_coul_model(md, lib.EmDee_coul_long())\n"
     "lib.upload(md, 'charges', np.array([-1.0, 0.5, 0.5]))\n"
     "lib.upload(md, 'box', np.array([10.0])); lib.upload(md, 'coordinates', R)",
     "Error in system initialization: a kspace solver is required, but has not been defined."),
    ("softcore_lambda",       # pair_softcore_cut.f90:72-74
     "lib.EmDee_pair_softcore_cut(1.0, 1.0, 1.5)",
     "Error in pair_softcore_cut setup: out-of-range parameter lambda."),
]


def run_case(so, code):
    src = PRELUDE.format(root=ROOT, so=so) + code + "\nprint('not reached', flush=True)\n"
    env = dict(os.environ, EMDEE_QUIET="1")
    return subprocess.run([sys.executable, "-c", src], capture_output=True, text=True, env=env, timeout=120)


def check(so, code, expected):
    r = run_case(so, code)
    assert r.returncode == 1, (r.returncode, r.stderr[-2000:])
    assert expected in r.stderr.splitlines(), r.stderr[-2000:]
    assert "not reached" not in r.stdout


@pytest.mark.parametrize("case,code,expected", CASES, ids=[c[0] for c in CASES])
def test_oracle_errors(oracle, case, code, expected):
    check(oracle.path, code, expected)


@pytest.mark.gpu
@pytest.mark.parametrize("case,code,expected", CASES, ids=[c[0] for c in CASES])
def test_gpu_errors(gpu, case, code, expected):
    check(gpu.path, code, expected)
