"""waveinput.dat loader and s-grid dump parity with the compiled reference.

Each fixture holds an input file, a program of C-ABI calls with the
reference's answers, and the text of every file the reference wrote during
the run (tests/golden/make_golden.py, make_abi_inputs): the reference's own
t008 (312 components with directions, bandwidth auto, s-grid with
ignore_at_init) and python-example inputs (ndir 1 spreading form, normalize,
amplify), a 10 x 5 spreading input with unsorted rows, individual_ramp rows,
and an s-grid with [lsgrid2vtk] dumps.  libcfdwm runs the same program in a
process of its own: the answers to 1e-9 of each output's scale (FP64), the
same files, spectral_components.dat line for line, and .vtu files with the
same markup and the same numbers at their printed precision."""
import re

import numpy as np
import pytest

from tests.conftest import scaled_err
from tests.test_abi import run_program

pytestmark = pytest.mark.gpu

FIXTURES = ["abi_t008.npz", "abi_example.npz", "abi_ndir.npz", "abi_iramp.npz", "abi_vtk.npz"]
NUM = re.compile(r"^[\s\-+0-9.eE]+$")


def _vtu_parts(text):
    """Markup lines (not pure numbers) and the numbers of each DataArray."""
    markup, arrays, cur, name = [], {}, None, None
    for line in text.splitlines():
        s = line.strip()
        if s and NUM.match(s):
            cur.extend(float(v) for v in s.split())
            continue
        markup.append(line)
        m = re.search(r'Name\s*=\s*"([^"]+)"', s)
        if s.startswith("<DataArray"):
            name = m.group(1) if m else "Points"
            cur = arrays.setdefault(name, [])
    return markup, {k: np.array(v) for k, v in arrays.items()}


@pytest.mark.parametrize("fixture", FIXTURES)
def test_input_programs_and_written_files(fixture):
    g, rc, out, written = run_program(fixture, files=True)
    assert rc == int(g["rc"]) == 0
    ref = g["out"]
    ops = g["op"]
    for op in np.unique(ops):
        sel = ops == op
        for c in range(4):
            err = scaled_err(out[sel, c], ref[sel, c])
            assert err <= 1e-9, (fixture, int(op), c, err)
    ref_files = {k[len("file__"):]: str(g[k]) for k in g if k.startswith("file__")}
    assert sorted(written) == sorted(ref_files), (sorted(written), sorted(ref_files))
    for name, text in ref_files.items():
        mine = written[name]
        if name.endswith("spectral_components.dat"):
            assert mine.splitlines() == text.splitlines(), name
            continue
        mk_ref, arr_ref = _vtu_parts(text)
        mk, arr = _vtu_parts(mine)
        assert mk == mk_ref, name
        assert sorted(arr) == sorted(arr_ref), name
        for key, a in arr_ref.items():
            b = arr[key]
            assert a.shape == b.shape, (name, key)
            if key in ("connectivity", "offsets", "types"):
                assert np.array_equal(a, b), (name, key)
            elif key == "Points":
                assert np.max(np.abs(a - b)) <= 1.5e-4, (name, key)          # printed with 4 decimals
            else:
                # printed with 6 significant digits: at most one unit of the last
                # digit apart where the FP64 values straddle a rounding boundary
                scale = max(np.max(np.abs(a)), 1e-300)
                assert np.all(np.abs(a - b) <= 1.01e-5 * np.abs(a) + 1e-12 * scale), (name, key)
        print(f"{fixture}: {name} matches", flush=True)
