"""The C ABI from a plain C program (tests/c_caller/c_caller.c, built by
__graft_entry__.build() with gcc against include/tsgpu.h): its synthesis,
run with a progress callback and through the batch API, equals the Python
mirror's result for the same example and seed."""
import os
import subprocess

import numpy as np
import pytest

import tsgpu as ts
from tsgpu import synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "c_caller", "c_caller")


def test_c_caller_matches_python(tmp_path):
    if not os.path.exists(BIN):
        pytest.fail("tests/c_caller/c_caller is not built (run __graft_entry__.build())")
    ex = synth.example(8100, 72, 56)
    src = tmp_path / "ex.rgba"
    src.write_bytes(ex.tobytes())
    prefix = str(tmp_path / "out")
    print("case: c_caller 72x56 -> 96^2 seed 8100", flush=True)
    r = subprocess.run([BIN, str(src), "72", "56", "96", "8100", prefix], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
    g = ts.Session.builder().add_example(ex).seed(8100).output_size(ts.Dims.square(96)).build().run()
    rgba = np.fromfile(prefix + ".rgba", np.uint8).reshape(96, 96, 4)
    coord = np.fromfile(prefix + ".coord", np.uint32).reshape(-1, 3)
    ids = np.fromfile(prefix + ".ids", np.uint32).reshape(-1, 2)
    assert np.array_equal(rgba, g.image)
    assert np.array_equal(coord, g.coord.reshape(-1, 3))
    assert np.array_equal(ids, g.ids.reshape(-1, 2))
    assert np.array_equal(np.fromfile(prefix + ".batch.rgba", np.uint8).reshape(96, 96, 4), g.image)
