"""PETSc binary and text writers: a fixture system's matrix written back must be
byte-identical to the reference's pristine file, a read vector written back
byte-identical to its source, and the text dump must hold the elements."""
import os

import numpy as np
import pytest

import quac_amd as qa
import recipes
from qbuild import build_system

pytestmark = pytest.mark.gpu

FX = os.path.join(os.path.dirname(__file__), "golden", "pristine_matrices")


@pytest.mark.parametrize("fam,nop,fl", recipes.FIXTURES)
def test_matrix_written_back_is_the_pristine_file(fam, nop, fl, tmp_path):
    qs, _, _ = build_system(recipes.fixture_recipe(fam, nop, fl))
    out = tmp_path / "m.bin"
    qs.matrix().write_petsc_binary(out)
    qs.destroy()
    with open(os.path.join(FX, f"{fam}_{nop}op_{fl}"), "rb") as f:
        want = f.read()
    assert out.read_bytes() == want


def test_vector_written_back_is_the_source_file(tmp_path):
    src = os.path.join(FX, "rand_dm.dat")
    x = qa.read_qvec_dm_binary(src)
    out = tmp_path / "v.bin"
    x.write_petsc_binary(out)
    with open(src, "rb") as f:
        want = f.read()
    assert out.read_bytes() == want
    txt = tmp_path / "v.txt"
    x.print_file(txt)
    rows = [r.split() for r in txt.read_text().strip().splitlines()]
    v = x.get().reshape(4, 4, order="F")
    assert len(rows) == 4
    for i in range(4):
        vals = [float(a) for a in rows[i]]
        got = np.array(vals[0::2]) + 1j * np.array(vals[1::2])
        assert np.abs(got - v[i]).max() < 1e-28 + 1e-15 * np.abs(v[i]).max()
    x.destroy()
