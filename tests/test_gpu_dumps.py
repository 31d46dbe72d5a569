"""Debug text dumps (SaveInfoTable, SaveParticleInfo; particle_simulation.cu:2709-2852)
written by the library, against the same tables formatted here from the
oracle's sort and from the read-back fields.

The reference writes with iostream setw and the default float format
(6 significant digits, printf %g); the table format below restates that
(widths 8/12 for the info table, 20 for the particle table).  The info table
holds integers only and must equal the oracle's byte for byte; the particle
table must equal the formatting of the library's own read-back values byte
for byte, and its values the oracle's within the one-step bars."""
import numpy as np
import pytest

import cudasphnew_amd as S
from cudasphnew_amd import Field
from oracle import oracle as O

pytestmark = pytest.mark.gpu
G = 9.81


def info_table(ids, keys, start, end):
    n = ids.shape[0]
    out = [f"Number of particles {n}", f"{'index':>8}{' id':>12}{' hash':>12}{' start':>12}{' end':>12}"]
    for i in range(start.shape[0]):
        row = f"{i:>8}"
        row += f"{ids[i]:>12}{keys[i]:>12}" if i < n else " " * 24
        row += " " * 12 if start[i] < 0 else f"{start[i]:>12}"
        row += " " * 12 if end[i] < 0 else f"{end[i]:>12}"
        out.append(row)
    return "\n".join(out) + "\n"


def particle_table(x, rho, p, a):
    cols = ["Index", "X", "Y", "Z", "Density", "Pressure", "Acc X", "Acc Y", "Acc Z"]
    out = ["".join(f"{c:>20}" for c in cols)]
    for i in range(x.shape[0]):
        vals = [x[i, 0], x[i, 1], x[i, 2], rho[i], p[i], a[i, 0], a[i, 1], a[i, 2]]
        out.append(f"{i:>20}" + "".join(f"{float(v):>20.6g}" for v in vals))
    return "\n".join(out) + "\n"


@pytest.fixture(scope="module")
def dam16k():
    return S.make_scene("dam16k")


def test_info_table_matches_oracle(dam16k, tmp_path):
    sim = S.ParticleSimulation(dam16k).Init()
    sim.Advance(1)
    path = tmp_path / "info.txt"
    sim.SaveInfoTable(str(path))
    keys, _ = O.keys(dam16k.params, dam16k.positions, real=1)
    sk, si, cs, ce = O.sort_cells(dam16k.params, keys)
    assert path.read_text() == info_table(si, sk, cs, ce)


def test_particle_table_format_and_values(dam16k, tmp_path):
    sim = S.ParticleSimulation(dam16k).Init()
    sim.Advance(1)
    path = tmp_path / "particles.txt"
    sim.SaveParticleInfo(str(path))
    text = path.read_text()
    x, rho = sim.read(Field.POSITION), sim.read(Field.DENSITY)
    p, a = sim.read(Field.PRESSURE), sim.read(Field.ACCEL)
    assert text == particle_table(x, rho, p, a)
    rows = np.array([line.split() for line in text.splitlines()[1:]], dtype=np.float64)
    assert rows.shape == (dam16k.n, 9) and np.array_equal(rows[:, 0], np.arange(dam16k.n))
    orc = O.OracleSim(dam16k, real=0, threads=16)
    orc.step(1)
    # 6 significant digits in the file: compare at that resolution
    assert np.allclose(rows[:, 1:4], orc.read(Field.POSITION), rtol=1e-5, atol=1e-6)
    assert np.allclose(rows[:, 4], orc.read(Field.DENSITY), rtol=2e-5)
    ao = orc.read(Field.ACCEL)
    scale = np.maximum(np.linalg.norm(ao, axis=1), G)[:, None]
    assert (np.abs(rows[:, 6:9] - ao) / scale).max() <= 2e-4
