"""Observability formats: per-level timers in the reference's time_log JSON layout and the text
dump of a stencil in the layout of the reference's operator<< (src/2d/stencil_op.cc:5-51)."""
import json
import re

import numpy as np
import pytest

from cedar_amd import gallery as G

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_timers_json_layout_and_counts(tmp_path):
    import cedar_amd
    so, b, _ = G.heterogeneous_9pt(255, seed=3, block=16, sigma=1.0, contrast=1e2)
    so, b = torch.as_tensor(so).cuda(), torch.as_tensor(b).cuda()
    kw = dict(relaxation="line-xy", nrelax_pre=1, nrelax_post=1, maxiter=30, tol=1e-10,
              precision="mixed")
    x0, h0 = cedar_amd.Solver(so, **kw).solve(b)
    s = cedar_amd.Solver(so, **kw)
    s.timers(True)
    x1, h1 = s.solve(b)
    assert np.array_equal(h0, h1) and torch.equal(x0, x1)  # timers change nothing
    path = tmp_path / "timings.json"
    s.save_timers(path)
    levels = json.loads(path.read_text())["levels"]
    assert len(levels) == s.nlevels
    cycles = len(h1)
    fine, coarsest = levels[-1], levels[0]
    assert set(fine) >= {"setup", "solve", "relaxation", "residual", "restrict", "interp-add"}
    assert int(fine["relaxation"]["count"]) == 2 * cycles
    assert int(fine["solve"]["count"]) == 1
    assert int(coarsest["coarse-solve"]["count"]) == cycles
    for lvl in levels[1:]:
        assert int(lvl["residual"]["count"]) == cycles
    phases = sum(float(v["time"]) for lvl in levels for k, v in lvl.items()
                 if k not in ("setup", "solve"))
    assert 0 < phases < float(fine["solve"]["time"])
    s.save_timers(path)  # reset: only the levels that logged something remain
    assert json.loads(path.read_text())["levels"] == []


@pytest.mark.parametrize("ns", [3, 5])
def test_stencil_dump_layout(tmp_path, ns):
    import cedar_amd
    rng = np.random.default_rng(ns)
    nx, ny = 4, 3
    so = np.zeros((ns, ny + 2, nx + 2))
    so[:, 1:-1, 1:-1] = rng.standard_normal((ns, ny, nx))
    path = tmp_path / "so.txt"
    cedar_amd.write_stencil(so, path)
    lines = path.read_text().splitlines()
    assert len(lines) == nx * ny
    for ln in lines:
        head, vals = ln[:16], ln[16:]
        idx, i, j = int(head[:4]), int(head[5:9]), int(head[11:15])
        assert idx == (j - 2) * (nx + 2) + i - 2
        # %.7e fields written back to back (one blank before the centre value); the entries
        # here are O(1), so every exponent has two digits
        got = np.array([float(v) for v in re.findall(r"-?\d\.\d{7}e[-+]\d\d", vals)])
        ii, jj = i - 1, j - 1
        if ns == 3:
            want = [-so[2, jj + 1, ii], -so[1, jj, ii], so[0, jj, ii], -so[1, jj, ii + 1],
                    -so[2, jj, ii]]
        else:
            want = [-so[4, jj + 1, ii], -so[2, jj + 1, ii], -so[3, jj + 1, ii + 1], -so[1, jj, ii],
                    so[0, jj, ii], -so[1, jj, ii + 1], -so[3, jj, ii], -so[2, jj, ii],
                    -so[4, jj, ii + 1]]
        assert np.allclose(got, want, rtol=1e-7, atol=0)
