"""bench.py's byte models and its one-GPU contract (no GPU needed).

The step-level model (B_P, B_A per stage) is pinned against the table of
bytes per cell-update it restates: TGV 128^3 FP32 1243.8 / 2767.7 B, FP64
2415.1 / 5390.2 B; TGV 64^3 FP32 1271.7 / 2823.4 B; cylinder 512 x 256
FP32 2408.6 / 5097.3 B (primal / primal+adjoint)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

SIZES = {  # nC, nF, nG
    'tgv128': (2097152, 6340608, 98304),
    'tgv64': (262144, 798720, 24576),
    'cylinder': (131072, 524800, 263168),
}


@pytest.mark.parametrize('case,s,primal,pa', [('tgv128', 4, 1243.8, 2767.7), ('tgv128', 8, 2415.1, 5390.2),
                                              ('tgv64', 4, 1271.7, 2823.4), ('cylinder', 4, 2408.6, 5097.3)])
def test_step_bytes_model_matches_table(case, s, primal, pa):
    p, a = bench.step_bytes_model(s, *SIZES[case])
    assert abs(p - primal) <= 0.051 and abs(a - pa) <= 0.051      # the table has one decimal


def test_face_adjoint_record_count():
    """Owner sums: one record per internal face (neighbour side), one per
    cell owning an internal face, two per cyclic face; without them two per
    face.  The difference is exactly (nIF - nO) records of 20 scalars, read
    once by the gather and written once by the face adjoint."""
    nC, nF, nG = SIZES['tgv128']
    nO = nC - 1
    for k in ('face_vjp_kernel', 'vjp_gather_kernel'):
        per_face = bench.kernel_bytes(k, 4, nC, nF, nG)
        summed = bench.kernel_bytes(k, 4, nC, nF, nG, nO)
        saved_idx = 4 * (nF - nG) if k == 'face_vjp_kernel' else 0   # one record index per internal face
        assert per_face - summed == ((nF - nG) - nO) * 20 * 4 + saved_idx


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2', '--steps', '1',
                        '--warmup', '0'], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert 'one GPU' in (r.stdout + r.stderr)


def test_kernel_rooflines():
    nC, nF, nG = SIZES['tgv128']
    prof = {'face_vjp_kernel': (4.3, 10), 'grad_kernel': (1.0, 10), 'unknown_kernel': (1.0, 10)}
    r = bench.kernel_rooflines(prof, 4, nC, nF, nG, nC - 1, 6000.0)
    assert set(r) == {'face_vjp_kernel', 'grad_kernel'}
    b = bench.kernel_bytes('face_vjp_kernel', 4, nC, nF, nG, nC - 1)
    assert r['face_vjp_kernel']['algorithmic_bytes'] == int(b)
    assert abs(r['face_vjp_kernel']['GBs'] - b / 0.43e-3 / 1e9) < 0.1
    assert abs(r['face_vjp_kernel']['frac_peak'] - b / 0.43e-3 / 8e12) < 1e-3
