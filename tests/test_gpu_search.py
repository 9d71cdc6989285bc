"""End-to-end parity of the drop-in search (mrc_obj -> mrc_set_vox_size ->
fastVEC -> search_map_fft) against the oracle's search on the same resampled
maps, on the 100^3 / 60^3 synthetic pair (-g 16 -s 7 -A 30 -N 10, N = 24)."""
import contextlib
import io

import numpy as np
import pytest

from oracle import vesper_ref as ref
from oracle.build_c import fastvec_c
from vesper_python_amd import synth
from vesper_python_amd import VESPER_1_prob_new as V
from vesper_python_amd import main as cli

pytestmark = pytest.mark.gpu
RTOL = 1e-4


@pytest.fixture(scope="module")
def pair_files(tmp_path_factory):
    d = tmp_path_factory.mktemp("pair100")
    p = synth.make_config_pair("pair100", 1001)
    paths = synth.write_pair(p, str(d / "p"))
    return paths, p


def _grid(m):
    return ref.MapGrid(data=np.asarray(m.data, np.float64), xwidth=np.float32(m.xwidth), orig=m.orig,
                       vec=m.vec, cent=m.cent, xdim=m.xdim, ave=m.ave, std=m.std, std_norm_ave=m.std_norm_ave)


def _resampled(paths, p, s=7.0, g=16.0):
    with contextlib.redirect_stdout(io.StringIO()):
        m1, m2, n1, n2 = cli._resample_pair(paths["A"], paths["B"], p.a.threshold, p.b.threshold, s, g)
        src1, src2 = m1, m2
        n1 = V.fastVEC(src1, n1, g)
        n2 = V.fastVEC(src2, n2, g)
    return src1, src2, n1, n2


def test_resample_and_fastvec_match_oracle(gpu, pair_files, fastvec_ref):
    paths, p = pair_files
    src1, src2, n1, n2 = _resampled(paths, p)
    assert n1.xdim == n2.xdim == 24
    for src, dst in ((src1, n1), (src2, n2)):
        g = ref.MapGrid(data=np.asarray(src.data), xwidth=src.xwidth, orig=src.orig)
        d_c, v_c = fastvec_c(fastvec_ref, g.data, g.xwidth, g.orig, dst.xdim, dst.xwidth, dst.orig, 16.0)
        assert np.array_equal(d_c != 0, dst.data != 0)
        assert np.allclose(dst.data, d_c, rtol=1e-12)
        assert np.max(np.abs(dst.vec - v_c)) <= 4e-7
        st = ref.map_stats(d_c)
        assert np.isclose(dst.ave, st["ave"], rtol=1e-12) and np.isclose(dst.std, st["std"], rtol=1e-12)


@pytest.mark.parametrize("mode", ["VecProduct", "CC"])
def test_search_top10_matches_oracle(gpu, pair_files, mode):
    paths, p = pair_files
    _, _, n1, n2 = _resampled(paths, p)
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        got = V.search_map_fft(n1, n2, TopN=10, ang=30.0, mode=mode)
    # the product's default ranking re-decides the rows near the cutoff and
    # the refinement rows in FP64; the oracle does the same
    exp = ref.search_map_fft(_grid(n1), _grid(n2), TopN=10, ang=30.0, mode=mode, prec="ref32", fp64_recheck=True)
    # refined top-10: rotations identical, translations identical, scores 1e-4
    assert [tuple(float(v) for v in r["angle"]) for r in got] == \
           [tuple(float(v) for v in r["angle"]) for r in exp["refined"]]
    for g, e in zip(got, exp["refined"]):
        assert tuple(g["vec_trans"]) == tuple(e["vec_trans"])
        assert abs(g["vec_score"] - e["vec_score"]) <= RTOL * abs(e["vec_score"])
        assert np.array_equal(g["data"], e["data"]) and np.array_equal(g["vec"], e["vec"])
    out = buf.getvalue()
    assert out.count("\n#") == 10 and out.count("Overlap=") == 10 and out.count("Score=") == 10
    assert "M0 Rotation= (" in out


def test_cli_orig_runs(gpu, pair_files):
    paths, p = pair_files
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        res = cli.main(["orig", "-a", paths["A"], "-b", paths["B"], "-t", str(p.a.threshold),
                        "-T", str(p.b.threshold), "-s", "7", "-A", "30", "-N", "5"])
    out = buf.getvalue()
    assert len(res) == 5
    assert "###Start Searching###" in out and "#MAP SUM=" in out and "Nvox= 24, 24, 24" in out


def test_cli_eval_mode_and_pdb_output(gpu, pair_files, tmp_path, monkeypatch):
    """-E scores the maps as placed (get_score at translation 0, no search);
    -S writes one CA/CB pseudo-atom PDB per refined model."""
    paths, p = pair_files
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        res = cli.main(["orig", "-a", paths["A"], "-b", paths["A"], "-t", str(p.a.threshold),
                        "-T", str(p.a.threshold), "-s", "7", "-E", "True"])
    out = buf.getvalue()
    assert res is None and "#For Evaluation Mode" in out and "Overlap=" in out and "Score=" in out
    # the same map at translation 0: every voxel overlaps
    ov = [l for l in out.splitlines() if l.startswith("Overlap=")][0]
    nm, total = ov.split()[1].split("/")
    assert nm == total
    monkeypatch.chdir(tmp_path)
    with contextlib.redirect_stdout(io.StringIO()):
        res = cli.main(["orig", "-a", paths["A"], "-b", paths["B"], "-t", str(p.a.threshold),
                        "-T", str(p.b.threshold), "-s", "7", "-A", "30", "-N", "3", "-S", "True",
                        "-o", "pdb_out"])
    pdbs = sorted((tmp_path / "pdb_out").glob("*.pdb"))
    assert len(res) == 3 and len(pdbs) == 3
    lines = pdbs[0].read_text().splitlines()
    assert lines and all(l.startswith("ATOM") for l in lines) and " CA  ALA" in lines[0] and " CB  ALA" in lines[1]
